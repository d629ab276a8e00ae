// nicr_nucleate.hip -- Ksp oxide insertion on the device.
//
// Reference semantics (Model::nucleate / valid_nucleus / stamp,
// mpi2d/NiCr2D_grains.cpp:493-516,543-550; mpi3d-slab/NiCr3D_grains.cpp:218-231,
// 687-724,752-759): a lexicographic (i, j, k) scan; a cell is stamped when
// p <= .01, o*cr > Ksp and no cell with p > .01 lies closer than r + gap
// (minimum image in Y/Z); each stamp turns every cell with p < .01 within r into
// hard oxide. Later candidates see earlier stamps, so the scan is sequential.
//
// Device form:
//   k_ksp_mark     parallel candidate test on the scan-start state (stamps only
//                  ever remove candidates: stamped cells get p = 1), per-tile
//                  counts and per-plane oxide counts
//   k_scan_counts  exclusive scan of the tile counts (one block)
//   k_ksp_compact  order-preserving compaction of candidate indices
//   k_ksp_resolve  one workgroup walks the candidates in order; each one is
//                  re-tested against the live state exactly as the reference
//                  does, with the neighbourhood search done by all lanes and
//                  skipped when provably empty (no oxide in the X-window at scan
//                  start and no accepted stamp within reach), then stamped in
//                  parallel.
// The distance predicates compare sqrt of exact integer sums (correctly
// rounded) with r + gap, the same decision as the reference's hypot / sqrt(pow)
// whenever r + gap is not within one ulp of sqrt(n) for an integer n.
#include "nicr_kernels.hpp"

namespace nicr {

struct NucPar {
    int x, y, z;           // 1-based stamp centre (k_stamp only)
    double r, gap;
    int max_centres;
};

__device__ __forceinline__ double min_image(double d, double period) { return d - period * round(d / period); }
__device__ __forceinline__ int bounded_reach(double radius, int n) { return radius >= n ? n : (int)ceil(radius); }

// Distinct offsets of a periodic neighbourhood (SL:220-231): returns the count
// and the first offset, such that offsets first..first+count-1 (mod n) cover it.
__device__ __forceinline__ void axis_span(int reach, int n, int& count, int& first) {
    if (2 * reach + 1 >= n) {
        count = n;
        first = 0;
    } else {
        count = 2 * reach + 1;
        first = -reach;
    }
}

// Parallel stamp of one centre by the whole block. Returns the changed-cell
// count (block total, valid on every thread). Accumulates nucO/nucCr per thread.
template <int DIM, class TE>
__device__ int block_stamp(const Par& P, double* o, double* cr, double* p, TE* eta, int x, int y, int z,
                           double r, DD& aO, DD& aC, int* sh_count) {
    const int rr = bounded_reach(r, P.nx);
    const int i0 = max(1, x - rr), i1 = min(P.nx, x + rr);
    int cy, fy, cz = 1, fz = 0;
    if (DIM == 2) {
        cy = P.ny;   // 2D stamps loop over every j (2D:506)
        fy = 0;
    } else {
        axis_span(bounded_reach(r, P.ny), P.ny, cy, fy);
        axis_span(bounded_reach(r, P.nz), P.nz, cz, fz);
    }
    const long long nbox = (long long)(i1 - i0 + 1) * cy * cz;
    if (threadIdx.x == 0) *sh_count = 0;
    __syncthreads();
    int mine = 0;
    for (long long q = threadIdx.x; q < nbox; q += blockDim.x) {
        const int di = (int)(q / ((long long)cy * cz));
        const int rem = (int)(q - (long long)di * cy * cz);
        const int dj = rem / cz, dk = rem - dj * cz;
        const int i = i0 + di;
        int j, k = 1;
        if (DIM == 2) j = 1 + dj;
        else {
            j = ((y - 1 + fy + dj) % P.ny + P.ny) % P.ny + 1;
            k = ((z - 1 + fz + dk) % P.nz + P.nz) % P.nz + 1;
        }
        const double dy = min_image((double)(j - y), (double)P.ny);
        double d2 = (double)((i - x) * (i - x)) + dy * dy;
        if (DIM == 3) {
            const double dz = min_image((double)(k - z), (double)P.nz);
            d2 = d2 + dz * dz;
        }
        const long long c = (long long)(i - 1) * P.nyz + (long long)(j - 1) * P.nz + (DIM == 3 ? k - 1 : 0);
        if (d2 <= r * r && p[c] < .01) {
            aO = dd_add(aO, Phys::oxideO - o[c]);
            aC = dd_add(aC, Phys::oxideCr - cr[c]);
            o[c] = Phys::oxideO;
            cr[c] = Phys::oxideCr;
            p[c] = 1.;
            for (int g = 0; g < P.n_eta; ++g) eta[(long long)g * P.fstride + c] = (TE)0;
            ++mine;
        }
    }
    if (mine) atomicAdd(sh_count, mine);
    __syncthreads();
    const int total = *sh_count;
    __syncthreads();
    return total;
}

template <int DIM, class TE>
__global__ __launch_bounds__(1024) void k_stamp(Par P, NucPar N, double* const* obuf, double* const* crbuf,
                                                double* const* pbuf, TE* eta, DevCtl* ctl, int* changed_out) {
    __shared__ int sh_count;
    __shared__ DD sh[16];
    const int cur = ctl->cur;
    DD aO{0., 0.}, aC{0., 0.};
    const int n = block_stamp<DIM, TE>(P, obuf[cur], crbuf[cur], pbuf[cur], eta, N.x, N.y, N.z, N.r, aO, aC, &sh_count);
    // block reduce (1024 threads -> 16 waves)
    for (int off = 32; off > 0; off >>= 1) {
        aO = dd_add(aO, DD{__shfl_down(aO.hi, off), __shfl_down(aO.lo, off)});
        aC = dd_add(aC, DD{__shfl_down(aC.hi, off), __shfl_down(aC.lo, off)});
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = aO;
    __syncthreads();
    DD tO{0., 0.};
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tO = dd_add(tO, sh[w]);
    __syncthreads();
    if (lane == 0) sh[wave] = aC;
    __syncthreads();
    if (threadIdx.x == 0) {
        DD tC{0., 0.};
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tC = dd_add(tC, sh[w]);
        DD b0{ctl->budget[0], ctl->budget[1]}, b1{ctl->budget[2], ctl->budget[3]};
        b0 = dd_add(b0, tO);
        b1 = dd_add(b1, tC);
        ctl->budget[0] = b0.hi;
        ctl->budget[1] = b0.lo;
        ctl->budget[2] = b1.hi;
        ctl->budget[3] = b1.lo;
        if (n) ctl->nuclei += 1;
        *changed_out = n;
    }
}

// Phase 1: candidate flags per tile (natural tile order == lexicographic order).
template <int DIM>
__global__ __launch_bounds__(TILE) void k_ksp_mark(Par P, double r, double* const* obuf, double* const* crbuf,
                                                   double* const* pbuf, const DevCtl* ctl, int* tile_count,
                                                   int* plane_oxide) {
    const int t = blockIdx.x;
    int x = t / P.nchunk;
    const int yz = (t - x * P.nchunk) * TILE + (int)threadIdx.x;
    int cand = 0;
    if (yz < P.nyz) {
        const int cur = ctl->cur;
        const long long c = (long long)x * P.nyz + yz;
        const double pp = pbuf[cur][c];
        const int xi = x + 1;
        if (pp > .01) atomicAdd(&plane_oxide[x], 1);
        cand = (pp <= .01 && obuf[cur][c] * crbuf[cur][c] > Phys::ksp && !(xi - r < 5. || xi + r > double(P.nx - 3)));
    }
    __shared__ int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    if (cand) atomicAdd(&sh, 1);
    __syncthreads();
    if (threadIdx.x == 0) tile_count[t] = sh;
}

// Exclusive scan of n counts in place (one block); total to *total.
__global__ __launch_bounds__(1024) void k_scan_counts(int* a, int n, int* total) {
    __shared__ int part[1024];
    const int per = (n + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        int v = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x == 0 ? 0 : part[threadIdx.x - 1];
    for (int i = b; i < e; ++i) {
        const int v = a[i];
        a[i] = run;
        run += v;
    }
    if (threadIdx.x == blockDim.x - 1) *total = part[blockDim.x - 1];
}

template <int DIM>
__global__ __launch_bounds__(TILE) void k_ksp_compact(Par P, double r, double* const* obuf, double* const* crbuf,
                                                      double* const* pbuf, const DevCtl* ctl, const int* tile_off,
                                                      int* list) {
    const int t = blockIdx.x;
    int x = t / P.nchunk;
    const int yz = (t - x * P.nchunk) * TILE + (int)threadIdx.x;
    int cand = 0;
    long long c = 0;
    if (yz < P.nyz) {
        const int cur = ctl->cur;
        c = (long long)x * P.nyz + yz;
        const double pp = pbuf[cur][c];
        const int xi = x + 1;
        cand = (pp <= .01 && obuf[cur][c] * crbuf[cur][c] > Phys::ksp && !(xi - r < 5. || xi + r > double(P.nx - 3)));
    }
    __shared__ int flags[TILE];
    flags[threadIdx.x] = cand;
    __syncthreads();
    if (cand) {
        int rank = 0;
        for (int i = 0; i < (int)threadIdx.x; ++i) rank += flags[i];
        list[tile_off[t] + rank] = (int)c;
    }
}

// Phase 2: ordered greedy resolve by one workgroup.
template <int DIM, class TE>
__global__ __launch_bounds__(1024) void k_ksp_resolve(Par P, NucPar N, double* const* obuf, double* const* crbuf,
                                                      double* const* pbuf, TE* eta_base, DevCtl* ctl,
                                                      const int* list, const int* count, const int* plane_oxide,
                                                      int* centres, int* cells_out) {
    __shared__ int sh_count, sh_flag, sh_go;
    __shared__ DD sh[16];
    const int cur = ctl->cur;
    double* o = obuf[cur];
    double* cr = crbuf[cur];
    double* p = pbuf[cur];
    TE* eta = eta_base;
    const double limit = N.r + N.gap;
    const int sr = bounded_reach(limit, P.nx);
    const int rr = bounded_reach(N.r, P.nx);
    int ncent = 0, total_cells = 0;
    long long nuclei = 0;
    DD aO{0., 0.}, aC{0., 0.};
    const int ncand = *count;
    int cy, fy, cz = 1, fz = 0;
    if (DIM == 2) {
        cy = P.ny;
        fy = 0;
        const int sr2 = (int)ceil(limit);   // 2D:497 (no periodic bound on X reach)
        (void)sr2;
    } else {
        axis_span(bounded_reach(limit, P.ny), P.ny, cy, fy);
        axis_span(bounded_reach(limit, P.nz), P.nz, cz, fz);
    }
    const int srx = DIM == 2 ? (int)ceil(limit) : sr;
    for (int q = 0; q < ncand; ++q) {
        const int c = list[q];
        const int x = c / P.nyz + 1;
        const int yz = c - (x - 1) * P.nyz;
        const int y = yz / P.nz + 1, z = DIM == 3 ? yz % P.nz + 1 : 0;
        if (threadIdx.x == 0) {
            // live re-test of the scan predicate (2D:548)
            const bool still = p[c] <= .01 && o[c] * cr[c] > Phys::ksp;
            int need = 0;
            if (still) {
                const int lo = max(1, x - srx), hi = min(P.nx, x + srx);
                for (int i = lo; i <= hi && !need; ++i) need = plane_oxide[i - 1] > 0;
                for (int a = 0; a < ncent && !need; ++a) need = abs(centres[a] - x) <= srx + rr;
                if (ncent >= N.max_centres) need = 1;
            }
            sh_go = still ? (need ? 2 : 1) : 0;
            sh_flag = 0;
        }
        __syncthreads();
        const int go = sh_go;
        if (go == 0) {
            __syncthreads();
            continue;
        }
        if (go == 2) {
            const int i0 = max(1, x - srx), i1 = min(P.nx, x + srx);
            const long long nbox = (long long)(i1 - i0 + 1) * cy * cz;
            for (long long w = threadIdx.x; w < nbox; w += blockDim.x) {
                const int di = (int)(w / ((long long)cy * cz));
                const int rem = (int)(w - (long long)di * cy * cz);
                const int dj = rem / cz, dk = rem - dj * cz;
                const int i = i0 + di;
                int j, k = 1;
                if (DIM == 2) j = 1 + dj;
                else {
                    j = ((y - 1 + fy + dj) % P.ny + P.ny) % P.ny + 1;
                    k = ((z - 1 + fz + dk) % P.nz + P.nz) % P.nz + 1;
                }
                const long long cc = (long long)(i - 1) * P.nyz + (long long)(j - 1) * P.nz + (DIM == 3 ? k - 1 : 0);
                if (p[cc] > .01) {
                    const double a = (double)(i - x), b = min_image((double)(j - y), (double)P.ny);
                    double d2 = a * a + b * b;
                    if (DIM == 3) {
                        const double e = min_image((double)(k - z), (double)P.nz);
                        d2 = d2 + e * e;
                    }
                    if (sqrt(d2) < limit) sh_flag = 1;
                }
            }
            __syncthreads();
        }
        const bool valid = sh_flag == 0;
        __syncthreads();
        if (!valid) continue;
        const int n = block_stamp<DIM, TE>(P, o, cr, p, eta, x, y, z, N.r, aO, aC, &sh_count);
        if (n) {
            ++nuclei;
            total_cells += n;
            if (threadIdx.x == 0 && ncent < N.max_centres) centres[ncent] = x;
            if (ncent < N.max_centres) ++ncent;
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) {
        aO = dd_add(aO, DD{__shfl_down(aO.hi, off), __shfl_down(aO.lo, off)});
        aC = dd_add(aC, DD{__shfl_down(aC.hi, off), __shfl_down(aC.lo, off)});
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = aO;
    __syncthreads();
    DD tO{0., 0.};
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tO = dd_add(tO, sh[w]);
    __syncthreads();
    if (lane == 0) sh[wave] = aC;
    __syncthreads();
    if (threadIdx.x == 0) {
        DD tC{0., 0.};
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tC = dd_add(tC, sh[w]);
        DD b0{ctl->budget[0], ctl->budget[1]}, b1{ctl->budget[2], ctl->budget[3]};
        b0 = dd_add(b0, tO);
        b1 = dd_add(b1, tC);
        ctl->budget[0] = b0.hi;
        ctl->budget[1] = b0.lo;
        ctl->budget[2] = b1.hi;
        ctl->budget[3] = b1.lo;
        ctl->nuclei += nuclei;
        *cells_out = total_cells;
    }
}

// Centre seed (initialize_center_oxide, mpi2d/NiCr2D_grains.cpp:517-542,
// mpi3d-slab/NiCr3D_grains.cpp:725-751): every cell centre with
// (i-x)^2 + (j-y)^2 (+ (k-z)^2) <= r^2 about the box centre x = (nx+1)/2, ...,
// no periodic image and no phi test, becomes hard oxide; eta is zeroed there.
// One block walks the seed's bounding box; the inserted O and Cr go to the
// nucO/nucCr budgets (double-double), the cell count to *cells_out. The caller
// has checked the radius and that no nucleus exists yet; a seed covering no cell
// changes nothing and records nothing.
template <int DIM, class TE>
__global__ __launch_bounds__(1024) void k_center_seed(Par P, double r, double* const* obuf, double* const* crbuf,
                                                      double* const* pbuf, TE* eta, DevCtl* ctl, int* cells_out) {
    __shared__ int sh_count;
    __shared__ DD sh[16];
    const int cur = ctl->cur;
    double* o = obuf[cur];
    double* cr = crbuf[cur];
    double* p = pbuf[cur];
    const double xc = 0.5 * ((double)P.nx + 1.), yc = 0.5 * ((double)P.ny + 1.), zc = 0.5 * ((double)P.nz + 1.);
    // bounding box of the seed, clipped to the mesh (1-based)
    const int i0 = max(1, (int)floor(xc - r)), i1 = min(P.nx, (int)ceil(xc + r));
    const int j0 = max(1, (int)floor(yc - r)), j1 = min(P.ny, (int)ceil(yc + r));
    const int k0 = DIM == 3 ? max(1, (int)floor(zc - r)) : 1, k1 = DIM == 3 ? min(P.nz, (int)ceil(zc + r)) : 1;
    const long long cj = j1 - j0 + 1, ck = k1 - k0 + 1;
    const long long nbox = (long long)(i1 - i0 + 1) * cj * ck;
    if (threadIdx.x == 0) sh_count = 0;
    __syncthreads();
    DD aO{0., 0.}, aC{0., 0.};
    int mine = 0;
    for (long long q = threadIdx.x; q < nbox; q += blockDim.x) {
        const int i = i0 + (int)(q / (cj * ck));
        const long long rem = q % (cj * ck);
        const int j = j0 + (int)(rem / ck), k = k0 + (int)(rem % ck);
        const double dx = i - xc, dy = j - yc;
        double d2 = dx * dx + dy * dy;
        if (DIM == 3) {
            const double dz = k - zc;
            d2 = d2 + dz * dz;
        }
        if (d2 <= r * r) {
            const long long c = (long long)(i - 1) * P.nyz + (long long)(j - 1) * P.nz + (DIM == 3 ? k - 1 : 0);
            aO = dd_add(aO, Phys::oxideO - o[c]);
            aC = dd_add(aC, Phys::oxideCr - cr[c]);
            o[c] = Phys::oxideO;
            cr[c] = Phys::oxideCr;
            p[c] = 1.;
            if (eta)
                for (int g = 0; g < P.n_eta; ++g) eta[(long long)g * P.fstride + c] = (TE)0;
            ++mine;
        }
    }
    if (mine) atomicAdd(&sh_count, mine);
    const DD tO = block_reduce_dd(aO, sh);
    const DD tC = block_reduce_dd(aC, sh);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int n = sh_count;
        if (n) {
            DD b0{ctl->budget[0], ctl->budget[1]}, b1{ctl->budget[2], ctl->budget[3]};
            b0 = dd_add(b0, tO);
            b1 = dd_add(b1, tC);
            ctl->budget[0] = b0.hi;
            ctl->budget[1] = b0.lo;
            ctl->budget[2] = b1.hi;
            ctl->budget[3] = b1.lo;
            ctl->nuclei += 1;
        }
        *cells_out = n;
    }
}

}  // namespace nicr
