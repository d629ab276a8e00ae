// nicr_grains.hip -- the per-cell parts of the 3D grain rows on the device
// (nicr_grain_terms, include/nicr.h). Included by nicr_api.hip.
//
// Model::grain_log and grain_energy (mpi3d-slab/NiCr3D_grains.cpp:546-590)
// reduce the eta fields at output time: extrema, the dominant grain of every
// metal cell, and the discrete grain energy, whose per-cell density needs every
// eta field at the cell and at its x+, y+, z+ neighbours. Those are computed
// here from the device state. The energy's per-cell term is binary64 in the
// host code's operation order (no contraction, pow(a, 2) as a * a), so the
// host's long double sum over the terms, in cell order, is the reference's;
// extrema and counts do not depend on order. What stays on the host are the
// order-dependent long double sums (sum of S, the masked GB sum, the
// per-grain amounts).

namespace nicr {

struct GrainPart {
    double lo, hi, slo, shi, gbmax;
};

// One thread per cell. term: the energy density times the cell volume.
// part: per-block extrema; counts: [0] unassigned metal cells, [g] cells whose
// dominant grain is g (1-based).
template <class TE, class TM>
__global__ __launch_bounds__(256) void k_grain_terms(Par P, const TE* __restrict__ eta, const TM* __restrict__ s2,
                                                     const TM* __restrict__ gb, const double* __restrict__ p,
                                                     int n_eta, double etaW, double etaA, double etaK,
                                                     double* __restrict__ term, GrainPart* __restrict__ part,
                                                     unsigned long long* __restrict__ counts) {
    __shared__ unsigned int hist[257];
    __shared__ GrainPart red[256];
    const int tid = threadIdx.x;
    for (int g = tid; g <= n_eta; g += 256) hist[g] = 0u;
    __syncthreads();
    const long long q = (long long)blockIdx.x * 256 + tid;
    // std::min / std::max as the host writes them: a NaN never replaces the running value
    auto mn = [](double a, double b) { return (b < a) ? b : a; };
    auto mx = [](double a, double b) { return (a < b) ? b : a; };
    GrainPart r{1., 0., __builtin_inf(), 0., 0.};
    if (q < P.ncell) {
        const int nz = P.nz, ny = P.ny;
        const long long plane = P.nyz;
        const int i = (int)(q / plane) + 1;
        const long long rem = q - (long long)(i - 1) * plane;
        const int j = (int)(rem / nz) + 1, k = (int)(rem - (long long)(j - 1) * nz) + 1;
        const double S = (double)s2[q], ph = p[q];
        r.slo = mn(r.slo, S);
        r.shi = mx(r.shi, S);
        r.gbmax = mx(r.gbmax, (double)gb[q]);
        // current_grain (SL:527-533)
        int id = 0;
        if (!(ph >= .5)) {
            double best = 1.e-12;
            for (int g = 0; g < n_eta; ++g) {
                const double v = (double)eta[(long long)g * P.fstride + q];
                if (v > best) {
                    best = v;
                    id = g + 1;
                }
            }
        }
        if (id) atomicAdd(&hist[id], 1u);
        else if (ph < .5) atomicAdd(&hist[0], 1u);
        // grain_energy's density (SL:546-565)
        double Ss = 0., pairs = 0.;
        for (int g = 0; g < n_eta; ++g) {
            const double e = (double)eta[(long long)g * P.fstride + q];
            const double v = e * e;
            pairs += Ss * v;
            Ss += v;
            r.lo = mn(r.lo, e);
            r.hi = mx(r.hi, e);
        }
        const double H = ph * ph * ph * (10. - 15. * ph + 6. * ph * ph), d = S - (1. - H);
        double density = 0.25 * etaW * d * d + etaW * pairs + 0.5 * etaA * H * S;
        const long long qjp = q + (j == ny ? -(long long)(ny - 1) * nz : (long long)nz);
        const long long qkp = q + (k == nz ? -(long long)(nz - 1) : 1LL);
        for (int g = 0; g < n_eta; ++g) {
            const TE* f = eta + (long long)g * P.fstride;
            const double e = (double)f[q];
            if (i < P.nx) {
                const double t = ((double)f[q + plane] - e) / P.dx;
                density += .5 * etaK * (t * t);
            }
            const double ty = ((double)f[qjp] - e) / P.dy;
            density += .5 * etaK * (ty * ty);
            const double tz = ((double)f[qkp] - e) / P.dz;
            density += .5 * etaK * (tz * tz);
        }
        term[q] = density * P.dx * P.dy * P.dz;
    }
    red[tid] = r;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            GrainPart a = red[tid];
            const GrainPart b = red[tid + s];
            a.lo = mn(a.lo, b.lo);
            a.hi = mx(a.hi, b.hi);
            a.slo = mn(a.slo, b.slo);
            a.shi = mx(a.shi, b.shi);
            a.gbmax = mx(a.gbmax, b.gbmax);
            red[tid] = a;
        }
        __syncthreads();
    }
    if (tid == 0) part[blockIdx.x] = red[0];
    for (int g = tid; g <= n_eta; g += 256)
        if (hist[g]) atomicAdd(&counts[g], (unsigned long long)hist[g]);
}

}  // namespace nicr

extern "C" int nicr_grain_terms(void* h, double* energy_terms, double extrema[5], int64_t* counts) {
    using namespace nicr;
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !energy_terms || !extrema || !counts) return fail(NICR_E_INVALID, "nicr_grain_terms: null argument");
    if (c->dim != 3 || !c->has_eta()) return fail(NICR_E_INVALID, "nicr_grain_terms: needs a 3D context with eta fields");
    if (c->cfg.n_eta > 256) return fail(NICR_E_INVALID, "nicr_grain_terms: at most 256 eta fields");
    if (c->mask_dirty) {
        do_mask(c);
        c->mask_dirty = false;
    }
    const long long nb = (c->ncell + 255) / 256;
    const int n = c->cfg.n_eta;
    if (!c->grain_term) {
        int rc = dalloc(c, (void**)&c->grain_term, (size_t)c->ncell * 8);
        if (!rc) rc = dalloc(c, (void**)&c->grain_part, (size_t)nb * sizeof(GrainPart));
        if (!rc) rc = dalloc(c, (void**)&c->grain_counts, 257 * sizeof(unsigned long long));
        if (rc) return rc;
    }
    const int mi = mask_index(c);
    const double* p = c->p[c->cur];
    HIPCHK(hipMemsetAsync(c->grain_counts, 0, (size_t)(n + 1) * sizeof(unsigned long long), c->stream));
    const Par& P = c->P;
    if (c->te == 4)
        hipLaunchKernelGGL((k_grain_terms<float, float>), dim3((unsigned)nb), dim3(256), 0, c->stream, P,
                           (const float*)c->eta[mi], (const float*)c->s2[mi], (const float*)c->gb[mi], p, n,
                           c->cfg.eta_W, c->cfg.eta_A, c->cfg.eta_K, c->grain_term, (GrainPart*)c->grain_part,
                           (unsigned long long*)c->grain_counts);
    else
        hipLaunchKernelGGL((k_grain_terms<double, double>), dim3((unsigned)nb), dim3(256), 0, c->stream, P,
                           (const double*)c->eta[mi], (const double*)c->s2[mi], (const double*)c->gb[mi], p, n,
                           c->cfg.eta_W, c->cfg.eta_A, c->cfg.eta_K, c->grain_term, (GrainPart*)c->grain_part,
                           (unsigned long long*)c->grain_counts);
    HIPCHK(hipGetLastError());
    std::vector<GrainPart> parts((size_t)nb);
    std::vector<unsigned long long> cnt((size_t)n + 1);
    HIPCHK(hipMemcpyAsync(energy_terms, c->grain_term, (size_t)c->ncell * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(parts.data(), c->grain_part, (size_t)nb * sizeof(GrainPart), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), c->grain_counts, cnt.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    GrainPart r{1., 0., __builtin_inf(), 0., 0.};
    for (const GrainPart& b : parts) {
        r.lo = (b.lo < r.lo) ? b.lo : r.lo;
        r.hi = (r.hi < b.hi) ? b.hi : r.hi;
        r.slo = (b.slo < r.slo) ? b.slo : r.slo;
        r.shi = (r.shi < b.shi) ? b.shi : r.shi;
        r.gbmax = (r.gbmax < b.gbmax) ? b.gbmax : r.gbmax;
    }
    extrema[0] = r.lo;
    extrema[1] = r.hi;
    extrema[2] = r.slo;
    extrema[3] = r.shi;
    extrema[4] = r.gbmax;
    for (int g = 0; g <= n; ++g) counts[g] = (int64_t)cnt[(size_t)g];
    return NICR_OK;
}
