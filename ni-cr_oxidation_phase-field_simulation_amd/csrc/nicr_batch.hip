// nicr_batch.hip -- ensembles of independent 2D instances (different mesh sizes
// and inlet O supply) stepped together. Included by nicr_api.hip.
//
// Every instance follows the single-instance 2D step exactly (binary64, the
// arithmetic of k_mu + k_trial: mu64_vals, mob_vals, trial_mm, finish_cell), with
// its own state buffers, accepted-buffer index, failure flag and budgets, so an
// instance gives bitwise the results of its own nicr context. One ensemble step
// is two launches over all instances:
//   k_batch_pde     a persistent grid over a flat list of tiles (TX rows of x by
//                   64 cells of y, of one instance), a contiguous run of tiles
//                   per block: o, cr, phi, gb of the tile and its halo into LDS,
//                   mu and M over the tile and a one-cell halo, then the trial of
//                   the tile; budget partials per run of one instance's tiles,
//                   first failing cell per instance;
//   k_batch_commit  one block per instance: that instance's all-or-nothing
//                   acceptance (k_commit per instance).
// Sequences of steps are captured once into a HIP graph and replayed, so a run of
// n steps costs one graph launch per chunk instead of 2n kernel launches. The
// tile height TX and the grid size are chosen per ensemble by an autotuner that
// times the candidates on dry steps: the kernels run in full but the commit
// neither accepts nor flags, so tuning leaves every instance unchanged.
#include <map>

namespace nicr {

struct BInst {
    int nx, ny;         // instance mesh (x: Neumann, y: periodic)
    long long off;      // first cell of the instance in the ensemble arrays
    int tile0, ntile;   // the instance's tiles in the tile list
    double layer_o;     // inlet O of the concentration BC
};
struct BTile {
    int inst, x0, y0;
};
struct BBufs {
    double* o[2];
    double* cr[2];
    double* p[2];
    double* gb;         // static GB mask (grains off: zero)
    double* partials;   // NPART per tile
    DevCtl* ctl;        // one per instance
    const BInst* inst;
    const BTile* tiles;
};

constexpr int BW = 64;   // tile width along y (one wave)

// The tile list, the instance table and each instance's buffer index and
// failure flag are uniform over a block and do not change during k_batch_pde
// (only k_batch_commit writes them). They are read through the scalar data
// cache (constant address space: s_load). Vector loads of them would share the
// in-order vmcnt counter with the state loads and stores, so every wait for a
// metadata word would also wait for the next tile's prefetch or the previous
// tile's stores.
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}
// Records field by field: a struct copy becomes a vector load (an aggregate
// copy is not kept in the constant address space, and there is no three-dword
// scalar load).
__device__ __forceinline__ BTile ld_tile(const BTile* p) {
    return BTile{ld_uniform(&p->inst), ld_uniform(&p->x0), ld_uniform(&p->y0)};
}
__device__ __forceinline__ BInst ld_inst(const BInst* p) {
    BInst I;
    I.nx = ld_uniform(&p->nx);
    I.ny = ld_uniform(&p->ny);
    I.off = ld_uniform(&p->off);
    I.tile0 = ld_uniform(&p->tile0);
    I.ntile = ld_uniform(&p->ntile);
    I.layer_o = ld_uniform(&p->layer_o);
    return I;
}

template <int TX>
struct BatchLds {
    static constexpr int W2 = BW + 4, H2 = TX + 4, W1 = BW + 2, H1 = TX + 2;
    static constexpr int P2 = W2 * H2, P1 = W1 * H1;
    static constexpr size_t bytes() {
        return size_t(P1) * sizeof(MuMob<double>) + size_t(2) * P2 * 8 + size_t(2) * P1 * 8 + size_t(W2) * 8;
    }
};

template <int TX>
__global__ __launch_bounds__(64 * TX) void k_batch_pde(Par P, BBufs B, int ntile) {
    using L = BatchLds<TX>;
    using MM = MuMob<double>;
    constexpr int NTH = 64 * TX;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MM* smm = reinterpret_cast<MM*>(lds_raw);                    // [P1] mu, M
    double* so = reinterpret_cast<double*>(smm + L::P1);         // [P2] o (inlet substituted)
    double* sc = so + L::P2;                                     // [P2] cr
    double* sp = sc + L::P2;                                     // [P1] phi
    double* sg = sp + L::P1;                                     // [P1] gb
    double* so0 = sg + L::P1;                                    // [W2] o of row x = 0 as stored
    const int tid = threadIdx.x;
    const int tx = tid >> 6, ty = tid & 63;
    // A block takes a contiguous run of tiles (the list is grouped by
    // instance) and sums the budgets of consecutive tiles of one instance in
    // registers: one block reduction per instance run, into the slot of the
    // run's first tile; the run's other slots hold zeros (the commit sums all
    // slots of an instance).
    const int per = (ntile + (int)gridDim.x - 1) / (int)gridDim.x;
    const int tbeg = (int)blockIdx.x * per, tend = min(ntile, tbeg + per);
    StepAcc acc;
    int run_inst = -1, run_slot = 0;
    auto flush = [&]() {
        if (run_inst >= 0) {
            block_partials(acc, B.partials + (long long)run_slot * NPART, &B.ctl[run_inst].pde_key);
            acc = StepAcc{};
        }
    };
    // the next tile's o, cr (halo 2) and phi, gb (halo 1) are loaded into
    // registers while the current tile computes
    constexpr int R2 = (L::P2 + NTH - 1) / NTH, R1 = (L::P1 + NTH - 1) / NTH;
    double st_o[R2], st_c[R2], st_p[R1], st_g[R1];
    auto fetch = [&](int t) {
        const BTile tile = ld_tile(B.tiles + t);
        const BInst I = ld_inst(B.inst + tile.inst);
        const int cur = ld_uniform(&B.ctl[tile.inst].cur);
        const double* o = B.o[cur] + I.off;
        const double* cr = B.cr[cur] + I.off;
        const double* p = B.p[cur] + I.off;
        const double* gb = B.gb + I.off;
        const int nx = I.nx, ny = I.ny;
        auto clampx = [&](int v) { return v < 0 ? 0 : (v >= nx ? nx - 1 : v); };
        auto wrapy = [&](int v) { return ((v % ny) + ny) % ny; };
        // o, cr over rows x0-2 .. x0+TX+1 (clamped: the X ghosts are copies),
        // columns y0-2 .. y0+65 (periodic); phi, gb over the one-cell halo.
        // Every lane loads (slots past the tile repeat its last cell; the
        // staging skips them) and the inlet value replaces O at the staging:
        // with no branch around the loads and no value written into their
        // registers, the compiler's waits do not serialise them.
#pragma unroll
        for (int k = 0; k < R2; ++k) {
            const int q = min(tid + k * NTH, L::P2 - 1);
            const int r = q / L::W2, cy = q - r * L::W2;
            const long long g = (long long)clampx(tile.x0 - 2 + r) * ny + wrapy(tile.y0 - 2 + cy);
            st_o[k] = o[g];
            st_c[k] = cr[g];
        }
#pragma unroll
        for (int k = 0; k < R1; ++k) {
            const int q = min(tid + k * NTH, L::P1 - 1);
            const int r = q / L::W1, cy = q - r * L::W1;
            const long long g = (long long)clampx(tile.x0 - 1 + r) * ny + wrapy(tile.y0 - 1 + cy);
            st_p[k] = p[g];
            st_g[k] = gb[g];
        }
    };
    // A tile's results wait in registers until the next tile's values are
    // staged: stored before the staging, they would enter the in-order vmcnt
    // count ahead of its last wait, which would then wait for them as well.
    double r_o = 0., r_c = 0., r_p = 0.;
    long long r_g = -1;   // the cell in the ensemble arrays; -1: nothing to store
    int r_nxt = 0;
    auto store_results = [&]() {
        if (r_g >= 0) {
            st_nt(&B.o[r_nxt][r_g], r_o);
            st_nt(&B.cr[r_nxt][r_g], r_c);
            st_nt(&B.p[r_nxt][r_g], r_p);
        }
        r_g = -1;
    };
    if (tbeg < tend) fetch(tbeg);
    for (int t = tbeg; t < tend; ++t) {
        const BTile tile = ld_tile(B.tiles + t);
        if (tile.inst != run_inst) {
            flush();
            run_inst = tile.inst;
            run_slot = t;
        } else if (tid < NPART) {
            B.partials[(long long)t * NPART + tid] = 0.;
        }
        // this tile's staged values into LDS (O of rows at or before the
        // inlet plane x = 0 is the inlet value), then the next tile's loads
        const bool enf = P.enforce_layer != 0;
        const BInst I = ld_inst(B.inst + tile.inst);
#pragma unroll
        for (int k = 0; k < R2; ++k) {
            const int q = tid + k * NTH;
            if (q < L::P2) {
                const int r = q / L::W2, v = tile.x0 - 2 + r;
                const int xs = max(0, min(v, I.nx - 1));   // the clamped row
                so[q] = (enf && xs == 0) ? I.layer_o : st_o[k];
                if (v == 0) so0[q - r * L::W2] = st_o[k];
                sc[q] = st_c[k];
            }
        }
#pragma unroll
        for (int k = 0; k < R1; ++k) {
            const int q = tid + k * NTH;
            if (q < L::P1) {
                sp[q] = st_p[k];
                sg[q] = st_g[k];
            }
        }
        store_results();
        __syncthreads();
        if (t + 1 < tend) fetch(t + 1);
        if (ld_uniform(&B.ctl[tile.inst].failed)) {
            __syncthreads();   // the LDS tile is rewritten by the next iteration
            continue;
        }
        const int cur = ld_uniform(&B.ctl[tile.inst].cur), nxt = cur ^ 1;
        const int nx = I.nx, ny = I.ny;
        auto clampx = [&](int v) { return v < 0 ? 0 : (v >= nx ? nx - 1 : v); };
        // mu and M over the tile and its halo. A halo row outside the mesh is
        // the ghost copy of its clamped row, mu included (2D:427), so it is
        // evaluated at the clamped row.
        for (int q = tid; q < L::P1; q += NTH) {
            const int r = q / L::W1, cy = q - r * L::W1;
            const int xr = clampx(tile.x0 - 1 + r);
            const int rr = xr - (tile.x0 - 2);          // row of xr in the halo-2 tile
            const int rp = xr - (tile.x0 - 1);          // row of xr in the halo-1 tile
            const int c2 = rr * L::W2 + cy + 1, c1 = rp * L::W1 + cy;
            double o7[7], c7[7];
            o7[SC] = so[c2];
            c7[SC] = sc[c2];
            o7[SXM] = so[c2 - L::W2];
            c7[SXM] = sc[c2 - L::W2];
            o7[SXP] = so[c2 + L::W2];
            c7[SXP] = sc[c2 + L::W2];
            o7[SYM] = so[c2 - 1];
            c7[SYM] = sc[c2 - 1];
            o7[SYP] = so[c2 + 1];
            c7[SYP] = sc[c2 + 1];
            o7[SZM] = o7[SZP] = o7[SC];
            c7[SZM] = c7[SZP] = c7[SC];
            smm[q] = mu_mob<2, PREC_FP64, double>(P, o7, c7, sp[c1], sg[c1]);
        }
        __syncthreads();
        // trial of the tile's cells
        const int x = tile.x0 + tx, y = tile.y0 + ty;
        if (x < nx && y < ny) {
            const int c1 = (tx + 1) * L::W1 + ty + 1, c2 = (tx + 2) * L::W2 + ty + 2;
            MM m[7];
            double p7[7];
            m[SC] = smm[c1];
            m[SXM] = smm[c1 - L::W1];
            m[SXP] = smm[c1 + L::W1];
            m[SYM] = smm[c1 - 1];
            m[SYP] = smm[c1 + 1];
            p7[SC] = sp[c1];
            p7[SXM] = sp[c1 - L::W1];
            p7[SXP] = sp[c1 + L::W1];
            p7[SYM] = sp[c1 - 1];
            p7[SYP] = sp[c1 + 1];
            const double oo = so[c2], cc = sc[c2];
            const long long c = (long long)x * ny + y;
            const double o_stored = x == 0 ? so0[ty + 2] : oo;
            double onv, crv, pnv, left_flux;
            trial_mm<2, PREC_FP64, double>(P, m, p7, oo, cc, x, 0., sg[c1], 0., onv, crv, pnv, left_flux);
            finish_cell_vals<PREC_FP64>(P, c, x, o_stored, oo, cc, p7[SC], onv, crv, pnv, left_flux, acc, I.layer_o);
            r_o = onv;
            r_c = crv;
            r_p = pnv;
            r_g = I.off + c;
        }
        r_nxt = nxt;
        __syncthreads();   // LDS is reused by the next tile
    }
    store_results();
    flush();
}

// All-or-nothing acceptance of each instance's step (k_commit per instance).
// dry: the step is discarded (tuning runs): failure keys are cleared, nothing
// else changes.
__global__ __launch_bounds__(256) void k_batch_commit(BBufs B, int dry) {
    const int i = blockIdx.x;
    DevCtl* ctl = B.ctl + i;
    if (ctl->failed) return;
    if (dry) {
        __syncthreads();
        if (threadIdx.x == 0) ctl->pde_key = NO_KEY;
        return;
    }
    if (ctl->pde_key != NO_KEY) {
        if (threadIdx.x == 0) {
            ctl->failed = 1;
            ctl->fail_kind = 1;
        }
        return;
    }
    const BInst I = B.inst[i];
    __shared__ DD sh[6 * 4];
    DD acc[6];
    for (int s = 0; s < 6; ++s) acc[s] = DD{0., 0.};
    for (int b = threadIdx.x; b < I.ntile; b += blockDim.x) {
        const double* w = B.partials + (long long)(I.tile0 + b) * NPART;
        for (int s = 0; s < 6; ++s) acc[s] = dd_add(acc[s], DD{w[2 * s], w[2 * s + 1]});
    }
    const DD r = block_reduce_dd6(acc, sh);
    // thread s < 5 adds sum s (bcO, ckO, ckCr, clipO, clipCr) into its budget
    // slot (6, 2, 3, 4, 5), thread 5 the clipped count
    const int s = threadIdx.x;
    if (s < 5) {
        const int k = s == 0 ? 6 : s + 1;
        DD v{ctl->budget[2 * k], ctl->budget[2 * k + 1]};
        v = dd_add(v, r);
        ctl->budget[2 * k] = v.hi;
        ctl->budget[2 * k + 1] = v.lo;
    } else if (s == 5) {
        ctl->clipped += (long long)r.hi;
    }
    if (s == 0) {
        ctl->cur ^= 1;
        ctl->steps_done += 1;
    }
}

__global__ void k_batch_reset(DevCtl* ctl, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ctl[i].failed = 0;
    ctl[i].steps_done = 0;
    ctl[i].fail_kind = 0;
    ctl[i].pde_key = NO_KEY;
    ctl[i].eta_key = NO_KEY;
}

// Concentration-BC row left in place by a rejected 2D step (as k_apply_layer).
__global__ void k_batch_apply_layer(BBufs B, int i) {
    const BInst I = B.inst[i];
    const DevCtl* ctl = B.ctl + i;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y < I.ny) B.o[ctl->cur][I.off + y] = I.layer_o;
}

}  // namespace nicr

namespace {

// tile heights with a compiled kernel
constexpr int BATCH_TX[] = {1, 2, 4, 8};

struct BatchCtx {
    nicr_config cfg{};
    Par P{};
    int n = 0;
    long long cells = 0, total = 0;   // interior cells; array length (instances padded)
    std::vector<BInst> inst;   // host copy (offsets and sizes)
    hipStream_t stream = nullptr;
    char* slab = nullptr;
    double* o[2] = {nullptr, nullptr};
    double* cr[2] = {nullptr, nullptr};
    double* p[2] = {nullptr, nullptr};
    double* gb = nullptr;
    double* partials = nullptr;
    DevCtl* ctl = nullptr;
    BInst* d_inst = nullptr;
    BTile* d_tiles[4] = {nullptr, nullptr, nullptr, nullptr};   // tile list per TX
    int ntile[4] = {0, 0, 0, 0};
    int tx_index = 2;                 // tile height in use (index into BATCH_TX)
    int grid = 0;                     // persistent grid size (0: resident blocks)
    bool use_graph = true;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;   // (steps, tx_index) -> graph
    std::vector<void*> allocs;
    size_t bytes = 0;
};

int balloc(BatchCtx* b, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(NICR_E_DEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    b->allocs.push_back(*p);
    b->bytes += bytes;
    return NICR_OK;
}

BBufs bbufs(BatchCtx* b) {
    BBufs B;
    for (int i = 0; i < 2; ++i) {
        B.o[i] = b->o[i];
        B.cr[i] = b->cr[i];
        B.p[i] = b->p[i];
    }
    B.gb = b->gb;
    B.partials = b->partials;
    B.ctl = b->ctl;
    B.inst = b->d_inst + (size_t)b->tx_index * b->n;   // the instance table of this tile height
    B.tiles = b->d_tiles[b->tx_index];
    return B;
}

// Resident blocks of k_batch_pde<TX> on the device (its persistent grid), or 0
// when a device query fails: a launch then takes one block per tile, which is
// always a valid grid. A failed LDS attribute surfaces as the launch's error.
template <int TX>
int batch_resident(BatchCtx* b) {
    static int resident = -1;
    if (resident < 0) {
        const size_t lds = BatchLds<TX>::bytes();
        int per_cu = 0, ncu = 0, dev = 0;
        const bool ok =
            hipFuncSetAttribute((const void*)k_batch_pde<TX>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                hipSuccess &&
            hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_pde<TX>, 64 * TX, lds) == hipSuccess &&
            per_cu > 0 && ncu > 0;
        resident = ok ? per_cu * ncu : 0;
    }
    return resident;
}
template <int TX>
void batch_pde_launch(BatchCtx* b, const BBufs& B, int ntile) {
    const int res = batch_resident<TX>(b);
    const int grid = std::max(1, std::min(ntile, b->grid > 0 ? b->grid : (res > 0 ? res : ntile)));
    hipLaunchKernelGGL(k_batch_pde<TX>, dim3(grid), dim3(64 * TX), BatchLds<TX>::bytes(), b->stream, b->P, B, ntile);
}
// one ensemble step (dry: discarded, for tuning)
void batch_step_launch(BatchCtx* b, int dry) {
    const BBufs B = bbufs(b);
    const int nt = b->ntile[b->tx_index];
    switch (BATCH_TX[b->tx_index]) {
        case 1: batch_pde_launch<1>(b, B, nt); break;
        case 2: batch_pde_launch<2>(b, B, nt); break;
        case 4: batch_pde_launch<4>(b, B, nt); break;
        default: batch_pde_launch<8>(b, B, nt); break;
    }
    hipLaunchKernelGGL(k_batch_commit, dim3(b->n), dim3(256), 0, b->stream, B, dry);
}
int batch_resident_for(BatchCtx* b, int txi) {
    switch (BATCH_TX[txi]) {
        case 1: return batch_resident<1>(b);
        case 2: return batch_resident<2>(b);
        case 4: return batch_resident<4>(b);
        default: return batch_resident<8>(b);
    }
}

// graph of `steps` ensemble steps with the current tiling
int batch_graph(BatchCtx* b, int steps, hipGraphExec_t* out) {
    const auto key = std::make_pair(steps, b->tx_index * 100000 + b->grid);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return NICR_OK;
    }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps; ++s) batch_step_launch(b, 0);
    HIPCHK(hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t ge;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(NICR_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->graphs[key] = ge;
    *out = ge;
    return NICR_OK;
}

void batch_drop_graphs(BatchCtx* b) {
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// A 2D binary64 context without eta fields (grains off or static) steps as a
// one-instance ensemble: k_batch_pde evaluates mu, M and the trial of each tile
// in one pass, one launch fewer than k_mu + k_trial, and k_batch_commit accepts
// or rejects. The fields are bitwise those of the two-kernel path (the same
// arithmetic in the same order); the budget partials are grouped by runs of
// tiles instead of by blocks, so the budgets can differ in their last bits.
// Small meshes are launch-bound: 128 x 128, 1000 steps, 20.5 -> 14.6 us per
// step on one box. The tables are built on the first step.
template <int TX>
void batch2d_pde(Ctx* c, const BBufs& B) {
    const int res = batch_resident<TX>(nullptr);
    const int grid = std::max(1, std::min(c->b1_ntile, res > 0 ? res : c->b1_ntile));
    hipLaunchKernelGGL(k_batch_pde<TX>, dim3(grid), dim3(64 * TX), BatchLds<TX>::bytes(), c->stream, c->P, B,
                       c->b1_ntile);
}
bool batch2d_step(Ctx* c) {
    if (c->dim != 2 || c->prec != PREC_FP64 || c->has_eta() || c->P.nx < 2 || c->b1_disabled) return false;
    if (!c->b1_inst) {
        const int nx = c->P.nx, ny = c->P.ny;
        const int tx = ((nx + 3) / 4) * ((ny + BW - 1) / BW) >= 1024 ? 4 : 2;   // enough tiles for the CUs
        std::vector<BTile> tiles;
        for (int x0 = 0; x0 < nx; x0 += tx)
            for (int y0 = 0; y0 < ny; y0 += BW) tiles.push_back(BTile{0, x0, y0});
        BInst I{};
        I.nx = nx;
        I.ny = ny;
        I.off = 0;
        I.tile0 = 0;
        I.ntile = (int)tiles.size();
        I.layer_o = c->P.layer_o;
        void *inst = nullptr, *tl = nullptr, *part = nullptr;
        int rc = dalloc(c, &inst, sizeof(BInst));
        if (rc == NICR_OK) rc = dalloc(c, &tl, tiles.size() * sizeof(BTile));
        if (rc == NICR_OK) rc = dalloc(c, &part, tiles.size() * NPART * sizeof(double));
        if (rc == NICR_OK && (hipMemcpy(inst, &I, sizeof(I), hipMemcpyHostToDevice) != hipSuccess ||
                              hipMemcpy(tl, tiles.data(), tiles.size() * sizeof(BTile), hipMemcpyHostToDevice) !=
                                  hipSuccess))
            rc = fail(NICR_E_DEVICE, "hipMemcpy of the 2D tile tables failed");
        if (rc != NICR_OK) {
            c->launch_error = std::string("2D step set-up failed: ") + nicr_last_error();
            return true;   // nothing of this step is queued; nicr_step reports it
        }
        c->b1_inst = inst;
        c->b1_tiles = tl;
        c->b1_partials = (double*)part;
        c->b1_ntile = I.ntile;
        c->b1_tx = tx;
    }
    BBufs B;
    for (int i = 0; i < 2; ++i) {
        B.o[i] = c->o[i];
        B.cr[i] = c->cr[i];
        B.p[i] = c->p[i];
    }
    B.gb = (double*)c->gb[0];
    B.partials = c->b1_partials;
    B.ctl = c->ctl;
    B.inst = (const BInst*)c->b1_inst;
    B.tiles = (const BTile*)c->b1_tiles;
    tmark_begin(c, T_TRIAL);
    if (c->b1_tx == 4)
        batch2d_pde<4>(c, B);
    else
        batch2d_pde<2>(c, B);
    tmark_end(c);
    tmark_begin(c, T_COMMIT);
    hipLaunchKernelGGL(k_batch_commit, dim3(1), dim3(256), 0, c->stream, B, 0);
    tmark_end(c);
    return true;
}

int batch_field_ptr(BatchCtx* b, int inst, int field, double** ptr) {
    if (inst < 0 || inst >= b->n) return fail(NICR_E_INVALID, "batch: instance index out of range");
    DevCtl h;
    HIPCHK(hipMemcpy(&h, b->ctl + inst, sizeof(h), hipMemcpyDeviceToHost));
    const long long off = b->inst[inst].off;
    switch (field) {
        case NICR_FIELD_O: *ptr = b->o[h.cur] + off; break;
        case NICR_FIELD_CR: *ptr = b->cr[h.cur] + off; break;
        case NICR_FIELD_PHI: *ptr = b->p[h.cur] + off; break;
        case NICR_FIELD_GB: *ptr = b->gb + off; break;
        default: return fail(NICR_E_INVALID, "batch: field must be O, CR, PHI or GB");
    }
    return NICR_OK;
}

}  // namespace

extern "C" {

int nicr_batch_create(const nicr_config* common, const nicr_batch_instance* in, int n, void** out) {
    if (!common || !in || n < 1 || !out) return fail(NICR_E_INVALID, "nicr_batch_create: bad argument");
    if (common->abi_version != NICR_ABI_VERSION) return fail(NICR_E_INVALID, "nicr_batch_create: ABI version mismatch");
    if (common->dim != 2) return fail(NICR_E_INVALID, "nicr_batch_create: ensembles are 2D");
    if (common->precision != NICR_PRECISION_FP64)
        return fail(NICR_E_INVALID, "nicr_batch_create: ensembles step in binary64");
    if (common->grains && common->grain_evolution)
        return fail(NICR_E_INVALID, "nicr_batch_create: evolving grains are not supported in ensembles");
    for (int i = 0; i < n; ++i)
        if (in[i].nx < 2 || in[i].ny < 1) return fail(NICR_E_INVALID, "nicr_batch_create: instance mesh too small");
    BatchCtx* b = new BatchCtx;
    b->cfg = *common;
    b->n = n;
    Par& P = b->P;
    P.n_eta = 0;
    P.bc = common->bc;
    P.ck_legacy = common->ck == NICR_CK_LEGACY;
    P.bounds_legacy = common->bounds == NICR_BOUNDS_LEGACY;
    P.grains = common->grains;
    P.oxy = common->oxygen_diffusion;
    P.has_eta = 0;
    P.eta_evolves = 0;
    P.enforce_layer = common->bc == NICR_BC_CONCENTRATION && common->oxygen_diffusion;
    P.dx = common->dx;
    P.dy = common->dy;
    P.dz = common->dz;
    P.dt = common->dt;
    P.ck_o_dt = Phys::ckO / P.dt;
    P.ck_cr_dt = Phys::ckCr / P.dt;
    P.gbfo = common->gb_factor_o;
    P.gbfc = common->gb_factor_cr;
    P.mu_res = common->mu_res;
    P.c_ref = common->surface_c_ref;
    P.layer_o = common->layer_o;
    P.nz = 1;
    // instances, each padded to 32 cells (256 bytes) with an odd number of
    // 256-byte units between starts (HBM channel spread, as in nicr_create)
    long long off = 0;
    for (int i = 0; i < n; ++i) {
        BInst I{};
        I.nx = in[i].nx;
        I.ny = in[i].ny;
        I.off = off;
        I.layer_o = in[i].layer_o;
        long long units = ((long long)I.nx * I.ny + 31) / 32;
        if (units % 2 == 0) ++units;
        off += units * 32;
        b->cells += (long long)I.nx * I.ny;
        b->inst.push_back(I);
    }
    b->total = off;
    int rc = NICR_OK;
    auto cleanup = [&]() {
        for (void* q : b->allocs) hipFree(q);
        if (b->stream) hipStreamDestroy(b->stream);
        delete b;
    };
    if (hipStreamCreate(&b->stream) != hipSuccess) {
        cleanup();
        return fail(NICR_E_DEVICE, "nicr_batch_create: hipStreamCreate failed");
    }
    const size_t fb = (size_t)b->total * 8;
    rc = balloc(b, (void**)&b->slab, 7 * fb);
    if (rc == NICR_OK) {
        for (int i = 0; i < 2; ++i) {
            b->o[i] = (double*)(b->slab + (3 * i + 0) * fb);
            b->cr[i] = (double*)(b->slab + (3 * i + 1) * fb);
            b->p[i] = (double*)(b->slab + (3 * i + 2) * fb);
        }
        b->gb = (double*)(b->slab + 6 * fb);
    }
    // tile lists for every compiled tile height; instance tiles are contiguous
    int max_tiles = 0;
    std::vector<std::vector<BTile>> lists(4);
    std::vector<std::vector<std::pair<int, int>>> ranges(4);
    for (int k = 0; k < 4; ++k) {
        const int TXk = BATCH_TX[k];
        for (int i = 0; i < n; ++i) {
            const int t0 = (int)lists[k].size();
            for (int x0 = 0; x0 < b->inst[i].nx; x0 += TXk)
                for (int y0 = 0; y0 < b->inst[i].ny; y0 += BW) lists[k].push_back(BTile{i, x0, y0});
            ranges[k].push_back({t0, (int)lists[k].size() - t0});
        }
        b->ntile[k] = (int)lists[k].size();
        max_tiles = std::max(max_tiles, b->ntile[k]);
        if (rc == NICR_OK) rc = balloc(b, (void**)&b->d_tiles[k], lists[k].size() * sizeof(BTile));
        if (rc == NICR_OK)
            rc = hipMemcpy(b->d_tiles[k], lists[k].data(), lists[k].size() * sizeof(BTile), hipMemcpyHostToDevice) ==
                         hipSuccess ? NICR_OK : fail(NICR_E_DEVICE, "nicr_batch_create: tile upload failed");
    }
    if (rc == NICR_OK) rc = balloc(b, (void**)&b->partials, (size_t)max_tiles * NPART * 8);
    if (rc == NICR_OK) rc = balloc(b, (void**)&b->ctl, (size_t)n * sizeof(DevCtl));
    if (rc == NICR_OK) rc = balloc(b, (void**)&b->d_inst, (size_t)n * 4 * sizeof(BInst));
    if (rc == NICR_OK) {
        // one instance table per tile height (tile ranges differ)
        std::vector<BInst> all;
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < n; ++i) {
                BInst I = b->inst[i];
                I.tile0 = ranges[k][i].first;
                I.ntile = ranges[k][i].second;
                all.push_back(I);
            }
        if (hipMemcpy(b->d_inst, all.data(), all.size() * sizeof(BInst), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(NICR_E_DEVICE, "nicr_batch_create: instance upload failed");
    }
    if (rc != NICR_OK) {
        std::string m = g_error;
        cleanup();
        return fail(NICR_E_DEVICE, m);
    }
    hipMemsetAsync(b->slab, 0, 7 * fb, b->stream);
    std::vector<DevCtl> h(n);
    for (auto& d : h) {
        d = DevCtl{};
        d.pde_key = NO_KEY;
        d.eta_key = NO_KEY;
    }
    hipMemcpyAsync(b->ctl, h.data(), n * sizeof(DevCtl), hipMemcpyHostToDevice, b->stream);
    if (hipStreamSynchronize(b->stream) != hipSuccess) {
        cleanup();
        return fail(NICR_E_DEVICE, "nicr_batch_create: initialisation failed");
    }
    *out = b;
    return NICR_OK;
}

int nicr_batch_destroy(void* h) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b) return NICR_OK;
    hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (void* q : b->allocs) hipFree(q);
    hipStreamDestroy(b->stream);
    delete b;
    return NICR_OK;
}

int nicr_batch_info(void* h, long long* cells, size_t* device_bytes) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b) return fail(NICR_E_INVALID, "null batch");
    if (cells) *cells = b->cells;
    if (device_bytes) *device_bytes = b->bytes;
    return NICR_OK;
}

int nicr_batch_upload_field(void* h, int inst, int field, const double* host) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || !host) return fail(NICR_E_INVALID, "null argument");
    double* d = nullptr;
    int rc = batch_field_ptr(b, inst, field, &d);
    if (rc) return rc;
    HIPCHK(hipMemcpy(d, host, (size_t)b->inst[inst].nx * b->inst[inst].ny * 8, hipMemcpyHostToDevice));
    return NICR_OK;
}

int nicr_batch_download_field(void* h, int inst, int field, double* host) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || !host) return fail(NICR_E_INVALID, "null argument");
    double* d = nullptr;
    int rc = batch_field_ptr(b, inst, field, &d);
    if (rc) return rc;
    HIPCHK(hipMemcpy(host, d, (size_t)b->inst[inst].nx * b->inst[inst].ny * 8, hipMemcpyDeviceToHost));
    return NICR_OK;
}

int nicr_batch_get_budget(void* h, int inst, nicr_budget* out) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || !out || inst < 0 || inst >= b->n) return fail(NICR_E_INVALID, "nicr_batch_get_budget: bad argument");
    DevCtl d;
    HIPCHK(hipMemcpy(&d, b->ctl + inst, sizeof(d), hipMemcpyDeviceToHost));
    double* dst[8] = {out->nucO, out->nucCr, out->ckO, out->ckCr, out->clipO, out->clipCr, out->bcO, out->bcCr};
    for (int s = 0; s < 8; ++s) {
        dst[s][0] = d.budget[2 * s];
        dst[s][1] = d.budget[2 * s + 1];
    }
    out->nuclei = d.nuclei;
    out->clipped = d.clipped;
    return NICR_OK;
}

int nicr_batch_set_budget(void* h, int inst, const nicr_budget* in) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || !in || inst < 0 || inst >= b->n) return fail(NICR_E_INVALID, "nicr_batch_set_budget: bad argument");
    DevCtl d;
    HIPCHK(hipMemcpy(&d, b->ctl + inst, sizeof(d), hipMemcpyDeviceToHost));
    const double* src[8] = {in->nucO, in->nucCr, in->ckO, in->ckCr, in->clipO, in->clipCr, in->bcO, in->bcCr};
    for (int s = 0; s < 8; ++s) {
        d.budget[2 * s] = src[s][0];
        d.budget[2 * s + 1] = src[s][1];
    }
    d.nuclei = in->nuclei;
    d.clipped = in->clipped;
    HIPCHK(hipMemcpy(b->ctl + inst, &d, sizeof(d), hipMemcpyHostToDevice));
    return NICR_OK;
}

int nicr_batch_set_tiling(void* h, int tx, int grid, int use_graph) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b) return fail(NICR_E_INVALID, "null batch");
    int k = -1;
    for (int i = 0; i < 4; ++i)
        if (BATCH_TX[i] == tx) k = i;
    if (k < 0) return fail(NICR_E_INVALID, "nicr_batch_set_tiling: tile height must be 1, 2, 4 or 8");
    if (grid < 0) return fail(NICR_E_INVALID, "nicr_batch_set_tiling: grid must be >= 0");
    b->tx_index = k;
    b->grid = grid;
    b->use_graph = use_graph != 0;
    return NICR_OK;
}

// Advance every instance by up to nsteps steps; an instance stops at its first
// rejected step (its state and budgets stay at the last accepted step, as for
// nicr_step). reports: n entries or NULL.
int nicr_batch_step(void* h, int nsteps, nicr_step_report* reports) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || nsteps < 0) return fail(NICR_E_INVALID, "nicr_batch_step: bad argument");
    hipLaunchKernelGGL(k_batch_reset, dim3((b->n + 255) / 256), dim3(256), 0, b->stream, b->ctl, b->n);
    if (b->use_graph && nsteps > 1) {
        const int chunk = std::min(nsteps, 64);
        hipGraphExec_t ge;
        int rc = batch_graph(b, chunk, &ge);
        if (rc) return rc;
        int done = 0;
        for (; done + chunk <= nsteps; done += chunk) HIPCHK(hipGraphLaunch(ge, b->stream));
        for (; done < nsteps; ++done) batch_step_launch(b, 0);
    } else {
        for (int s = 0; s < nsteps; ++s) batch_step_launch(b, 0);
    }
    HIPCHK(hipGetLastError());
    std::vector<DevCtl> hc(b->n);
    HIPCHK(hipMemcpyAsync(hc.data(), b->ctl, b->n * sizeof(DevCtl), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->n; ++i) {
        nicr_step_report r{};
        r.steps_done = hc[i].steps_done;
        r.status = NICR_STEP_OK;
        if (hc[i].failed) {
            const BInst& I = b->inst[i];
            const int nxt = hc[i].cur ^ 1;
            const long long q = (long long)hc[i].pde_key;
            double v[3];
            HIPCHK(hipMemcpy(&v[0], b->o[nxt] + I.off + q, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&v[1], b->cr[nxt] + I.off + q, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&v[2], b->p[nxt] + I.off + q, 8, hipMemcpyDeviceToHost));
            r.o = v[0];
            r.cr = v[1];
            r.phi = v[2];
            r.status = (std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])) ? NICR_STEP_OUT_OF_BOUNDS
                                                                                            : NICR_STEP_NONFINITE;
            r.i = (int)(q / I.ny) + 1;
            r.j = (int)(q % I.ny) + 1;
            // mpi2d keeps the concentration-BC reset of a rejected step (2D:552)
            if (b->P.enforce_layer) {
                hipLaunchKernelGGL(k_batch_apply_layer, dim3((I.ny + 255) / 256), dim3(256), 0, b->stream, bbufs(b), i);
                HIPCHK(hipStreamSynchronize(b->stream));
            }
        }
        if (reports) reports[i] = r;
    }
    return NICR_OK;
}

// Time every (tile height, grid) candidate on `steps` dry ensemble steps (state
// unchanged) and keep the fastest. Returns the choice and its ms per step.
int nicr_batch_autotune(void* h, int steps, int* tx_out, int* grid_out, double* ms_out) {
    BatchCtx* b = static_cast<BatchCtx*>(h);
    if (!b || steps < 1) return fail(NICR_E_INVALID, "nicr_batch_autotune: bad argument");
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    int best_k = b->tx_index, best_grid = b->grid;
    const int keep_k = b->tx_index, keep_grid = b->grid;
    for (int k = 0; k < 4; ++k) {
        const int resident = batch_resident_for(b, k);
        const int grids[3] = {resident / 2, resident, 2 * resident};
        for (int gi = 0; gi < 3; ++gi) {
            b->tx_index = k;
            b->grid = std::max(1, grids[gi]);
            hipLaunchKernelGGL(k_batch_reset, dim3((b->n + 255) / 256), dim3(256), 0, b->stream, b->ctl, b->n);
            batch_step_launch(b, 1);   // warm-up
            HIPCHK(hipEventRecord(e0, b->stream));
            for (int s = 0; s < steps; ++s) batch_step_launch(b, 1);
            HIPCHK(hipEventRecord(e1, b->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms / steps < best) {
                best = ms / steps;
                best_k = k;
                best_grid = b->grid;
            }
        }
    }
    (void)keep_k;
    (void)keep_grid;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->tx_index = best_k;
    b->grid = best_grid;
    if (tx_out) *tx_out = BATCH_TX[best_k];
    if (grid_out) *grid_out = best_grid;
    if (ms_out) *ms_out = best;
    return NICR_OK;
}

}  // extern "C"
