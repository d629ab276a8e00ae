// nicr_api.hip -- C ABI of libnicr.so (declared in include/nicr.h).
//
// Owns the device state of one simulation (the storage of class Model,
// mpi2d/NiCr2D_grains.cpp:194-207, mpi3d-slab/NiCr3D_grains.cpp:336-355) and
// queues the step kernels of nicr_step.hip on the context's own HIP stream.
#include "../../include/nicr.h"
#include "nicr_kernels.hpp"
#include "nicr_step.hip"
#include "nicr_nucleate.hip"
#include "host_threads.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace nicr;

namespace {

thread_local std::string g_error;
constexpr double PI = 3.14159265358979323846;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(NICR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

struct TimerSlot {
    const char* name;
    double total_ms = 0.;
    long long count = 0;
};

// timer slots: one record per kernel launch
enum { T_MU, T_TRIAL, T_PDE_MARCH, T_ETA, T_ETA_MARCH, T_ETA_ROWS, T_COMMIT, T_NSLOT };

struct Ctx {
    nicr_config cfg{};
    Par P{};
    int dim = 3, prec = PREC_FP64;
    long long ncell = 0;
    size_t te = 8, tm = 8;  // element sizes of eta/mask and mu scratch
    char* slab = nullptr;   // state arrays (see nicr_create)
    size_t slab_bytes = 0;
    hipStream_t stream = nullptr;
    double* o[2] = {nullptr, nullptr};
    double* cr[2] = {nullptr, nullptr};
    double* p[2] = {nullptr, nullptr};
    void* mu_o = nullptr;
    void* mu_cr = nullptr;
    void* eta[2] = {nullptr, nullptr};
    void* s2[2] = {nullptr, nullptr};
    void* pairs[2] = {nullptr, nullptr};
    void* gb[2] = {nullptr, nullptr};
    float* ecoef = nullptr;     // eta-force factor per cell (mixed, marching kernels; see Bufs)
    double* partials = nullptr;
    DevCtl* ctl = nullptr;
    double* stat_part = nullptr;
    double* out4[4] = {nullptr, nullptr, nullptr, nullptr};  // refresh outputs muO, muCr, mo, mc
    double* dev_obuf = nullptr;  // device copy of {o0, o1, cr0, cr1, p0, p1}
    float* stage = nullptr;      // pinned host copy of one binary32 field (first conversion on)
    double* grain_term = nullptr;  // nicr_grain_terms scratch (allocated on first use)
    void* grain_part = nullptr;
    void* grain_counts = nullptr;
    int* nuc_tiles = nullptr;    // nucleation scratch (allocated on first use)
    int* nuc_planes = nullptr;
    int* nuc_list = nullptr;
    int* nuc_misc = nullptr;     // [0] candidate count, [1] changed/cells out
    int* nuc_centres = nullptr;
    int g_trial = 0, g_eta = 0, g_stat = 0;
    int partial_slots = 0;         // budget partial slots allocated (one per block of the trial kernel)
    int pde_ty = 0;                // Y tile height of the marching PDE kernel (NICR_PDE_TY: 2, 4, 6, 8;
                                   // 0: 8 in binary64, 4 in mixed precision, the faster of each at 256^3)
    int pde_nseg = 0;              // X segments of the marching PDE kernel (NICR_PDE_NSEG; 0: auto)
    int cur = 0;                 // host mirror of ctl->cur
    int device = 0;              // the HIP device current at nicr_create
    bool mask_dirty = true;
    bool pairs_stale = false;      // the last eta pass of a step leaves P unwritten
    bool refresh_valid = false;
    bool march_disabled = false;   // NICR_KERNEL_VARIANT=basic selects the unblocked kernels
    int eta_ty = 8;                // Y tile height of the marching eta kernel (NICR_ETA_TY: 2, 4, 8)
    int eta_nseg = 0;              // X segments per column (NICR_ETA_NSEG; 0: one wave of resident blocks)
    int eta_ng = 16;               // max eta fields per marching pass (NICR_ETA_NG: 8, 16, 32)
    bool eta_rows_disabled = false;  // NICR_ETA_ROWS=0 keeps the 64-wide tiles where whole rows would fit
    bool eta_march64_fp64 = true;    // NICR_ETA_MARCH64=0: binary64 with more than 8 fields unblocked off row meshes
    // 2D binary64 step through the ensemble kernels (batch2d_step, set up on
    // the first step); NICR_BATCH2D=0 keeps k_mu + k_trial
    bool b1_disabled = false;
    void* b1_inst = nullptr;      // one-instance table (BInst)
    void* b1_tiles = nullptr;     // its tile list (BTile)
    double* b1_partials = nullptr;
    int b1_ntile = 0, b1_tx = 0;
    std::string launch_error;        // set-up failure of a marching kernel (reported by nicr_step)
    size_t bytes = 0;
    std::vector<void*> allocs;
    bool timing = false;
    TimerSlot timers[T_NSLOT] = {{"k_mu"},        {"k_trial"},    {"k_pde_march"}, {"k_eta"},
                                 {"k_eta_march"}, {"k_eta_rows"}, {"k_commit"}};
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, int>> pending;  // (slot, event index of start)
    size_t evnext = 0;

    bool has_eta() const { return cfg.grains && cfg.grain_evolution && cfg.n_eta > 0; }
    bool eta_evolves() const { return has_eta() && (dim == 2 || cfg.eta_L > 0.); }
};

int dalloc(Ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(NICR_E_DEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    c->bytes += bytes;
    return NICR_OK;
}

template <class TE, class TM>
Bufs<TE, TM> bufs(Ctx* c) {
    Bufs<TE, TM> b;
    for (int i = 0; i < 2; ++i) {
        b.o[i] = c->o[i];
        b.cr[i] = c->cr[i];
        b.p[i] = c->p[i];
        b.eta[i] = (TE*)c->eta[i];
        b.s2[i] = (TE*)c->s2[i];
        b.pairs[i] = (TE*)c->pairs[i];
        b.gb[i] = (TE*)c->gb[i];
    }
    b.mu_o = (TM*)c->mu_o;
    b.mu_cr = (TM*)c->mu_cr;
    b.ecoef = c->ecoef;
    b.partials = c->partials;
    b.ctl = c->ctl;
    return b;
}

// ---- timing with HIP events on the context stream ----
void tmark_begin(Ctx* c, int slot) {
    if (!c->timing) return;
    if (c->evnext + 2 > c->evpool.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) == hipSuccess) c->evpool.push_back(e);
        }
    }
    hipEventRecord(c->evpool[c->evnext], c->stream);
    c->pending.push_back({slot, (int)c->evnext});
    c->evnext += 2;
}
void tmark_end(Ctx* c) {
    if (!c->timing) return;
    hipEventRecord(c->evpool[c->pending.back().second + 1], c->stream);
}
void tcollect(Ctx* c) {
    if (!c->timing) return;
    for (auto& pr : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->evpool[pr.second], c->evpool[pr.second + 1]) == hipSuccess) {
            c->timers[pr.first].total_ms += ms;
            c->timers[pr.first].count += 1;
        }
    }
    c->pending.clear();
    c->evnext = 0;
}

// ---- launches, dispatched over (DIM, PREC) ----
template <int DIM, int PREC, class TE, class TM>
void launch_mask(Ctx* c, int pairs_only) {
    hipLaunchKernelGGL((k_mask<DIM, PREC, TE, TM>), dim3(c->P.ntiles), dim3(TILE), 0, c->stream, c->P,
                       bufs<TE, TM>(c), c->eta_evolves() ? -1 : 0, pairs_only);
}
// X segments per column tile for a marching kernel: every block marches nx/nseg
// planes after a prologue of about `lead` planes, and the blocks run in
// ceil(tiles*nseg/resident) waves; pick the nseg with the least planes per CU slot.
static int pick_nseg(int tiles, int resident, int nx, double lead) {
    int best = 1;
    double best_cost = 1e300;
    const int hi = std::min(nx, std::max(1, 8 * resident / std::max(1, tiles) + 1));
    for (int nseg = 1; nseg <= hi; ++nseg) {
        const int xs = (nx + nseg - 1) / nseg;
        const int ns = (nx + xs - 1) / xs;
        const long long blocks = (long long)tiles * ns;
        const double waves = (double)((blocks + resident - 1) / resident);
        const double cost = waves * (xs + lead);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = ns;
        }
    }
    return best;
}

// Raise a marching kernel's dynamic LDS limit and count the blocks the chip
// holds at once. Every HIP result is checked: a launch whose LDS the runtime
// refused would otherwise be enqueued with a size it cannot run, so a failure
// is kept in c->launch_error and reported by the calling entry point.
static int prepare_kernel(Ctx* c, const void* fn, int threads, size_t lds) {
    const size_t lds_cap = 160 * 1024;
    if (lds > lds_cap || threads < 64 || threads > 1024) {
        c->launch_error = "kernel launch shape out of range (threads " + std::to_string(threads) + ", LDS " +
                          std::to_string(lds) + " bytes)";
        return 0;
    }
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 0, ncu = 0, dev = 0;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds);
    if (e != hipSuccess || per_cu < 1 || ncu < 1) {
        c->launch_error = std::string("kernel set-up failed: ") +
                          (e != hipSuccess ? hipGetErrorString(e) : "no block fits a compute unit");
        return 0;
    }
    return per_cu * ncu;
}

// Whole-row variant (k_eta_rows): blocks of 512 / nz rows of nz cells.
template <int PREC, int NG, int NZ, class TE, class TM>
void launch_eta_rows_ng(Ctx* c, const Bufs<TE, TM>& B, int g0) {
    const Par& P = c->P;
    constexpr int TYR = 512 / NZ;
    const int tiles = (P.ny + TYR - 1) / TYR;
    const size_t lds = (size_t)NG * (TYR + 2) * NZ * sizeof(TE);
    static int resident = 0;
    if (!resident) resident = prepare_kernel(c, (const void*)k_eta_rows<PREC, NG, NZ, TE, TM>, 512, lds);
    if (!resident) return;
    int nseg = pick_nseg(tiles, resident, P.nx, 1.);
    if (c->eta_nseg > 0) nseg = std::min(c->eta_nseg, P.nx);
    const int XS = (P.nx + nseg - 1) / nseg;
    nseg = (P.nx + XS - 1) / XS;
    tmark_begin(c, T_ETA_ROWS);
    hipLaunchKernelGGL((k_eta_rows<PREC, NG, NZ, TE, TM>), dim3(tiles * nseg), dim3(512), lds, c->stream, P, B, XS,
                       g0);
    tmark_end(c);
}
// Whether k_eta_rows applies: nz of 64, 128 or 256, the halo rows x fields split
// evenly over the 512 / nz thread rows, and the tile fits LDS.
bool eta_rows_apply(const Ctx* c, int ng, size_t te) {
    const int nz = c->P.nz;
    if (c->eta_rows_disabled || (nz != 64 && nz != 128 && nz != 256)) return false;
    const int ty = 512 / nz;
    return (2 * ng) % ty == 0 && (size_t)ng * (ty + 2) * nz * te <= 160 * 1024;
}

// X-marching eta kernel for 3D meshes: one pass over fields g0 .. g0+NG-1
// (NG <= 32; more fields take several passes).
template <int PREC, int NG, int TY, class TE, class TM>
void launch_eta_march_ng(Ctx* c, const Bufs<TE, TM>& B, int g0) {
    const Par& P = c->P;
    if (eta_rows_apply(c, NG, sizeof(TE))) {
        // the thread rows of a block, 512 / nz, must divide 2 * NG
        if constexpr ((2 * NG) % 2 == 0) {
            if (P.nz == 256) return launch_eta_rows_ng<PREC, NG, 256>(c, B, g0);
        }
        if constexpr ((2 * NG) % 4 == 0) {
            if (P.nz == 128) return launch_eta_rows_ng<PREC, NG, 128>(c, B, g0);
        }
        if constexpr ((2 * NG) % 8 == 0) {
            if (P.nz == 64) return launch_eta_rows_ng<PREC, NG, 64>(c, B, g0);
        }
    }

    const int nTz = (P.nz + 63) / 64, nTy = (P.ny + TY - 1) / TY;
    const int tiles = nTz * nTy;
    const size_t lds = (size_t)NG * (TY + 2) * 66 * sizeof(TE);
    static int resident = 0;   // blocks of this instantiation that fit on the chip at once
    if (!resident) resident = prepare_kernel(c, (const void*)k_eta_march<PREC, NG, TY, TE, TM>, 64 * TY, lds);
    if (!resident) return;
    // each segment loads one extra plane of every field at its start
    int nseg = pick_nseg(tiles, resident, P.nx, 1.);
    if (c->eta_nseg > 0) nseg = std::min(c->eta_nseg, P.nx);
    const int XS = (P.nx + nseg - 1) / nseg;
    nseg = (P.nx + XS - 1) / XS;
    tmark_begin(c, T_ETA_MARCH);
    hipLaunchKernelGGL((k_eta_march<PREC, NG, TY, TE, TM>), dim3(tiles * nseg), dim3(64 * TY), lds, c->stream, P, B,
                       XS, g0);
    tmark_end(c);
}
template <int PREC, int TY, class TE, class TM>
void launch_eta_march_ty(Ctx* c, const Bufs<TE, TM>& B, int cap) {
    // Passes of exactly NG fields each: the largest NG <= cap (passes keep the
    // per-XCD plane working set inside L2, and the registers of a binary64 pass
    // in bounds) that fits the fields left, then smaller powers of two for the
    // remainder.
    const int n = c->P.n_eta;
    for (int g0 = 0; g0 < n;) {
        const int left = n - g0;
        if (left >= 32 && cap >= 32) launch_eta_march_ng<PREC, 32, TY>(c, B, g0), g0 += 32;
        else if (left >= 16 && cap >= 16) launch_eta_march_ng<PREC, 16, TY>(c, B, g0), g0 += 16;
        else if (left >= 8) launch_eta_march_ng<PREC, 8, TY>(c, B, g0), g0 += 8;
        else if (left >= 4) launch_eta_march_ng<PREC, 4, TY>(c, B, g0), g0 += 4;
        else if (left >= 2) launch_eta_march_ng<PREC, 2, TY>(c, B, g0), g0 += 2;
        else launch_eta_march_ng<PREC, 1, TY>(c, B, g0), g0 += 1;
    }
}
template <int PREC, class TE, class TM>
bool launch_eta_march(Ctx* c, const Bufs<TE, TM>& B) {
    const int n = c->P.n_eta;
    // binary64 passes hold at most 8 fields (register planes of more limit
    // occupancy to one wave per SIMD)
    const int cap = PREC == PREC_FP64 ? std::min(c->eta_ng, 8) : c->eta_ng;
    // the kernel addresses the fields of one pass with 32-bit byte offsets
    int ng_max = 1;
    while (ng_max * 2 <= std::min(n, cap) && ng_max < 32) ng_max *= 2;
    if (c->march_disabled || c->dim != 3 || c->P.nz < 32 ||
        (unsigned long long)c->P.fstride * sizeof(TE) * (unsigned long long)ng_max >= (1ull << 32))
        return false;
    if constexpr (PREC == PREC_FP64) {
        // more than 8 fields in binary64: passes of 8, on whole-row or 64-wide
        // tiles (64-wide passes are 1.9x faster than the unblocked kernel at
        // 256x256x512 and 256x256x320 with 32 fields)
        if (n > 8 && !c->eta_march64_fp64 && !eta_rows_apply(c, 8, sizeof(TE))) return false;
        launch_eta_march_ty<PREC, 2>(c, B, cap);
    } else {
        if (c->eta_ty == 2) launch_eta_march_ty<PREC, 2>(c, B, cap);
        else if (c->eta_ty == 8) launch_eta_march_ty<PREC, 8>(c, B, cap);
        else launch_eta_march_ty<PREC, 4>(c, B, cap);
    }
    return true;
}

// Fused X-marching mu + trial kernel for 3D meshes. Returns the number of budget
// partial slots it wrote (one per block).
template <int PREC, int TY, class TE, class TM>
int launch_pde_march_ty(Ctx* c, const Bufs<TE, TM>& B) {
    const Par& P = c->P;
    const int nTz = (P.nz + 63) / 64, nTy = (P.ny + TY - 1) / TY;
    const int tiles = nTz * nTy;
    const size_t lds = PdeLds<PREC, TY>::bytes();
    static int resident = 0;
    if (!resident) resident = prepare_kernel(c, (const void*)k_pde_march<PREC, TY, TE, TM>, 64 * TY, lds);
    if (!resident) return -2;
    // the prologue of a segment costs about three planes (4 plane loads, one mu plane)
    int nseg = pick_nseg(tiles, resident, P.nx, 3.);
    if (c->pde_nseg > 0) nseg = std::min(c->pde_nseg, P.nx);
    while ((long long)tiles * nseg > c->partial_slots && nseg > 1) --nseg;
    const int XS = (P.nx + nseg - 1) / nseg;
    nseg = (P.nx + XS - 1) / XS;
    if ((long long)tiles * nseg > c->partial_slots) return -1;
    tmark_begin(c, T_PDE_MARCH);
    hipLaunchKernelGGL((k_pde_march<PREC, TY, TE, TM>), dim3(tiles * nseg), dim3(64 * TY), lds, c->stream, P, B, XS);
    tmark_end(c);
    return tiles * nseg;
}
bool pde_march_applies(const Ctx* c) { return !c->march_disabled && c->dim == 3 && c->P.nz >= 32; }
template <int PREC, class TE, class TM>
int launch_pde_march(Ctx* c, const Bufs<TE, TM>& B) {
    switch (c->pde_ty) {
        case 2: return launch_pde_march_ty<PREC, 2>(c, B);
        case 6: return launch_pde_march_ty<PREC, 6>(c, B);
        case 8: return launch_pde_march_ty<PREC, 8>(c, B);
        default: return launch_pde_march_ty<PREC, 4>(c, B);
    }
}

bool batch2d_step(Ctx* c);   // nicr_batch.hip
template <int DIM, int PREC, class TE, class TM>
void launch_step(Ctx* c) {
    if constexpr (DIM == 2 && PREC == PREC_FP64) {
        if (batch2d_step(c)) return;
    }
    auto B = bufs<TE, TM>(c);
    int nparts = -1;
    if constexpr (DIM == 3) {
        if (pde_march_applies(c)) nparts = launch_pde_march<PREC, TE, TM>(c, B);
    }
    if (!c->launch_error.empty()) return;   // reported by nicr_step; nothing of this step is queued
    if (nparts < 0) {
        B.ecoef = nullptr;   // written only by the marching PDE kernel
        tmark_begin(c, T_MU);
        hipLaunchKernelGGL((k_mu<DIM, PREC, TE, TM>), dim3(c->P.ntiles), dim3(TILE), 0, c->stream, c->P, B);
        tmark_end(c);
        tmark_begin(c, T_TRIAL);
        hipLaunchKernelGGL((k_trial<DIM, PREC, TE, TM>), dim3(c->g_trial), dim3(TILE), 0, c->stream, c->P, B);
        tmark_end(c);
        nparts = c->g_trial;
    }
    if (c->eta_evolves()) {
        bool done = false;
        if constexpr (DIM == 3) done = launch_eta_march<PREC, TE, TM>(c, B);
        if (!c->launch_error.empty()) return;
        if (!done) {
            tmark_begin(c, T_ETA);
            hipLaunchKernelGGL((k_eta<DIM, PREC, TE, TM>), dim3(c->g_eta), dim3(TILE), 0, c->stream, c->P, B);
            tmark_end(c);
        }
    }
    tmark_begin(c, T_COMMIT);
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(1024), 0, c->stream, bufs<double, double>(c), nparts);
    tmark_end(c);
}
template <int DIM, class TE, class TM>
void launch_refresh_out(Ctx* c) {
    hipLaunchKernelGGL((k_refresh_out<DIM, TE, TM>), dim3(c->P.ntiles), dim3(TILE), 0, c->stream, c->P,
                       bufs<TE, TM>(c), c->out4[0], c->out4[1], c->out4[2], c->out4[3]);
}

int ensure_nuc_scratch(Ctx* c) {
    if (c->nuc_list) return NICR_OK;
    int rc = dalloc(c, (void**)&c->nuc_tiles, (size_t)c->P.ntiles * sizeof(int));
    if (!rc) rc = dalloc(c, (void**)&c->nuc_planes, (size_t)c->P.nx * sizeof(int));
    if (!rc) rc = dalloc(c, (void**)&c->nuc_list, (size_t)c->ncell * sizeof(int));
    if (!rc) rc = dalloc(c, (void**)&c->nuc_misc, 4 * sizeof(int));
    if (!rc) rc = dalloc(c, (void**)&c->nuc_centres, 4096 * sizeof(int));
    return rc;
}

template <int DIM, class TE>
void launch_stamp(Ctx* c, NucPar N) {
    double* const* d = (double* const*)c->dev_obuf;
    TE* eta = c->has_eta() ? (TE*)c->eta[c->eta_evolves() ? c->cur : 0] : nullptr;
    hipLaunchKernelGGL((k_stamp<DIM, TE>), dim3(1), dim3(1024), 0, c->stream, c->P, N, d, d + 2, d + 4, eta,
                       c->ctl, c->nuc_misc + 1);
}
// Scratch of one Ksp scan: per-tile candidate counts (then offsets), per-plane
// oxide counts, [0] candidate total and [1] inserted cells, accepted centres,
// the ordered candidate list.
struct NucScratch {
    int* tiles;
    int* planes;
    int* misc;
    int* centres;
    int* list;
};

__global__ void k_fill_int(int* a, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <int DIM, class TE>
void launch_nucleate(Ctx* c, NucPar N, const NucScratch& S, bool fill_kernel) {
    double* const* d = (double* const*)c->dev_obuf;
    TE* eta = c->has_eta() ? (TE*)c->eta[c->eta_evolves() ? c->cur : 0] : nullptr;
    if (fill_kernel)
        hipLaunchKernelGGL(k_fill_int, dim3((c->P.nx + 255) / 256), dim3(256), 0, c->stream, S.planes, c->P.nx, 0);
    else
        hipMemsetAsync(S.planes, 0, (size_t)c->P.nx * sizeof(int), c->stream);
    hipLaunchKernelGGL((k_ksp_mark<DIM>), dim3(c->P.ntiles), dim3(TILE), 0, c->stream, c->P, N.r, d, d + 2, d + 4,
                       (const DevCtl*)c->ctl, S.tiles, S.planes);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, c->stream, S.tiles, c->P.ntiles, S.misc);
    hipLaunchKernelGGL((k_ksp_compact<DIM>), dim3(c->P.ntiles), dim3(TILE), 0, c->stream, c->P, N.r, d, d + 2,
                       d + 4, (const DevCtl*)c->ctl, (const int*)S.tiles, S.list);
    hipLaunchKernelGGL((k_ksp_resolve<DIM, TE>), dim3(1), dim3(1024), 0, c->stream, c->P, N, d, d + 2, d + 4, eta,
                       c->ctl, (const int*)S.list, (const int*)S.misc, (const int*)S.planes, S.centres, S.misc + 1);
}

void do_mask(Ctx* c, int pairs_only = 0) {
    if (!c->has_eta()) return;
    if (c->prec == PREC_FP64) {
        if (c->dim == 2) launch_mask<2, PREC_FP64, double, double>(c, pairs_only);
        else launch_mask<3, PREC_FP64, double, double>(c, pairs_only);
    } else {
        if (c->dim == 2) launch_mask<2, PREC_MIXED, float, float>(c, pairs_only);
        else launch_mask<3, PREC_MIXED, float, float>(c, pairs_only);
    }
    c->pairs_stale = false;
}
void do_step(Ctx* c) {
    // --frozen-eta-cache off (SL:346-349,599-603): a frozen eta mask is rebuilt
    // by every step's refresh instead of only after an edit; same values
    if (c->has_eta() && !c->eta_evolves() && !c->cfg.frozen_eta_cache) do_mask(c);
    if (c->prec == PREC_FP64) {
        if (c->dim == 2) launch_step<2, PREC_FP64, double, double>(c);
        else launch_step<3, PREC_FP64, double, double>(c);
    } else {
        if (c->dim == 2) launch_step<2, PREC_MIXED, float, float>(c);
        else launch_step<3, PREC_MIXED, float, float>(c);
    }
}
void do_refresh_out(Ctx* c) {
    if (c->prec == PREC_FP64) {
        if (c->dim == 2) launch_refresh_out<2, double, double>(c);
        else launch_refresh_out<3, double, double>(c);
    } else {
        if (c->dim == 2) launch_refresh_out<2, float, float>(c);
        else launch_refresh_out<3, float, float>(c);
    }
}

// One 16-byte element per thread, one short block per 4 KiB: the fastest
// streaming copy measured on MI355X (6.2-6.4 TB/s of read + write for 1 GiB;
// grid-stride loops over a few thousand resident blocks reach only 4.2-4.6).
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    using V4 = float __attribute__((ext_vector_type(4)));
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(reinterpret_cast<const V4*>(a)[i], reinterpret_cast<V4*>(b) + i);
}

__global__ void k_reset_ctl(DevCtl* ctl) {
    ctl->failed = 0;
    ctl->steps_done = 0;
    ctl->fail_kind = 0;
    ctl->pde_key = NO_KEY;
    ctl->eta_key = NO_KEY;
}

int sync_ctl(Ctx* c, DevCtl* h) {
    HIPCHK(hipMemcpyAsync(h, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->cur = h->cur;
    return NICR_OK;
}

void* eta_ptr(Ctx* c, int which, int g) {
    return (char*)c->eta[which] + (size_t)g * c->P.fstride * c->te;
}
int mask_index(Ctx* c) { return c->eta_evolves() ? c->cur : 0; }

}  // namespace

extern "C" {

const char* nicr_last_error(void) { return g_error.c_str(); }

int nicr_config_default(nicr_config* cfg, int dim) {
    if (!cfg || (dim != 2 && dim != 3)) return fail(NICR_E_INVALID, "nicr_config_default: dim must be 2 or 3");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->abi_version = NICR_ABI_VERSION;
    cfg->dim = dim;
    cfg->nx = 151;
    cfg->ny = 101;
    cfg->nz = dim == 3 ? 101 : 1;
    cfg->steps = 0;
    cfg->output_every = 5000;
    cfg->left_edge_sites = 0;
    cfg->random_seed = 20260909ull;
    cfg->dx = cfg->dy = cfg->dz = 1.e-8;
    cfg->dt = 2e-10;
    cfg->grain_diameter_grid = 50.;
    cfg->gb_fwhm_grid = 4.;
    cfg->gb_factor_o = 28.2;
    cfg->gb_factor_cr = 563.;
    cfg->radius_grid = 5.5;
    cfg->exclusion_gap_grid = 6.;
    cfg->nucleation_period = 1.e-7;
    cfg->mu_res = 364688.;
    cfg->surface_c_ref = 0.0023;
    cfg->bc = NICR_BC_CONCENTRATION;
    cfg->ck = NICR_CK_LEGACY;
    cfg->bounds = NICR_BOUNDS_STOP;
    cfg->nucleation = NICR_NUC_KSP;
    cfg->grains = 1;
    cfg->oxygen_diffusion = 1;
    cfg->grain_evolution = 0;
    cfg->frozen_eta_cache = 1;
    cfg->eta_W = 1.e4;
    cfg->eta_kappa = 0.;
    cfg->eta_L = 0.2;
    cfg->eta_oxide_penalty = -1.;
    cfg->max_eta_fields = 512;
    cfg->n_eta = 0;
    cfg->precision = NICR_PRECISION_FP64;
    cfg->layer_o = Phys::layerO;
    return nicr_config_derive(cfg);
}

// eta_K, eta_A, eta_transport_scale (2D:88-103, SL:88-103)
int nicr_config_derive(nicr_config* c) {
    if (!c) return fail(NICR_E_INVALID, "null config");
    const double ref_w = c->gb_fwhm_grid * c->dx * std::sqrt(PI) / (2. * std::sqrt(std::log(2.)));
    double K;
    if (c->eta_kappa > 0.) K = c->eta_kappa;
    else {
        const double ell = 3. * ref_w / 8.;
        K = 2. * c->eta_W * ell * ell;
    }
    const double ell = std::sqrt(K / (2. * c->eta_W));
    c->eta_K = K;
    c->eta_A = c->eta_oxide_penalty < 0. ? c->eta_W : c->eta_oxide_penalty;
    c->eta_scale = ref_w / (8. * ell / 3.);
    return NICR_OK;
}

int nicr_create(const nicr_config* cfg, void** out) {
    if (!cfg || !out) return fail(NICR_E_INVALID, "nicr_create: null argument");
    if (cfg->abi_version != NICR_ABI_VERSION) return fail(NICR_E_INVALID, "nicr_create: ABI version mismatch");
    if (cfg->dim != 2 && cfg->dim != 3) return fail(NICR_E_INVALID, "nicr_create: dim must be 2 or 3");
    if (cfg->nx < 1 || cfg->ny < 1 || (cfg->dim == 3 && cfg->nz < 1))
        return fail(NICR_E_INVALID, "nicr_create: mesh dimensions must be positive");
    if (cfg->precision != NICR_PRECISION_FP64 && cfg->precision != NICR_PRECISION_MIXED)
        return fail(NICR_E_INVALID, "nicr_create: unknown precision");
    if (cfg->n_eta < 0 || (cfg->grains && cfg->grain_evolution && cfg->n_eta == 0 && cfg->max_eta_fields < 0))
        return fail(NICR_E_INVALID, "nicr_create: bad n_eta");
    const long long nyz = (long long)cfg->ny * (cfg->dim == 3 ? cfg->nz : 1);
    if (nyz > (1ll << 30) || (long long)cfg->nx * ((nyz + TILE - 1) / TILE) > (1ll << 30))
        return fail(NICR_E_INVALID, "nicr_create: mesh too large for 32-bit tile indexing");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(NICR_E_DEVICE, "nicr_create: no HIP device is available");

    Ctx* c = new Ctx;
    c->cfg = *cfg;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(NICR_E_DEVICE, "nicr_create: hipGetDevice failed");
    }
    if (const char* kv = std::getenv("NICR_KERNEL_VARIANT")) c->march_disabled = std::strcmp(kv, "basic") == 0;
    if (const char* ty = std::getenv("NICR_ETA_TY")) c->eta_ty = std::atoi(ty);
    if (const char* ns = std::getenv("NICR_ETA_NSEG")) c->eta_nseg = std::atoi(ns);
    if (const char* ng = std::getenv("NICR_ETA_NG")) c->eta_ng = std::atoi(ng);
    if (const char* t = std::getenv("NICR_PDE_TY")) c->pde_ty = std::atoi(t);
    if (const char* t = std::getenv("NICR_ETA_ROWS")) c->eta_rows_disabled = std::atoi(t) == 0;
    if (const char* t = std::getenv("NICR_ETA_MARCH64")) c->eta_march64_fp64 = std::atoi(t) != 0;
    if (const char* t = std::getenv("NICR_BATCH2D")) c->b1_disabled = std::atoi(t) == 0;
    if (const char* t = std::getenv("NICR_PDE_NSEG")) c->pde_nseg = std::atoi(t);
    if (c->cfg.dim == 2) c->cfg.nz = 1;
    if (!(c->cfg.grains && c->cfg.grain_evolution)) c->cfg.n_eta = 0;
    c->dim = c->cfg.dim;
    c->prec = c->cfg.precision;
    c->te = c->prec == PREC_FP64 ? 8 : 4;
    c->tm = c->te;
    if (c->pde_ty == 0) c->pde_ty = c->prec == PREC_FP64 ? 8 : 4;
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz;
    c->ncell = (long long)nx * ny * nz;
    Par& P = c->P;
    P.nx = nx;
    P.ny = ny;
    P.nz = nz;
    P.nyz = ny * nz;
    P.nchunk = (P.nyz + TILE - 1) / TILE;
    P.ntiles = nx * P.nchunk;
    P.ncell = c->ncell;
    P.n_eta = c->cfg.n_eta;
    P.bc = c->cfg.bc;
    P.ck_legacy = c->cfg.ck == NICR_CK_LEGACY;
    P.bounds_legacy = c->cfg.bounds == NICR_BOUNDS_LEGACY;
    P.grains = c->cfg.grains;
    P.oxy = c->cfg.oxygen_diffusion;
    P.has_eta = c->has_eta();
    P.eta_evolves = c->eta_evolves();
    P.enforce_layer = c->cfg.bc == NICR_BC_CONCENTRATION && c->cfg.oxygen_diffusion;
    P.dx = c->cfg.dx;
    P.dy = c->cfg.dy;
    P.dz = c->cfg.dz;
    P.dt = c->cfg.dt;
    P.ck_o_dt = Phys::ckO / P.dt;
    P.ck_cr_dt = Phys::ckCr / P.dt;
    P.gbfo = c->cfg.gb_factor_o;
    P.gbfc = c->cfg.gb_factor_cr;
    P.mu_res = c->cfg.mu_res;
    P.c_ref = c->cfg.surface_c_ref;
    P.layer_o = c->cfg.layer_o;
    P.etaW = c->cfg.eta_W;
    P.etaK = c->cfg.eta_K;
    P.etaL = c->cfg.eta_L;
    P.etaA = c->cfg.eta_A;
    P.eta_scale = c->cfg.eta_scale;
    P.fdx2i = (float)(1. / (P.dx * P.dx));
    P.fdy2i = (float)(1. / (P.dy * P.dy));
    P.fdz2i = (float)(1. / (P.dz * P.dz));
    P.fdxi = (float)(1. / P.dx);
    P.fdyi = (float)(1. / P.dy);
    P.fdzi = (float)(1. / P.dz);

    int rc = NICR_OK;
    auto cleanup = [&]() {
        for (void* p : c->allocs) hipFree(p);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        cleanup();
        return fail(NICR_E_DEVICE, "nicr_create: hipStreamCreate failed");
    }
    const size_t n8 = (size_t)c->ncell * 8, nte = (size_t)c->ncell * c->te;
    // State arrays share one slab. Each starts 256-byte aligned and the distance
    // between consecutive arrays, and between consecutive eta fields, is an odd
    // multiple of 256 bytes: arrays read at the same cell index then begin on
    // different HBM channels. With the power-of-two sizes of cubic meshes they
    // would otherwise all map the same cell to the same channel, which costs
    // about a sixth of the streaming bandwidth.
    auto stagger = [](size_t bytes) {
        size_t b = (bytes + 255) / 256;
        if (b % 2 == 0) ++b;
        return b * 256;
    };
    P.fstride = (long long)(stagger(nte) / c->te);
    const int nbuf = c->eta_evolves() ? 2 : 1;
    size_t slab_bytes = 6 * stagger(n8) + 2 * stagger((size_t)c->ncell * c->tm);
    for (int i = 0; i < nbuf; ++i)
        slab_bytes += (c->has_eta() ? stagger((size_t)P.fstride * c->te * c->cfg.n_eta) + 2 * stagger(nte) : 0)
                    + stagger(nte);
    char* slab = nullptr;
    rc = dalloc(c, (void**)&slab, slab_bytes);
    if (rc == NICR_OK) {
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            void* q = slab + off;
            off += stagger(bytes);
            return q;
        };
        for (int i = 0; i < 2; ++i) {
            c->o[i] = (double*)carve(n8);
            c->cr[i] = (double*)carve(n8);
            c->p[i] = (double*)carve(n8);
        }
        c->mu_o = carve((size_t)c->ncell * c->tm);
        c->mu_cr = carve((size_t)c->ncell * c->tm);
        for (int i = 0; i < nbuf; ++i) {
            if (c->has_eta()) {
                c->eta[i] = carve((size_t)P.fstride * c->te * c->cfg.n_eta);
                c->s2[i] = carve(nte);
                c->pairs[i] = carve(nte);
            }
            c->gb[i] = carve(nte);
        }
        c->slab = slab;
        c->slab_bytes = slab_bytes;
    }
    if (nbuf == 1) {
        c->eta[1] = c->eta[0];
        c->s2[1] = c->s2[0];
        c->pairs[1] = c->pairs[0];
        c->gb[1] = c->gb[0];
    }
    // persistent grids: a multiple of 8 blocks (one run per XCD)
    c->g_trial = std::min(P.ntiles, 2048);
    if (c->g_trial > 8) c->g_trial &= ~7;
    c->g_eta = std::min(P.ntiles, 2048);
    if (c->g_eta > 8) c->g_eta &= ~7;
    c->g_stat = 512;
    c->partial_slots = std::max(c->g_trial, 8192);
    if (rc == NICR_OK) rc = dalloc(c, (void**)&c->partials, (size_t)c->partial_slots * NPART * 8);
    const char* ec = std::getenv("NICR_ETA_COEF");
    if (rc == NICR_OK && c->prec == PREC_MIXED && c->eta_evolves() && pde_march_applies(c) && !(ec && ec[0] == '0'))
        rc = dalloc(c, (void**)&c->ecoef, (size_t)c->ncell * sizeof(float));
    if (rc == NICR_OK) rc = dalloc(c, (void**)&c->ctl, sizeof(DevCtl));
    if (rc == NICR_OK) rc = dalloc(c, (void**)&c->stat_part, (size_t)c->g_stat * 4 * 8);
    if (rc == NICR_OK) rc = dalloc(c, (void**)&c->dev_obuf, 6 * sizeof(double*));
    if (rc != NICR_OK) {
        std::string m = g_error;
        cleanup();
        return fail(NICR_E_DEVICE, m);
    }
    // zero state, control block
    hipMemsetAsync(c->slab, 0, c->slab_bytes, c->stream);
    DevCtl h{};
    h.cur = 0;
    h.pde_key = NO_KEY;
    h.eta_key = NO_KEY;
    hipMemcpyAsync(c->ctl, &h, sizeof(h), hipMemcpyHostToDevice, c->stream);
    double* ob[6] = {c->o[0], c->o[1], c->cr[0], c->cr[1], c->p[0], c->p[1]};
    hipMemcpyAsync(c->dev_obuf, ob, sizeof(ob), hipMemcpyHostToDevice, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        cleanup();
        return fail(NICR_E_DEVICE, "nicr_create: initialisation failed");
    }
    *out = c;
    return NICR_OK;
}

int nicr_destroy(void* h) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return NICR_OK;
    hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) hipFree(p);
    for (hipEvent_t e : c->evpool) hipEventDestroy(e);
    if (c->stage) hipHostFree(c->stage);
    hipStreamDestroy(c->stream);
    delete c;
    return NICR_OK;
}

int nicr_device_bytes(void* h, size_t* bytes) {
    if (!h || !bytes) return fail(NICR_E_INVALID, "null argument");
    *bytes = static_cast<Ctx*>(h)->bytes;
    return NICR_OK;
}

// Host layout == device layout (interior cells, reference storage order).
// Binary32 fields pass through one pinned host buffer of ncell floats, kept
// until nicr_destroy; every transfer waits for its copy, so calls on one
// context never share it in flight.
static int pinned_stage(Ctx* c) {
    if (!c->stage) HIPCHK(hipHostMalloc((void**)&c->stage, (size_t)c->ncell * 4, hipHostMallocDefault));
    return NICR_OK;
}

int nicr_upload_field(void* h, int field, const double* src) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !src) return fail(NICR_E_INVALID, "nicr_upload_field: null argument");
    const size_t n8 = (size_t)c->ncell * 8;
    void* dst = nullptr;
    bool convert = false;
    switch (field) {
    case NICR_FIELD_O: dst = c->o[c->cur]; break;
    case NICR_FIELD_CR: dst = c->cr[c->cur]; break;
    case NICR_FIELD_PHI: dst = c->p[c->cur]; break;
    case NICR_FIELD_GB:
        dst = c->gb[mask_index(c)];
        convert = c->te == 4;
        if (c->has_eta()) c->mask_dirty = true;
        break;
    default:
        if (field >= NICR_FIELD_ETA0 && field < NICR_FIELD_ETA0 + c->cfg.n_eta) {
            dst = eta_ptr(c, mask_index(c), field - NICR_FIELD_ETA0);
            convert = c->te == 4;
            c->mask_dirty = true;
        } else {
            return fail(NICR_E_INVALID, "nicr_upload_field: field cannot be uploaded");
        }
    }
    c->refresh_valid = false;
    if (convert) {
        // rounded to binary32 on host threads into pinned memory, one DMA copy
        int rc = pinned_stage(c);
        if (rc) return rc;
        float* st = c->stage;
        nicr_threads::parallel_chunks((size_t)c->ncell, size_t(1) << 18, [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) st[i] = (float)src[i];
        });
        HIPCHK(hipMemcpyAsync(dst, st, (size_t)c->ncell * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(dst, src, n8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return NICR_OK;
}

int nicr_download_field(void* h, int field, double* dstp) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !dstp) return fail(NICR_E_INVALID, "nicr_download_field: null argument");
    const void* src = nullptr;
    bool convert = false;
    switch (field) {
    case NICR_FIELD_O: src = c->o[c->cur]; break;
    case NICR_FIELD_CR: src = c->cr[c->cur]; break;
    case NICR_FIELD_PHI: src = c->p[c->cur]; break;
    case NICR_FIELD_GB: src = c->gb[mask_index(c)]; convert = c->te == 4; break;
    case NICR_FIELD_ETA_S2:
    case NICR_FIELD_ETA_PAIRS:
        if (!c->has_eta()) return fail(NICR_E_INVALID, "nicr_download_field: no eta fields");
        src = field == NICR_FIELD_ETA_S2 ? c->s2[mask_index(c)] : c->pairs[mask_index(c)];
        convert = c->te == 4;
        break;
    case NICR_FIELD_MU_O:
    case NICR_FIELD_MU_CR:
    case NICR_FIELD_M_O:
    case NICR_FIELD_M_CR:
        if (!c->refresh_valid) return fail(NICR_E_STATE, "nicr_download_field: call nicr_refresh first");
        src = c->out4[field - NICR_FIELD_MU_O];
        break;
    default:
        if (field >= NICR_FIELD_ETA0 && field < NICR_FIELD_ETA0 + c->cfg.n_eta) {
            src = eta_ptr(c, mask_index(c), field - NICR_FIELD_ETA0);
            convert = c->te == 4;
        } else {
            return fail(NICR_E_INVALID, "nicr_download_field: unknown field");
        }
    }
    if ((field == NICR_FIELD_GB || field == NICR_FIELD_ETA_S2 || field == NICR_FIELD_ETA_PAIRS) && c->mask_dirty) {
        do_mask(c);
        c->mask_dirty = false;
    }
    if (field == NICR_FIELD_ETA_PAIRS && c->pairs_stale) do_mask(c, 1);
    if (convert) {
        int rc = pinned_stage(c);
        if (rc) return rc;
        const float* st = c->stage;
        HIPCHK(hipMemcpyAsync(c->stage, src, (size_t)c->ncell * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        nicr_threads::parallel_chunks((size_t)c->ncell, size_t(1) << 18, [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) dstp[i] = (double)st[i];
        });
    } else {
        HIPCHK(hipMemcpyAsync(dstp, src, (size_t)c->ncell * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return NICR_OK;
}

int nicr_set_budget(void* h, const nicr_budget* b) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !b) return fail(NICR_E_INVALID, "null argument");
    DevCtl hc;
    int rc = sync_ctl(c, &hc);
    if (rc) return rc;
    const double* src[8] = {b->nucO, b->nucCr, b->ckO, b->ckCr, b->clipO, b->clipCr, b->bcO, b->bcCr};
    for (int s = 0; s < 8; ++s) {
        hc.budget[2 * s] = src[s][0];
        hc.budget[2 * s + 1] = src[s][1];
    }
    hc.nuclei = b->nuclei;
    hc.clipped = b->clipped;
    HIPCHK(hipMemcpyAsync(c->ctl, &hc, sizeof(hc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NICR_OK;
}

int nicr_get_budget(void* h, nicr_budget* b) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !b) return fail(NICR_E_INVALID, "null argument");
    DevCtl hc;
    int rc = sync_ctl(c, &hc);
    if (rc) return rc;
    double* dst[8] = {b->nucO, b->nucCr, b->ckO, b->ckCr, b->clipO, b->clipCr, b->bcO, b->bcCr};
    for (int s = 0; s < 8; ++s) {
        dst[s][0] = hc.budget[2 * s];
        dst[s][1] = hc.budget[2 * s + 1];
    }
    b->nuclei = hc.nuclei;
    b->clipped = hc.clipped;
    return NICR_OK;
}

int nicr_refresh(void* h) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    if (!c->out4[0]) {
        for (int i = 0; i < 4; ++i) {
            int rc = dalloc(c, (void**)&c->out4[i], (size_t)c->ncell * 8);
            if (rc) return rc;
        }
    }
    if (c->mask_dirty) {
        do_mask(c);
        c->mask_dirty = false;
    }
    do_refresh_out(c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->refresh_valid = true;
    return NICR_OK;
}

int nicr_step(void* h, int nsteps, nicr_step_report* rep) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || nsteps < 0) return fail(NICR_E_INVALID, "nicr_step: bad argument");
    if (c->mask_dirty) {
        do_mask(c);
        c->mask_dirty = false;
    }
    hipLaunchKernelGGL(k_reset_ctl, dim3(1), dim3(1), 0, c->stream, c->ctl);
    for (int s = 0; s < nsteps && c->launch_error.empty(); ++s) do_step(c);
    if (nsteps > 0 && c->eta_evolves()) c->pairs_stale = true;
    HIPCHK(hipGetLastError());
    if (!c->launch_error.empty()) {
        hipStreamSynchronize(c->stream);
        return fail(NICR_E_DEVICE, "nicr_step: " + c->launch_error);
    }
    DevCtl hc;
    int rc = sync_ctl(c, &hc);
    if (rc) return rc;
    tcollect(c);
    c->refresh_valid = false;
    nicr_step_report r{};
    r.steps_done = hc.steps_done;
    r.status = NICR_STEP_OK;
    if (hc.failed) {
        const int nxt = hc.cur ^ 1;
        const long long nyz = (long long)c->P.nyz;
        if (hc.fail_kind == 1) {
            const long long q = (long long)hc.pde_key;
            double v[3];
            HIPCHK(hipMemcpy(&v[0], c->o[nxt] + q, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&v[1], c->cr[nxt] + q, 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&v[2], c->p[nxt] + q, 8, hipMemcpyDeviceToHost));
            r.o = v[0];
            r.cr = v[1];
            r.phi = v[2];
            r.status = (std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])) ? NICR_STEP_OUT_OF_BOUNDS
                                                                                            : NICR_STEP_NONFINITE;
            r.i = (int)(q / nyz) + 1;
            const long long yz = q % nyz;
            r.j = (int)(yz / c->P.nz) + 1;
            r.k = c->dim == 3 ? (int)(yz % c->P.nz) + 1 : 0;
        } else {
            // the key orders (field, cell) as the reference's loops do: g * ncell + cell;
            // the field itself sits at g * fstride (padded) in the trial buffer
            const long long key = (long long)hc.eta_key;
            const long long g = key / c->ncell, q = key % c->ncell;
            r.status = NICR_STEP_ETA_RANGE;
            r.grain = (int)g + 1;
            const char* src = (const char*)eta_ptr(c, nxt, (int)g) + (size_t)q * c->te;
            if (c->te == 8) {
                HIPCHK(hipMemcpy(&r.eta_value, src, 8, hipMemcpyDeviceToHost));
            } else {
                float f;
                HIPCHK(hipMemcpy(&f, src, 4, hipMemcpyDeviceToHost));
                r.eta_value = f;
            }
            r.i = (int)(q / nyz) + 1;
            const long long yz = q % nyz;
            r.j = (int)(yz / c->P.nz) + 1;
            r.k = c->dim == 3 ? (int)(yz % c->P.nz) + 1 : 0;
        }
        // mpi2d keeps the concentration-BC reset of a rejected step (2D:552).
        if (c->dim == 2 && c->P.enforce_layer) {
            hipLaunchKernelGGL(k_apply_layer, dim3((c->P.nyz + 255) / 256), dim3(256), 0, c->stream, c->P,
                               (double* const*)c->dev_obuf, (const DevCtl*)c->ctl);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    if (rep) *rep = r;
    return NICR_OK;
}

int nicr_stamp(void* h, int x, int y, int z, int* changed) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    if (x < 1 || x > c->P.nx || y < 1 || y > c->P.ny || (c->dim == 3 && (z < 1 || z > c->P.nz)))
        return fail(NICR_E_INVALID, "nicr_stamp: centre outside the mesh");
    int rc = ensure_nuc_scratch(c);
    if (rc) return rc;
    NucPar N{x, y, c->dim == 3 ? z : 1, c->cfg.radius_grid, c->cfg.exclusion_gap_grid, 4096};
    if (c->dim == 2) {
        if (c->te == 8) launch_stamp<2, double>(c, N);
        else launch_stamp<2, float>(c, N);
    } else {
        if (c->te == 8) launch_stamp<3, double>(c, N);
        else launch_stamp<3, float>(c, N);
    }
    HIPCHK(hipGetLastError());
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, c->nuc_misc + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n && c->has_eta()) c->mask_dirty = true;
    c->refresh_valid = false;
    if (changed) *changed = n;
    return NICR_OK;
}

int nicr_nucleate(void* h, int* inserted) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    if (inserted) *inserted = 0;
    if (c->cfg.nucleation != NICR_NUC_KSP) return NICR_OK;   // 2D:545
    int rc = ensure_nuc_scratch(c);
    if (rc) return rc;
    NucPar N{0, 0, 0, c->cfg.radius_grid, c->cfg.exclusion_gap_grid, 4096};
    const NucScratch S{c->nuc_tiles, c->nuc_planes, c->nuc_misc, c->nuc_centres, c->nuc_list};
    if (c->dim == 2) {
        if (c->te == 8) launch_nucleate<2, double>(c, N, S, false);
        else launch_nucleate<2, float>(c, N, S, false);
    } else {
        if (c->te == 8) launch_nucleate<3, double>(c, N, S, false);
        else launch_nucleate<3, float>(c, N, S, false);
    }
    HIPCHK(hipGetLastError());
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, c->nuc_misc + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n && c->has_eta()) c->mask_dirty = true;
    c->refresh_valid = false;
    if (inserted) *inserted = n;
    return NICR_OK;
}

// ---- Ksp scan without device allocation (nicr_nucleate_noalloc) ----
// The scan's scratch lives in one file-scope device array per GPU instead of
// buffers the context allocates on its first scan. Its size caps the mesh:
// tiles + planes + 4 + centres + cells ints (each part 256-byte aligned).
constexpr size_t NOALLOC_INTS = size_t(8) << 20;   // 32 MiB per GPU, held until the process ends
static __device__ __align__(256) int scratch[NOALLOC_INTS];
static void* g_scratch_ptr[64] = {};   // device address of `scratch` on each GPU (hipGetSymbolAddress)

static size_t align64(size_t n) { return (n + 63) & ~size_t(63); }

static int noalloc_scratch(Ctx* c, NucScratch* S) {
    const size_t need = align64((size_t)c->P.ntiles) + align64((size_t)c->P.nx) + 64 + align64(4096) +
                        align64((size_t)c->ncell);
    if (need > NOALLOC_INTS)
        return fail(NICR_E_INVALID, "nicr_nucleate_noalloc: a mesh of " + std::to_string(c->ncell) + " cells needs " +
                                        std::to_string(need * 4) + " bytes of scan scratch; the static buffer holds " +
                                        std::to_string(NOALLOC_INTS * 4) + " (use nicr_nucleate)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(NICR_E_DEVICE, "nicr_nucleate_noalloc: device index out of range");
    if (dev != c->device)
        return fail(NICR_E_STATE, "nicr_nucleate_noalloc: the current device is not the context's");
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    int* base = static_cast<int*>(g_scratch_ptr[dev]);
    size_t off = 0;
    auto take = [&](size_t n) {
        int* q = base + off;
        off += align64(n);
        return q;
    };
    S->tiles = take((size_t)c->P.ntiles);
    S->planes = take((size_t)c->P.nx);
    S->misc = take(64);
    S->centres = take(4096);
    S->list = take((size_t)c->ncell);
    return NICR_OK;
}

int nicr_nucleate_noalloc(void* h, int* inserted) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    if (inserted) *inserted = 0;
    if (c->cfg.nucleation != NICR_NUC_KSP) return NICR_OK;   // 2D:545
    NucScratch S{};
    int rc = noalloc_scratch(c, &S);
    if (rc) return rc;
    NucPar N{0, 0, 0, c->cfg.radius_grid, c->cfg.exclusion_gap_grid, 4096};
    if (c->dim == 2) {
        if (c->te == 8) launch_nucleate<2, double>(c, N, S, true);
        else launch_nucleate<2, float>(c, N, S, true);
    } else {
        if (c->te == 8) launch_nucleate<3, double>(c, N, S, true);
        else launch_nucleate<3, float>(c, N, S, true);
    }
    HIPCHK(hipGetLastError());
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, S.misc + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n && c->has_eta()) c->mask_dirty = true;
    c->refresh_valid = false;
    if (inserted) *inserted = n;
    return NICR_OK;
}

int nicr_center_seed(void* h, int* cells) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    if (cells) *cells = 0;
    const double r = c->cfg.radius_grid;
    const int nmin = c->dim == 3 ? std::min({c->P.nx, c->P.ny, c->P.nz}) : std::min(c->P.nx, c->P.ny);
    if (r >= 0.5 * (double(nmin) - 1.))
        return fail(NICR_E_INVALID, c->dim == 3 ? "Center seed is too large: --radius-grid must be less than (min(nx,ny,nz)-1)/2."
                                                : "Center seed is too large: --radius-grid must be less than (min(nx,ny)-1)/2.");
    DevCtl hc;
    int rc = sync_ctl(c, &hc);
    if (rc) return rc;
    if (hc.nuclei != 0) return fail(NICR_E_STATE, "The center oxide has already been initialized.");
    rc = ensure_nuc_scratch(c);
    if (rc) return rc;
    double* const* d = (double* const*)c->dev_obuf;
    if (c->dim == 2) {
        if (c->te == 8)
            hipLaunchKernelGGL((k_center_seed<2, double>), dim3(1), dim3(1024), 0, c->stream, c->P, r, d, d + 2, d + 4,
                               c->has_eta() ? (double*)c->eta[mask_index(c)] : nullptr, c->ctl, c->nuc_misc + 1);
        else
            hipLaunchKernelGGL((k_center_seed<2, float>), dim3(1), dim3(1024), 0, c->stream, c->P, r, d, d + 2, d + 4,
                               c->has_eta() ? (float*)c->eta[mask_index(c)] : nullptr, c->ctl, c->nuc_misc + 1);
    } else {
        if (c->te == 8)
            hipLaunchKernelGGL((k_center_seed<3, double>), dim3(1), dim3(1024), 0, c->stream, c->P, r, d, d + 2, d + 4,
                               c->has_eta() ? (double*)c->eta[mask_index(c)] : nullptr, c->ctl, c->nuc_misc + 1);
        else
            hipLaunchKernelGGL((k_center_seed<3, float>), dim3(1), dim3(1024), 0, c->stream, c->P, r, d, d + 2, d + 4,
                               c->has_eta() ? (float*)c->eta[mask_index(c)] : nullptr, c->ctl, c->nuc_misc + 1);
    }
    HIPCHK(hipGetLastError());
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, c->nuc_misc + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n == 0) return fail(NICR_E_INVALID, "Center seed covers no cell centres: increase --radius-grid.");
    if (c->has_eta()) c->mask_dirty = true;
    c->refresh_valid = false;
    if (cells) *cells = n;
    return NICR_OK;
}

int nicr_field_stats(void* h, int field, nicr_stats* out) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !out) return fail(NICR_E_INVALID, "null argument");
    const double* src = nullptr;
    int mode = 0;
    switch (field) {
    case NICR_FIELD_O: src = c->o[c->cur]; break;
    case NICR_FIELD_CR: src = c->cr[c->cur]; break;
    case NICR_FIELD_PHI: src = c->p[c->cur]; break;
    default: return fail(NICR_E_INVALID, "nicr_field_stats: O, CR or PHI only");
    }
    hipLaunchKernelGGL((k_stats<double>), dim3(c->g_stat), dim3(256), 0, c->stream, src, (const double*)nullptr,
                       c->ncell, mode, c->stat_part);
    std::vector<double> part((size_t)c->g_stat * 4);
    HIPCHK(hipMemcpyAsync(part.data(), c->stat_part, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    long double s = 0;
    double mn = INFINITY, mx = -INFINITY;
    for (int b = 0; b < c->g_stat; ++b) {
        s += (long double)part[4 * b] + (long double)part[4 * b + 1];
        mn = std::min(mn, part[4 * b + 2]);
        mx = std::max(mx, part[4 * b + 3]);
    }
    out->sum[0] = (double)s;
    out->sum[1] = (double)(s - (long double)out->sum[0]);
    out->min = mn;
    out->max = mx;
    return NICR_OK;
}

int nicr_sum_h_phi(void* h, double out_sum[2]) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !out_sum) return fail(NICR_E_INVALID, "null argument");
    hipLaunchKernelGGL((k_stats<double>), dim3(c->g_stat), dim3(256), 0, c->stream, (const double*)c->p[c->cur],
                       (const double*)nullptr, c->ncell, 1, c->stat_part);
    std::vector<double> part((size_t)c->g_stat * 4);
    HIPCHK(hipMemcpyAsync(part.data(), c->stat_part, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    long double s = 0;
    for (int b = 0; b < c->g_stat; ++b) s += (long double)part[4 * b] + (long double)part[4 * b + 1];
    out_sum[0] = (double)s;
    out_sum[1] = (double)(s - (long double)out_sum[0]);
    return NICR_OK;
}

int nicr_set_timing(void* h, int enabled) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    c->timing = enabled != 0;
    return NICR_OK;
}
int nicr_reset_kernel_times(void* h) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    for (auto& t : c->timers) {
        t.total_ms = 0.;
        t.count = 0;
    }
    return NICR_OK;
}
int nicr_kernel_times(void* h, char* buf, size_t len) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !buf || len == 0) return fail(NICR_E_INVALID, "null argument");
    std::string s;
    int n = 0;
    for (auto& t : c->timers) {
        if (!t.count) continue;
        char line[128];
        std::snprintf(line, sizeof line, "%s=%.6f:%lld;", t.name, t.total_ms / t.count, t.count);
        s += line;
        ++n;
    }
    std::snprintf(buf, len, "%s", s.c_str());
    return n;
}
// Plain device copy of `bytes` bytes, `reps` times, on the context stream: the
// bandwidth reference that rocprof byte counters and roofline fractions are
// calibrated against. Returns the mean duration per copy in *ms.
int nicr_copy_probe(void* h, size_t bytes, int reps, double* ms) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c || !ms || reps < 1 || bytes < 1024) return fail(NICR_E_INVALID, "nicr_copy_probe: bad argument");
    const size_t n4 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n4 * 16));
    HIPCHK(hipMalloc(&b, n4 * 16));
    HIPCHK(hipMemsetAsync(a, 0, n4 * 16, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const unsigned grid = (unsigned)((n4 + 255) / 256);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, (const float4*)a, (float4*)b, n4);
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, (const float4*)a, (float4*)b, n4);
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return NICR_OK;
}

int nicr_sync(void* h) {
    Ctx* c = static_cast<Ctx*>(h);
    if (!c) return fail(NICR_E_INVALID, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return NICR_OK;
}

}  // extern "C"

#include "nicr_init.hip"
#include "nicr_grains.hip"
#include "nicr_batch.hip"
#include "nicr_nn.hip"
