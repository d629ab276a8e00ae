// nicr_kernels.hpp -- CDNA4 (gfx950) kernels of the Ni-Cr phase-field time step.
//
// The per-step path of the reference programs (Model::refresh + Model::step +
// Model::eta_trial, mpi2d/NiCr2D_grains.cpp:328-359,408-450,551-597 and
// mpi3d-slab/NiCr3D_grains.cpp:495-526,592-638,761-826) is split into:
//
//   k_mask    eta -> S, P, gb                       (rebuild_eta_mask, 2D:328-340)
//   k_mu      o, cr, phi -> mu_O, mu_Cr             (refresh hot loop 1, 2D:418-426)
//   k_trial   fused mobility + six-face flux + phi  (refresh loop 2 + step trial, 2D:430-433,556-591)
//             + Ck + finite/bounds + budget partials
//   k_eta     eta trial for all grains + next-step S, P, gb   (eta_trial, 2D:347-359)
//   k_commit  all-or-nothing acceptance: reduces budget partials, flips the
//             state index on the device (collective_error + swap, 2D:594-596)
//
// Storage: interior cells only, no halo. Index ((x*ny)+y)*nz+z with Z fastest
// (2D: nz == 1, Y fastest). Halo semantics of ghosts() (2D:154-169, SL:173-191)
// are reproduced by index mapping: every ghost in the reference is an exact copy
// of one interior cell (X: Neumann copy of the first/last plane; Y, Z: periodic
// wrap), so reading that interior cell returns the identical value.
//
// Double buffering: o, cr, phi (and eta with its mask when eta evolves) live in
// two buffers. The kernels read the current index `cur` from the device control
// block, so a rejected step simply leaves `cur` unchanged and any number of
// steps can be queued (or graph-captured) without host round trips. A rejected
// step sets a sticky `failed` flag; later kernels return immediately.
//
// Precision: PREC_FP64 follows the reference's binary64 operation sequence with
// FMA contraction off (file compiled with -ffp-contract=off), so results are
// bitwise equal to the reference. PREC_MIXED evaluates mu, Laplacians, fluxes and
// the phi/eta forces in binary32 (FMA allowed), keeps O/Cr/phi in binary64, keeps
// thresholds, Ck, the oxygen inlet flux and all budgets in binary64 and stores
// eta and its mask in binary32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace nicr {

constexpr int PREC_FP64 = 0;
constexpr int PREC_MIXED = 1;

constexpr int TILE = 256;        // cells of one (x, yz-chunk) tile = threads per block
constexpr int NPART = 12;        // doubles per block of step budget partials
constexpr unsigned long long NO_KEY = ~0ull;

// Physics constants (2D:65-71, SL:73-79).
struct Phys {
    static constexpr double Vm = 1., W = 1e4, ko = 1e-9, kc = 7.5e-9, kp = 1.7e-10, L = 0.2;
    static constexpr double bmo = 2.21e-16, bmc = 3.5436e-20, boo = 2.21e-18, boc = 3.5436e-21;
    static constexpr double ckO = 2e-10, ckCr = 2e-10, ksp = 1.28e-7;
    static constexpr double bulkO = 0., bulkCr = 0.1986, layerO = 0.0023, layerCr = 0.182;
    static constexpr double oxideO = 0.56300203, oxideCr = 0.35677915;
};

// Device control block: one per context.
struct DevCtl {
    int cur;                      // index of the accepted state buffers
    int failed;                   // sticky: a step was rejected
    int steps_done;               // accepted steps since the host last cleared it
    int fail_kind;                // 1 PDE trial, 2 eta trial
    unsigned long long pde_key;   // min failing cell (reference loop order)
    unsigned long long eta_key;   // min failing g*ncell + cell
    double budget[16];            // nucO, nucCr, ckO, ckCr, clipO, clipCr, bcO, bcCr as (hi, lo)
    long long nuclei, clipped;
};

// Scalar parameters passed by value.
struct Par {
    int nx, ny, nz, nyz;          // nz == 1 in 2D
    int nchunk, ntiles;           // tiles of TILE cells along the flattened yz plane
    long long ncell;
    long long fstride;            // elements from one eta field to the next (>= ncell, see nicr_create)
    int n_eta, bc, ck_legacy, bounds_legacy;
    int grains, oxy, has_eta, eta_evolves;
    int enforce_layer, pad0;
    double layer_o;               // inlet O of the concentration BC (Phys::layerO by default)
    double dx, dy, dz, dt;
    double gbfo, gbfc, mu_res, c_ref;
    double etaW, etaK, etaL, etaA, eta_scale;
    float fdx2i, fdy2i, fdz2i, fdxi, fdyi, fdzi;   // binary32 reciprocals for PREC_MIXED
    double ck_o_dt, ck_cr_dt;     // Ck_O / dt, Ck_Cr / dt (PREC_MIXED)
};

// Buffers. TE is the eta/mask storage type (double, or float for PREC_MIXED);
// TM the mu scratch type.
template <class TE, class TM>
struct Bufs {
    double* o[2];
    double* cr[2];
    double* p[2];
    TM* mu_o;
    TM* mu_cr;
    TE* eta[2];     // n_eta fields each, field g at g * fstride
    TE* s2[2];
    TE* pairs[2];
    TE* gb[2];      // static mask: gb[0] == gb[1]
    // PREC_MIXED with the marching kernels: the per-cell factor of the eta force,
    // W_eta (S - (1 - h)) + 2 W_eta S + A_eta h of the step's input state,
    // written by the PDE kernel (which holds S and phi of every cell) for the
    // eta passes; nullptr: the eta kernels form it from phi and S themselves
    float* ecoef;
    double* partials;
    DevCtl* ctl;
};

// XCD-aware tile order: the 8 XCDs of MI355X receive blocks round-robin, so
// block b runs on XCD b % 8. Give each XCD a contiguous run of tiles so that
// neighbouring planes share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int total) {
    const int x = b & 7, q = total >> 3, r = total & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// ---------------------------------------------------------------------------
// binary64 restatements (2D:79,152,174-191; SL:87,171,198-215)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double h64(double p) { return p * p * p * (10. - 15. * p + 6. * p * p); }
__device__ __forceinline__ double fmetal64(double o, double cr) {
    return (6000000. * o * o + 583264.303945031 * cr * cr + 3200000. * o * cr
            - 298432. * o - 97045.64 * cr - 61732.2974223055) / Phys::Vm;
}
__device__ __forceinline__ double foxide64(double o, double cr) {
    return (12000000. * o * o + 7000000. * cr * cr - 13800000. * o * cr
            - 8880000. * o + 2680000. * cr + 1860990.) / Phys::Vm;
}
__device__ __forceinline__ double mu_metal_o64(double o, double cr) {
    return (12000000. * o + 3200000. * cr - 298432.) / Phys::Vm;
}
__device__ __forceinline__ double mobO64(const Par& P, double conc, double p, double gb) {
    if (!P.oxy) return 0.;
    double hh = h64(p);
    const double active_gb = P.grains ? gb : 0.;
    return conc * ((1. - hh) * Phys::bmo * (1. + (P.gbfo - 1.) * active_gb) + hh * Phys::boo);
}
__device__ __forceinline__ double mobCr64(const Par& P, double conc, double p, double gb) {
    double hh = h64(p);
    const double active_gb = P.grains ? gb : 0.;
    return conc * ((1. - hh) * Phys::bmc * (1. + (P.gbfc - 1.) * active_gb) + hh * Phys::boc);
}

// binary32 forms for PREC_MIXED (same polynomials; FMA contraction allowed).
__device__ __forceinline__ float h32(float p) {
#pragma clang fp contract(fast)
    return p * p * p * (10.f - 15.f * p + 6.f * p * p);
}
__device__ __forceinline__ float mobO32(const Par& P, float conc, float hh, float gb) {
#pragma clang fp contract(fast)
    const float active_gb = P.grains ? gb : 0.f;
    return conc * ((1.f - hh) * float(Phys::bmo) * (1.f + float(P.gbfo - 1.) * active_gb) + hh * float(Phys::boo));
}
__device__ __forceinline__ float mobCr32(const Par& P, float conc, float hh, float gb) {
#pragma clang fp contract(fast)
    const float active_gb = P.grains ? gb : 0.f;
    return conc * ((1.f - hh) * float(Phys::bmc) * (1.f + float(P.gbfc - 1.) * active_gb) + hh * float(Phys::boc));
}

// ---------------------------------------------------------------------------
// double-double accumulation for budgets (the reference uses long double)
// ---------------------------------------------------------------------------
struct DD { double hi, lo; };
__device__ __forceinline__ DD dd_add(DD a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + err;
    const double hi = s + lo;
    return DD{hi, lo - (hi - s)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = a.lo + b.lo + err;
    const double hi = s + lo;
    return DD{hi, lo - (hi - s)};
}

// Raw buffer access (gfx9 buffer resource, dword 3 = 32-bit data format): the
// address is base + soffset (scalar) + voffset (lane), so one lane offset can
// serve many fields whose distance is a uniform scalar.
constexpr int BUF_DWORD3 = 0x00020000;
// s_waitcnt immediate (gfx9 encoding) for vmcnt(0) with the other counters
// not waited on: expcnt 7 (bits 6:4), lgkmcnt 15 (bits 11:8), vmcnt 0 (bits
// 3:0 and 15:14)
constexpr int WAIT_VM0 = 0x0F70;
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
// Stores carry the non-temporal hint (cache-policy bit 1): the step writes
// each output once and reads it only in the next step, so the lines should not
// displace the inputs that neighbouring blocks are about to read from L2.
constexpr int BUF_NT = 2;
__device__ __forceinline__ void buf_store_raw(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, BUF_NT);
}
__device__ __forceinline__ void buf_store_raw(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double v) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, (int)voff, (int)soff, BUF_NT);
}
// the same hint for plain pointer stores
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
template <class T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, T v) {
    buf_store_raw(r, voff, soff, v);
}

// Neighbour mapping that reproduces ghosts(): X Neumann, Y/Z periodic.
struct Nbr {
    long long c, xm, xp, ym, yp, zm, zp;
    int x;
};
template <int DIM>
__device__ __forceinline__ Nbr neighbours(const Par& P, int x, int y, int z) {
    Nbr n;
    const long long plane = P.nyz;
    const long long c = (long long)x * plane + (long long)y * P.nz + z;
    n.c = c;
    n.x = x;
    n.xm = x == 0 ? c : c - plane;
    n.xp = x == P.nx - 1 ? c : c + plane;
    n.ym = y == 0 ? c + (long long)(P.ny - 1) * P.nz : c - P.nz;
    n.yp = y == P.ny - 1 ? c - (long long)(P.ny - 1) * P.nz : c + P.nz;
    if (DIM == 3) {
        n.zm = z == 0 ? c + (P.nz - 1) : c - 1;
        n.zp = z == P.nz - 1 ? c - (P.nz - 1) : c + 1;
    } else {
        n.zm = n.zp = c;
    }
    return n;
}

}  // namespace nicr
