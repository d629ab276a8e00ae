// nicr_step.hip -- step kernels (see nicr_kernels.hpp for the design notes).
#include "nicr_kernels.hpp"

namespace nicr {

// Tile -> cell coordinates. Returns false for padding lanes of the last chunk.
__device__ __forceinline__ bool tile_cell(const Par& P, int tile, int& x, int& y, int& z) {
    x = tile / P.nchunk;
    const int yz = (tile - x * P.nchunk) * TILE + (int)threadIdx.x;
    if (yz >= P.nyz) return false;
    y = yz / P.nz;
    z = yz - y * P.nz;
    return true;
}

// ---------------------------------------------------------------------------
// k_mask: S = sum eta^2, P = sum_{a<b} eta_a^2 eta_b^2 (running form of
// 2D:332), gb = scale * min(1, 16 P)  (rebuild_eta_mask, 2D:328-340, SL:495-507)
// ---------------------------------------------------------------------------
// pairs_only: write P alone (the step's last eta pass leaves P unwritten, as
// nothing in the step reads it; downloads rebuild it here without touching the
// S and gb that the next step reads).
template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_mask(Par P, Bufs<TE, TM> B, int which, int pairs_only) {
    int x, y, z;
    const int t = xcd_remap(blockIdx.x, P.ntiles);
    if (!tile_cell(P, t, x, y, z)) return;
    const int e = which < 0 ? B.ctl->cur : which;
    const long long c = (long long)x * P.nyz + (long long)y * P.nz + z;
    const TE* eta = B.eta[e];
    if constexpr (PREC == PREC_FP64) {
        double S = 0., pr = 0.;
        for (int g = 0; g < P.n_eta; ++g) {
            const double v = (double)eta[(long long)g * P.fstride + c];
            const double q = v * v;
            pr += S * q;
            S += q;
        }
        const double s16 = 16. * pr;
        B.pairs[e][c] = (TE)pr;
        if (pairs_only) return;
        B.s2[e][c] = (TE)S;
        B.gb[e][c] = (TE)(P.eta_scale * (s16 < 1. ? s16 : 1.));
    } else {
        float S = 0.f, pr = 0.f;
        for (int g = 0; g < P.n_eta; ++g) {
            const float v = (float)eta[(long long)g * P.fstride + c];
            const float q = v * v;
            pr += S * q;
            S += q;
        }
        const float s16 = 16.f * pr;
        B.pairs[e][c] = (TE)pr;
        if (pairs_only) return;
        B.s2[e][c] = (TE)S;
        B.gb[e][c] = (TE)((float)P.eta_scale * (s16 < 1.f ? s16 : 1.f));
    }
}

// ---------------------------------------------------------------------------
// k_mu: chemical potentials (refresh hot loop 1, 2D:418-426, SL:604-612).
// `enforce` substitutes the concentration-BC inlet value for o(1,.) as
// enforce_legacy_layer() does in place before refresh() (2D:408-411,552).
// PREC_MIXED stores mu + const offsets (only differences of mu enter the flux).
// ---------------------------------------------------------------------------
constexpr double MU_O_OFFSET = 298432.;     // mu_O(metal) = ... - 298432
constexpr double MU_CR_OFFSET = 97045.64;   // mu_Cr(metal) = ... - 97045.64

template <int DIM>
__device__ __forceinline__ double lap64(const double* a, const Nbr& n, double ac, const Par& P,
                                         double axm, double axp) {
    const double xy = (axp - 2. * ac + axm) / (P.dx * P.dx)
                    + (a[n.yp] - 2. * ac + a[n.ym]) / (P.dy * P.dy);
    if (DIM == 2) return xy;
    return xy + (a[n.zp] - 2. * ac + a[n.zm]) / (P.dz * P.dz);
}

// Seven-point stencil values: centre, x-, x+, y-, y+, z-, z+ (2D: z entries
// equal the centre and are not used). o carries the inlet substitution.
enum { SC = 0, SXM = 1, SXP = 2, SYM = 3, SYP = 4, SZM = 5, SZP = 6 };

// mu_O, mu_Cr at the centre from o, cr at 7 points and phi at the centre
// (refresh hot loop 1, 2D:418-426, SL:604-612), binary64 in reference order.
template <int DIM>
__device__ __forceinline__ void mu64_vals(const Par& P, const double* o7, const double* c7, double pc, double& muO,
                                          double& muCr) {
    const double oc = o7[SC], cc = c7[SC];
    double hh = h64(pc);
    double lo = (o7[SXP] - 2. * oc + o7[SXM]) / (P.dx * P.dx) + (o7[SYP] - 2. * oc + o7[SYM]) / (P.dy * P.dy);
    if (DIM == 3) lo = lo + (o7[SZP] - 2. * oc + o7[SZM]) / (P.dz * P.dz);
    double lc = (c7[SXP] - 2. * cc + c7[SXM]) / (P.dx * P.dx) + (c7[SYP] - 2. * cc + c7[SYM]) / (P.dy * P.dy);
    if (DIM == 3) lc = lc + (c7[SZP] - 2. * cc + c7[SZM]) / (P.dz * P.dz);
    muO = (1. - hh) * mu_metal_o64(oc, cc)
        + hh * (24000000. * oc - 13800000. * cc - 8880000.) / Phys::Vm
        - 2. * Phys::ko * lo;
    muCr = (1. - hh) * (1166528.60789006 * cc + 3200000. * oc - 97045.64) / Phys::Vm
         + hh * (14000000. * cc - 13800000. * oc + 2680000.) / Phys::Vm
         - 2. * Phys::kc * lc;
}

// binary32 form: differences of the binary64 state are formed in binary64, the
// rest in binary32; mu is stored as mu_O + 298432 and mu_Cr + 97045.64 (the
// metal constants cancel in every face difference).
template <int DIM>
__device__ __forceinline__ void mu32_vals(const Par& P, const double* o7, const double* c7, double pc, float& muO,
                                          float& muCr) {
#pragma clang fp contract(fast)
    const double oc = o7[SC], cc = c7[SC];
    float lapo = (float)(o7[SXP] - 2. * oc + o7[SXM]) * P.fdx2i + (float)(o7[SYP] - 2. * oc + o7[SYM]) * P.fdy2i;
    float lapc = (float)(c7[SXP] - 2. * cc + c7[SXM]) * P.fdx2i + (float)(c7[SYP] - 2. * cc + c7[SYM]) * P.fdy2i;
    if (DIM == 3) {
        lapo += (float)(o7[SZP] - 2. * oc + o7[SZM]) * P.fdz2i;
        lapc += (float)(c7[SZP] - 2. * cc + c7[SZM]) * P.fdz2i;
    }
    const float hh = h32((float)pc);
    const float fo = (float)oc, fc = (float)cc;
    muO = (1.f - hh) * (12000000.f * fo + 3200000.f * fc) + hh * (24000000.f * fo - 13800000.f * fc - 8581568.f)
        - float(2. * Phys::ko) * lapo;
    muCr = (1.f - hh) * (1166528.60789006f * fc + 3200000.f * fo) + hh * (14000000.f * fc - 13800000.f * fo + 2777045.64f)
         - float(2. * Phys::kc) * lapc;
}

// The per-cell factor of the binary32 eta force, F_g = eta_g (coef - 2 W_eta
// eta_g^2) - K lap(eta_g), with coef = W_eta (S - (1 - h)) + 2 W_eta S + A_eta h
// (grain_local_deta, 2D:100-101, SL:108-109).
__device__ __forceinline__ float eta_coef32(const Par& P, float S, float pphi) {
#pragma clang fp contract(fast)
    const float hphi = h32(pphi);
    return (float)P.etaW * (S - (1.f - hphi)) + 2.f * (float)P.etaW * S + (float)P.etaA * hphi;
}

// o, cr at the 7 points of a cell from global memory, with the inlet substitution.
template <int DIM>
__device__ __forceinline__ void gather_oc(const Par& P, const double* o, const double* cr, const Nbr& n, bool enf,
                                          double* o7, double* c7) {
    const long long idx[7] = {n.c, n.xm, n.xp, n.ym, n.yp, n.zm, n.zp};
    const int xm = n.x == 0 ? 0 : n.x - 1, xp = n.x == P.nx - 1 ? n.x : n.x + 1;
    const int qx[7] = {n.x, xm, xp, n.x, n.x, n.x, n.x};
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        if (DIM == 2 && s >= SZM) {
            o7[s] = o7[SC];
            c7[s] = c7[SC];
            continue;
        }
        o7[s] = (enf && qx[s] == 0) ? P.layer_o : o[idx[s]];
        c7[s] = cr[idx[s]];
    }
}

template <int DIM>
__device__ __forceinline__ void mu64_at(const Par& P, const double* o, const double* cr, const double* p,
                                         const Nbr& n, bool enforce, double& muO, double& muCr) {
    double o7[7], c7[7];
    gather_oc<DIM>(P, o, cr, n, enforce, o7, c7);
    mu64_vals<DIM>(P, o7, c7, p[n.c], muO, muCr);
}

template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_mu(Par P, Bufs<TE, TM> B) {
    const DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur = ctl->cur;
    int x, y, z;
    const int t = xcd_remap(blockIdx.x, P.ntiles);
    if (!tile_cell(P, t, x, y, z)) return;
    const Nbr n = neighbours<DIM>(P, x, y, z);
    double o7[7], c7[7];
    gather_oc<DIM>(P, B.o[cur], B.cr[cur], n, P.enforce_layer != 0, o7, c7);
    const double pc = B.p[cur][n.c];
    if constexpr (PREC == PREC_FP64) {
        double muO, muCr;
        mu64_vals<DIM>(P, o7, c7, pc, muO, muCr);
        B.mu_o[n.c] = (TM)muO;
        B.mu_cr[n.c] = (TM)muCr;
    } else {
        float muO, muCr;
        mu32_vals<DIM>(P, o7, c7, pc, muO, muCr);
        B.mu_o[n.c] = (TM)muO;
        B.mu_cr[n.c] = (TM)muCr;
    }
}

// ---------------------------------------------------------------------------
// block reduction of the step budget partials
// ---------------------------------------------------------------------------
__device__ __forceinline__ DD block_reduce_dd(DD v, DD* sh) {
    // wave-level first (64 lanes), then across the 4 waves of the block
    for (int off = 32; off > 0; off >>= 1) {
        DD o;
        o.hi = __shfl_down(v.hi, off);
        o.lo = __shfl_down(v.lo, off);
        v = dd_add(v, o);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    DD r{0., 0.};
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = dd_add(r, sh[w]);
    return r;
}

// Six double-double block sums at once. Each value goes through the same
// operations in the same order as in block_reduce_dd (the shuffle tree, then
// the waves in order from a zero sum), so the results are bitwise those of six
// block_reduce_dd calls; the six chains interleave, one barrier pair serves
// all of them, and thread s (s < 6) sums value s across the waves. sh holds
// 6 * (blockDim.x / 64) entries. Returns value threadIdx.x on threads 0..5.
__device__ __forceinline__ DD block_reduce_dd6(const DD (&in)[6], DD* sh) {
    DD v[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) v[s] = in[s];
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            DD o;
            o.hi = __shfl_down(v[s].hi, off);
            o.lo = __shfl_down(v[s].lo, off);
            v[s] = dd_add(v[s], o);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    __syncthreads();   // sh may still be read by a previous reduction
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < 6; ++s) sh[s * nw + wave] = v[s];
    }
    __syncthreads();
    DD r{0., 0.};
    if (threadIdx.x < 6)
        for (int w = 0; w < nw; ++w) r = dd_add(r, sh[threadIdx.x * nw + w]);
    return r;
}

// ---------------------------------------------------------------------------
// k_trial: mobilities + conservative six-face update of O and Cr, Allen-Cahn
// phi update, legacy Ck, finite check and bounds (2D:430-433,435-450,553-591;
// SL:616-638,771-809). Persistent grid over tiles; one budget partial per block.
// ---------------------------------------------------------------------------
// mu_O, mu_Cr and the mobilities M_O, M_Cr of one cell (refresh loops 1 and 2,
// 2D:418-433, SL:604-619). Every face of the trial uses the same per-cell M, and
// the ghosts are copies, so M(ghost) = M(copy).
template <class T>
struct MuMob {
    T mo, mc, Mo, Mc;
};
template <int PREC>
using MuT = typename std::conditional<PREC == PREC_FP64, double, float>::type;   // mu, M arithmetic type

template <int PREC, class T>
__device__ __forceinline__ void mob_vals(const Par& P, double o, double cr, double p, double g, MuMob<T>& r) {
    if constexpr (PREC == PREC_FP64) {
        r.Mo = mobO64(P, o, p, g);
        r.Mc = mobCr64(P, cr, p, g);
    } else {
        const float hh = h32((float)p);
        r.Mo = P.oxy ? mobO32(P, (float)o, hh, (float)g) : 0.f;
        r.Mc = mobCr32(P, (float)cr, hh, (float)g);
    }
}

template <int DIM, int PREC, class T>
__device__ __forceinline__ MuMob<T> mu_mob(const Par& P, const double* o7, const double* c7, double pc, double g) {
    MuMob<T> r;
    if constexpr (PREC == PREC_FP64) {
        mu64_vals<DIM>(P, o7, c7, pc, r.mo, r.mc);
    } else {
        mu32_vals<DIM>(P, o7, c7, pc, r.mo, r.mc);
    }
    mob_vals<PREC, T>(P, o7[SC], c7[SC], pc, g, r);
    return r;
}

// The flux through the face between cells a (lower index) and b along an axis
// of spacing hh, as div_flux writes it (2D:444-448, SL:630-636): the same
// expression for both cells that share the face.
__device__ __forceinline__ double face64(double Ma, double Mb, double ua, double ub, double hh) {
    return -.5 * (Ma + Mb) * (ub - ua) / hh;
}
// The face fluxes of one cell and species: x-, x+, y-, y+, z-, z+ (left,
// right, bottom, top, back, front).
struct Faces {
    double xm, xp, ym, yp, zm, zp;
};

// Binary64 trial of one cell from its face fluxes: the mu-BC inlet face
// (2D:435-442,446), div_flux (2D:449, SL:637), the O and Cr updates and the phi
// force and update (2D:556-566, SL:774-784). uo: mu_O at the cell, g_c: gb at
// the cell (inlet face only).
template <int DIM>
__device__ __forceinline__ void trial_faces64(const Par& P, Faces fo, const Faces& fc, const double* p7, double oo,
                                              double cc, int x, double S, double g_c, double uo, double& onv,
                                              double& crv, double& pnv, double& left_flux) {
    const double pp = p7[SC];
    left_flux = 0.;
    double divO = 0.;
    if (P.oxy) {
        if (x == 0 && P.bc == 1) {
            const double mf = mobO64(P, P.c_ref, pp, g_c);
            left_flux = 2. * mf * (P.mu_res - uo) / P.dx;
            fo.xm = left_flux;
        }
        double d = (fo.xm - fo.xp) / P.dx + (fo.ym - fo.yp) / P.dy;
        if (DIM == 3) d = d + (fo.zm - fo.zp) / P.dz;
        divO = d;
    }
    double divC = (fc.xm - fc.xp) / P.dx + (fc.ym - fc.yp) / P.dy;
    if (DIM == 3) divC = divC + (fc.zm - fc.zp) / P.dz;
    onv = P.oxy ? oo + P.dt * divO : oo;
    crv = cc + P.dt * divC;
    const double hp = 30. * pp * pp * (pp - 1.) * (pp - 1.);
    double lp = (p7[SXP] - 2. * pp + p7[SXM]) / (P.dx * P.dx) + (p7[SYP] - 2. * pp + p7[SYM]) / (P.dy * P.dy);
    if (DIM == 3) lp = lp + (p7[SZP] - 2. * pp + p7[SZM]) / (P.dz * P.dz);
    double fp = Phys::W * (4. * pp * pp * pp - 6. * pp * pp + 2. * pp)
              + hp * (foxide64(oo, cc) - fmetal64(oo, cc)) - 2. * Phys::kp * lp;
    if (P.has_eta) fp += 0.5 * hp * (P.etaW * (S - (1. - h64(pp))) + P.etaA * S);
    pnv = pp - P.dt * Phys::L * fp;
}

// Trial values of one cell from mu, M at its 7 points, phi at its 7 points and
// o, cr at the cell: div_flux with the mu-BC inlet face (2D:435-450,
// SL:621-638), the phi force and update (2D:562-566, SL:780-784).
// g_c: gb at the cell, m64O: binary64 mu_O at the cell (both for the inlet face
// only; m64O only in PREC_MIXED).
template <int DIM, int PREC, class T>
__device__ __forceinline__ void trial_mm(const Par& P, const MuMob<T>* m, const double* p7, double oo, double cc,
                                         int x, double S, double g_c, double m64O, double& onv, double& crv,
                                         double& pnv, double& left_flux) {
    const double pp = p7[SC];
    left_flux = 0.;
    if constexpr (PREC == PREC_FP64) {
        const double moc = m[SC].Mo, mcc = m[SC].Mc;
        const double uo = m[SC].mo, uc = m[SC].mc;
        Faces fo{0., 0., 0., 0., 0., 0.}, fc{0., 0., 0., 0., 0., 0.};
        if (P.oxy) {
            fo.xm = face64(m[SXM].Mo, moc, m[SXM].mo, uo, P.dx);
            fo.xp = face64(moc, m[SXP].Mo, uo, m[SXP].mo, P.dx);
            fo.ym = face64(m[SYM].Mo, moc, m[SYM].mo, uo, P.dy);
            fo.yp = face64(moc, m[SYP].Mo, uo, m[SYP].mo, P.dy);
            if (DIM == 3) {
                fo.zm = face64(m[SZM].Mo, moc, m[SZM].mo, uo, P.dz);
                fo.zp = face64(moc, m[SZP].Mo, uo, m[SZP].mo, P.dz);
            }
        }
        fc.xm = face64(m[SXM].Mc, mcc, m[SXM].mc, uc, P.dx);
        fc.xp = face64(mcc, m[SXP].Mc, uc, m[SXP].mc, P.dx);
        fc.ym = face64(m[SYM].Mc, mcc, m[SYM].mc, uc, P.dy);
        fc.yp = face64(mcc, m[SYP].Mc, uc, m[SYP].mc, P.dy);
        if (DIM == 3) {
            fc.zm = face64(m[SZM].Mc, mcc, m[SZM].mc, uc, P.dz);
            fc.zp = face64(mcc, m[SZP].Mc, uc, m[SZP].mc, P.dz);
        }
        trial_faces64<DIM>(P, fo, fc, p7, oo, cc, x, S, g_c, uo, onv, crv, pnv, left_flux);
    } else {
#pragma clang fp contract(fast)
        const float hc = h32((float)pp);
        const float moc = m[SC].Mo, mcc = m[SC].Mc;
        const float uo = m[SC].mo, uc = m[SC].mc;
        float sumO = 0.f, sumC = 0.f;   // sum over faces of M_face * (mu_nb - mu_c) / h^2
        auto face = [&](int q, float inv2, bool with_o) {
            if (P.oxy && with_o) sumO += 0.5f * (moc + m[q].Mo) * (m[q].mo - uo) * inv2;
            sumC += 0.5f * (mcc + m[q].Mc) * (m[q].mc - uc) * inv2;
        };
        face(SXP, P.fdx2i, true);
        const bool inlet = x == 0 && P.bc == 1;
        face(SXM, P.fdx2i, !inlet);
        if (inlet && P.oxy) {
            // left face replaced by the binary64 inlet flux (2D:435-442,446)
            const double mf = mobO64(P, P.c_ref, pp, g_c);
            left_flux = 2. * mf * (P.mu_res - m64O) / P.dx;
            sumO += (float)(left_flux / P.dx);
        }
        face(SYP, P.fdy2i, true);
        face(SYM, P.fdy2i, true);
        if (DIM == 3) {
            face(SZP, P.fdz2i, true);
            face(SZM, P.fdz2i, true);
        }
        onv = P.oxy ? oo + P.dt * (double)sumO : oo;
        crv = cc + P.dt * (double)sumC;
        const float fpp = (float)pp;
        const float hp = 30.f * fpp * fpp * (fpp - 1.f) * (fpp - 1.f);
        const float lp = (float)(p7[SXP] - 2. * pp + p7[SXM]) * P.fdx2i + (float)(p7[SYP] - 2. * pp + p7[SYM]) * P.fdy2i
                       + (DIM == 3 ? (float)(p7[SZP] - 2. * pp + p7[SZM]) * P.fdz2i : 0.f);
        const float fo = (float)oo, fcr = (float)cc;
        // foxide - fmetal, expanded (2D:174-181)
        const float dfx = 6000000.f * fo * fo + (7000000.f - 583264.303945031f) * fcr * fcr - 17000000.f * fo * fcr
                        - 8581568.f * fo + 2777045.64f * fcr + float(1860990. + 61732.2974223055);
        float fp = float(Phys::W) * (4.f * fpp * fpp * fpp - 6.f * fpp * fpp + 2.f * fpp) + hp * dfx
                 - float(2. * Phys::kp) * lp;
        if (P.has_eta) fp += 0.5f * hp * ((float)P.etaW * ((float)S - (1.f - hc)) + (float)P.etaA * (float)S);
        pnv = pp - P.dt * Phys::L * (double)fp;
    }
}

// Per-block accumulators of the step budget.
struct StepAcc {
    DD bc{0., 0.}, ckO{0., 0.}, ckCr{0., 0.}, clO{0., 0.}, clCr{0., 0.};
    double n_clip = 0.;
    unsigned long long bad = NO_KEY;
};

// Legacy Ck (2D:567-571), finite check (2D:572-573), bounds: clamp + phi snap
// (2D:574-580) or rejection (2D:581-589), inlet budget (2D:408-411,553) with
// inlet O layer_o. binary64 in both precisions. Writes the trial values (raw
// at a failing cell).
// binary64 h(phi) with fused multiply-adds (mixed precision only)
__device__ __forceinline__ double h64_fma(double p) {
#pragma clang fp contract(fast)
    return p * p * p * (10. + p * (-15. + 6. * p));
}

// finish_cell_vals leaves the final values in onv, crv, pnv for the caller to
// store; finish_cell stores them.
template <int PREC = PREC_FP64>
__device__ __forceinline__ void finish_cell_vals(const Par& P, long long c, int x, double o_stored, double oo,
                                                 double cc, double pp, double& onv, double& crv, double& pnv,
                                                 double left_flux, StepAcc& a, double layer_o) {
    if (P.ck_legacy) {
        double dO, dCr;
        if constexpr (PREC == PREC_FP64) {
            dO = -Phys::ckO * ((Phys::oxideO - onv) * h64(pnv) - (Phys::oxideO - oo) * h64(pp)) / P.dt;
            dCr = -Phys::ckCr * ((Phys::oxideCr - crv) * h64(pnv) - (Phys::oxideCr - cc) * h64(pp)) / P.dt;
        } else {
            // mixed: still binary64, with Ck/dt as one factor; each product is
            // rounded on its own, so an unchanged cell gives exactly zero
            const double hn = h64_fma(pnv), ho = h64_fma(pp);
            dO = -P.ck_o_dt * (__dmul_rn(Phys::oxideO - onv, hn) - __dmul_rn(Phys::oxideO - oo, ho));
            dCr = -P.ck_cr_dt * (__dmul_rn(Phys::oxideCr - crv, hn) - __dmul_rn(Phys::oxideCr - cc, ho));
        }
        onv += dO;
        crv += dCr;
        // adding a zero leaves a double-double sum bitwise unchanged (its parts
        // are never -0), so the common zero terms (no change of h(phi), no
        // clip) skip the addition
        if (dO != 0.) a.ckO = dd_add(a.ckO, dO);
        if (dCr != 0.) a.ckCr = dd_add(a.ckCr, dCr);
    }
    if (x == 0) {
        if (P.enforce_layer) a.bc = dd_add(a.bc, layer_o - o_stored);
        if (P.bc == 1 && P.oxy) a.bc = dd_add(a.bc, P.dt * left_flux / P.dx);
    }
    if (!isfinite(onv) || !isfinite(crv) || !isfinite(pnv)) {
        a.bad = a.bad < (unsigned long long)c ? a.bad : (unsigned long long)c;
    } else if (P.bounds_legacy) {
        const double oc = onv < 0. ? 0. : (1. < onv ? 1. : onv);
        const double ccr = crv < 0. ? 0. : (1. < crv ? 1. : crv);
        if (oc != onv || ccr != crv) {
            a.clO = dd_add(a.clO, oc - onv);
            a.clCr = dd_add(a.clCr, ccr - crv);
            a.n_clip += 1.;
        }
        onv = oc;
        crv = ccr;
        if (pnv > .99999) pnv = 1.;
        if (pnv < 1.e-5) pnv = 0.;
    } else if (onv < -1.e-12 || crv < -1.e-12 || onv + crv > 1. + 1.e-12 || pnv < -1.e-12 || pnv > 1. + 1.e-12) {
        a.bad = a.bad < (unsigned long long)c ? a.bad : (unsigned long long)c;
    }
}
template <int PREC = PREC_FP64>
__device__ __forceinline__ void finish_cell(const Par& P, long long c, int x, double o_stored, double oo, double cc,
                                            double pp, double onv, double crv, double pnv, double left_flux,
                                            StepAcc& a, double* on, double* crn, double* pn, double layer_o) {
    finish_cell_vals<PREC>(P, c, x, o_stored, oo, cc, pp, onv, crv, pnv, left_flux, a, layer_o);
    st_nt(&on[c], onv);
    st_nt(&crn[c], crv);
    st_nt(&pn[c], pnv);
}

// Block-reduce the accumulators into the partial slot w (NPART doubles) and
// the failing cell into *key.
__device__ __forceinline__ void block_partials(const StepAcc& a, double* w, unsigned long long* key) {
    if (a.bad != NO_KEY) atomicMin(key, a.bad);
    __shared__ DD sh[6 * 16];
    const DD v[6] = {a.bc, a.ckO, a.ckCr, a.clO, a.clCr, DD{a.n_clip, 0.}};
    const DD r = block_reduce_dd6(v, sh);
    // slots: bcO, ckO, ckCr, clipO, clipCr as (hi, lo), then the clipped count and 0
    if (threadIdx.x < 6) {
        w[2 * threadIdx.x] = r.hi;
        w[2 * threadIdx.x + 1] = threadIdx.x < 5 ? r.lo : 0.;
    }
}
// ... into this block's partial slot.
__device__ __forceinline__ void write_partials(const StepAcc& a, double* partials, DevCtl* ctl) {
    block_partials(a, partials + (long long)blockIdx.x * NPART, &ctl->pde_key);
}

template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_trial(Par P, Bufs<TE, TM> B) {
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur = ctl->cur, nxt = cur ^ 1;
    const double* o = B.o[cur];
    const double* cr = B.cr[cur];
    const double* p = B.p[cur];
    const int me = P.eta_evolves ? cur : 0;
    const TE* gb = B.gb[me];
    const TE* s2 = B.s2[me];
    const bool enf = P.enforce_layer != 0;
    StepAcc acc;
    const int G = gridDim.x;
    for (int vb = blockIdx.x; vb < P.ntiles; vb += G) {
        const int t = xcd_remap(vb, P.ntiles);
        int x, y, z;
        if (!tile_cell(P, t, x, y, z)) continue;
        const Nbr n = neighbours<DIM>(P, x, y, z);
        const long long c = n.c;
        using T = MuT<PREC>;
        double o7[7], c7[7], p7[7];
        MuMob<T> m[7];
        gather_oc<DIM>(P, o, cr, n, enf, o7, c7);
        const long long idx[7] = {n.c, n.xm, n.xp, n.ym, n.yp, n.zm, n.zp};
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (DIM == 2 && q >= SZM) break;
            p7[q] = p[idx[q]];
            m[q].mo = (T)B.mu_o[idx[q]];
            m[q].mc = (T)B.mu_cr[idx[q]];
            mob_vals<PREC, T>(P, o7[q], c7[q], p7[q], (double)gb[idx[q]], m[q]);
        }
        double m64O = 0.;
        if constexpr (PREC != PREC_FP64) {
            if (x == 0 && P.bc == 1 && P.oxy) {
                double m64C;
                mu64_vals<DIM>(P, o7, c7, p7[SC], m64O, m64C);
            }
        }
        const double S = P.has_eta ? (double)s2[c] : 0.;
        double onv, crv, pnv, left_flux;
        trial_mm<DIM, PREC, T>(P, m, p7, o7[SC], c7[SC], x, S, (double)gb[c], m64O, onv, crv, pnv, left_flux);
        finish_cell<PREC>(P, c, x, o[c], o7[SC], c7[SC], p7[SC], onv, crv, pnv, left_flux, acc, B.o[nxt], B.cr[nxt],
                    B.p[nxt], P.layer_o);
    }
    write_partials(acc, B.partials, ctl);
}

// ---------------------------------------------------------------------------
// k_pde_march (3D): k_mu and k_trial fused, so mu never makes an HBM round trip
// and every o, cr, phi, gb value is fetched from HBM once. A block owns a 64 (z)
// x TY (y) column tile and marches along X over XS planes. LDS holds
//   o, cr      ring of 3 planes with a two-cell Y/Z halo,
//   phi        ring of 3 planes, gb ring of 2 planes, one-cell halo,
//   mu, M      2 planes with a one-cell halo, evaluated once per cell,
// and the own cell's values of plane x-1 sit in registers. Iteration x:
//   A  global loads of o, cr (plane x+3), phi, gb (plane x+2) into registers
//   B  mu, M of plane x+1 over the tile and its halo
//   -- barrier --
//   C  trial of plane x (needs only mu, M, phi of the neighbours)
//   D  the registers of A into the ring slots that B and C no longer read
//   -- barrier --
// so the HBM latency of A is hidden behind B and C. The ring offsets of each
// thread are fixed for the block and computed once. The binary64 arithmetic is
// that of k_mu + k_trial (mu64_vals, mob_vals, trial_mm), so both give bitwise
// equal results.
// ---------------------------------------------------------------------------
template <int PREC, int TY>
struct PdeLds {
    static constexpr int NTH = 64 * TY;
    static constexpr int W2 = 68, H2 = TY + 4, W1 = 66, H1 = TY + 2;
    static constexpr int P2 = H2 * W2, P1 = H1 * W1;
    static constexpr int R2 = (P2 + NTH - 1) / NTH, R1 = (P1 + NTH - 1) / NTH;   // ring entries per thread
    using T = MuT<PREC>;
    static constexpr size_t bytes() {
        return size_t(2) * P1 * sizeof(MuMob<T>) + size_t(3) * P2 * 8 * 2 + size_t(3) * P1 * 8
             + size_t(2) * P1 * sizeof(T);
    }
};

template <int PREC, int TY, class TE, class TM>
__global__ __launch_bounds__(64 * TY) void k_pde_march(Par P, Bufs<TE, TM> B, int XS) {
    using L = PdeLds<PREC, TY>;
    using T = typename L::T;
    using MM = MuMob<T>;
    constexpr int NTH = L::NTH, P1 = L::P1, P2 = L::P2, W1 = L::W1, W2 = L::W2, R1 = L::R1, R2 = L::R2;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    MM* smm = reinterpret_cast<MM*>(lds_raw);                  // [2][P1] mu, M
    double* so = reinterpret_cast<double*>(smm + 2 * P1);      // [3][P2] o
    double* sc = so + 3 * P2;                                  // [3][P2] cr
    double* sp = sc + 3 * P2;                                  // [3][P1] phi
    T* sg = reinterpret_cast<T*>(sp + 3 * P1);                 // [2][P1] gb
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur = ctl->cur, nxt = cur ^ 1;
    const int tid = threadIdx.x;
    const int tz = tid & 63, ty = tid >> 6;
    const int nTz = (P.nz + 63) >> 6, nTy = (P.ny + TY - 1) / TY, nSeg = (P.nx + XS - 1) / XS;
    const int b = xcd_remap(blockIdx.x, nTz * nTy * nSeg);
    const int seg = b / (nTz * nTy);
    const int rem = b - seg * (nTz * nTy);
    const int tyT = rem / nTz, tzT = rem - tyT * nTz;
    const int z0 = tzT * 64, y0 = tyT * TY;
    const int y = y0 + ty, z = z0 + tz;
    const bool active = y < P.ny && z < P.nz;
    const int x0 = seg * XS, x1 = min(P.nx, x0 + XS);
    const long long plane = P.nyz;
    const double* o = B.o[cur];
    const double* cr = B.cr[cur];
    const double* p = B.p[cur];
    const int me = P.eta_evolves ? cur : 0;
    const TE* gb = B.gb[me];
    const TE* s2 = B.s2[me];
    const bool enf = P.enforce_layer != 0;
    auto wrapy = [&](int v) { return ((v % P.ny) + P.ny) % P.ny; };
    auto wrapz = [&](int v) { return ((v % P.nz) + P.nz) % P.nz; };
    auto clampx = [&](int v) { return v < 0 ? 0 : (v >= P.nx ? P.nx - 1 : v); };

    // Ring entries of this thread: entry r is cell tid + r*NTH of the halo-2
    // (o, cr) or halo-1 (phi, gb, mu, M) plane tile. Global in-plane offsets with
    // the Y/Z wrap applied, and for halo-1 entries the index into halo-2 planes.
    int off2[R2], off1[R1], q2of[R1];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int q = tid + r * NTH, ly = q / W2, lz = q - ly * W2;
        off2[r] = q < P2 ? wrapy(y0 - 2 + ly) * P.nz + wrapz(z0 - 2 + lz) : 0;
    }
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        const int q = tid + r * NTH, ly = q / W1, lz = q - ly * W1;
        off1[r] = q < P1 ? wrapy(y0 - 1 + ly) * P.nz + wrapz(z0 - 1 + lz) : 0;
        q2of[r] = (ly + 1) * W2 + (lz + 1);
    }
    auto in2 = [&](int r) { return (r + 1) * NTH <= P2 || tid + r * NTH < P2; };
    auto in1 = [&](int r) { return (r + 1) * NTH <= P1 || tid + r * NTH < P1; };

    // staging registers: o, cr of one plane, phi, gb of one plane
    double ro[R2], rc[R2], rp[R1];
    T rg[R1];
    auto fetch_oc = [&](int X) {   // X clamped: the X ghosts are copies
        const int xc = clampx(X);
        const double* bo = o + (long long)xc * plane;
        const double* bc = cr + (long long)xc * plane;
        const bool sub = enf && xc == 0;
#pragma unroll
        for (int r = 0; r < R2; ++r)
            if (in2(r)) {
                ro[r] = sub ? P.layer_o : bo[off2[r]];
                rc[r] = bc[off2[r]];
            }
    };
    auto store_oc = [&](int slot) {
#pragma unroll
        for (int r = 0; r < R2; ++r)
            if (in2(r)) {
                so[slot * P2 + tid + r * NTH] = ro[r];
                sc[slot * P2 + tid + r * NTH] = rc[r];
            }
    };
    auto fetch_pg = [&](int X) {
        const int xc = clampx(X);
        const double* bp = p + (long long)xc * plane;
        const TE* bg = gb + (long long)xc * plane;
#pragma unroll
        for (int r = 0; r < R1; ++r)
            if (in1(r)) {
                rp[r] = bp[off1[r]];
                rg[r] = (T)bg[off1[r]];
            }
    };
    auto store_pg = [&](int ps, int gs) {
#pragma unroll
        for (int r = 0; r < R1; ++r)
            if (in1(r)) {
                sp[ps * P1 + tid + r * NTH] = rp[r];
                sg[gs * P1 + tid + r * NTH] = rg[r];
            }
    };
    // mu, M of one plane over the tile and its one-cell halo. Ring slots: o, cr
    // of the plane (sx) and of its X neighbours (sm, sq), phi (ps), gb (gs);
    // output mu, M slot ms.
    auto mm_plane = [&](int sm, int sx, int sq, int ps, int gs, int ms) {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            if (!in1(r)) continue;
            const int q = tid + r * NTH, q2 = q2of[r];
            double o7[7], c7[7];
            o7[SC] = so[sx * P2 + q2];
            c7[SC] = sc[sx * P2 + q2];
            o7[SXM] = so[sm * P2 + q2];
            c7[SXM] = sc[sm * P2 + q2];
            o7[SXP] = so[sq * P2 + q2];
            c7[SXP] = sc[sq * P2 + q2];
            o7[SYM] = so[sx * P2 + q2 - W2];
            c7[SYM] = sc[sx * P2 + q2 - W2];
            o7[SYP] = so[sx * P2 + q2 + W2];
            c7[SYP] = sc[sx * P2 + q2 + W2];
            o7[SZM] = so[sx * P2 + q2 - 1];
            c7[SZM] = sc[sx * P2 + q2 - 1];
            o7[SZP] = so[sx * P2 + q2 + 1];
            c7[SZP] = sc[sx * P2 + q2 + 1];
            smm[ms * P1 + q] = mu_mob<3, PREC, T>(P, o7, c7, sp[ps * P1 + q], (double)sg[gs * P1 + q]);
        }
    };

    // Prologue. o, cr of plane x0-1+k sit in ring slot k % 3; phi of plane x0+k
    // in slot k % 3; gb and mu, M of plane x0+k in slot k & 1.
    fetch_oc(x0 - 1);
    store_oc(0);
    fetch_oc(x0);
    store_oc(1);
    fetch_oc(x0 + 1);
    store_oc(2);
    fetch_pg(x0);
    store_pg(0, 0);
    fetch_pg(x0 + 1);
    store_pg(1, 1);
    __syncthreads();
    fetch_oc(x0 + 2);
    mm_plane(0, 1, 2, 0, 0, 0);
    // own cell of plane x0-1 (for x0 == 0 the X ghost, a copy of plane 0, mu
    // included: 2D:427) straight from global memory
    MM m_prev;
    double p_prev;
    {
        const int yq = y < P.ny ? y : y % P.ny, zq = z < P.nz ? z : z % P.nz;
        const Nbr n = neighbours<3>(P, x0 > 0 ? x0 - 1 : 0, yq, zq);
        double o7[7], c7[7];
        gather_oc<3>(P, o, cr, n, enf, o7, c7);
        p_prev = p[n.c];
        m_prev = mu_mob<3, PREC, T>(P, o7, c7, p_prev, (double)gb[n.c]);
    }
    __syncthreads();
    store_oc(0);   // plane x0+2 over plane x0-1
    __syncthreads();
    const int own1 = (ty + 1) * W1 + (tz + 1);
    const int own2 = (ty + 2) * W2 + (tz + 2);
    const long long yz = (long long)(y < P.ny ? y : 0) * P.nz + (z < P.nz ? z : 0);
    int so_x = 1, so_x1 = 2, so_x2 = 0;   // o, cr slots of planes x, x+1, x+2
    int sp_x = 0, sp_x1 = 1, sp_x2 = 2;   // phi slots of planes x, x+1 and of the next plane loaded
    StepAcc acc;
    // S of the own cell, one plane ahead: its load is issued before the plane
    // prefetch, so using it never waits for the prefetch (loads retire in order)
    const bool sload = active && P.has_eta;
    TE s_next = sload ? s2[(long long)x0 * plane + yz] : (TE)0;
    for (int x = x0; x < x1; ++x) {
        const int m0 = (x - x0) & 1, m1 = m0 ^ 1;   // gb, mu/M slots of planes x (and x+2), x+1
        const double S = (double)s_next;
        if (sload && x + 1 < x1) s_next = s2[(long long)(x + 1) * plane + yz];
        // A: planes x+3 (o, cr) and x+2 (phi, gb) into registers
        fetch_oc(x + 3);
        fetch_pg(x + 2);
        const long long c = (long long)x * plane + yz;
        // B: mu, M of plane x+1; o, cr of the own cell of plane x
        mm_plane(so_x, so_x1, so_x2, sp_x1, m1, m1);
        const double oo = so[so_x * P2 + own2], cc = sc[so_x * P2 + own2];
        double m64O = 0.;
        if constexpr (PREC != PREC_FP64) {
            if (x == 0 && P.bc == 1 && P.oxy) {
                // binary64 mu_O for the inlet face; the x-1 ghost of plane 0 is a copy
                double o7[7], c7[7], m64C;
                o7[SC] = o7[SXM] = oo;
                c7[SC] = c7[SXM] = cc;
                o7[SXP] = so[so_x1 * P2 + own2];
                c7[SXP] = sc[so_x1 * P2 + own2];
                o7[SYM] = so[so_x * P2 + own2 - W2];
                c7[SYM] = sc[so_x * P2 + own2 - W2];
                o7[SYP] = so[so_x * P2 + own2 + W2];
                c7[SYP] = sc[so_x * P2 + own2 + W2];
                o7[SZM] = so[so_x * P2 + own2 - 1];
                c7[SZM] = sc[so_x * P2 + own2 - 1];
                o7[SZP] = so[so_x * P2 + own2 + 1];
                c7[SZP] = sc[so_x * P2 + own2 + 1];
                mu64_vals<3>(P, o7, c7, sp[sp_x * P1 + own1], m64O, m64C);
            }
        }
        __syncthreads();
        // C: trial of the own cell of plane x
        MM m[7];
        double p7[7];
        m[SC] = smm[m0 * P1 + own1];
        m[SYM] = smm[m0 * P1 + own1 - W1];
        m[SYP] = smm[m0 * P1 + own1 + W1];
        m[SZM] = smm[m0 * P1 + own1 - 1];
        m[SZP] = smm[m0 * P1 + own1 + 1];
        m[SXP] = smm[m1 * P1 + own1];
        if (x + 1 == P.nx) {
            // the x+1 ghost of the last plane is a copy of it, mu included
            m[SXP].mo = m[SC].mo;
            m[SXP].mc = m[SC].mc;
        }
        m[SXM] = m_prev;
        p7[SC] = sp[sp_x * P1 + own1];
        p7[SYM] = sp[sp_x * P1 + own1 - W1];
        p7[SYP] = sp[sp_x * P1 + own1 + W1];
        p7[SZM] = sp[sp_x * P1 + own1 - 1];
        p7[SZP] = sp[sp_x * P1 + own1 + 1];
        p7[SXP] = sp[sp_x1 * P1 + own1];
        p7[SXM] = p_prev;
        // the cell's results are stored after the staging below: the staging
        // waits for this plane's prefetch loads, and stores issued before it
        // would be waited for as well
        float ecoef_v = 0.f;
        if constexpr (PREC != PREC_FP64) ecoef_v = eta_coef32(P, (float)S, (float)p7[SC]);
        double onv = 0., crv = 0., pnv = 0.;
        if (active) {
            const double g_c = x == 0 ? (double)gb[c] : 0.;
            const double o_stored = x == 0 ? o[c] : oo;
            double left_flux;
            trial_mm<3, PREC, T>(P, m, p7, oo, cc, x, S, g_c, m64O, onv, crv, pnv, left_flux);
            finish_cell_vals<PREC>(P, c, x, o_stored, oo, cc, p7[SC], onv, crv, pnv, left_flux, acc, P.layer_o);
        }
        m_prev = m[SC];
        p_prev = p7[SC];
        // D: staged planes into the slots of planes x (o, cr; gb) and x-1 (phi)
        store_oc(so_x);
        store_pg(sp_x2, m0);
        if (active) {
            st_nt(&B.o[nxt][c], onv);
            st_nt(&B.cr[nxt][c], crv);
            st_nt(&B.p[nxt][c], pnv);
            if constexpr (PREC != PREC_FP64) {
                if (B.ecoef) B.ecoef[c] = ecoef_v;
            }
        }
        __syncthreads();
        const int t = so_x;
        so_x = so_x1;
        so_x1 = so_x2;
        so_x2 = t;
        const int u = sp_x;
        sp_x = sp_x1;
        sp_x1 = sp_x2;
        sp_x2 = u;
    }
    write_partials(acc, B.partials, ctl);
}

// ---------------------------------------------------------------------------
// k_commit: all-or-nothing acceptance (collective_error + swap, 2D:594-596;
// budget restore on rejection, SL:812-820). One block, after every other
// kernel of the step.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void commit_body(const double* partials, DevCtl* ctl, int nparts) {
    const unsigned long long pk = __hip_atomic_load(&ctl->pde_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long ek = __hip_atomic_load(&ctl->eta_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (pk != NO_KEY || ek != NO_KEY) {
        if (threadIdx.x == 0) {
            ctl->failed = 1;
            ctl->fail_kind = pk != NO_KEY ? 1 : 2;
        }
        return;
    }
    __shared__ DD sh[6 * 16];
    DD acc[6];
    for (int s = 0; s < 6; ++s) acc[s] = DD{0., 0.};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        const double* w = partials + (long long)b * NPART;
        for (int s = 0; s < 6; ++s) acc[s] = dd_add(acc[s], DD{w[2 * s], w[2 * s + 1]});
    }
    const DD r = block_reduce_dd6(acc, sh);
    // partial slots: bcO, ckO, ckCr, clipO, clipCr, clipped
    // budget slots:  nucO0 nucCr1 ckO2 ckCr3 clipO4 clipCr5 bcO6 bcCr7
    // thread s < 5 adds sum s into its budget slot, thread 5 the clipped count
    const int s = threadIdx.x;
    if (s < 5) {
        const int k = s == 0 ? 6 : s + 1;
        DD cur{ctl->budget[2 * k], ctl->budget[2 * k + 1]};
        cur = dd_add(cur, r);
        ctl->budget[2 * k] = cur.hi;
        ctl->budget[2 * k + 1] = cur.lo;
    } else if (s == 5) {
        ctl->clipped += (long long)r.hi;
    }
    if (s == 0) {
        ctl->cur ^= 1;
        ctl->steps_done += 1;
    }
}

__global__ __launch_bounds__(1024) void k_commit(Bufs<double, double> B, int nparts) {
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    commit_body(B.partials, ctl, nparts);
}

// ---------------------------------------------------------------------------
// k_eta: eta trial for every grain (2D:347-359, SL:514-526), then the next
// step's mask from the trial values (same running sums as 2D:332).
// ---------------------------------------------------------------------------
template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_eta(Par P, Bufs<TE, TM> B) {
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur = ctl->cur, nxt = cur ^ 1;
    unsigned long long bad = NO_KEY;
    const int G = gridDim.x;
    for (int vb = blockIdx.x; vb < P.ntiles; vb += G) {
        const int t = xcd_remap(vb, P.ntiles);
        int x, y, z;
        if (!tile_cell(P, t, x, y, z)) continue;
        const Nbr n = neighbours<DIM>(P, x, y, z);
        const long long c = n.c;
        const double pp = B.p[cur][c];
        const TE* ein = B.eta[cur];
        TE* eout = B.eta[nxt];
        if constexpr (PREC == PREC_FP64) {
            const double S = (double)B.s2[cur][c];
            double S2 = 0., pr = 0.;
            const double dtl = P.dt * P.etaL;
            for (int g = 0; g < P.n_eta; ++g) {
                const long long off = (long long)g * P.fstride;
                const TE* e = ein + off;
                const double ec = (double)e[c];
                const double lp = lap64<DIM>((const double*)e, n, ec, P, (double)e[n.xm], (double)e[n.xp]);
                const double dF = ec * (P.etaW * (S - (1. - h64(pp)) + 2. * (S - ec * ec)) + P.etaA * h64(pp));
                const double v = ec - dtl * (dF - P.etaK * lp);
                if (!isfinite(v) || v < -1.e-12 || v > 1. + 1.e-12) {
                    const unsigned long long key = (unsigned long long)g * (unsigned long long)P.ncell + (unsigned long long)c;
                    bad = bad < key ? bad : key;
                }
                eout[off + c] = (TE)v;
                const double q = v * v;
                pr += S2 * q;
                S2 += q;
            }
            const double s16 = 16. * pr;
            B.s2[nxt][c] = (TE)S2;
            B.pairs[nxt][c] = (TE)pr;
            B.gb[nxt][c] = (TE)(P.eta_scale * (s16 < 1. ? s16 : 1.));
        } else {
#pragma clang fp contract(fast)
            const float S = (float)B.s2[cur][c];
            const float hphi = h32((float)pp);
            const float base = (float)P.etaW * (S - (1.f - hphi)) + 2.f * (float)P.etaW * S;
            const float A = (float)P.etaA * hphi;
            const float dtl = (float)(P.dt * P.etaL);
            const float K = (float)P.etaK;
            float S2 = 0.f, pr = 0.f;
            for (int g = 0; g < P.n_eta; ++g) {
                const long long off = (long long)g * P.fstride;
                const TE* e = ein + off;
                const float ec = (float)e[c];
                float lp = ((float)e[n.xp] - 2.f * ec + (float)e[n.xm]) * P.fdx2i
                         + ((float)e[n.yp] - 2.f * ec + (float)e[n.ym]) * P.fdy2i;
                if (DIM == 3) lp += ((float)e[n.zp] - 2.f * ec + (float)e[n.zm]) * P.fdz2i;
                const float dF = ec * (base - 2.f * (float)P.etaW * ec * ec + A);
                const float v = ec - dtl * (dF - K * lp);
                if (!isfinite(v) || v < -1.e-12f || v > 1.f + 1.e-12f) {
                    const unsigned long long key = (unsigned long long)g * (unsigned long long)P.ncell + (unsigned long long)c;
                    bad = bad < key ? bad : key;
                }
                eout[off + c] = (TE)v;
                const float q = v * v;
                pr += S2 * q;
                S2 += q;
            }
            const float s16 = 16.f * pr;
            B.s2[nxt][c] = (TE)S2;
            B.pairs[nxt][c] = (TE)pr;
            B.gb[nxt][c] = (TE)((float)P.eta_scale * (s16 < 1.f ? s16 : 1.f));
        }
    }
    if (bad != NO_KEY) atomicMin(&ctl->eta_key, bad);
}

// Trial value of field g at one cell for the X-marching eta kernels
// (eta_trial, 2D:347-359, SL:514-526): the Y/Z neighbours sit in LDS (one plane
// tile of PL values per field; o*: offsets of the cell's neighbours in it), the
// X neighbours and the cell itself in registers. binary64: the operation order
// of k_eta (and of the reference); mixed: binary32 with the cell's force factor.
// Per-cell inputs of the eta trial besides eta itself: phi and S of the
// step's input state (binary64) or the cell's force factor (mixed), and the
// running mask sums S, P of the fields of earlier passes. Loaded ahead of the
// plane's barrier so that their latency overlaps the LDS staging.
template <int PREC, class TE>
struct EtaCell {
    using V = typename std::conditional<PREC == PREC_FP64, double, float>::type;
    double pp, S, dtl;          // binary64
    float coef, fdtl, K, W2;    // binary32
    V S2, pr;                   // running sums entering this pass
};
// Per-cell arrays of the eta kernels are addressed by the cell's byte offset
// cb = cell * sizeof(TE) as a 32-bit value (the launcher guarantees that the
// fields of one pass, and so every cell offset, stay below 2^32 bytes): a
// scalar base plus one 32-bit lane offset instead of a 64-bit address per array.
// These are the cell's own inputs (phi, S, the force factor, the running sums),
// read by this block only: non-temporal loads, so that they do not displace
// the eta rows, which the neighbouring blocks read again as their halo (about
// 1 % off the binary64 pass, 0.5 % off the binary32 pass, same box).
template <class T>
__device__ __forceinline__ T ld_off(const T* base, unsigned ob) {
    return __builtin_nontemporal_load(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + ob));
}
template <class T>
__device__ __forceinline__ void st_nt_off(T* base, unsigned ob, T v) {
    st_nt(reinterpret_cast<T*>(reinterpret_cast<char*>(base) + ob), v);
}
template <int PREC, class TE, class TM>
__device__ __forceinline__ EtaCell<PREC, TE> eta_cell_coefs(const Par& P, const Bufs<TE, TM>& B, unsigned cb,
                                                            int cur_i, int nxt_i, bool first_pass) {
    using V = typename EtaCell<PREC, TE>::V;
    EtaCell<PREC, TE> e;
    if constexpr (PREC == PREC_FP64) {
        static_assert(sizeof(TE) == sizeof(double), "binary64 eta: phi and the mask share the cell offset");
        e.pp = ld_off(B.p[cur_i], cb);
        e.S = (double)ld_off(B.s2[cur_i], cb);
        e.dtl = P.dt * P.etaL;
    } else {
        static_assert(sizeof(TE) == sizeof(float), "mixed eta: the force factor shares the cell offset");
        if (B.ecoef) {
            e.coef = ld_off(B.ecoef, cb);
        } else {
            const long long c = (long long)(cb / sizeof(TE));
            e.coef = eta_coef32(P, (float)B.s2[cur_i][c], (float)B.p[cur_i][c]);
        }
        e.fdtl = (float)(P.dt * P.etaL);
        e.K = (float)P.etaK;
        e.W2 = 2.f * (float)P.etaW;
    }
    e.S2 = first_pass ? (V)0 : (V)ld_off(B.s2[nxt_i], cb);
    e.pr = first_pass ? (V)0 : (V)ld_off(B.pairs[nxt_i], cb);
    return e;
}
template <int PREC, class TE>
__device__ __forceinline__ typename std::conditional<PREC == PREC_FP64, double, float>::type
eta_value(const Par& P, const EtaCell<PREC, TE>& e, const TE* L, int oym, int oyp, int ozm, int ozp, TE pv, TE cc,
          TE nv) {
    if constexpr (PREC == PREC_FP64) {
        const double ec = (double)cc;
        const double eyp = (double)L[oyp];
        const double eym = (double)L[oym];
        const double ezp = (double)L[ozp];
        const double ezm = (double)L[ozm];
        const double xy = ((double)nv - 2. * ec + (double)pv) / (P.dx * P.dx) + (eyp - 2. * ec + eym) / (P.dy * P.dy);
        const double lp = xy + (ezp - 2. * ec + ezm) / (P.dz * P.dz);
        const double dF = ec * (P.etaW * (e.S - (1. - h64(e.pp)) + 2. * (e.S - ec * ec)) + P.etaA * h64(e.pp));
        return ec - e.dtl * (dF - P.etaK * lp);
    } else {
        // binary32 with every rounding spelled out (no contraction left to the
        // compiler), so that every kernel that evaluates a cell gives the same
        // bits: the row and the 64-wide marching kernels are bitwise equal
        const float ec = (float)cc;
        const float ax = __fadd_rn(__fmaf_rn(-2.f, ec, (float)nv), (float)pv);
        const float ay = __fadd_rn(__fmaf_rn(-2.f, ec, (float)L[oyp]), (float)L[oym]);
        const float az = __fadd_rn(__fmaf_rn(-2.f, ec, (float)L[ozp]), (float)L[ozm]);
        const float lp = __fmaf_rn(az, P.fdz2i, __fmaf_rn(ay, P.fdy2i, __fmul_rn(ax, P.fdx2i)));
        const float dF = __fmul_rn(ec, __fmaf_rn(-ec, __fmul_rn(e.W2, ec), e.coef));
        return __fmaf_rn(-e.fdtl, __fmaf_rn(-lp, e.K, dF), ec);
    }
}

// eta trial of all NG fields of one cell and the next-step mask sums (2D:332).
// The range test (2D:351: reject a non-finite value or one outside
// [-1e-12, 1 + 1e-12]) is one comparison pair per field; a cell that fails it
// recomputes its values (same inputs, same code) to find its first failing
// field for the report, so the common path carries no key arithmetic.
template <int PREC, int NG, class TE, class TM>
__device__ __forceinline__ void eta_cell_update(const Par& P, const Bufs<TE, TM>& B, const EtaCell<PREC, TE>& e,
                                                const TE* lds, int PL, int oym, int oyp, int ozm, int ozp,
                                                const TE* prev, const TE* cv, const TE* nxv, unsigned cbo, int nxt_i,
                                                bool last_pass, long long key0, __amdgpu_buffer_rsrc_t rout,
                                                unsigned fs, unsigned long long& bad, bool active) {
    using T = TE;
    using V = typename EtaCell<PREC, TE>::V;
    constexpr V lo = (V)-1.e-12, hi = (V)1. + (V)1.e-12;
    // Every lane computes; a lane past the mesh edge of a partial tile (whose
    // values are not the trial) stores through an offset beyond the buffer
    // resource's range, which the hardware discards, so the common path has no
    // divergent branch around the field loop.
    const unsigned cb = active ? cbo : 0xFFFFFFF0u;
    const long long c = (long long)(cbo / sizeof(T));   // the cell, for the report key
    V S2 = e.S2, pr = e.pr;
    bool ok = true;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const V v = eta_value<PREC>(P, e, lds + g * PL, oym, oyp, ozm, ozp, prev[g], cv[g], nxv[g]);
        if constexpr (PREC == PREC_FP64) {
            // binary64 keeps the key per field: the recomputation of the slow
            // path would hold more registers than the comparison costs
            if (active && !(v >= lo && v <= hi)) {
                const unsigned long long key = (unsigned long long)(key0 + (long long)g * P.ncell + c);
                bad = bad < key ? bad : key;
            }
        } else {
            ok = ok & (v >= lo) & (v <= hi);   // false for NaN as well
        }
        buf_store<T>(rout, cb, (unsigned)g * fs, (TE)v);
        if constexpr (PREC == PREC_FP64) {
            const double q = v * v;
            pr += S2 * q;
            S2 += q;
        } else {
            const float q = __fmul_rn(v, v);
            pr = __fmaf_rn(S2, q, pr);
            S2 = __fadd_rn(S2, q);
        }
        // keep the scheduler from hoisting the LDS reads of all fields at
        // once (register pressure): groups of four fields
        if ((g & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    // Materialise the running sums before the slow path: otherwise the compiler
    // sinks their chain below it, which keeps all NG trial values live across
    // the slow path and pushes the kernel past its register budget.
    if constexpr (PREC != PREC_FP64) asm volatile("" : "+v"(S2), "+v"(pr));
    if (PREC != PREC_FP64 && active && !ok) {
        for (int g = 0; g < NG; ++g) {
            const V v = eta_value<PREC>(P, e, lds + g * PL, oym, oyp, ozm, ozp, prev[g], cv[g], nxv[g]);
            if (!(v >= lo && v <= hi)) {
                const unsigned long long key = (unsigned long long)(key0 + (long long)g * P.ncell + c);
                bad = bad < key ? bad : key;
                break;
            }
        }
    }
    // The mask sums: S, then the next pass's running sum P or, after the last
    // pass, gb. Stored through buffer resources, lanes past the mesh edge out
    // of range as above, so that no branch separates these stores from the
    // next plane's loads (a branch there makes the compiler's wait for those
    // loads wait for the stores as well).
    const int nb = (int)((unsigned)P.ncell * (unsigned)sizeof(TE));
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(B.s2[nxt_i], 0, nb, BUF_DWORD3);
    const __amdgpu_buffer_rsrc_t rq =
        __builtin_amdgcn_make_buffer_rsrc(last_pass ? B.gb[nxt_i] : B.pairs[nxt_i], 0, nb, BUF_DWORD3);
    const V s16 = (V)16 * pr;
    const V q = last_pass ? (V)P.eta_scale * (s16 < (V)1 ? s16 : (V)1) : pr;
    buf_store<TE>(rs, cb, 0u, (TE)S2);
    buf_store<TE>(rq, cb, 0u, (TE)q);
}

// ---------------------------------------------------------------------------
// k_eta_march: the same eta trial, restructured so each eta value crosses HBM
// once. A block owns a 64 (z) x TY (y) column tile of all NG eta fields and
// marches along X over a segment of XS planes. Planes x-1, x, x+1 of every
// field sit in registers (X neighbours), plane x with a one-cell Y/Z halo sits
// in LDS (Y/Z neighbours). Arithmetic is identical to k_eta (same operation
// order in binary64), so both kernels give bitwise equal results.
// ---------------------------------------------------------------------------
// Fields are processed in passes of NG (g0 = first field of this pass) so that
// the plane working set of one XCD fits its L2; the running sums S and P of the
// next-step mask are carried between passes in the output mask buffers, which
// keeps the accumulation order, and so the result, identical to one pass.
template <int PREC, int NG, int TY, class TE, class TM>
__global__ __launch_bounds__(64 * TY) void k_eta_march(Par P, Bufs<TE, TM> B, int XS, int g0) {
    using T = TE;
    constexpr int W = 66;                 // LDS row: 64 interior + 2 halo
    constexpr int PL = (TY + 2) * W;      // one field's plane tile
    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur_i = ctl->cur, nxt_i = cur_i ^ 1;
    const int tz = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int nTz = (P.nz + 63) >> 6, nTy = (P.ny + TY - 1) / TY, nSeg = (P.nx + XS - 1) / XS;
    const int b = xcd_remap(blockIdx.x, nTz * nTy * nSeg);
    const int seg = b / (nTz * nTy);
    const int rem = b - seg * (nTz * nTy);
    const int tyT = rem / nTz, tzT = rem - tyT * nTz;
    const int z0 = tzT * 64, y0 = tyT * TY;
    const int z = z0 + tz, y = y0 + ty;
    const bool active = z < P.nz && y < P.ny;
    const int x0 = seg * XS, x1 = min(P.nx, x0 + XS);
    const long long plane = P.nyz;
    // Lanes past the mesh edge of a partial tile hold the periodically wrapped
    // cell, so the LDS tile row/column after the last active one is exactly
    // the Y/Z periodic neighbour; they compute and store nothing.
    const long long rc = (long long)(y % P.ny) * P.nz + (z % P.nz);
    // Every pass holds exactly NG fields (the launcher splits n_eta so), which
    // keeps the field loops free of guards and fully unrolled.
    const bool first_pass = g0 == 0, last_pass = g0 + NG >= P.n_eta;
    const long long key0 = (long long)g0 * P.ncell;
    // Buffer loads and stores over the NG fields of this pass (the launcher
    // guarantees NG * fstride * sizeof(TE) < 2^32): the field offset goes in the
    // scalar offset and the cell offset in one lane register shared by all
    // fields, so no 64-bit address is kept per field.
    const unsigned fs = (unsigned)P.fstride * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<TE*>(B.eta[cur_i] + (long long)g0 * P.fstride), 0, (int)(fs * (unsigned)NG), BUF_DWORD3);
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(
        B.eta[nxt_i] + (long long)g0 * P.fstride, 0, (int)(fs * (unsigned)NG), BUF_DWORD3);
    auto cell_bytes = [](long long cell) { return (unsigned)cell * (unsigned)sizeof(T); };

    T prev[NG], cv[NG], nx_[NG];
    {
        const int xp = x0 == 0 ? 0 : x0 - 1;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            prev[g] = buf_load<T>(rin, cell_bytes(xp * plane + rc), (unsigned)g * fs);
            cv[g] = buf_load<T>(rin, cell_bytes((long long)x0 * plane + rc), (unsigned)g * fs);
        }
    }
    // halo cells of the tile: row y0-1, row y0+TY (64 each), columns z0-1, z0+64 (TY each)
    const int nhalo = 128 + 2 * TY;
    unsigned long long bad = NO_KEY;
    for (int x = x0; x < x1; ++x) {
        const long long xo = (long long)x * plane;
        const long long xn = (long long)(x == P.nx - 1 ? x : x + 1) * plane;
        // issue the next plane's loads early (register prefetch)
#pragma unroll
        for (int g = 0; g < NG; ++g) nx_[g] = buf_load<T>(rin, cell_bytes(xn + rc), (unsigned)g * fs);
        // interior of plane x into LDS
#pragma unroll
        for (int g = 0; g < NG; ++g) lds[g * PL + (ty + 1) * W + tz + 1] = cv[g];
        // halo of plane x: work items (halo cell, chunk of 8 fields); each item
        // issues 8 independent loads so that many are in flight per lane
        constexpr int GC = NG < 8 ? NG : 8;
        constexpr int nchunk_g = (NG + GC - 1) / GC;
        for (int qq = threadIdx.x; qq < nhalo * nchunk_g; qq += 64 * TY) {
            const int gk = qq / nhalo;
            const int hh = qq - gk * nhalo;
            int ly, lz, gy, gz;
            if (hh < 128) {
                const int side = hh >> 6;   // 0: row y0-1, 1: row y0+TY
                lz = (hh & 63) + 1;
                ly = side ? TY + 1 : 0;
                gy = side ? y0 + TY : y0 - 1;
                gz = z0 + (hh & 63);
            } else {
                const int q = hh - 128;
                const int side = q / TY;    // 0: column z0-1, 1: column z0+64
                ly = q - side * TY + 1;
                lz = side ? 65 : 0;
                gy = y0 + ly - 1;
                gz = side ? z0 + 64 : z0 - 1;
            }
            // clip to the mesh: a tile row/column beyond the mesh edge is never read
            // by an active lane except through the periodic wrap below
            if (gy >= P.ny && gy != y0 + TY) continue;
            if (gz >= P.nz && gz != z0 + 64) continue;
            gy = ((gy % P.ny) + P.ny) % P.ny;
            gz = ((gz % P.nz) + P.nz) % P.nz;
            const long long hc = xo + (long long)gy * P.nz + gz;
            T hv[GC];
#pragma unroll
            for (int u = 0; u < GC; ++u) {
                const int g = gk * GC + u;
                hv[u] = buf_load<T>(rin, cell_bytes(hc) + (unsigned)g * fs, 0u);
            }
#pragma unroll
            for (int u = 0; u < GC; ++u) {
                const int g = gk * GC + u;
                lds[g * PL + ly * W + lz] = hv[u];
            }
        }
        const EtaCell<PREC, TE> ec = eta_cell_coefs<PREC>(P, B, cell_bytes(xo + rc), cur_i, nxt_i, first_pass);
        __syncthreads();
        eta_cell_update<PREC, NG>(P, B, ec, lds, PL, ty * W + tz + 1, (ty + 2) * W + tz + 1, (ty + 1) * W + tz,
                                  (ty + 1) * W + tz + 2, prev, cv, nx_, cell_bytes(xo + rc), nxt_i, last_pass, key0,
                                  rout, fs, bad, active);
        __syncthreads();
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            prev[g] = cv[g];
            cv[g] = nx_[g];
        }
    }
    if (bad != NO_KEY) atomicMin(&ctl->eta_key, bad);
}

// ---------------------------------------------------------------------------
// k_eta_rows: the X-marching eta kernel for meshes whose Z rows fit one block
// (nz = NZ in 64, 128, 256; TY * NZ = 512). A block owns TY whole Z rows, so the
// Z neighbours wrap inside its LDS tile and, besides the interior, only the two
// Y halo rows are loaded, as whole contiguous rows. With 64-wide tiles every Z
// halo cell is a separate cache-line request: as many requests as the interior,
// most of them missing L2. Same arithmetic as k_eta_march (eta_cell_update).
// ---------------------------------------------------------------------------
// Waves per SIMD the register allocation must allow: the binary32 pass of 16
// fields is held to 128 VGPRs, so two 512-thread blocks share a CU (measured
// faster than one block with the unconstrained 139 VGPRs); others unconstrained.
template <int PREC, int NG>
constexpr int eta_rows_waves() { return PREC == PREC_MIXED && NG == 16 ? 4 : 1; }
template <int PREC, int NG, int NZ, class TE, class TM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(eta_rows_waves<PREC, NG>()))) void k_eta_rows(
    Par P, Bufs<TE, TM> B, int XS, int g0) {
    using T = TE;
    constexpr int TY = 512 / NZ;
    static_assert((2 * NG) % TY == 0, "halo rows x fields must split evenly over the thread rows");
    constexpr int HC = 2 * NG / TY;          // halo loads per thread: (halo row, field) pairs
    extern __shared__ unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur_i = ctl->cur, nxt_i = cur_i ^ 1;
    constexpr int nz = NZ;
    constexpr int PL = (TY + 2) * nz;        // per field: halo row y0-1, TY rows, halo row y0+TY
    const int tid = threadIdx.x;
    const int ty = tid / nz, z = tid - ty * nz;
    const int nTy = (P.ny + TY - 1) / TY, nSeg = (P.nx + XS - 1) / XS;
    const int b = xcd_remap(blockIdx.x, nTy * nSeg);
    const int seg = b / nTy, tyT = b - seg * nTy;
    const int y0 = tyT * TY, y = y0 + ty;
    const bool active = y < P.ny;
    const int x0 = seg * XS, x1 = min(P.nx, x0 + XS);
    const long long plane = P.nyz;
    // rows past the mesh edge of a partial tile hold the periodically wrapped
    // row, so the tile row after the last active one is its Y neighbour
    const long long rc = (long long)(y % P.ny) * nz + z;
    const bool first_pass = g0 == 0, last_pass = g0 + NG >= P.n_eta;
    const long long key0 = (long long)g0 * P.ncell;
    // buffer addressing as in k_eta_march
    const unsigned fs = (unsigned)P.fstride * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rin = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<TE*>(B.eta[cur_i] + (long long)g0 * P.fstride), 0, (int)(fs * (unsigned)NG), BUF_DWORD3);
    const __amdgpu_buffer_rsrc_t rout = __builtin_amdgcn_make_buffer_rsrc(
        B.eta[nxt_i] + (long long)g0 * P.fstride, 0, (int)(fs * (unsigned)NG), BUF_DWORD3);
    auto cell_bytes = [](long long cell) { return (unsigned)cell * (unsigned)sizeof(T); };
    // Cell byte offsets are 32-bit (below 2^32, see above): the own cell's
    // offset within a plane, and a plane's offset, uniform over the block.
    const unsigned rcb = cell_bytes(rc);
    auto plane_bytes = [&](int X) { return (unsigned)__builtin_amdgcn_readfirstlane((int)cell_bytes((long long)X * plane)); };
    // halo work of this thread: pairs ty + k*TY = (halo row: bit 0, field: the
    // rest). TY is even, so the row is ty & 1 for every k (one byte offset and
    // one LDS row per thread) and the field is (ty >> 1) + k * TY / 2.
    static_assert(TY % 2 == 0, "the halo row of a thread must not depend on k");
    const long long hrow = (ty & 1) ? (long long)((y0 + TY) % P.ny) * nz + z
                                    : (long long)(((y0 - 1) % P.ny + P.ny) % P.ny) * nz + z;
    const unsigned hrow_b = cell_bytes(hrow);
    const int hlds = ((ty & 1) ? (TY + 1) * nz : 0) + z;
    // LDS offsets of the cell's Y/Z neighbours within a field tile
    const int own = (ty + 1) * nz + z;
    const int oym = own - nz, oyp = own + nz;
    const int ozm = (ty + 1) * nz + (z == 0 ? nz - 1 : z - 1);
    const int ozp = (ty + 1) * nz + (z == nz - 1 ? 0 : z + 1);

    T prev[NG], cv[NG], nx_[NG];
    {
        const int xp = x0 == 0 ? 0 : x0 - 1;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            prev[g] = buf_load<T>(rin, plane_bytes(xp) + rcb, (unsigned)g * fs);
            cv[g] = buf_load<T>(rin, plane_bytes(x0) + rcb, (unsigned)g * fs);
        }
    }
    // Halo rows of plane X (a wave lies inside one row, as nz is a multiple of
    // 64: ty, and so the field of each (row, field) pair, is uniform over the
    // wave). In binary32 they are loaded one plane ahead, together with the
    // next plane's interior: the Y-neighbour blocks load the same rows as their
    // interior at that moment, so the request finds the line in L2 (or merges
    // with the neighbour's miss) instead of fetching it from HBM a plane later.
    // Binary64 passes of up to 8 fields do the same (about 114 VGPRs, still two
    // resident blocks per CU; 2 % faster than loading them in their plane).
    constexpr bool AHEAD = PREC == PREC_MIXED || NG <= 8;
    T hv[HC];
    auto load_halo = [&](int X) {
        const unsigned vo = plane_bytes(X) + hrow_b;
#pragma unroll
        for (int k = 0; k < HC; ++k) {
            const int g = __builtin_amdgcn_readfirstlane((ty >> 1) + k * (TY / 2));
            hv[k] = buf_load<T>(rin, vo, (unsigned)g * fs);
        }
    };
    if (AHEAD) load_halo(x0);
    // The prologue's loads are needed by the first plane anyway: waiting for
    // them here keeps them out of the compiler's count at the loop head, which
    // would otherwise make every plane wait for its own stores there.
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    unsigned long long bad = NO_KEY;
    for (int x = x0; x < x1; ++x) {
        const unsigned cb = plane_bytes(x) + rcb;   // the own cell of plane x
        const int xq = x == P.nx - 1 ? x : x + 1;
        const unsigned nb = plane_bytes(xq) + rcb;  // ... of plane xq
        // binary32: the cell's own inputs of plane x, first of this iteration's loads
        EtaCell<PREC, TE> ec;
        if (AHEAD) ec = eta_cell_coefs<PREC>(P, B, cb, cur_i, nxt_i, first_pass);
        if (!AHEAD) {
            load_halo(x);
#pragma unroll
            for (int g = 0; g < NG; ++g) nx_[g] = buf_load<T>(rin, nb, (unsigned)g * fs);
        }
        // plane x (interior and halo) into LDS, then the next plane's interior
        // and halo rows into the freed registers
#pragma unroll
        for (int g = 0; g < NG; ++g) lds[g * PL + own] = cv[g];
#pragma unroll
        for (int k = 0; k < HC; ++k) lds[((ty >> 1) + k * (TY / 2)) * PL + hlds] = hv[k];
        if (AHEAD) {
#pragma unroll
            for (int g = 0; g < NG; ++g) nx_[g] = buf_load<T>(rin, nb, (unsigned)g * fs);
            // also after the segment's last plane (plane xq exists; the rows
            // are cached and unused): a branch around these loads would make
            // the waits for the loads above wait for them as well
            load_halo(xq);
        }
        __syncthreads();
        if (!AHEAD) ec = eta_cell_coefs<PREC>(P, B, cb, cur_i, nxt_i, first_pass);
        eta_cell_update<PREC, NG>(P, B, ec, lds, PL, oym, oyp, ozm, ozp, prev, cv, nx_, cb, nxt_i, last_pass,
                                  key0, rout, fs, bad, active);
        __syncthreads();
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            prev[g] = cv[g];
            cv[g] = nx_[g];
        }
    }
    if (bad != NO_KEY) atomicMin(&ctl->eta_key, bad);
}

// ---------------------------------------------------------------------------
// Output-time refresh (Model::refresh before vtk(), 2D:600): binary64 mu and M
// at every interior cell, no inlet substitution.
// ---------------------------------------------------------------------------
template <int DIM, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_refresh_out(Par P, Bufs<TE, TM> B, double* muO, double* muCr,
                                                      double* mo, double* mc) {
    int x, y, z;
    const int t = xcd_remap(blockIdx.x, P.ntiles);
    if (!tile_cell(P, t, x, y, z)) return;
    const int cur = B.ctl->cur;
    const int me = P.eta_evolves ? cur : 0;
    const Nbr n = neighbours<DIM>(P, x, y, z);
    double a, b;
    mu64_at<DIM>(P, B.o[cur], B.cr[cur], B.p[cur], n, false, a, b);
    muO[n.c] = a;
    muCr[n.c] = b;
    const double g = (double)B.gb[me][n.c];
    mo[n.c] = mobO64(P, B.o[cur][n.c], B.p[cur][n.c], g);
    mc[n.c] = mobCr64(P, B.cr[cur][n.c], B.p[cur][n.c], g);
}

// Concentration-BC reset left in place by a rejected 2D step (2D:552 runs
// enforce_legacy_layer before the trial and nothing restores it).
__global__ void k_apply_layer(Par P, double* const* obuf, const DevCtl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.nyz) obuf[ctl->cur][i] = P.layer_o;
}

// ---------------------------------------------------------------------------
// Field statistics: double-double sum, min, max (Model::sum 2D:229-232 and
// log() extrema 2D:702-705). mode 0: value, 1: h(value), 2: 1 - a - b (Ni).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_stats(const T* a, const double* b2, long long n, int mode, double* part) {
    DD s{0., 0.};
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double v = (double)a[i];
        if (mode == 1) v = h64(v);
        else if (mode == 2) v = 1. - v - b2[i];
        s = dd_add(s, v);
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    __shared__ DD sh[4];
    __shared__ double smn[256], smx[256];
    const DD r = block_reduce_dd(s, sh);
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            smn[threadIdx.x] = smn[threadIdx.x + off] < smn[threadIdx.x] ? smn[threadIdx.x + off] : smn[threadIdx.x];
            smx[threadIdx.x] = smx[threadIdx.x + off] > smx[threadIdx.x] ? smx[threadIdx.x + off] : smx[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = r.hi;
        part[blockIdx.x * 4 + 1] = r.lo;
        part[blockIdx.x * 4 + 2] = smn[0];
        part[blockIdx.x * 4 + 3] = smx[0];
    }
}

}  // namespace nicr
