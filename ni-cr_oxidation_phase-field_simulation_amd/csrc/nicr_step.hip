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
template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_mask(Par P, Bufs<TE, TM> B, int which) {
    int x, y, z;
    const int t = xcd_remap(blockIdx.x, P.ntiles);
    if (!tile_cell(P, t, x, y, z)) return;
    const int e = which < 0 ? B.ctl->cur : which;
    const long long c = (long long)x * P.nyz + (long long)y * P.nz + z;
    const TE* eta = B.eta[e];
    if constexpr (PREC == PREC_FP64) {
        double S = 0., pr = 0.;
        for (int g = 0; g < P.n_eta; ++g) {
            const double v = (double)eta[(long long)g * P.ncell + c];
            const double q = v * v;
            pr += S * q;
            S += q;
        }
        const double s16 = 16. * pr;
        B.s2[e][c] = (TE)S;
        B.pairs[e][c] = (TE)pr;
        B.gb[e][c] = (TE)(P.eta_scale * (s16 < 1. ? s16 : 1.));
    } else {
        float S = 0.f, pr = 0.f;
        for (int g = 0; g < P.n_eta; ++g) {
            const float v = (float)eta[(long long)g * P.ncell + c];
            const float q = v * v;
            pr += S * q;
            S += q;
        }
        const float s16 = 16.f * pr;
        B.s2[e][c] = (TE)S;
        B.pairs[e][c] = (TE)pr;
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

template <int DIM>
__device__ __forceinline__ void mu64_at(const Par& P, const double* o, const double* cr, const double* p,
                                         const Nbr& n, bool enforce, double& muO, double& muCr) {
    const bool sub = enforce;
    const double oc = (sub && n.x == 0) ? Phys::layerO : o[n.c];
    const double oxm = (sub && n.x <= 1) ? Phys::layerO : o[n.xm];
    const double oxp = (sub && n.x == 0 && P.nx == 1) ? Phys::layerO : o[n.xp];
    const double cc = cr[n.c];
    double hh = h64(p[n.c]);
    // y/z neighbours share plane x; they are substituted iff the centre is.
    double lo;
    {
        const double oyp = (sub && n.x == 0) ? Phys::layerO : o[n.yp];
        const double oym = (sub && n.x == 0) ? Phys::layerO : o[n.ym];
        double xy = (oxp - 2. * oc + oxm) / (P.dx * P.dx) + (oyp - 2. * oc + oym) / (P.dy * P.dy);
        if (DIM == 3) {
            const double ozp = (sub && n.x == 0) ? Phys::layerO : o[n.zp];
            const double ozm = (sub && n.x == 0) ? Phys::layerO : o[n.zm];
            xy = xy + (ozp - 2. * oc + ozm) / (P.dz * P.dz);
        }
        lo = xy;
    }
    muO = (1. - hh) * mu_metal_o64(oc, cc)
        + hh * (24000000. * oc - 13800000. * cc - 8880000.) / Phys::Vm
        - 2. * Phys::ko * lo;
    const double lc = lap64<DIM>(cr, n, cc, P, cr[n.xm], cr[n.xp]);
    muCr = (1. - hh) * (1166528.60789006 * cc + 3200000. * oc - 97045.64) / Phys::Vm
         + hh * (14000000. * cc - 13800000. * oc + 2680000.) / Phys::Vm
         - 2. * Phys::kc * lc;
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
    const double* o = B.o[cur];
    const double* cr = B.cr[cur];
    const double* p = B.p[cur];
    const bool enf = P.enforce_layer != 0;
    if constexpr (PREC == PREC_FP64) {
        double muO, muCr;
        mu64_at<DIM>(P, o, cr, p, n, enf, muO, muCr);
        B.mu_o[n.c] = (TM)muO;
        B.mu_cr[n.c] = (TM)muCr;
    } else {
#pragma clang fp contract(fast)
        const double oc = (enf && x == 0) ? Phys::layerO : o[n.c];
        const double oxm = (enf && x <= 1) ? Phys::layerO : o[n.xm];
        const double oxp = o[n.xp];
        const double onb_y = (enf && x == 0) ? Phys::layerO : o[n.yp];
        const double onb_ym = (enf && x == 0) ? Phys::layerO : o[n.ym];
        const double cc = cr[n.c];
        double lo2 = (oxp - 2. * oc + oxm);
        double lo2y = (onb_y - 2. * oc + onb_ym);
        double lc2 = (cr[n.xp] - 2. * cc + cr[n.xm]);
        double lc2y = (cr[n.yp] - 2. * cc + cr[n.ym]);
        float lapo = (float)lo2 * P.fdx2i + (float)lo2y * P.fdy2i;
        float lapc = (float)lc2 * P.fdx2i + (float)lc2y * P.fdy2i;
        if (DIM == 3) {
            const double oz = (enf && x == 0) ? Phys::layerO : o[n.zp];
            const double ozm = (enf && x == 0) ? Phys::layerO : o[n.zm];
            lapo += (float)(oz - 2. * oc + ozm) * P.fdz2i;
            lapc += (float)(cr[n.zp] - 2. * cc + cr[n.zm]) * P.fdz2i;
        }
        const float hh = h32((float)p[n.c]);
        const float fo = (float)oc, fc = (float)cc;
        // mu_O + 298432 and mu_Cr + 97045.64, expanded so the metal constant cancels.
        const float muO = (1.f - hh) * (12000000.f * fo + 3200000.f * fc)
                        + hh * (24000000.f * fo - 13800000.f * fc - 8581568.f)
                        - float(2. * Phys::ko) * lapo;
        const float muCr = (1.f - hh) * (1166528.60789006f * fc + 3200000.f * fo)
                         + hh * (14000000.f * fc - 13800000.f * fo + 2777045.64f)
                         - float(2. * Phys::kc) * lapc;
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

// ---------------------------------------------------------------------------
// k_trial: mobilities + conservative six-face update of O and Cr, Allen-Cahn
// phi update, legacy Ck, finite check and bounds (2D:430-433,435-450,553-591;
// SL:616-638,771-809). Persistent grid over tiles; one budget partial per block.
// ---------------------------------------------------------------------------
template <int DIM, int PREC, class TE, class TM>
__global__ __launch_bounds__(TILE) void k_trial(Par P, Bufs<TE, TM> B) {
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    const int cur = ctl->cur, nxt = cur ^ 1;
    const double* o = B.o[cur];
    const double* cr = B.cr[cur];
    const double* p = B.p[cur];
    double* on = B.o[nxt];
    double* crn = B.cr[nxt];
    double* pn = B.p[nxt];
    const int me = P.eta_evolves ? cur : 0;
    const TE* gb = B.gb[me];
    const TE* s2 = B.s2[me];
    const TM* muO = B.mu_o;
    const TM* muCr = B.mu_cr;
    const bool enf = P.enforce_layer != 0;

    DD a_bc{0., 0.}, a_ckO{0., 0.}, a_ckCr{0., 0.}, a_clO{0., 0.}, a_clCr{0., 0.};
    double n_clip = 0.;
    unsigned long long bad = NO_KEY;

    const int G = gridDim.x;
    for (int vb = blockIdx.x; vb < ((P.ntiles + G - 1) / G) * G; vb += G) {
        if (vb >= P.ntiles) break;
        const int t = xcd_remap(vb, P.ntiles);
        int x, y, z;
        if (!tile_cell(P, t, x, y, z)) continue;
        const Nbr n = neighbours<DIM>(P, x, y, z);
        const long long c = n.c;
        const double o_stored = o[c];
        const double oo = (enf && x == 0) ? Phys::layerO : o_stored;
        const double cc = cr[c], pp = p[c];
        double onv, crv, pnv;
        double left_flux = 0.;
        if constexpr (PREC == PREC_FP64) {
            // o with the inlet substitution, at the 7 stencil points
            auto O = [&](long long q, int qx) { return (enf && qx == 0) ? Phys::layerO : o[q]; };
            const int xm = x == 0 ? 0 : x - 1, xp = x == P.nx - 1 ? x : x + 1;
            const double o_xm = O(n.xm, xm), o_xp = O(n.xp, xp), o_ym = O(n.ym, x), o_yp = O(n.yp, x);
            const double o_zm = DIM == 3 ? O(n.zm, x) : 0., o_zp = DIM == 3 ? O(n.zp, x) : 0.;
            const double g_c = (double)gb[c];
            // mobilities (refresh loop 2, 2D:430-433): ghosts are copies, so M(ghost) = M(copy)
            const double moc = mobO64(P, oo, pp, g_c);
            const double mcc = mobCr64(P, cc, pp, g_c);
            const double mo_xm = mobO64(P, o_xm, p[n.xm], (double)gb[n.xm]);
            const double mo_xp = mobO64(P, o_xp, p[n.xp], (double)gb[n.xp]);
            const double mo_ym = mobO64(P, o_ym, p[n.ym], (double)gb[n.ym]);
            const double mo_yp = mobO64(P, o_yp, p[n.yp], (double)gb[n.yp]);
            const double mc_xm = mobCr64(P, cr[n.xm], p[n.xm], (double)gb[n.xm]);
            const double mc_xp = mobCr64(P, cr[n.xp], p[n.xp], (double)gb[n.xp]);
            const double mc_ym = mobCr64(P, cr[n.ym], p[n.ym], (double)gb[n.ym]);
            const double mc_yp = mobCr64(P, cr[n.yp], p[n.yp], (double)gb[n.yp]);
            const double uo = (double)muO[c], uc = (double)muCr[c];
            // div_flux (2D:443-450, SL:629-638)
            double divO = 0.;
            if (P.oxy) {
                const double right = -.5 * (moc + mo_xp) * ((double)muO[n.xp] - uo) / P.dx;
                double left = -.5 * (mo_xm + moc) * (uo - (double)muO[n.xm]) / P.dx;
                if (x == 0 && P.bc == 1) {
                    // left_flux (2D:435-442): positive face-mobility closure
                    const double mf = mobO64(P, P.c_ref, pp, g_c);
                    left_flux = 2. * mf * (P.mu_res - uo) / P.dx;
                    left = left_flux;
                }
                const double top = -.5 * (moc + mo_yp) * ((double)muO[n.yp] - uo) / P.dy;
                const double bottom = -.5 * (mo_ym + moc) * (uo - (double)muO[n.ym]) / P.dy;
                double d = (left - right) / P.dx + (bottom - top) / P.dy;
                if (DIM == 3) {
                    const double mo_zm = mobO64(P, o_zm, p[n.zm], (double)gb[n.zm]);
                    const double mo_zp = mobO64(P, o_zp, p[n.zp], (double)gb[n.zp]);
                    const double front = -.5 * (moc + mo_zp) * ((double)muO[n.zp] - uo) / P.dz;
                    const double back = -.5 * (mo_zm + moc) * (uo - (double)muO[n.zm]) / P.dz;
                    d = d + (back - front) / P.dz;
                }
                divO = d;
            }
            double divC;
            {
                const double right = -.5 * (mcc + mc_xp) * ((double)muCr[n.xp] - uc) / P.dx;
                const double left = -.5 * (mc_xm + mcc) * (uc - (double)muCr[n.xm]) / P.dx;
                const double top = -.5 * (mcc + mc_yp) * ((double)muCr[n.yp] - uc) / P.dy;
                const double bottom = -.5 * (mc_ym + mcc) * (uc - (double)muCr[n.ym]) / P.dy;
                double d = (left - right) / P.dx + (bottom - top) / P.dy;
                if (DIM == 3) {
                    const double mc_zm = mobCr64(P, cr[n.zm], p[n.zm], (double)gb[n.zm]);
                    const double mc_zp = mobCr64(P, cr[n.zp], p[n.zp], (double)gb[n.zp]);
                    const double front = -.5 * (mcc + mc_zp) * ((double)muCr[n.zp] - uc) / P.dz;
                    const double back = -.5 * (mc_zm + mcc) * (uc - (double)muCr[n.zm]) / P.dz;
                    d = d + (back - front) / P.dz;
                }
                divC = d;
            }
            onv = P.oxy ? oo + P.dt * divO : oo;
            crv = cc + P.dt * divC;
            const double hp = 30. * pp * pp * (pp - 1.) * (pp - 1.);
            const double lp = lap64<DIM>(p, n, pp, P, p[n.xm], p[n.xp]);
            double fp = Phys::W * (4. * pp * pp * pp - 6. * pp * pp + 2. * pp)
                      + hp * (foxide64(oo, cc) - fmetal64(oo, cc)) - 2. * Phys::kp * lp;
            if (P.has_eta) {
                const double S = (double)s2[c];
                const double hpp = 30. * pp * pp * (pp - 1.) * (pp - 1.);
                fp += 0.5 * hpp * (P.etaW * (S - (1. - h64(pp))) + P.etaA * S);
            }
            pnv = pp - P.dt * Phys::L * fp;
        } else {
#pragma clang fp contract(fast)
            const int xm = x == 0 ? 0 : x - 1;
            auto O = [&](long long q, int qx) { return (enf && qx == 0) ? Phys::layerO : o[q]; };
            const float hc = h32((float)pp);
            const float g_c = (float)gb[c];
            const float moc = P.oxy ? mobO32(P, (float)oo, hc, g_c) : 0.f;
            const float mcc = mobCr32(P, (float)cc, hc, g_c);
            const float uo = (float)muO[c], uc = (float)muCr[c];
            float sumO = 0.f, sumC = 0.f;   // sum over faces of M_face * (mu_nb - mu_c)
            auto face = [&](long long q, int qx, float inv2) {
                const float hq = h32((float)p[q]);
                const float gq = (float)gb[q];
                if (P.oxy) {
                    const float mq = mobO32(P, (float)O(q, qx), hq, gq);
                    sumO += 0.5f * (moc + mq) * ((float)muO[q] - uo) * inv2;
                }
                const float mq2 = mobCr32(P, (float)cr[q], hq, gq);
                sumC += 0.5f * (mcc + mq2) * ((float)muCr[q] - uc) * inv2;
            };
            face(n.xp, x == P.nx - 1 ? x : x + 1, P.fdx2i);
            if (!(x == 0 && P.bc == 1)) {
                face(n.xm, xm, P.fdx2i);
            } else {
                // left face replaced by the binary64 inlet flux (2D:435-442,446)
                const float hq = h32((float)p[n.xm]);
                const float mq2 = mobCr32(P, (float)cr[n.xm], hq, (float)gb[n.xm]);
                sumC += 0.5f * (mcc + mq2) * ((float)muCr[n.xm] - uc) * P.fdx2i;
                if (P.oxy) {
                    double m64O, m64C;
                    mu64_at<DIM>(P, o, cr, p, n, enf, m64O, m64C);
                    const double mf = mobO64(P, P.c_ref, pp, (double)g_c);
                    left_flux = 2. * mf * (P.mu_res - m64O) / P.dx;
                    sumO += (float)(left_flux / P.dx);
                }
            }
            face(n.yp, x, P.fdy2i);
            face(n.ym, x, P.fdy2i);
            if (DIM == 3) {
                face(n.zp, x, P.fdz2i);
                face(n.zm, x, P.fdz2i);
            }
            onv = P.oxy ? oo + P.dt * (double)sumO : oo;
            crv = cc + P.dt * (double)sumC;
            const float fpp = (float)pp;
            const float hp = 30.f * fpp * fpp * (fpp - 1.f) * (fpp - 1.f);
            const float lp = (float)(p[n.xp] - 2. * pp + p[n.xm]) * P.fdx2i
                           + (float)(p[n.yp] - 2. * pp + p[n.ym]) * P.fdy2i
                           + (DIM == 3 ? (float)(p[n.zp] - 2. * pp + p[n.zm]) * P.fdz2i : 0.f);
            const float fo = (float)oo, fcr = (float)cc;
            // foxide - fmetal, expanded (2D:174-181)
            const float dfx = 6000000.f * fo * fo + (7000000.f - 583264.303945031f) * fcr * fcr
                            - 17000000.f * fo * fcr - 8581568.f * fo + 2777045.64f * fcr
                            + float(1860990. + 61732.2974223055);
            float fp = float(Phys::W) * (4.f * fpp * fpp * fpp - 6.f * fpp * fpp + 2.f * fpp)
                     + hp * dfx - float(2. * Phys::kp) * lp;
            if (P.has_eta) {
                const float S = (float)s2[c];
                fp += 0.5f * hp * ((float)P.etaW * (S - (1.f - hc)) + (float)P.etaA * S);
            }
            pnv = pp - P.dt * Phys::L * (double)fp;
        }
        // legacy Ck (2D:567-571), binary64 in both precisions
        if (P.ck_legacy) {
            double dO = -Phys::ckO * ((Phys::oxideO - onv) * h64(pnv) - (Phys::oxideO - oo) * h64(pp)) / P.dt;
            double dCr = -Phys::ckCr * ((Phys::oxideCr - crv) * h64(pnv) - (Phys::oxideCr - cc) * h64(pp)) / P.dt;
            onv += dO;
            crv += dCr;
            a_ckO = dd_add(a_ckO, dO);
            a_ckCr = dd_add(a_ckCr, dCr);
        }
        if (x == 0) {
            if (enf) a_bc = dd_add(a_bc, Phys::layerO - o_stored);
            if (P.bc == 1 && P.oxy) a_bc = dd_add(a_bc, P.dt * left_flux / P.dx);
        }
        if (!isfinite(onv) || !isfinite(crv) || !isfinite(pnv)) {
            bad = bad < (unsigned long long)c ? bad : (unsigned long long)c;
        } else if (P.bounds_legacy) {
            // std::clamp then phi snap (2D:574-580)
            const double oc = onv < 0. ? 0. : (1. < onv ? 1. : onv);
            const double ccr = crv < 0. ? 0. : (1. < crv ? 1. : crv);
            a_clO = dd_add(a_clO, oc - onv);
            a_clCr = dd_add(a_clCr, ccr - crv);
            if (oc != onv || ccr != crv) n_clip += 1.;
            onv = oc;
            crv = ccr;
            if (pnv > .99999) pnv = 1.;
            if (pnv < 1.e-5) pnv = 0.;
        } else if (onv < -1.e-12 || crv < -1.e-12 || onv + crv > 1. + 1.e-12 ||
                   pnv < -1.e-12 || pnv > 1. + 1.e-12) {
            bad = bad < (unsigned long long)c ? bad : (unsigned long long)c;
        }
        on[c] = onv;
        crn[c] = crv;
        pn[c] = pnv;
    }
    if (bad != NO_KEY) atomicMin(&ctl->pde_key, bad);
    __shared__ DD sh[TILE / 64];
    const DD r0 = block_reduce_dd(a_bc, sh);
    const DD r1 = block_reduce_dd(a_ckO, sh);
    const DD r2 = block_reduce_dd(a_ckCr, sh);
    const DD r3 = block_reduce_dd(a_clO, sh);
    const DD r4 = block_reduce_dd(a_clCr, sh);
    const DD r5 = block_reduce_dd(DD{n_clip, 0.}, sh);
    if (threadIdx.x == 0) {
        double* w = B.partials + (long long)blockIdx.x * NPART;
        w[0] = r0.hi; w[1] = r0.lo; w[2] = r1.hi; w[3] = r1.lo; w[4] = r2.hi; w[5] = r2.lo;
        w[6] = r3.hi; w[7] = r3.lo; w[8] = r4.hi; w[9] = r4.lo; w[10] = r5.hi; w[11] = 0.;
    }
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
                const long long off = (long long)g * P.ncell;
                const TE* e = ein + off;
                const double ec = (double)e[c];
                const double lp = lap64<DIM>((const double*)e, n, ec, P, (double)e[n.xm], (double)e[n.xp]);
                const double dF = ec * (P.etaW * (S - (1. - h64(pp)) + 2. * (S - ec * ec)) + P.etaA * h64(pp));
                const double v = ec - dtl * (dF - P.etaK * lp);
                if (!isfinite(v) || v < -1.e-12 || v > 1. + 1.e-12) {
                    const unsigned long long key = (unsigned long long)off + (unsigned long long)c;
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
                const long long off = (long long)g * P.ncell;
                const TE* e = ein + off;
                const float ec = (float)e[c];
                float lp = ((float)e[n.xp] - 2.f * ec + (float)e[n.xm]) * P.fdx2i
                         + ((float)e[n.yp] - 2.f * ec + (float)e[n.ym]) * P.fdy2i;
                if (DIM == 3) lp += ((float)e[n.zp] - 2.f * ec + (float)e[n.zm]) * P.fdz2i;
                const float dF = ec * (base - 2.f * (float)P.etaW * ec * ec + A);
                const float v = ec - dtl * (dF - K * lp);
                if (!isfinite(v) || v < -1.e-12f || v > 1.f + 1.e-12f) {
                    const unsigned long long key = (unsigned long long)off + (unsigned long long)c;
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
    const TE* ein = B.eta[cur_i] + (long long)g0 * P.ncell;
    TE* eout = B.eta[nxt_i] + (long long)g0 * P.ncell;
    const int ng = min(NG, P.n_eta - g0);
    const bool first_pass = g0 == 0, last_pass = g0 + NG >= P.n_eta;
    const long long key0 = (long long)g0 * P.ncell;

    T prev[NG], cv[NG], nx_[NG];
    {
        const int xp = x0 == 0 ? 0 : x0 - 1;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (g < ng) {
                prev[g] = ein[(long long)g * P.ncell + xp * plane + rc];
                cv[g] = ein[(long long)g * P.ncell + (long long)x0 * plane + rc];
            } else {
                prev[g] = T(0);
                cv[g] = T(0);
            }
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
        for (int g = 0; g < NG; ++g)
            nx_[g] = g < ng ? ein[(long long)g * P.ncell + xn + rc] : T(0);
        // interior of plane x into LDS
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g < ng) lds[g * PL + (ty + 1) * W + tz + 1] = cv[g];
        // halo of plane x: work items (halo cell, chunk of 8 fields); each item
        // issues 8 independent loads so that many are in flight per lane
        constexpr int GC = NG < 8 ? NG : 8;
        const int nchunk_g = (ng + GC - 1) / GC;
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
                hv[u] = g < ng ? ein[(long long)g * P.ncell + hc] : T(0);
            }
#pragma unroll
            for (int u = 0; u < GC; ++u) {
                const int g = gk * GC + u;
                if (g < ng) lds[g * PL + ly * W + lz] = hv[u];
            }
        }
        __syncthreads();
        if (active) {
            const int lym = ty;
            const int lzp = tz + 2, lzm = tz;
            const long long c = xo + rc;
            const double pp = B.p[cur_i][c];
            if constexpr (PREC == PREC_FP64) {
                const double S = (double)B.s2[cur_i][c];
                double S2 = first_pass ? 0. : (double)B.s2[nxt_i][c];
                double pr = first_pass ? 0. : (double)B.pairs[nxt_i][c];
                const double dtl = P.dt * P.etaL;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    if (g >= ng) break;
                    const T* L = lds + g * PL;
                    const double ec = (double)cv[g];
                    const double eyp = (double)L[(ty + 2) * W + tz + 1];
                    const double eym = (double)L[lym * W + tz + 1];
                    const double ezp = (double)L[(ty + 1) * W + lzp];
                    const double ezm = (double)L[(ty + 1) * W + lzm];
                    const double xy = ((double)nx_[g] - 2. * ec + (double)prev[g]) / (P.dx * P.dx)
                                    + (eyp - 2. * ec + eym) / (P.dy * P.dy);
                    const double lp = xy + (ezp - 2. * ec + ezm) / (P.dz * P.dz);
                    const double dF = ec * (P.etaW * (S - (1. - h64(pp)) + 2. * (S - ec * ec)) + P.etaA * h64(pp));
                    const double v = ec - dtl * (dF - P.etaK * lp);
                    if (!isfinite(v) || v < -1.e-12 || v > 1. + 1.e-12) {
                        const unsigned long long key = (unsigned long long)(key0 + (long long)g * P.ncell + c);
                        bad = bad < key ? bad : key;
                    }
                    eout[(long long)g * P.ncell + c] = (TE)v;
                    const double q = v * v;
                    pr += S2 * q;
                    S2 += q;
                }
                B.s2[nxt_i][c] = (TE)S2;
                B.pairs[nxt_i][c] = (TE)pr;
                if (last_pass) {
                    const double s16 = 16. * pr;
                    B.gb[nxt_i][c] = (TE)(P.eta_scale * (s16 < 1. ? s16 : 1.));
                }
            } else {
#pragma clang fp contract(fast)
                const float S = (float)B.s2[cur_i][c];
                const float hphi = h32((float)pp);
                const float base = (float)P.etaW * (S - (1.f - hphi)) + 2.f * (float)P.etaW * S;
                const float A = (float)P.etaA * hphi;
                const float dtl = (float)(P.dt * P.etaL);
                const float K = (float)P.etaK;
                const float W2 = 2.f * (float)P.etaW;
                float S2 = first_pass ? 0.f : (float)B.s2[nxt_i][c];
                float pr = first_pass ? 0.f : (float)B.pairs[nxt_i][c];
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    if (g >= ng) break;
                    const T* L = lds + g * PL;
                    const float ec = (float)cv[g];
                    const float eyp = (float)L[(ty + 2) * W + tz + 1];
                    const float eym = (float)L[lym * W + tz + 1];
                    const float ezp = (float)L[(ty + 1) * W + lzp];
                    const float ezm = (float)L[(ty + 1) * W + lzm];
                    const float lp = ((float)nx_[g] - 2.f * ec + (float)prev[g]) * P.fdx2i
                                   + (eyp - 2.f * ec + eym) * P.fdy2i + (ezp - 2.f * ec + ezm) * P.fdz2i;
                    const float dF = ec * (base - W2 * ec * ec + A);
                    const float v = ec - dtl * (dF - K * lp);
                    if (!isfinite(v) || v < -1.e-12f || v > 1.f + 1.e-12f) {
                        const unsigned long long key = (unsigned long long)(key0 + (long long)g * P.ncell + c);
                        bad = bad < key ? bad : key;
                    }
                    eout[(long long)g * P.ncell + c] = (TE)v;
                    const float q = v * v;
                    pr += S2 * q;
                    S2 += q;
                }
                B.s2[nxt_i][c] = (TE)S2;
                B.pairs[nxt_i][c] = (TE)pr;
                if (last_pass) {
                    const float s16 = 16.f * pr;
                    B.gb[nxt_i][c] = (TE)((float)P.eta_scale * (s16 < 1.f ? s16 : 1.f));
                }
            }
        }
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
// k_commit: all-or-nothing acceptance (collective_error + swap, 2D:594-596;
// budget restore on rejection, SL:812-820). One block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_commit(Bufs<double, double> B, int nparts) {
    DevCtl* ctl = B.ctl;
    if (ctl->failed) return;
    if (ctl->pde_key != NO_KEY || ctl->eta_key != NO_KEY) {
        if (threadIdx.x == 0) {
            ctl->failed = 1;
            ctl->fail_kind = ctl->pde_key != NO_KEY ? 1 : 2;
        }
        return;
    }
    __shared__ DD sh[16];
    DD acc[6];
    for (int s = 0; s < 6; ++s) acc[s] = DD{0., 0.};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        const double* w = B.partials + (long long)b * NPART;
        for (int s = 0; s < 6; ++s) acc[s] = dd_add(acc[s], DD{w[2 * s], w[2 * s + 1]});
    }
    DD r[6];
    for (int s = 0; s < 6; ++s) r[s] = block_reduce_dd(acc[s], sh);
    if (threadIdx.x == 0) {
        // partial slots: bcO, ckO, ckCr, clipO, clipCr, clipped
        // budget slots:  nucO0 nucCr1 ckO2 ckCr3 clipO4 clipCr5 bcO6 bcCr7
        const int map[5] = {6, 2, 3, 4, 5};
        for (int s = 0; s < 5; ++s) {
            DD cur{ctl->budget[2 * map[s]], ctl->budget[2 * map[s] + 1]};
            cur = dd_add(cur, r[s]);
            ctl->budget[2 * map[s]] = cur.hi;
            ctl->budget[2 * map[s] + 1] = cur.lo;
        }
        ctl->clipped += (long long)r[5].hi;
        ctl->cur ^= 1;
        ctl->steps_done += 1;
    }
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
    if (i < P.nyz) obuf[ctl->cur][i] = Phys::layerO;
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
