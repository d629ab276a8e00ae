// nicr_nn.hip -- learned surrogate: a 4-level 3D U-Net (BF16 weights and
// activations, FP32 accumulation) trained on the GPU by its own forward and
// backward kernels on trajectories of the plain step, and surrogate jumps that
// the binary64 step accepts or rejects. Included by nicr_api.hip (one library).
// No counterpart in the reference.
//
// Network (channels per level 32, 64, 128, 256; 3x3x3 convolutions, ReLU; X
// zero padded, Y/Z periodic):
//   x0 (O, Cr, phi, gb normalised; 4 of 32 channels)
//   e1 = conv(conv(x0))          p1 = pool(e1)       level 0: 32
//   e2 = conv(conv(p1))          p2 = pool(e2)       level 1: 64
//   e3 = conv(conv(p2))          p3 = pool(e3)       level 2: 128
//   b4 = conv(conv(p3))                              level 3: 256
//   d3 = conv(conv([up(b4), e3]))                    level 2: 384 -> 128
//   d2 = conv(conv([up(d3), e2]))                    level 1: 192 -> 64
//   d1 = conv(conv([up(d2), e1]))                    level 0: 96 -> 32
//   out = head(d1)   FP32 1x1 projection onto the 3 normalised 50-step deltas
// Parameters are one FP32 array: per convolution W[27][cout][cin] then b[cout],
// in the order above (e1a e1b e2a e2b e3a e3b b4a b4b d3a d3b d2a d2b d1a d1b),
// then the head W[3][32] and b[3].
#include "../../include/nicr_nn.h"
#include "nicr_nn_kernels.hpp"

#include <algorithm>
#include <cstdlib>
#include <random>

namespace nicr {
namespace nn {

constexpr int N_CONV = 14;
constexpr int HEAD_IN = 32;
constexpr int N_FIELD = 4;            // O, Cr, phi, mask
constexpr int N_OUT = 3;              // changes of O, Cr, phi
constexpr int N_IN = N_FIELD + N_OUT; // stored per volume: the fields and the plain update of O, Cr, phi
constexpr int N_STAT = 7;             // mean O, std O, mean Cr, std Cr, rms update O, Cr, phi

// ---- launch checks ----
// Every kernel launch of the network is checked on the host before it is
// enqueued: a grid of 1 .. 2^31 - 1 blocks in X and 1 in Y/Z, a block of 1 ..
// the kernel's own maximum threads (its launch bounds), and static plus
// dynamic LDS within the 160 KiB of a CU. A launch that fails a check is not
// enqueued; the first failure is kept (per host thread) and the entry point
// that made the launch returns it as NICR_E_INVALID.
constexpr size_t LDS_PER_CU = 160 * 1024;
struct LaunchCheck {
    bool failed = false;
    std::string msg;
};
thread_local LaunchCheck g_launch;

struct KernelLimits {
    const void* fn;
    int max_threads;
    size_t static_lds;
};

bool launch_ok(const void* fn, dim3 grid, dim3 block, size_t dyn_lds, const char* name) {
    static thread_local std::vector<KernelLimits> cache;
    const KernelLimits* kl = nullptr;
    for (const auto& k : cache)
        if (k.fn == fn) kl = &k;
    if (!kl) {
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, fn) != hipSuccess) {
            if (!g_launch.failed) {
                g_launch.failed = true;
                g_launch.msg = std::string("launch check: no attributes for ") + name;
            }
            return false;
        }
        cache.push_back({fn, a.maxThreadsPerBlock, a.sharedSizeBytes});
        kl = &cache.back();
    }
    const unsigned long long threads = (unsigned long long)block.x * block.y * block.z;
    const char* why = nullptr;
    if (grid.x < 1 || grid.x > 0x7fffffffu || grid.y != 1 || grid.z != 1) why = "grid";
    else if (threads < 1 || threads > (unsigned long long)kl->max_threads) why = "block";
    else if (kl->static_lds + dyn_lds > LDS_PER_CU) why = "LDS";
    if (!why) return true;
    if (!g_launch.failed) {
        g_launch.failed = true;
        g_launch.msg = std::string("launch check: ") + name + ": " + why + " out of range (grid " +
                       std::to_string(grid.x) + "x" + std::to_string(grid.y) + "x" + std::to_string(grid.z) +
                       ", block " + std::to_string(threads) + " of max " + std::to_string(kl->max_threads) +
                       ", LDS " + std::to_string(kl->static_lds) + " + " + std::to_string(dyn_lds) + " B)";
    }
    return false;
}

// the sticky launch failure of this thread as the entry point's result (cleared)
int launch_status() {
    if (!g_launch.failed) return NICR_OK;
    g_launch.failed = false;
    return fail(NICR_E_INVALID, g_launch.msg);
}

#define NN_LAUNCH(fn, grid, block, lds, stream, ...)                                                  \
    do {                                                                                              \
        if (::nicr::nn::launch_ok((const void*)(fn), dim3(grid), dim3(block), (size_t)(lds), #fn))    \
            hipLaunchKernelGGL(fn, dim3(grid), dim3(block), lds, stream, __VA_ARGS__);                \
    } while (0)

// ---- head, loss, inputs, optimiser ----

// out[p][c] = bh[c] + sum_ci D[p][ci] Wh[c][ci] (c < 3; out row pitch 4).
__global__ void k_head(int P, const bf16* __restrict__ D, const float* __restrict__ Wh, const float* __restrict__ bh,
                       float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float s0 = bh[0], s1 = bh[1], s2 = bh[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        F8 f;
        f.from(bf8_load(D + (long long)p * HEAD_IN + 8 * q));
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int ci = 8 * q + i;
            s0 += f.v[i] * Wh[ci];
            s1 += f.v[i] * Wh[HEAD_IN + ci];
            s2 += f.v[i] * Wh[2 * HEAD_IN + ci];
        }
    }
    *reinterpret_cast<float4*>(out + 4LL * p) = make_float4(s0, s1, s2, 0.f);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Loss L = (1/3) sum_p sum_c (out - t)^2 and its gradient through the head:
// gD[p][ci] = (sum_c g_c Wh[c][ci]) * (D[p][ci] > 0) with g_c = 2 (out - t) / 3;
// dWh[c][ci] += sum_p g_c D[p][ci]; dbh[c] += sum_p g_c; loss[0] += sum (out - t)^2.
constexpr int HB_THREADS = 256;
__global__ __launch_bounds__(HB_THREADS) void k_head_bwd(int P, const bf16* __restrict__ D,
                                                         const float* __restrict__ out, const float* __restrict__ tgt,
                                                         const float* __restrict__ Wh, bf16* __restrict__ gD,
                                                         float* __restrict__ dWh, float* __restrict__ dbh,
                                                         float* __restrict__ loss) {
    constexpr int NACC = N_OUT * HEAD_IN + N_OUT + 1;
    __shared__ float red[HB_THREADS / 64][NACC];
    float acc[NACC];
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = 0.f;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const float4 o = *reinterpret_cast<const float4*>(out + 4LL * p);
        const float4 t = *reinterpret_cast<const float4*>(tgt + 4LL * p);
        const float e0 = o.x - t.x, e1 = o.y - t.y, e2 = o.z - t.z;
        const float g0 = (2.f / 3.f) * e0, g1 = (2.f / 3.f) * e1, g2 = (2.f / 3.f) * e2;
        acc[NACC - 1] += e0 * e0 + e1 * e1 + e2 * e2;
        acc[N_OUT * HEAD_IN + 0] += g0;
        acc[N_OUT * HEAD_IN + 1] += g1;
        acc[N_OUT * HEAD_IN + 2] += g2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            F8 f, gg;
            f.from(bf8_load(D + (long long)p * HEAD_IN + 8 * q));
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int ci = 8 * q + i;
                acc[ci] += g0 * f.v[i];
                acc[HEAD_IN + ci] += g1 * f.v[i];
                acc[2 * HEAD_IN + ci] += g2 * f.v[i];
                const float gv = g0 * Wh[ci] + g1 * Wh[HEAD_IN + ci] + g2 * Wh[2 * HEAD_IN + ci];
                gg.v[i] = f.v[i] > 0.f ? gv : 0.f;
            }
            bf8_store(gD + (long long)p * HEAD_IN + 8 * q, gg.to());
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NACC; ++i) {
        const float v = wave_sum(acc[i]);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NACC; i += blockDim.x) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < HB_THREADS / 64; ++w) v += red[w][i];
        if (i < N_OUT * HEAD_IN) atomicAdd(&dWh[i], v);
        else if (i < NACC - 1) atomicAdd(&dbh[i - N_OUT * HEAD_IN], v);
        else atomicAdd(loss, v);
    }
}

// Input and output normalisation. The network sees each volume relative to its
// own statistics. Input channels: 0, 1 the deviations of O and Cr from their
// volume means over their volume standard deviations (formed in binary64
// before rounding to BF16, so the small deviations that drive the change keep
// their precision); 2, 3 phi and the mask under fixed scales; 4..6 the plain
// one-step update u of O, Cr, phi at the volume (the step just before the
// jump) over its volume RMS; 7..10 constant fields carrying log10 std O, mean
// Cr, log10 rms u_O, log10 rms u_Cr under fixed scales. The network predicts
// the correction to the linear extrapolation J u of the update over the J
// steps of a jump: change_c = J u_c + out_c J rms(u_c) gain_c (a field whose
// update is zero does not change).
constexpr int N_CST = 4;
struct VolNorm {
    double mean[N_FIELD];          // channel c = (field_c - mean_c) * inv_c, c < 4
    double inv[N_FIELD];
    double uinv[N_OUT];            // channel 4 + c = u_c * uinv_c
    float cst[N_CST];              // channels 7..10
    double jump;                   // J
    double out_scale[N_OUT];       // change_c = J u_c + out_c * out_scale_c
    double out_inv[N_OUT];         // target_c = (change_c - J u_c) * out_inv_c (0 where out_scale is 0)
};

// Network input of one volume.
template <class T0, class T3, class TU>
__global__ void k_input(int V, const T0* __restrict__ f0, const T0* __restrict__ f1, const T0* __restrict__ f2,
                        const T3* __restrict__ f3, const TU* __restrict__ u0, const TU* __restrict__ u1,
                        const TU* __restrict__ u2, VolNorm vn, bf16* __restrict__ x0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= V) return;
    bf16x8 v, w;
    v[0] = (bf16)(float)(((double)f0[p] - vn.mean[0]) * vn.inv[0]);
    v[1] = (bf16)(float)(((double)f1[p] - vn.mean[1]) * vn.inv[1]);
    v[2] = (bf16)(float)(((double)f2[p] - vn.mean[2]) * vn.inv[2]);
    v[3] = (bf16)(float)(((double)f3[p] - vn.mean[3]) * vn.inv[3]);
    v[4] = (bf16)(float)((double)u0[p] * vn.uinv[0]);
    v[5] = (bf16)(float)((double)u1[p] * vn.uinv[1]);
    v[6] = (bf16)(float)((double)u2[p] * vn.uinv[2]);
    v[7] = (bf16)vn.cst[0];
    w[0] = (bf16)vn.cst[1];
    w[1] = (bf16)vn.cst[2];
    w[2] = (bf16)vn.cst[3];
#pragma unroll
    for (int i = 3; i < 8; ++i) w[i] = (bf16)0.f;
    uint4* dst = reinterpret_cast<uint4*>(x0 + 32LL * p);
    dst[0] = __builtin_bit_cast(uint4, v);
    dst[1] = __builtin_bit_cast(uint4, w);
    dst[2] = dst[3] = make_uint4(0u, 0u, 0u, 0u);
}

// A symmetry of the step applied to a training sample: Y and Z are periodic
// and the stencils are mirror-symmetric, so a sample shifted by (sy, sz) and
// mirrored in Y (fy) or Z (fz) is again a trajectory sample of the step; X,
// which holds the inlet, is kept. Position p of the transformed volume reads
// cell src(p) of the recorded one.
struct SampleMap {
    int ny, nz, sy, sz, fy, fz;
    __device__ __forceinline__ int src(int p) const {
        const int nyz = ny * nz, x = p / nyz, r = p - x * nyz, y = r / nz, z = r - y * nz;
        int ys = (fy ? ny - 1 - y : y) + sy, zs = (fz ? nz - 1 - z : z) + sz;
        ys -= ys >= ny ? ny : 0;
        zs -= zs >= nz ? nz : 0;
        return x * nyz + ys * nz + zs;
    }
};

// Network input of one recorded (binary32) training volume through a map.
__global__ void k_input_map(int V, const float* __restrict__ in, SampleMap m, VolNorm vn, bf16* __restrict__ x0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= V) return;
    const int q = m.src(p);
    bf16x8 v, w;
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (bf16)(float)(((double)in[(long long)c * V + q] - vn.mean[c]) * vn.inv[c]);
#pragma unroll
    for (int c = 0; c < N_OUT; ++c) v[4 + c] = (bf16)(float)((double)in[(long long)(4 + c) * V + q] * vn.uinv[c]);
    v[7] = (bf16)vn.cst[0];
    w[0] = (bf16)vn.cst[1];
    w[1] = (bf16)vn.cst[2];
    w[2] = (bf16)vn.cst[3];
#pragma unroll
    for (int i = 3; i < 8; ++i) w[i] = (bf16)0.f;
    uint4* dst = reinterpret_cast<uint4*>(x0 + 32LL * p);
    dst[0] = __builtin_bit_cast(uint4, v);
    dst[1] = __builtin_bit_cast(uint4, w);
    dst[2] = dst[3] = make_uint4(0u, 0u, 0u, 0u);
}

// Training target of one volume: tgt[p][c] = (d_c - J u_c) * out_inv_c, row pitch 4.
__global__ void k_target(int V, const float* __restrict__ d, const float* __restrict__ u, SampleMap m, VolNorm vn,
                         float* __restrict__ tgt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= V) return;
    const int q = m.src(p);
    float t[N_OUT];
#pragma unroll
    for (int c = 0; c < N_OUT; ++c)
        t[c] = (float)(((double)d[(long long)c * V + q] - vn.jump * (double)u[(long long)c * V + q]) * vn.out_inv[c]);
    *reinterpret_cast<float4*>(tgt + 4LL * p) = make_float4(t[0], t[1], t[2], 0.f);
}

// ---- fused passes of a surrogate jump (binary64 state, block partial sums) ----

constexpr int SUR_BLOCKS = 1024;   // blocks of the fused passes; partial sums: SUR_BLOCKS x 8

// block sum of NV binary64 values into part[blockIdx.x * 8 + k]
template <int NV>
__device__ __forceinline__ void block_sums8(double (&v)[NV], double* __restrict__ part) {
    __shared__ double red[NV][4];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV)
        part[blockIdx.x * 8 + threadIdx.x] =
            red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// Before a jump: u = x0 - x_-1 for O, Cr, phi; backup of x0; sums of O, O^2,
// Cr, Cr^2, u_O^2, u_Cr^2, u_phi^2. The first n_hold cells (the inlet plane x
// = 0 under the concentration boundary, whose O every step resets to the
// inlet value before its update, 2D:408-411) keep u_O = 0: their O after a
// step is the inlet value plus one step's change, so a jump leaves it as it
// is; extrapolating it would carry the reset of a check step (when the state
// before it came from a jump) into the next jump.
__global__ __launch_bounds__(256) void k_sur_pre(long long n, long long n_hold, const double* __restrict__ o,
                                                 const double* __restrict__ cr, const double* __restrict__ ph,
                                                 const double* __restrict__ op, const double* __restrict__ crp,
                                                 const double* __restrict__ php, double* __restrict__ u0,
                                                 double* __restrict__ u1, double* __restrict__ u2,
                                                 double* __restrict__ b0, double* __restrict__ b1,
                                                 double* __restrict__ b2, double* __restrict__ part) {
    double v[7] = {0., 0., 0., 0., 0., 0., 0.};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double a = o[i], c = cr[i], p = ph[i];
        const double d0 = i < n_hold ? 0. : a - op[i], d1 = c - crp[i], d2 = p - php[i];
        u0[i] = d0;
        u1[i] = d1;
        u2[i] = d2;
        b0[i] = a;
        b1[i] = c;
        b2[i] = p;
        v[0] += a;
        v[1] += a * a;
        v[2] += c;
        v[3] += c * c;
        v[4] += d0 * d0;
        v[5] += d1 * d1;
        v[6] += d2 * d2;
    }
    block_sums8<7>(v, part);
}

// The proposal x0 + J u + out * out_scale written to the state; sums of the
// new O and Cr and min/max of O, Cr, phi (max as min of the negation).
__global__ __launch_bounds__(256) void k_sur_apply(long long n, const float* __restrict__ out, VolNorm vn,
                                                   const double* __restrict__ u0, const double* __restrict__ u1,
                                                   const double* __restrict__ u2, double* __restrict__ o,
                                                   double* __restrict__ cr, double* __restrict__ ph,
                                                   double* __restrict__ part) {
    double s[2] = {0., 0.};
    double mn[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float4 w = *reinterpret_cast<const float4*>(out + 4 * i);
        const double a = o[i] + vn.jump * u0[i] + (double)w.x * vn.out_scale[0];
        const double c = cr[i] + vn.jump * u1[i] + (double)w.y * vn.out_scale[1];
        const double p = ph[i] + vn.jump * u2[i] + (double)w.z * vn.out_scale[2];
        o[i] = a;
        cr[i] = c;
        ph[i] = p;
        s[0] += a;
        s[1] += c;
        mn[0] = fmin(mn[0], a);
        mn[1] = fmin(mn[1], -a);
        mn[2] = fmin(mn[2], c);
        mn[3] = fmin(mn[3], -c);
        mn[4] = fmin(mn[4], p);
        mn[5] = fmin(mn[5], -p);
    }
    // non-finite values poison the sums and show in the bounds test on the host
    __shared__ double red[8][4];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double x = k < 2 ? s[k] : mn[k - 2];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_xor(x, off, 64);
            x = k < 2 ? x + y : fmin(x, y);
        }
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        const double* r = red[k];
        part[blockIdx.x * 8 + k] = k < 2 ? r[0] + r[1] + r[2] + r[3] : fmin(fmin(r[0], r[1]), fmin(r[2], r[3]));
    }
}

__global__ void k_scale2(long long n, double* __restrict__ a, double fa, double* __restrict__ b, double fb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] *= fa;
        b[i] *= fb;
    }
}

// Trapezoid residuals of O and Cr: R = (xh - x0) - (J/2)(u + (xh1 - xh));
// sums of R^2 and (xh - x0)^2 for each.
__global__ __launch_bounds__(256) void k_sur_trap(long long n, const double* __restrict__ ho,
                                                  const double* __restrict__ h1o, const double* __restrict__ xo,
                                                  const double* __restrict__ uo, const double* __restrict__ hc,
                                                  const double* __restrict__ h1c, const double* __restrict__ xc,
                                                  const double* __restrict__ uc, double halfJ,
                                                  double* __restrict__ part) {
    double v[4] = {0., 0., 0., 0.};
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double d0 = ho[i] - xo[i], r0 = d0 - halfJ * (uo[i] + (h1o[i] - ho[i]));
        const double d1 = hc[i] - xc[i], r1 = d1 - halfJ * (uc[i] + (h1c[i] - hc[i]));
        v[0] += r0 * r0;
        v[1] += d0 * d0;
        v[2] += r1 * r1;
        v[3] += d1 * d1;
    }
    block_sums8<4>(v, part);
}

// Evolving grains, before the check step of a jump: the eta fields (n_eta
// fields of ncell cells, field g at g * fs) move to the linear extrapolation
// eta_h = eta_0 + J (eta_0 - eta_-1) in place in `cur` (eta_-1 in `prev`),
// eta_0 goes to `bk`; the mask (S, P, gb) of `cur` goes to bk_mask and is
// rebuilt from eta_h with the arithmetic of k_mask. part: min eta_h and
// min(-eta_h) per block (columns 0, 1).
__global__ __launch_bounds__(256) void k_sur_eta_pre(long long ncell, long long fs, int n_eta, double J,
                                                     double eta_scale, const double* __restrict__ prev,
                                                     const double* __restrict__ cur, double* __restrict__ out,
                                                     double* __restrict__ s2, double* __restrict__ pairs,
                                                     double* __restrict__ gb, double* __restrict__ bk_mask,
                                                     double* __restrict__ part) {
    double mn = INFINITY, mx = INFINITY;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ncell;
         i += (long long)gridDim.x * blockDim.x) {
        bk_mask[i] = s2[i];
        bk_mask[ncell + i] = pairs[i];
        bk_mask[2 * ncell + i] = gb[i];
        double S = 0., pr = 0.;
        for (int g = 0; g < n_eta; ++g) {
            const long long k = (long long)g * fs + i;
            const double a = prev[k], b = cur[k];
            const double h = b + J * (b - a);
            out[k] = h;
            mn = fmin(mn, h);
            mx = fmin(mx, -h);
            const double q = h * h;
            pr += S * q;
            S += q;
        }
        const double s16 = 16. * pr;
        s2[i] = S;
        pairs[i] = pr;
        gb[i] = eta_scale * (s16 < 1. ? s16 : 1.);
    }
    __shared__ double red[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double x = k == 0 ? mn : mx;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off, 64));
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const double* r = red[threadIdx.x];
        part[blockIdx.x * 8 + threadIdx.x] = fmin(fmin(r[0], r[1]), fmin(r[2], r[3]));
    }
}

// Trapezoid residual of the extrapolated eta fields: with d = eta_h - eta_0 =
// J u_0, R = d - (J/2)(d / J + (eta_h1 - eta_h)); sums of R^2 and d^2 over all
// fields (columns 0, 1).
__global__ __launch_bounds__(256) void k_sur_eta_trap(long long ncell, long long fs, int n_eta, double J,
                                                      const double* __restrict__ h, const double* __restrict__ h1,
                                                      const double* __restrict__ x0, double* __restrict__ part) {
    double v[2] = {0., 0.};
    const double halfJ = 0.5 * J;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ncell;
         i += (long long)gridDim.x * blockDim.x)
        for (int g = 0; g < n_eta; ++g) {
            const long long k = (long long)g * fs + i;
            const double hv = h[k];
            const double d = hv - x0[k];
            const double r = d - halfJ * (d / J + (h1[k] - hv));
            v[0] += r * r;
            v[1] += d * d;
        }
    block_sums8<2>(v, part);
}

// acc += sum (d - J u)^2 (float arrays; binary64 sums)
__global__ __launch_bounds__(256) void k_corr_sq(long long n, const float* __restrict__ d, const float* __restrict__ u,
                                                 double J, double* __restrict__ acc) {
    __shared__ double red[4];
    double s = 0.;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double r = (double)d[i] - J * (double)u[i];
        s += r * r;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(acc, red[0] + red[1] + red[2] + red[3]);
}

// acc[0] += sum a, acc[1] += sum a^2 (binary64), for volume statistics.
__global__ __launch_bounds__(256) void k_moments_d(long long n, const double* __restrict__ a,
                                                   double* __restrict__ acc) {
    __shared__ double red[2][4];
    double s = 0., t = 0.;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double v = a[i];
        s += v;
        t += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        t += __shfl_xor(t, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s;
        red[1][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(&acc[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

__global__ void k_sub(long long n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

// Moments of a float array: acc[0] += sum x, acc[1] += sum x^2 (binary64).
__global__ __launch_bounds__(256) void k_moments(long long n, const float* __restrict__ x, double* __restrict__ acc) {
    __shared__ double red[2][4];
    double s = 0., t = 0.;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double v = x[i];
        s += v;
        t += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        t += __shfl_xor(t, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s;
        red[1][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(&acc[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

template <class T>
__global__ void k_to_float(long long n, const T* __restrict__ a, float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)a[i];
}
// out = float(a - b)
__global__ void k_delta_float(long long n, const double* __restrict__ a, const double* __restrict__ b,
                              float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)(a[i] - b[i]);
}

// Global gradient norm: acc[0] += sum g^2.
__global__ __launch_bounds__(256) void k_grad_sq(long long n, const float* __restrict__ g, double* __restrict__ acc) {
    __shared__ double red[4];
    double s = 0.;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        s += (double)g[i] * (double)g[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(acc, red[0] + red[1] + red[2] + red[3]);
}

// Gradient clipping factor from the norm (clip[0] = ||g||^2 of this iteration):
// scale = min(1, 3 ema / ||g||), with ema a running mean of the clipped norms
// (clip[1]; 0 before the first iteration). clip[2] receives the scale.
__global__ void k_clip_factor(double* clip) {
    const double nrm = sqrt(clip[0]);
    double ema = clip[1];
    if (!(ema > 0.)) ema = nrm;
    const double lim = 3. * ema;
    double f = nrm > lim && nrm > 0. ? lim / nrm : 1.;
    if (!(nrm == nrm) || nrm > 1e300) f = 0.;   // a non-finite gradient leaves the parameters unchanged
    clip[2] = f;
    if (f > 0.) clip[1] = 0.99 * ema + 0.01 * (nrm * f);
    clip[0] = 0.;
}

// Adam (bias-corrected) on the FP32 master parameters, gradient scaled by *gscale.
__global__ void k_adam(long long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps,
                       const double* __restrict__ gscale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sc = (float)*gscale;
    if (sc == 0.f) return;
    const float gi = g[i] * sc;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 copies of a convolution's weights: Wb[tap][co][ci] and the data-gradient
// form Wt[taps - 1 - tap][ci][co].
__global__ void k_cast_w(int taps, int cout, int cin, const float* __restrict__ W, bf16* __restrict__ Wb,
                         bf16* __restrict__ Wt) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)taps * cout * cin;
    if (i >= n) return;
    const int ci = (int)(i % cin), co = (int)((i / cin) % cout), tap = (int)(i / ((long long)cin * cout));
    const bf16 b = (bf16)W[i];
    Wb[i] = b;
    Wt[((long long)(taps - 1 - tap) * cin + ci) * cout + co] = b;
}

// ---- network ----

struct ConvL {
    int cin, cout, level;
    bool relu;
    long long w_off, b_off;   // into the parameter array
    bf16* Wb = nullptr;
    bf16* Wt = nullptr;
    bf16 *x = nullptr, *y = nullptr;      // input and output activations
    bf16 *gx = nullptr, *gy = nullptr;    // their gradients (gy: pre-activation)
};

inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

struct Net {
    int nbmax = 1, S[3] = {0, 0, 0};
    bool train = false;
    bool augment = false;   // training samples through random Y/Z shifts and mirrors
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    size_t bytes = 0;
    long long nparam = 0, head_w = 0, head_b = 0;
    float *param = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    int adam_t = 0;
    ConvL conv[N_CONV];
    // activations
    bf16 *x0, *e1a, *e1, *p1, *e2a, *e2, *p2, *e3a, *e3, *p3, *b4a, *b4, *u3, *d3a, *d3, *u2, *d2a, *d2, *u1, *d1a,
        *d1;
    float *out = nullptr, *tgt = nullptr;
    // gradients
    bf16 *g_d1, *g_d1a, *g_u1, *g_d2, *g_d2a, *g_u2, *g_d3, *g_d3a, *g_u3, *g_b4, *g_b4a, *g_p3, *g_e3, *g_e3a, *g_p2,
        *g_e2, *g_e2a, *g_p1, *g_e1, *g_e1a;
    float* loss_acc = nullptr;   // one slot per training iteration
    int loss_cap = 0;
    double* dacc = nullptr;      // reduction scratch
    double* clip = nullptr;      // gradient clipping: norm^2 of the iteration, running mean, factor
    nicr_unet_norm norm{};
    std::vector<double> ds_stat;   // per training sample: mean and std of O and Cr
    // training set: inputs [S][4][V] and 50-step deltas [S][3][V] (FP32)
    int ds_cap = 0, ds_n = 0;
    float *ds_in = nullptr, *ds_d = nullptr;
    // binary64 state backup of a surrogate jump (O, Cr, phi)
    double* bk[3] = {nullptr, nullptr, nullptr};
    // evolving grains: eta before a jump (the model's field layout) and its mask S, P, gb
    double* bk_eta = nullptr;
    double* bk_mask = nullptr;
    long long bk_eta_elems = 0, bk_mask_cells = 0;
    double* du[3] = {nullptr, nullptr, nullptr};   // plain update of O, Cr, phi before a jump
    double* part = nullptr;                          // block partial sums of the fused jump passes
    std::vector<double> part_host;
    int ds_jump = 0;                                 // steps of the recorded changes
    long long bk_cells = 0;
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double fwd_ms_sum = 0.;
    int fwd_count = 0;

    long long V(int l) const { return (long long)(S[0] >> l) * (S[1] >> l) * (S[2] >> l); }
};

int nalloc(Net* n, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(NICR_E_DEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    n->allocs.push_back(*p);
    n->bytes += bytes;
    return NICR_OK;
}
template <class T>
int nalloc_t(Net* n, T** p, long long count) {
    void* q = nullptr;
    int rc = nalloc(n, &q, (size_t)count * sizeof(T));
    *p = static_cast<T*>(q);
    return rc;
}

void net_free(Net* n) {
    for (void* p : n->allocs) hipFree(p);
    n->allocs.clear();
    if (n->ev[0]) hipEventDestroy(n->ev[0]);
    if (n->ev[1]) hipEventDestroy(n->ev[1]);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
}

Geo geo(const Net* n, int nb, int level, int cin, int cout, int taps) {
    Geo g;
    g.nb = nb;
    g.s0 = n->S[0] >> level;
    g.s1 = n->S[1] >> level;
    g.s2 = n->S[2] >> level;
    g.cin = cin;
    g.cout = cout;
    g.taps = taps;
    return g;
}

// Halo-tile kernel settings: compute units of the current device (one
// resident block each: 152 KB of LDS) and whether the kernel may be used
// (NICR_CONV_HALO=0 keeps the per-tap kernel everywhere).
struct HaloCfg {
    int ncu = 0;
    bool enabled = true;
    bool half_k = true;   // NICR_CONV_HALFK=0: the input layer computes all 32 channels
    bool attr_set = false;
};
static HaloCfg g_halo;

static bool halo_ready() {
    if (!g_halo.attr_set) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            return false;
        g_halo.ncu = ncu;
        const char* e = getenv("NICR_CONV_HALO");
        g_halo.enabled = !(e && e[0] == '0');
        const char* hk = getenv("NICR_CONV_HALFK");
        g_halo.half_k = !(hk && hk[0] == '0');
        if (hipFuncSetAttribute((const void*)k_conv_halo<true, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                HC_LDS_BYTES) != hipSuccess ||
            hipFuncSetAttribute((const void*)k_conv_halo<false, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                HC_LDS_BYTES) != hipSuccess ||
            hipFuncSetAttribute((const void*)k_conv_halo<true, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                HC_LDS_BYTES) != hipSuccess ||
            hipFuncSetAttribute((const void*)k_conv_halo<false, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                HC_LDS_BYTES) != hipSuccess)
            g_halo.enabled = false;
        g_halo.attr_set = true;
    }
    return g_halo.enabled;
}

bool halo_applies(const Geo& g) { return g.s0 % HT_X == 0 && g.s1 % HT_Y == 0 && g.s2 % HT_Z == 0 && halo_ready(); }

// Xc: coarse source of channels 0 .. g.cu - 1 (halo kernel only). half_k: X
// has 32 channels of which 16..31 are zero (the network's input tensor).
void launch_conv(Net* n, const Geo& g, const bf16* X, const bf16* W, const float* bias, bf16* Y, bool relu,
                 const bf16* Xc = nullptr, bool half_k = false) {
    if (halo_applies(g)) {
        const int units = g.nb * (g.s0 / HT_X) * (g.s1 / HT_Y) * (g.s2 / HT_Z) * (g.cout / 32);
        const int nblock = std::min(units, g_halo.ncu);
        const int upb = (units + nblock - 1) / nblock;
        const int grid = (units + upb - 1) / upb;
        const bool hk = half_k && g.cin == 32 && g.cu == 0 && g_halo.half_k;
        if (relu && hk)
            NN_LAUNCH((k_conv_halo<true, true>), dim3(grid), dim3(HC_THREADS), HC_LDS_BYTES, n->stream, g, X, W,
                      bias, Y, units, upb, Xc);
        else if (hk)
            NN_LAUNCH((k_conv_halo<false, true>), dim3(grid), dim3(HC_THREADS), HC_LDS_BYTES, n->stream, g, X, W,
                      bias, Y, units, upb, Xc);
        else if (relu)
            NN_LAUNCH((k_conv_halo<true, false>), dim3(grid), dim3(HC_THREADS), HC_LDS_BYTES, n->stream, g, X, W,
                      bias, Y, units, upb, Xc);
        else
            NN_LAUNCH((k_conv_halo<false, false>), dim3(grid), dim3(HC_THREADS), HC_LDS_BYTES, n->stream, g, X, W,
                      bias, Y, units, upb, Xc);
        return;
    }
    const int P = g.positions();
    const unsigned mb = (unsigned)((P + CBM - 1) / CBM);
    if (g.cout % 64 == 0) {
        const dim3 grid(mb * (g.cout / 64));
        if (relu) NN_LAUNCH((k_conv<64, true>), grid, dim3(256), 0, n->stream, g, X, W, bias, Y);
        else NN_LAUNCH((k_conv<64, false>), grid, dim3(256), 0, n->stream, g, X, W, bias, Y);
    } else {
        const dim3 grid(mb * (g.cout / 32));
        if (relu) NN_LAUNCH((k_conv<32, true>), grid, dim3(256), 0, n->stream, g, X, W, bias, Y);
        else NN_LAUNCH((k_conv<32, false>), grid, dim3(256), 0, n->stream, g, X, W, bias, Y);
    }
}

void launch_wgrad(Net* n, const Geo& g, const bf16* X, const bf16* dY, float* dW) {
    const int P = g.positions();
    const int tiles = (g.cout / 32) * (g.cin / 32);
    int nchunk = std::max(1, 4096 / (tiles * g.taps));
    int chunk = (P + nchunk - 1) / nchunk;
    chunk = std::max(128, (chunk + 127) / 128 * 128);
    nchunk = (P + chunk - 1) / chunk;
    NN_LAUNCH(k_conv_wgrad, dim3((unsigned)(tiles * g.taps * nchunk)), dim3(256), 0, n->stream, g, X, dY, dW,
                       chunk);
}

void launch_bias_grad(Net* n, int P, int C, const bf16* G, float* dB) {
    const int cg = C / 32;
    int nchunk = std::max(1, std::min(2048 / cg, (P + 1023) / 1024));
    const int chunk = (P + nchunk - 1) / nchunk;
    nchunk = (P + chunk - 1) / chunk;
    NN_LAUNCH(k_bias_grad, dim3((unsigned)(cg * nchunk)), dim3(256), 0, n->stream, P, C, chunk, G, dB);
}

int net_build(Net* n) {
    // convolution table: (cin, cout, level)
    const int tab[N_CONV][3] = {{32, 32, 0},   {32, 32, 0},   {32, 64, 1},   {64, 64, 1},  {64, 128, 2},
                                {128, 128, 2}, {128, 256, 3}, {256, 256, 3}, {384, 128, 2}, {128, 128, 2},
                                {192, 64, 1},  {64, 64, 1},   {96, 32, 0},   {32, 32, 0}};
    long long off = 0;
    for (int i = 0; i < N_CONV; ++i) {
        ConvL& L = n->conv[i];
        L.cin = tab[i][0];
        L.cout = tab[i][1];
        L.level = tab[i][2];
        L.relu = true;
        L.w_off = off;
        off += 27LL * L.cin * L.cout;
        L.b_off = off;
        off += L.cout;
    }
    n->head_w = off;
    off += N_OUT * HEAD_IN;
    n->head_b = off;
    off += N_OUT;
    n->nparam = off;
    int rc = 0;
    if ((rc = nalloc_t(n, &n->param, n->nparam)) || (rc = nalloc_t(n, &n->grad, n->nparam)) ||
        (rc = nalloc_t(n, &n->dacc, 64)))
        return rc;
    if (n->train && ((rc = nalloc_t(n, &n->am, n->nparam)) || (rc = nalloc_t(n, &n->av, n->nparam)) ||
                     (rc = nalloc_t(n, &n->clip, 4))))
        return rc;
    if (n->train) {
        HIPCHK(hipMemsetAsync(n->am, 0, n->nparam * 4, n->stream));
        HIPCHK(hipMemsetAsync(n->av, 0, n->nparam * 4, n->stream));
        HIPCHK(hipMemsetAsync(n->clip, 0, 4 * 8, n->stream));
    }
    for (int i = 0; i < N_CONV; ++i) {
        ConvL& L = n->conv[i];
        const long long nw = 27LL * L.cin * L.cout;
        if ((rc = nalloc_t(n, &L.Wb, nw)) || (rc = nalloc_t(n, &L.Wt, nw))) return rc;
    }
    const int nb = n->nbmax;
    auto act = [&](bf16** p, int level, int C) { return nalloc_t(n, p, (long long)nb * n->V(level) * C); };
    if ((rc = act(&n->x0, 0, 32)) || (rc = act(&n->e1a, 0, 32)) || (rc = act(&n->e1, 0, 32)) ||
        (rc = act(&n->p1, 1, 32)) || (rc = act(&n->e2a, 1, 64)) || (rc = act(&n->e2, 1, 64)) ||
        (rc = act(&n->p2, 2, 64)) || (rc = act(&n->e3a, 2, 128)) || (rc = act(&n->e3, 2, 128)) ||
        (rc = act(&n->p3, 3, 128)) || (rc = act(&n->b4a, 3, 256)) || (rc = act(&n->b4, 3, 256)) ||
        (rc = act(&n->u3, 2, 384)) || (rc = act(&n->d3a, 2, 128)) || (rc = act(&n->d3, 2, 128)) ||
        (rc = act(&n->u2, 1, 192)) || (rc = act(&n->d2a, 1, 64)) || (rc = act(&n->d2, 1, 64)) ||
        (rc = act(&n->u1, 0, 96)) || (rc = act(&n->d1a, 0, 32)) || (rc = act(&n->d1, 0, 32)))
        return rc;
    if ((rc = nalloc_t(n, &n->out, 4LL * nb * n->V(0)))) return rc;
    if (n->train) {
        if ((rc = nalloc_t(n, &n->tgt, 4LL * nb * n->V(0))) || (rc = act(&n->g_d1, 0, 32)) ||
            (rc = act(&n->g_d1a, 0, 32)) || (rc = act(&n->g_u1, 0, 96)) || (rc = act(&n->g_d2, 1, 64)) ||
            (rc = act(&n->g_d2a, 1, 64)) || (rc = act(&n->g_u2, 1, 192)) || (rc = act(&n->g_d3, 2, 128)) ||
            (rc = act(&n->g_d3a, 2, 128)) || (rc = act(&n->g_u3, 2, 384)) || (rc = act(&n->g_b4, 3, 256)) ||
            (rc = act(&n->g_b4a, 3, 256)) || (rc = act(&n->g_p3, 3, 128)) || (rc = act(&n->g_e3, 2, 128)) ||
            (rc = act(&n->g_e3a, 2, 128)) || (rc = act(&n->g_p2, 2, 64)) || (rc = act(&n->g_e2, 1, 64)) ||
            (rc = act(&n->g_e2a, 1, 64)) || (rc = act(&n->g_p1, 1, 32)) || (rc = act(&n->g_e1, 0, 32)) ||
            (rc = act(&n->g_e1a, 0, 32)))
            return rc;
    }
    // activation and gradient wiring of each convolution
    ConvL* c = n->conv;
    auto wire = [](ConvL& L, bf16* x, bf16* y, bf16* gx, bf16* gy) {
        L.x = x;
        L.y = y;
        L.gx = gx;
        L.gy = gy;
    };
    wire(c[0], n->x0, n->e1a, nullptr, n->g_e1a);
    wire(c[1], n->e1a, n->e1, n->g_e1a, n->g_e1);
    wire(c[2], n->p1, n->e2a, n->g_p1, n->g_e2a);
    wire(c[3], n->e2a, n->e2, n->g_e2a, n->g_e2);
    wire(c[4], n->p2, n->e3a, n->g_p2, n->g_e3a);
    wire(c[5], n->e3a, n->e3, n->g_e3a, n->g_e3);
    wire(c[6], n->p3, n->b4a, n->g_p3, n->g_b4a);
    wire(c[7], n->b4a, n->b4, n->g_b4a, n->g_b4);
    wire(c[8], n->u3, n->d3a, n->g_u3, n->g_d3a);
    wire(c[9], n->d3a, n->d3, n->g_d3a, n->g_d3);
    wire(c[10], n->u2, n->d2a, n->g_u2, n->g_d2a);
    wire(c[11], n->d2a, n->d2, n->g_d2a, n->g_d2);
    wire(c[12], n->u1, n->d1a, n->g_u1, n->g_d1a);
    wire(c[13], n->d1a, n->d1, n->g_d1a, n->g_d1);
    n->norm = nicr_unet_norm{};
    n->norm.phi_scale = n->norm.gb_scale = 1.f;
    for (int i = 0; i < N_CST; ++i) n->norm.c_scale[i] = 1.f;
    for (int i = 0; i < N_OUT; ++i) n->norm.gain[i] = 1.f;
    return NICR_OK;
}

void cast_weights(Net* n) {
    for (int i = 0; i < N_CONV; ++i) {
        ConvL& L = n->conv[i];
        const long long nw = 27LL * L.cin * L.cout;
        NN_LAUNCH(k_cast_w, dim3(nblk(nw)), dim3(256), 0, n->stream, 27, L.cout, L.cin, n->param + L.w_off,
                           L.Wb, L.Wt);
    }
}

// He-normal weights (fan-in 27 x input channels that carry data), zero biases;
// the head starts at zero so the first prediction is "no change".
int init_params(Net* n, uint64_t seed) {
    std::vector<float> h((size_t)n->nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (int i = 0; i < N_CONV; ++i) {
        const ConvL& L = n->conv[i];
        const int cin_real = i == 0 ? N_IN : L.cin;
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * cin_real)));
        for (int t = 0; t < 27; ++t)
            for (int co = 0; co < L.cout; ++co)
                for (int ci = 0; ci < L.cin; ++ci)
                    h[(size_t)(L.w_off + ((long long)t * L.cout + co) * L.cin + ci)] = ci < cin_real ? nd(rng) : 0.f;
    }
    HIPCHK(hipMemcpyAsync(n->param, h.data(), h.size() * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    cast_weights(n);
    return NICR_OK;
}

void forward(Net* n, int nb) {
    hipStream_t s = n->stream;
    ConvL* c = n->conv;
    auto conv = [&](int i) {
        const ConvL& L = c[i];
        // layer 0 reads the input tensor, whose channels 16..31 are zero
        launch_conv(n, geo(n, nb, L.level, L.cin, L.cout, 27), L.x, L.Wb, n->param + L.b_off, L.y, L.relu, nullptr,
                    i == 0);
    };
    auto pool = [&](int level, int C, const bf16* X, bf16* Y) {
        const long long nq = (long long)nb * n->V(level + 1);
        NN_LAUNCH(k_pool, dim3(nblk(nq * (C / 8))), dim3(256), 0, s, (int)nq, n->S[0] >> level,
                           n->S[1] >> level, n->S[2] >> level, C, X, Y);
    };
    auto upcat = [&](int level, int Cu, int Cs, const bf16* L, const bf16* Sk, bf16* U) {
        const long long np = (long long)nb * n->V(level);
        NN_LAUNCH(k_up_cat, dim3(nblk(np * ((Cu + Cs) / 8))), dim3(256), 0, s, (int)np, n->S[0] >> level,
                           n->S[1] >> level, n->S[2] >> level, Cu, Cs, L, Sk, U);
    };
    // decoder convolution i on [up(L), Sk]: without training buffers and where
    // the halo kernel applies, the concatenation is read in place (never stored)
    auto upcat_conv = [&](int i, int Cu, const bf16* L, const bf16* Sk) {
        const ConvL& Lc = c[i];
        Geo g = geo(n, nb, Lc.level, Lc.cin, Lc.cout, 27);
        if (!n->train && halo_applies(g)) {
            g.cu = Cu;
            launch_conv(n, g, Sk, Lc.Wb, n->param + Lc.b_off, Lc.y, Lc.relu, L);
        } else {
            upcat(Lc.level, Cu, Lc.cin - Cu, L, Sk, Lc.x);
            conv(i);
        }
    };
    if (n->timing) hipEventRecord(n->ev[0], s);
    conv(0);
    conv(1);
    pool(0, 32, n->e1, n->p1);
    conv(2);
    conv(3);
    pool(1, 64, n->e2, n->p2);
    conv(4);
    conv(5);
    pool(2, 128, n->e3, n->p3);
    conv(6);
    conv(7);
    upcat_conv(8, 256, n->b4, n->e3);
    conv(9);
    upcat_conv(10, 128, n->d3, n->e2);
    conv(11);
    upcat_conv(12, 64, n->d2, n->e1);
    conv(13);
    const long long P0 = (long long)nb * n->V(0);
    NN_LAUNCH(k_head, dim3(nblk(P0)), dim3(256), 0, s, (int)P0, n->d1, n->param + n->head_w,
                       n->param + n->head_b, n->out);
    if (n->timing) hipEventRecord(n->ev[1], s);
}

// Gradients of L = (1/3) sum (out - tgt)^2 into n->grad (accumulated; the
// caller zeroes them); the loss sum goes to *loss_slot.
void backward(Net* n, int nb, float* loss_slot) {
    hipStream_t s = n->stream;
    ConvL* c = n->conv;
    const long long P0 = (long long)nb * n->V(0);
    NN_LAUNCH(k_head_bwd, dim3((unsigned)std::min<long long>(1024, nblk(P0))), dim3(HB_THREADS), 0, s,
                       (int)P0, n->d1, n->out, n->tgt, n->param + n->head_w, n->g_d1, n->grad + n->head_w,
                       n->grad + n->head_b, loss_slot);
    // conv i backward: weight and bias gradients from gy (pre-activation
    // gradient) and x; data gradient gx = conv(gy, Wt) (masked by the caller's
    // ReLU where x is a ReLU output of a single consumer).
    auto conv_bwd = [&](int i, bool data, bool mask_x) {
        const ConvL& L = c[i];
        const Geo g = geo(n, nb, L.level, L.cin, L.cout, 27);
        launch_wgrad(n, g, L.x, L.gy, n->grad + L.w_off);
        launch_bias_grad(n, g.positions(), L.cout, L.gy, n->grad + L.b_off);
        if (!data) return;
        const Geo gt = geo(n, nb, L.level, L.cout, L.cin, 27);
        launch_conv(n, gt, L.gy, L.Wt, nullptr, L.gx, false);
        if (mask_x) {
            const long long n8 = (long long)g.positions() * L.cin / 8;
            NN_LAUNCH(k_relu_mask, dim3(nblk(n8)), dim3(256), 0, s, n8, L.x, L.gx);
        }
    };
    auto up_bwd = [&](int level, int Cu, int Ct, const bf16* dU, const bf16* Lo, bf16* dL) {
        const long long nq = (long long)nb * n->V(level + 1);
        NN_LAUNCH(k_up_bwd, dim3(nblk(nq * (Cu / 8))), dim3(256), 0, s, (int)nq, n->S[0] >> level,
                           n->S[1] >> level, n->S[2] >> level, Cu, Ct, dU, Lo, dL);
    };
    auto pool_skip_bwd = [&](int level, int C, int Cu, int Ct, const bf16* dP, const bf16* dU, const bf16* E,
                             bf16* dE) {
        const long long np = (long long)nb * n->V(level);
        NN_LAUNCH(k_pool_skip_bwd, dim3(nblk(np * (C / 8))), dim3(256), 0, s, (int)np, n->S[0] >> level,
                           n->S[1] >> level, n->S[2] >> level, C, Cu, Ct, dP, dU, E, dE);
    };
    conv_bwd(13, true, true);                        // -> g_d1a
    conv_bwd(12, true, false);                       // -> g_u1 [P0][96]
    up_bwd(0, 64, 96, n->g_u1, n->d2, n->g_d2);      // -> g_d2
    conv_bwd(11, true, true);
    conv_bwd(10, true, false);                       // -> g_u2 [P1][192]
    up_bwd(1, 128, 192, n->g_u2, n->d3, n->g_d3);
    conv_bwd(9, true, true);
    conv_bwd(8, true, false);                        // -> g_u3 [P2][384]
    up_bwd(2, 256, 384, n->g_u3, n->b4, n->g_b4);
    conv_bwd(7, true, true);
    conv_bwd(6, true, false);                        // -> g_p3
    pool_skip_bwd(2, 128, 256, 384, n->g_p3, n->g_u3, n->e3, n->g_e3);
    conv_bwd(5, true, true);
    conv_bwd(4, true, false);                        // -> g_p2
    pool_skip_bwd(1, 64, 128, 192, n->g_p2, n->g_u2, n->e2, n->g_e2);
    conv_bwd(3, true, true);
    conv_bwd(2, true, false);                        // -> g_p1
    pool_skip_bwd(0, 32, 64, 96, n->g_p1, n->g_u1, n->e1, n->g_e1);
    conv_bwd(1, true, true);
    conv_bwd(0, false, false);
}

// One Adam update with global-norm clipping (no host synchronisation: the
// norm, the clip factor and its running mean stay on the device).
void adam_step(Net* n, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++n->adam_t;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)n->adam_t));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)n->adam_t));
    NN_LAUNCH(k_grad_sq, dim3(512), dim3(256), 0, n->stream, n->nparam, n->grad, n->clip);
    NN_LAUNCH(k_clip_factor, dim3(1), dim3(1), 0, n->stream, n->clip);
    NN_LAUNCH(k_adam, dim3(nblk(n->nparam)), dim3(256), 0, n->stream, n->nparam, n->param, n->grad, n->am,
                       n->av, lr, b1, b2, c1, c2, eps, (const double*)(n->clip + 2));
    cast_weights(n);
}

// volume statistics (mean O, std O, mean Cr, std Cr, rms u_O, rms u_Cr, rms
// u_phi) -> the kernels' normalisation
VolNorm vol_norm(const nicr_unet_norm& nm, const double st[N_STAT]) {
    VolNorm v{};
    const double mo = st[0], so = st[1], mc = st[2], sc = st[3];
    v.mean[0] = mo;
    v.inv[0] = so > 0. ? 1. / so : 0.;
    v.mean[1] = mc;
    v.inv[1] = sc > 0. ? 1. / sc : 0.;
    v.mean[2] = nm.phi_mean;
    v.inv[2] = 1. / nm.phi_scale;
    v.mean[3] = nm.gb_mean;
    v.inv[3] = 1. / nm.gb_scale;
    for (int k = 0; k < N_OUT; ++k) v.uinv[k] = st[4 + k] > 0. ? 1. / st[4 + k] : 0.;
    const double c[N_CST] = {std::log10(std::max(so, 1e-300)), mc, std::log10(std::max(st[4], 1e-300)),
                             std::log10(std::max(st[5], 1e-300))};
    for (int k = 0; k < N_CST; ++k) v.cst[k] = (float)((c[k] - nm.c_mean[k]) / nm.c_scale[k]);
    v.jump = (double)nm.jump;
    for (int k = 0; k < N_OUT; ++k) {
        v.out_scale[k] = v.jump * st[4 + k] * (double)nm.gain[k];
        v.out_inv[k] = v.out_scale[k] > 0. ? 1. / v.out_scale[k] : 0.;
    }
    return v;
}

// sums s[0..3] = sum O, sum O^2, sum Cr, sum Cr^2 and s[4..6] = sum u_c^2 -> statistics
void stat_from_sums(const double s[N_STAT], double V, double st[N_STAT]) {
    for (int f = 0; f < 2; ++f) {
        const double m = s[2 * f] / V;
        st[2 * f] = m;
        st[2 * f + 1] = std::sqrt(std::max(0., s[2 * f + 1] / V - m * m));
    }
    for (int k = 0; k < N_OUT; ++k) st[4 + k] = std::sqrt(s[4 + k] / V);
}

// statistics of a binary64 volume and its update on the device
int device_stats(Net* n, const double* o, const double* cr, double* const u[N_OUT], long long V, double st[N_STAT]) {
    HIPCHK(hipMemsetAsync(n->dacc, 0, 10 * 8, n->stream));
    NN_LAUNCH(k_moments_d, dim3(512), dim3(256), 0, n->stream, V, o, n->dacc);
    NN_LAUNCH(k_moments_d, dim3(512), dim3(256), 0, n->stream, V, cr, n->dacc + 2);
    for (int k = 0; k < N_OUT; ++k)
        NN_LAUNCH(k_moments_d, dim3(512), dim3(256), 0, n->stream, V, (const double*)u[k], n->dacc + 4 + 2 * k);
    double a[10];
    HIPCHK(hipMemcpyAsync(a, n->dacc, 10 * 8, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    const double s[N_STAT] = {a[0], a[1], a[2], a[3], a[5], a[7], a[9]};
    stat_from_sums(s, (double)V, st);
    return NICR_OK;
}

// network input and target of batch slot `slot` from training sample `si`,
// through the symmetry map m (the identity: the sample as recorded)
void load_sample(Net* n, int slot, int si, bool target, SampleMap m) {
    const long long V = n->V(0);
    const float* in = n->ds_in + (long long)si * N_IN * V;
    const VolNorm vn = vol_norm(n->norm, &n->ds_stat[N_STAT * (size_t)si]);
    NN_LAUNCH(k_input_map, dim3(nblk(V)), dim3(256), 0, n->stream, (int)V, in, m, vn,
                       n->x0 + (long long)slot * V * 32);
    if (target)
        NN_LAUNCH(k_target, dim3(nblk(V)), dim3(256), 0, n->stream, (int)V,
                           n->ds_d + (long long)si * N_OUT * V, in + 4 * V, m, vn, n->tgt + 4LL * slot * V);
}
SampleMap identity_map(const Net* n) { return SampleMap{n->S[1], n->S[2], 0, 0, 0, 0}; }

int check_ctx_shape(Net* n, Ctx* c) {
    if (c->dim != 3 || c->P.nx != n->S[0] || c->P.ny != n->S[1] || c->P.nz != n->S[2])
        return fail(NICR_E_INVALID, "surrogate: model mesh differs from the network's volume");
    if (c->prec != PREC_FP64) return fail(NICR_E_INVALID, "surrogate: the model must step in binary64");
    return NICR_OK;
}

// network input from the model's current state (binary64 O, Cr, phi, the
// mask, and the plain update held in n->du) with the volume statistics st;
// *vn receives the volume's normalisation
int input_from_ctx(Net* n, Ctx* c, int slot, const double st[N_STAT], VolNorm* vn) {
    if (c->mask_dirty) {
        do_mask(c);
        c->mask_dirty = false;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    const long long V = n->V(0);
    const int cur = c->cur;
    *vn = vol_norm(n->norm, st);
    bf16* x0 = n->x0 + (long long)slot * V * 32;
    if (c->te == 8)
        NN_LAUNCH((k_input<double, double, double>), dim3(nblk(V)), dim3(256), 0, n->stream, (int)V,
                           c->o[cur], c->cr[cur], c->p[cur], (const double*)c->gb[mask_index(c)], n->du[0], n->du[1],
                           n->du[2], *vn, x0);
    else
        NN_LAUNCH((k_input<double, float, double>), dim3(nblk(V)), dim3(256), 0, n->stream, (int)V,
                           c->o[cur], c->cr[cur], c->p[cur], (const float*)c->gb[mask_index(c)], n->du[0], n->du[1],
                           n->du[2], *vn, x0);
    return NICR_OK;
}

}  // namespace nn
}  // namespace nicr

using namespace nicr::nn;

extern "C" {

static int nicr_unet_create_impl(int s0, int s1, int s2, int max_batch, int train, int max_samples, uint64_t seed, void** net) {
    if (!net) return fail(NICR_E_INVALID, "nicr_unet_create: null handle");
    *net = nullptr;
    if (s0 <= 0 || s1 <= 0 || s2 <= 0 || s0 % 8 || s1 % 8 || s2 % 8)
        return fail(NICR_E_INVALID, "nicr_unet_create: volume sides must be positive multiples of 8");
    if (max_batch < 1 || max_samples < 0) return fail(NICR_E_INVALID, "nicr_unet_create: bad batch or sample count");
    const long long V = (long long)s0 * s1 * s2;
    if (V * max_batch * 384 >= (1LL << 40) || V * max_batch >= (1LL << 31) / 4)
        return fail(NICR_E_INVALID, "nicr_unet_create: volume x batch too large");
    Net* n = new Net();
    n->S[0] = s0;
    n->S[1] = s1;
    n->S[2] = s2;
    n->nbmax = max_batch;
    n->train = train != 0;
    int rc = 0;
    hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(NICR_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    if (!rc && (hipEventCreate(&n->ev[0]) != hipSuccess || hipEventCreate(&n->ev[1]) != hipSuccess))
        rc = fail(NICR_E_DEVICE, "hipEventCreate failed");
    if (!rc) rc = net_build(n);
    if (!rc && max_samples > 0) {
        n->ds_cap = max_samples;
        if ((rc = nalloc_t(n, &n->ds_in, (long long)max_samples * N_IN * V)) == 0)
            rc = nalloc_t(n, &n->ds_d, (long long)max_samples * N_OUT * V);
    }
    if (!rc) rc = init_params(n, seed);
    if (rc) {
        net_free(n);
        return rc;
    }
    *net = n;
    return NICR_OK;
}

int nicr_unet_destroy(void* h) {
    if (h) net_free(static_cast<Net*>(h));
    return NICR_OK;
}

int nicr_unet_param_count(void* h, long long* count) {
    Net* n = static_cast<Net*>(h);
    if (!n || !count) return fail(NICR_E_INVALID, "null argument");
    *count = n->nparam;
    return NICR_OK;
}

int nicr_unet_device_bytes(void* h, size_t* bytes) {
    Net* n = static_cast<Net*>(h);
    if (!n || !bytes) return fail(NICR_E_INVALID, "null argument");
    *bytes = n->bytes;
    return NICR_OK;
}

int nicr_unet_get_params(void* h, float* host) {
    Net* n = static_cast<Net*>(h);
    if (!n || !host) return fail(NICR_E_INVALID, "null argument");
    HIPCHK(hipMemcpyAsync(host, n->param, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return NICR_OK;
}

static int nicr_unet_set_params_impl(void* h, const float* host) {
    Net* n = static_cast<Net*>(h);
    if (!n || !host) return fail(NICR_E_INVALID, "null argument");
    HIPCHK(hipMemcpyAsync(n->param, host, n->nparam * 4, hipMemcpyHostToDevice, n->stream));
    cast_weights(n);
    HIPCHK(hipStreamSynchronize(n->stream));
    n->adam_t = 0;
    if (n->train) {
        HIPCHK(hipMemsetAsync(n->am, 0, n->nparam * 4, n->stream));
        HIPCHK(hipMemsetAsync(n->av, 0, n->nparam * 4, n->stream));
        HIPCHK(hipMemsetAsync(n->clip, 0, 4 * 8, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
    }
    return NICR_OK;
}

int nicr_unet_get_grads(void* h, float* host) {
    Net* n = static_cast<Net*>(h);
    if (!n || !host) return fail(NICR_E_INVALID, "null argument");
    HIPCHK(hipMemcpyAsync(host, n->grad, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return NICR_OK;
}

int nicr_unet_set_norm(void* h, const nicr_unet_norm* nm) {
    Net* n = static_cast<Net*>(h);
    if (!n || !nm) return fail(NICR_E_INVALID, "null argument");
    if (!(nm->phi_scale > 0.f && nm->gb_scale > 0.f)) return fail(NICR_E_INVALID, "nicr_unet_set_norm: scales must be > 0");
    for (int k = 0; k < N_CST; ++k)
        if (!(nm->c_scale[k] > 0.f)) return fail(NICR_E_INVALID, "nicr_unet_set_norm: scales must be > 0");
    for (int k = 0; k < N_OUT; ++k)
        if (!(nm->gain[k] >= 0.f)) return fail(NICR_E_INVALID, "nicr_unet_set_norm: gains must be >= 0");
    if (nm->stat_planes < 0) return fail(NICR_E_INVALID, "nicr_unet_set_norm: stat_planes must be >= 0");
    n->norm = *nm;
    return NICR_OK;
}

int nicr_unet_get_norm(void* h, nicr_unet_norm* nm) {
    Net* n = static_cast<Net*>(h);
    if (!n || !nm) return fail(NICR_E_INVALID, "null argument");
    *nm = n->norm;
    return NICR_OK;
}

static int nicr_unet_forward_impl(void* h, int nb, const float* in, float* out) {
    Net* n = static_cast<Net*>(h);
    if (!n || !in || !out) return fail(NICR_E_INVALID, "null argument");
    if (nb < 1 || nb > n->nbmax) return fail(NICR_E_INVALID, "nicr_unet_forward: batch exceeds the network's maximum");
    const long long V = n->V(0);
    float* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, (size_t)V * N_IN * 4));
    int rc = NICR_OK;
    for (int b = 0; b < nb && !rc; ++b) {
        const float* src = in + (size_t)b * N_IN * V;
        double s[N_STAT] = {0., 0., 0., 0., 0., 0., 0.}, st[N_STAT];
        for (int f = 0; f < 2; ++f)
            for (long long p = 0; p < V; ++p) {
                const double v = src[(size_t)f * V + p];
                s[2 * f] += v;
                s[2 * f + 1] += v * v;
            }
        for (int k = 0; k < N_OUT; ++k)
            for (long long p = 0; p < V; ++p) {
                const double v = src[(size_t)(N_FIELD + k) * V + p];
                s[4 + k] += v * v;
            }
        stat_from_sums(s, (double)V, st);
        if (hipMemcpyAsync(tmp, src, (size_t)V * N_IN * 4, hipMemcpyHostToDevice, n->stream) != hipSuccess)
            rc = fail(NICR_E_DEVICE, "nicr_unet_forward: upload failed");
        NN_LAUNCH((k_input<float, float, float>), dim3(nblk(V)), dim3(256), 0, n->stream, (int)V, tmp,
                           tmp + V, tmp + 2 * V, tmp + 3 * V, tmp + 4 * V, tmp + 5 * V, tmp + 6 * V,
                           vol_norm(n->norm, st), n->x0 + (long long)b * V * 32);
        if (hipStreamSynchronize(n->stream) != hipSuccess) rc = fail(NICR_E_DEVICE, "nicr_unet_forward: input");
    }
    if (!rc) {
        forward(n, nb);
        std::vector<float> o4((size_t)4 * nb * V);
        if (hipMemcpyAsync(o4.data(), n->out, o4.size() * 4, hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
            hipStreamSynchronize(n->stream) != hipSuccess)
            rc = fail(NICR_E_DEVICE, "nicr_unet_forward: download failed");
        for (int b = 0; b < nb && !rc; ++b)
            for (int c = 0; c < N_OUT; ++c)
                for (long long p = 0; p < V; ++p) out[((size_t)b * N_OUT + c) * V + p] = o4[4 * ((size_t)b * V + p) + c];
    }
    hipFree(tmp);
    return rc;
}

static int nicr_unet_loss_backward_impl(void* h, int nb, const float* in, const float* target, double* loss) {
    Net* n = static_cast<Net*>(h);
    if (!n || !in || !target || !loss) return fail(NICR_E_INVALID, "null argument");
    if (!n->train) return fail(NICR_E_STATE, "nicr_unet_loss_backward: network created without training buffers");
    if (nb < 1 || nb > n->nbmax) return fail(NICR_E_INVALID, "nicr_unet_loss_backward: bad batch");
    std::vector<float> out((size_t)nb * N_OUT * n->V(0));
    int rc = nicr_unet_forward(h, nb, in, out.data());
    if (rc) return rc;
    const long long V = n->V(0);
    std::vector<float> t4((size_t)4 * nb * V, 0.f);
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < N_OUT; ++c)
            for (long long p = 0; p < V; ++p) t4[4 * ((size_t)b * V + p) + c] = target[((size_t)b * N_OUT + c) * V + p];
    HIPCHK(hipMemcpyAsync(n->tgt, t4.data(), t4.size() * 4, hipMemcpyHostToDevice, n->stream));
    HIPCHK(hipMemsetAsync(n->grad, 0, n->nparam * 4, n->stream));
    HIPCHK(hipMemsetAsync(n->dacc, 0, 64 * 8, n->stream));
    backward(n, nb, reinterpret_cast<float*>(n->dacc));
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, n->dacc, 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    *loss = l / (3.0 * (double)nb * V);
    return NICR_OK;
}

int nicr_unet_set_timing(void* h, int on) {
    Net* n = static_cast<Net*>(h);
    if (!n) return fail(NICR_E_INVALID, "null argument");
    n->timing = on != 0;
    n->fwd_ms_sum = 0.;
    n->fwd_count = 0;
    return NICR_OK;
}

int nicr_unet_forward_ms(void* h, double* ms) {
    Net* n = static_cast<Net*>(h);
    if (!n || !ms) return fail(NICR_E_INVALID, "null argument");
    *ms = n->fwd_count ? n->fwd_ms_sum / n->fwd_count : 0.;
    return NICR_OK;
}

// One convolution of the network's kernels on host arrays (float32 in, rounded
// to BF16 on upload): op 0 forward y = act(conv(x, w) + b) [P][cout]; op 1 data
// gradient dx = conv^T(dy, w) [P][cin]; op 2 weight gradient dw [27][cout][cin]
// from x and dy. x [P][cin], dy [P][cout], w [27][cout][cin], b [cout]; P = nb
// s0 s1 s2 positions in the network's cell order.
static int nicr_nn_conv_check_impl(int op, int nb, int s0, int s1, int s2, int cin, int cout, int relu, const float* x,
                       const float* w, const float* b, const float* dy, float* out) {
    if (op < 0 || op > 2 || nb < 1 || s0 < 1 || s1 < 1 || s2 < 1 || cin % 32 || cout % 32 || cin < 32 || cout < 32 ||
        !w || !out)
        return fail(NICR_E_INVALID, "nicr_nn_conv_check: bad argument");
    if ((op != 1 && !x) || (op != 0 && !dy) || (op == 0 && !b))
        return fail(NICR_E_INVALID, "nicr_nn_conv_check: missing operand");
    const long long P = (long long)nb * s0 * s1 * s2;
    if (P * std::max(cin, cout) >= (1LL << 31)) return fail(NICR_E_INVALID, "nicr_nn_conv_check: too large");
    Net tmp;
    Net* n = &tmp;
    HIPCHK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    n->S[0] = s0;
    n->S[1] = s1;
    n->S[2] = s2;
    const long long nw = 27LL * cin * cout;
    auto to_bf = [](const float* src, long long cnt) {
        std::vector<bf16> v((size_t)cnt);
        for (long long i = 0; i < cnt; ++i) v[(size_t)i] = (bf16)src[i];
        return v;
    };
    bf16 *dx = nullptr, *dyb = nullptr, *wb = nullptr, *wt = nullptr, *yb = nullptr;
    float *wf = nullptr, *bf = nullptr, *dwf = nullptr;
    int rc = 0;
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (!rc && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, n->stream) != hipSuccess)
            rc = fail(NICR_E_DEVICE, "nicr_nn_conv_check: upload");
    };
    if ((rc = nalloc_t(n, &wf, nw)) || (rc = nalloc_t(n, &wb, nw)) || (rc = nalloc_t(n, &wt, nw)) ||
        (rc = nalloc_t(n, &bf, cout)) || (rc = nalloc_t(n, &dx, P * cin)) || (rc = nalloc_t(n, &dyb, P * cout)) ||
        (rc = nalloc_t(n, &yb, P * cout)) || (rc = nalloc_t(n, &dwf, nw))) {
    } else {
        up(wf, w, nw * 4);
        if (!rc)
            NN_LAUNCH(k_cast_w, dim3(nblk(nw)), dim3(256), 0, n->stream, 27, cout, cin, wf, wb, wt);
        std::vector<bf16> hx, hdy;
        if (x) {
            hx = to_bf(x, P * cin);
            up(dx, hx.data(), hx.size() * 2);
        }
        if (dy) {
            hdy = to_bf(dy, P * cout);
            up(dyb, hdy.data(), hdy.size() * 2);
        }
        if (b) up(bf, b, cout * 4);
        std::vector<bf16> hb;
        if (!rc) {
            if (op == 0) {
                launch_conv(n, geo(n, nb, 0, cin, cout, 27), dx, wb, bf, yb, relu != 0);
                hb.resize((size_t)(P * cout));
                if (hipMemcpyAsync(hb.data(), yb, hb.size() * 2, hipMemcpyDeviceToHost, n->stream) != hipSuccess)
                    rc = fail(NICR_E_DEVICE, "download");
            } else if (op == 1) {
                launch_conv(n, geo(n, nb, 0, cout, cin, 27), dyb, wt, nullptr, dx, false);
                hb.resize((size_t)(P * cin));
                if (hipMemcpyAsync(hb.data(), dx, hb.size() * 2, hipMemcpyDeviceToHost, n->stream) != hipSuccess)
                    rc = fail(NICR_E_DEVICE, "download");
            } else {
                if (hipMemsetAsync(dwf, 0, nw * 4, n->stream) != hipSuccess) rc = fail(NICR_E_DEVICE, "memset");
                launch_wgrad(n, geo(n, nb, 0, cin, cout, 27), dx, dyb, dwf);
                if (hipMemcpyAsync(out, dwf, nw * 4, hipMemcpyDeviceToHost, n->stream) != hipSuccess)
                    rc = fail(NICR_E_DEVICE, "download");
            }
        }
        if (!rc && hipStreamSynchronize(n->stream) != hipSuccess) rc = fail(NICR_E_DEVICE, "nicr_nn_conv_check: run");
        if (!rc && hipGetLastError() != hipSuccess) rc = fail(NICR_E_DEVICE, "nicr_nn_conv_check: launch");
        for (size_t i = 0; !rc && op != 2 && i < hb.size(); ++i) out[i] = (float)hb[i];
    }
    for (void* p : n->allocs) hipFree(p);
    n->allocs.clear();
    hipStreamDestroy(n->stream);
    n->stream = nullptr;
    return rc;
}

static int sur_buffers(Net* n);

// ---- training set ----

static int nicr_sur_record_impl(void* h, void* model, int jump, nicr_step_report* rep) {
    Net* n = static_cast<Net*>(h);
    Ctx* c = static_cast<Ctx*>(model);
    if (!n || !c || jump < 1) return fail(NICR_E_INVALID, "nicr_sur_record: bad argument");
    int rc = check_ctx_shape(n, c);
    if (rc) return rc;
    if (n->ds_n >= n->ds_cap) return fail(NICR_E_STATE, "nicr_sur_record: training set is full");
    if (n->ds_n > 0 && jump != n->ds_jump)
        return fail(NICR_E_INVALID, "nicr_sur_record: every sample of a training set has the same jump");
    const long long V = n->V(0);
    if ((rc = sur_buffers(n))) return rc;
    if (rep) *rep = nicr_step_report{};
    // the plain step whose update enters the input
    nicr_step_report r{};
    if ((rc = nicr_step(c, 1, &r))) return rc;
    if (r.status != NICR_STEP_OK) {
        if (rep) *rep = r;
        return NICR_OK;
    }
    const int cur = c->cur;
    NN_LAUNCH(k_sub, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->o[cur], c->o[cur ^ 1], n->du[0]);
    NN_LAUNCH(k_sub, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->cr[cur], c->cr[cur ^ 1], n->du[1]);
    NN_LAUNCH(k_sub, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->p[cur], c->p[cur ^ 1], n->du[2]);
    float* in = n->ds_in + (long long)n->ds_n * N_IN * V;
    NN_LAUNCH(k_to_float<double>, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->o[cur], in);
    NN_LAUNCH(k_to_float<double>, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->cr[cur], in + V);
    NN_LAUNCH(k_to_float<double>, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->p[cur], in + 2 * V);
    if (c->te == 8)
        NN_LAUNCH(k_to_float<double>, dim3(nblk(V)), dim3(256), 0, n->stream, V,
                           (const double*)c->gb[mask_index(c)], in + 3 * V);
    else
        HIPCHK(hipMemcpyAsync(in + 3 * V, c->gb[mask_index(c)], V * 4, hipMemcpyDeviceToDevice, n->stream));
    for (int k = 0; k < N_OUT; ++k)
        NN_LAUNCH(k_to_float<double>, dim3(nblk(V)), dim3(256), 0, n->stream, V, (const double*)n->du[k],
                           in + (N_FIELD + k) * V);
    HIPCHK(hipMemcpyAsync(n->bk[0], c->o[cur], V * 8, hipMemcpyDeviceToDevice, n->stream));
    HIPCHK(hipMemcpyAsync(n->bk[1], c->cr[cur], V * 8, hipMemcpyDeviceToDevice, n->stream));
    HIPCHK(hipMemcpyAsync(n->bk[2], c->p[cur], V * 8, hipMemcpyDeviceToDevice, n->stream));
    double st[N_STAT];
    if ((rc = device_stats(n, c->o[cur], c->cr[cur], n->du, V, st))) return rc;
    if ((rc = nicr_step(c, jump, &r))) return rc;
    if (rep) *rep = r;
    if (r.status != NICR_STEP_OK) return NICR_OK;   // not recorded
    const int nc = c->cur;
    float* d = n->ds_d + (long long)n->ds_n * N_OUT * V;
    NN_LAUNCH(k_delta_float, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->o[nc], n->bk[0], d);
    NN_LAUNCH(k_delta_float, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->cr[nc], n->bk[1], d + V);
    NN_LAUNCH(k_delta_float, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->p[nc], n->bk[2], d + 2 * V);
    HIPCHK(hipStreamSynchronize(n->stream));
    n->ds_stat.resize(N_STAT * (size_t)(n->ds_n + 1));
    for (int k = 0; k < N_STAT; ++k) n->ds_stat[N_STAT * (size_t)n->ds_n + k] = st[k];
    n->ds_jump = jump;
    ++n->ds_n;
    return NICR_OK;
}

int nicr_sur_samples(void* h, int* count) {
    Net* n = static_cast<Net*>(h);
    if (!n || !count) return fail(NICR_E_INVALID, "null argument");
    *count = n->ds_n;
    return NICR_OK;
}

// Normalisation from the training set: phi and mask means and standard
// deviations over all samples, means and standard deviations of the volume
// statistics over the samples, and output gains that give each normalised
// correction unit mean square (0 for a field that never changes).
static int nicr_sur_fit_norm_impl(void* h) {
    Net* n = static_cast<Net*>(h);
    if (!n || n->ds_n < 1) return fail(NICR_E_STATE, "nicr_sur_fit_norm: empty training set");
    const long long V = n->V(0);
    const int S = n->ds_n;
    const double J = (double)n->ds_jump;
    // per sample: moments of phi and gb, squared corrections of the three fields
    constexpr int NM = 7;
    std::vector<double> m((size_t)S * NM);
    for (int s = 0; s < S; ++s) {
        HIPCHK(hipMemsetAsync(n->dacc, 0, NM * 8, n->stream));
        const float* in = n->ds_in + (long long)s * N_IN * V;
        for (int c = 0; c < 2; ++c)
            NN_LAUNCH(k_moments, dim3(256), dim3(256), 0, n->stream, V, in + (2 + c) * V, n->dacc + 2 * c);
        for (int c = 0; c < N_OUT; ++c)
            NN_LAUNCH(k_corr_sq, dim3(256), dim3(256), 0, n->stream, V,
                               n->ds_d + ((long long)s * N_OUT + c) * V, in + (N_FIELD + c) * V, J, n->dacc + 4 + c);
        HIPCHK(hipMemcpyAsync(&m[(size_t)s * NM], n->dacc, NM * 8, hipMemcpyDeviceToHost, n->stream));
    }
    HIPCHK(hipStreamSynchronize(n->stream));
    nicr_unet_norm nm{};
    auto mean_std = [](double s1, double s2, double cnt, float* mean, float* sd) {
        const double mu = s1 / cnt, var = std::max(0., s2 / cnt - mu * mu);
        *mean = (float)mu;
        *sd = var > 1e-30 ? (float)std::sqrt(var) : 1.f;
    };
    double a[4] = {0., 0., 0., 0.};
    for (int s = 0; s < S; ++s)
        for (int k = 0; k < 4; ++k) a[k] += m[(size_t)s * NM + k];
    mean_std(a[0], a[1], (double)V * S, &nm.phi_mean, &nm.phi_scale);
    mean_std(a[2], a[3], (double)V * S, &nm.gb_mean, &nm.gb_scale);
    double cs[N_CST][2] = {};
    for (int s = 0; s < S; ++s) {
        const double* st = &n->ds_stat[N_STAT * (size_t)s];
        const double c[N_CST] = {std::log10(std::max(st[1], 1e-300)), st[2], std::log10(std::max(st[4], 1e-300)),
                                 std::log10(std::max(st[5], 1e-300))};
        for (int k = 0; k < N_CST; ++k) {
            cs[k][0] += c[k];
            cs[k][1] += c[k] * c[k];
        }
    }
    for (int k = 0; k < N_CST; ++k) mean_std(cs[k][0], cs[k][1], (double)S, &nm.c_mean[k], &nm.c_scale[k]);
    double g[N_OUT] = {0., 0., 0.};
    long long cnt[N_OUT] = {0, 0, 0};
    for (int s = 0; s < S; ++s) {
        const double* st = &n->ds_stat[N_STAT * (size_t)s];
        for (int k = 0; k < N_OUT; ++k) {
            const double scale = J * st[4 + k];
            if (scale > 0.) {
                g[k] += m[(size_t)s * NM + 4 + k] / (scale * scale);
                cnt[k] += V;
            }
        }
    }
    for (int k = 0; k < N_OUT; ++k) {
        const double rms = cnt[k] ? std::sqrt(g[k] / (double)cnt[k]) : 0.;
        nm.gain[k] = rms > 1e-300 ? (float)rms : 0.f;
    }
    nm.jump = n->ds_jump;
    nm.stat_planes = n->S[0];
    return nicr_unet_set_norm(h, &nm);
}

// `iters` Adam iterations on random batches of `batch` training samples. The
// per-iteration losses (mean squared error of the normalised deltas) go to
// losses[iters] when given.
int nicr_sur_set_augment(void* h, int on) {
    Net* n = static_cast<Net*>(h);
    if (!n) return fail(NICR_E_INVALID, "nicr_sur_set_augment: null network");
    n->augment = on != 0;
    return NICR_OK;
}

static int nicr_sur_train_impl(void* h, int iters, int batch, float lr, uint64_t seed, float* losses) {
    Net* n = static_cast<Net*>(h);
    if (!n || iters < 0) return fail(NICR_E_INVALID, "nicr_sur_train: bad argument");
    if (!n->train) return fail(NICR_E_STATE, "nicr_sur_train: network created without training buffers");
    if (n->ds_n < 1) return fail(NICR_E_STATE, "nicr_sur_train: empty training set");
    if (batch < 1 || batch > n->nbmax) return fail(NICR_E_INVALID, "nicr_sur_train: bad batch");
    if (iters > n->loss_cap) {
        float* p = nullptr;
        int rc = nalloc_t(n, &p, iters);
        if (rc) return rc;
        n->loss_acc = p;
        n->loss_cap = iters;
    }
    if (iters == 0) return NICR_OK;
    HIPCHK(hipMemsetAsync(n->loss_acc, 0, (size_t)iters * 4, n->stream));
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> pick(0, n->ds_n - 1);
    std::uniform_int_distribution<int> shy(0, n->S[1] - 1), shz(0, n->S[2] - 1), flip(0, 1);
    for (int it = 0; it < iters; ++it) {
        for (int b = 0; b < batch; ++b) {
            const int si = pick(rng);
            SampleMap m = identity_map(n);
            if (n->augment) {
                m.sy = shy(rng);
                m.sz = shz(rng);
                m.fy = flip(rng);
                m.fz = flip(rng);
            }
            load_sample(n, b, si, true, m);
        }
        forward(n, batch);
        HIPCHK(hipMemsetAsync(n->grad, 0, n->nparam * 4, n->stream));
        backward(n, batch, n->loss_acc + it);
        adam_step(n, lr);
        if (g_launch.failed) break;   // reported by the entry point
    }
    HIPCHK(hipGetLastError());
    if (losses) {
        HIPCHK(hipMemcpyAsync(losses, n->loss_acc, (size_t)iters * 4, hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
        const double denom = 3.0 * (double)batch * n->V(0);
        for (int i = 0; i < iters; ++i) losses[i] = (float)(losses[i] / denom);
    }
    HIPCHK(hipStreamSynchronize(n->stream));
    return NICR_OK;
}

// Mean squared error of the normalised deltas over training samples [first,
// first + count) (no parameter update).
static int nicr_sur_eval_impl(void* h, int first, int count, double* mse) {
    Net* n = static_cast<Net*>(h);
    if (!n || !mse || first < 0 || count < 1 || first + count > n->ds_n)
        return fail(NICR_E_INVALID, "nicr_sur_eval: bad sample range");
    if (!n->train) return fail(NICR_E_STATE, "nicr_sur_eval: network created without training buffers");
    HIPCHK(hipMemsetAsync(n->dacc, 0, 64 * 8, n->stream));
    for (int s = first; s < first + count; ++s) {
        load_sample(n, 0, s, true, identity_map(n));
        forward(n, 1);
        HIPCHK(hipMemsetAsync(n->grad, 0, n->nparam * 4, n->stream));
        backward(n, 1, reinterpret_cast<float*>(n->dacc));
    }
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, n->dacc, 4, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    *mse = l / (3.0 * (double)count * n->V(0));
    return NICR_OK;
}

// ---- surrogate-accelerated stepping ----

static int sur_buffers(Net* n) {
    if (n->bk[0]) return NICR_OK;
    const long long V = n->V(0);
    int rc = 0;
    for (int i = 0; i < 3; ++i)
        if ((rc = nalloc_t(n, &n->bk[i], V))) return rc;
    for (int i = 0; i < N_OUT; ++i)
        if ((rc = nalloc_t(n, &n->du[i], V))) return rc;
    if ((rc = nalloc_t(n, &n->part, (long long)SUR_BLOCKS * 8))) return rc;
    n->part_host.resize((size_t)SUR_BLOCKS * 8);
    n->bk_cells = V;
    return NICR_OK;
}

static void nfree(Net* n, void* p, size_t bytes) {
    if (!p) return;
    auto it = std::find(n->allocs.begin(), n->allocs.end(), p);
    if (it != n->allocs.end()) n->allocs.erase(it);
    hipFree(p);
    n->bytes -= bytes;
}

// backups of the eta fields (the model's layout, padding included) and of the
// mask for jumps over evolving grains
static int sur_eta_buffers(Net* n, Ctx* c) {
    const long long need = (long long)c->cfg.n_eta * c->P.fstride;
    int rc = 0;
    if (n->bk_eta_elems < need) {
        nfree(n, n->bk_eta, (size_t)n->bk_eta_elems * 8);
        n->bk_eta = nullptr;
        n->bk_eta_elems = 0;
        if ((rc = nalloc_t(n, &n->bk_eta, need))) return rc;
        n->bk_eta_elems = need;
    }
    if (n->bk_mask_cells < c->ncell) {
        nfree(n, n->bk_mask, (size_t)n->bk_mask_cells * 3 * 8);
        n->bk_mask = nullptr;
        n->bk_mask_cells = 0;
        if ((rc = nalloc_t(n, &n->bk_mask, 3 * c->ncell))) return rc;
        n->bk_mask_cells = c->ncell;
    }
    return NICR_OK;
}

// block partial sums of a fused pass -> host (columns 0..nv-1); min columns
// (from `first_min` on) are reduced by min instead of sum
static int sur_partials(Net* n, int nv, long double* sum, int first_min = 8) {
    HIPCHK(hipMemcpyAsync(n->part_host.data(), n->part, (size_t)SUR_BLOCKS * 8 * 8, hipMemcpyDeviceToHost, n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    for (int k = 0; k < nv; ++k) sum[k] = k < first_min ? 0.L : (long double)INFINITY;
    for (int b = 0; b < SUR_BLOCKS; ++b)
        for (int k = 0; k < nv; ++k) {
            const double v = n->part_host[(size_t)b * 8 + k];
            if (k < first_min) sum[k] += (long double)v;
            else if (v < sum[k] || std::isnan(v)) sum[k] = v;
        }
    return NICR_OK;
}

// During nicr_sur_advance the net's spare eta buffer trades places with the
// context's current one at every jump. On every exit the context gets its own
// two buffers back (the state copied over if it sits in the spare) and the
// net its spare, so neither holds the other's memory between calls.
struct SpareEta {
    Net* n;
    Ctx* c;
    void*& spare;
    void* own0;
    void* own1;
    size_t bytes;
    ~SpareEta() {
        if (!bytes) return;
        for (int i = 0; i < 2; ++i)
            if (c->eta[i] != own0 && c->eta[i] != own1) {
                (void)hipStreamSynchronize(n->stream);
                (void)hipStreamSynchronize(c->stream);
                (void)hipMemcpy(spare, c->eta[i], bytes, hipMemcpyDeviceToDevice);
                std::swap(c->eta[i], spare);
            }
        n->bk_eta = static_cast<double*>(spare);
    }
};

static int nicr_sur_advance_impl(void* h, void* model, int nsteps, const nicr_sur_options* opt, nicr_sur_report* rep) {
    Net* n = static_cast<Net*>(h);
    Ctx* c = static_cast<Ctx*>(model);
    if (!n || !c || !opt || !rep || nsteps < 0) return fail(NICR_E_INVALID, "nicr_sur_advance: bad argument");
    int rc = check_ctx_shape(n, c);
    if (rc) return rc;
    // A jump replaces `jump` plain steps, and with them those steps' O and Cr
    // sources: the inlet (2D:408-411,435-442), legacy Ck (2D:567-571) and
    // clipping (2D:574-580). They are carried over the jump by the trapezoid
    // rule from the step before it and the check step after it, into the
    // budgets and, with conserve_mass, into the O and Cr totals. Ksp scans are
    // not run here: like nicr_step, a call never reaches one, and the caller
    // ends each call at the next scan (2D:1222).
    const int J = opt->jump;
    if (J < 2) return fail(NICR_E_INVALID, "nicr_sur_advance: jump must be >= 2");
    if (J != n->norm.jump) return fail(NICR_E_INVALID, "nicr_sur_advance: jump differs from the network's");
    if (!(opt->residual_tol > 0.)) return fail(NICR_E_INVALID, "nicr_sur_advance: residual_tol must be > 0");
    if ((rc = sur_buffers(n))) return rc;
    // evolving grains: eta jumps by linear extrapolation and is checked like O and Cr
    const bool evo = c->eta_evolves();
    if (evo && c->te != 8) return fail(NICR_E_INVALID, "nicr_sur_advance: eta must be binary64");
    if (evo && (rc = sur_eta_buffers(n, c))) return rc;
    void* spare = n->bk_eta;
    const SpareEta restore{n, c, spare, c->eta[0], c->eta[1], evo ? (size_t)c->cfg.n_eta * c->P.fstride * 8 : 0};
    const long long V = n->V(0);
    *rep = nicr_sur_report{};
    const bool conserve = opt->conserve_mass != 0;
    int left = nsteps;
    nicr_step_report r{};
    auto plain = [&](int k) -> int {
        if (k <= 0) return NICR_OK;
        int e = nicr_step(c, k, &r);
        if (e) return e;
        rep->plain_steps += r.steps_done;
        left -= r.steps_done;
        if (r.status != NICR_STEP_OK) {
            rep->step = r;
            rep->status = r.status;
        }
        return NICR_OK;
    };
    // Per-step sources: the change of each budget slot over one step (slots
    // nucO, nucCr, ckO, ckCr, clipO, clipCr, bcO, bcCr; even slots are O) and of
    // the clipped-cell count. src_last holds those of the last step taken,
    // which is also the step that gives u0.
    constexpr int NSLOT = 8;
    auto budget_at = [](const DevCtl& h, int k) { return (long double)h.budget[2 * k] + (long double)h.budget[2 * k + 1]; };
    auto sources = [&](const DevCtl& a, const DevCtl& b, long double* d) {
        for (int k = 0; k < NSLOT; ++k) d[k] = budget_at(b, k) - budget_at(a, k);
        d[NSLOT] = (long double)(b.clipped - a.clipped);
    };
    auto species_sum = [&](const long double* d, int sp) {
        long double t = 0.L;
        for (int k = sp; k < NSLOT; k += 2) t += d[k];
        return t;
    };
    long double src_last[NSLOT + 1] = {};
    // plain steps whose last one is followed by a jump: its sources are kept
    auto plain_tracked = [&](int k) -> int {
        if (k <= 0) return NICR_OK;
        int e = plain(k - 1);
        if (e || rep->status != NICR_STEP_OK) return e;
        DevCtl a, b;
        if ((e = sync_ctl(c, &a))) return e;
        if ((e = plain(1))) return e;
        if ((e = sync_ctl(c, &b))) return e;
        sources(a, b, src_last);
        return NICR_OK;
    };
    if (opt->settle_steps < 0) return fail(NICR_E_INVALID, "nicr_sur_advance: settle_steps must be >= 0");
    if (opt->fallback_steps < 0) return fail(NICR_E_INVALID, "nicr_sur_advance: fallback_steps must be >= 0");
    if (opt->min_jump < 0) return fail(NICR_E_INVALID, "nicr_sur_advance: min_jump must be >= 0");
    // shortest last jump of a call (none unless 2 <= min_jump < jump)
    const int min_jump = opt->min_jump >= 2 && opt->min_jump < J ? opt->min_jump : J;
    const int fallback = opt->fallback_steps > 0 ? opt->fallback_steps : J;
    // NICR_SUR_TRACE=1: one line per jump on stderr (steps left, jump length,
    // residuals of O, Cr, eta, the decision)
    const char* tr = std::getenv("NICR_SUR_TRACE");
    const bool trace = tr && tr[0] == '1';
    // the update u0 of a jump is the difference of the two state buffers, so
    // they must hold consecutive states: true after the check step of an
    // accepted jump and after plain steps, not on entry. On entry the first
    // cycle takes one plain step for them plus the settle steps: the state
    // handed in may itself follow a jump's check step (a run loop ends each
    // call at a Ksp scan, often right after a jump), and a proposal made one
    // step after a check step takes that step's transient into u0.
    bool consecutive = false;
    while (left > 0 && rep->status == NICR_STEP_OK) {
        const int pre = consecutive ? opt->settle_steps : 1 + opt->settle_steps;
        // this cycle's jump: a full one, or a last shorter one of the steps left
        int Jc = J;
        if (left < J + 1 + pre) {
            Jc = left - 1 - pre;
            if (Jc < min_jump) {
                if ((rc = plain(left))) return rc;
                break;
            }
        }
        // plain steps before the jump (none when the last step already gives u0)
        if ((rc = plain_tracked(pre))) return rc;
        if (rep->status != NICR_STEP_OK) break;
        DevCtl saved;
        if ((rc = sync_ctl(c, &saved))) return rc;   // also: the step has finished
        const int cur = c->cur;
        // one pass: u0, backup of x0, volume statistics and O/Cr totals
        const long long n_hold = c->P.enforce_layer ? (long long)n->S[1] * n->S[2] : 0;
        NN_LAUNCH(k_sur_pre, dim3(SUR_BLOCKS), dim3(256), 0, n->stream, V, n_hold, c->o[cur], c->cr[cur], c->p[cur],
                           c->o[cur ^ 1], c->cr[cur ^ 1], c->p[cur ^ 1], n->du[0], n->du[1], n->du[2], n->bk[0],
                           n->bk[1], n->bk[2], n->part);
        long double ps[7];
        if ((rc = sur_partials(n, 7, ps))) return rc;
        const double m0[2] = {(double)ps[0], (double)ps[2]};
        const double sums[N_STAT] = {(double)ps[0], (double)ps[1], (double)ps[2], (double)ps[3],
                                     (double)ps[4], (double)ps[5], (double)ps[6]};
        double st[N_STAT];
        // the network's statistics: over the whole volume, or over its first
        // stat_planes X planes (cells in X-major order, so a prefix)
        const long long nstat = n->norm.stat_planes > 0 && n->norm.stat_planes < n->S[0]
                                    ? (long long)n->norm.stat_planes * n->S[1] * n->S[2] : V;
        if (nstat < V) {
            if ((rc = device_stats(n, c->o[cur], c->cr[cur], n->du, nstat, st))) return rc;
        } else {
            stat_from_sums(sums, (double)V, st);
        }
        // proposal x0 + J u0 + network correction, in place in the current buffers
        VolNorm vn;
        if ((rc = input_from_ctx(n, c, 0, st, &vn))) return rc;
        if (Jc != J) {
            // a shorter jump: Jc u0 and the correction scaled by (Jc / J)^2
            const double r = (double)Jc / (double)J;
            vn.jump = (double)Jc;
            for (int k = 0; k < N_OUT; ++k) vn.out_scale[k] *= r * r;
        }
        forward(n, 1);
        NN_LAUNCH(k_sur_apply, dim3(SUR_BLOCKS), dim3(256), 0, n->stream, V, n->out, vn, n->du[0], n->du[1],
                           n->du[2], c->o[cur], c->cr[cur], c->p[cur], n->part);
        long double pa[8];
        if ((rc = sur_partials(n, 8, pa, 2))) return rc;
        if (n->timing) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, n->ev[0], n->ev[1]) == hipSuccess) {
                n->fwd_ms_sum += ms;
                ++n->fwd_count;
            }
        }
        c->refresh_valid = false;
        // O and Cr to their totals at x0 plus `jump` steps at the sources of
        // the step before the jump (a closed box without clipping: the totals
        // at x0); the check step's sources refine this below
        double f[2] = {1., 1.};
        const double s1[2] = {(double)pa[0], (double)pa[1]};
        const double target[2] = {(double)((long double)m0[0] + (long double)Jc * species_sum(src_last, 0)),
                                  (double)((long double)m0[1] + (long double)Jc * species_sum(src_last, 1))};
        if (conserve) {
            for (int k = 0; k < 2; ++k)
                if (s1[k] > 0. && target[k] > 0.) f[k] = target[k] / s1[k];
            NN_LAUNCH(k_scale2, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->o[cur], f[0], c->cr[cur], f[1]);
            HIPCHK(hipStreamSynchronize(n->stream));
        }
        const double omin = (double)pa[2] * f[0], omax = -(double)pa[3] * f[0];
        const double cmin = (double)pa[4] * f[1], cmax = -(double)pa[5] * f[1];
        const double pmin = (double)pa[6], pmax = -(double)pa[7];
        const double lo = opt->min_value, hi = opt->max_value;
        bool ok = std::isfinite(s1[0]) && std::isfinite(s1[1]) && omin >= lo && omax <= hi && cmin >= lo &&
                  cmax <= hi && pmin >= lo && pmax <= hi;
        if (evo) {
            // eta_h into the spare buffer, which then trades places with the
            // current one: the context steps from eta_h and the spare holds
            // eta_0 for the residual and a rollback, with no copy. Its mask in
            // the current mask buffers (the old one backed up); eta_-1 (the
            // other buffer) is overwritten by the check step
            const Par& P = c->P;
            NN_LAUNCH(k_sur_eta_pre, dim3(SUR_BLOCKS), dim3(256), 0, n->stream, c->ncell, P.fstride,
                               c->cfg.n_eta, (double)Jc, P.eta_scale, (const double*)c->eta[cur ^ 1],
                               (const double*)c->eta[cur], (double*)spare, (double*)c->s2[cur],
                               (double*)c->pairs[cur], (double*)c->gb[cur], n->bk_mask, n->part);
            std::swap(c->eta[cur], spare);
            long double pe[2];
            if ((rc = sur_partials(n, 2, pe, 0))) return rc;
            ok = ok && (double)pe[0] >= lo && -(double)pe[1] <= hi;
        }
        double ro = INFINITY, rcr = INFINITY, reta = evo ? INFINITY : 0.;
        if (ok) {
            nicr_step_report rr{};
            if ((rc = nicr_step(c, 1, &rr))) return rc;
            if (rr.status != NICR_STEP_OK) ok = false;
        }
        if (ok) {
            const int k = c->cur;   // proposal in k ^ 1, one step after it in k
            NN_LAUNCH(k_sur_trap, dim3(SUR_BLOCKS), dim3(256), 0, n->stream, V, c->o[k ^ 1], c->o[k],
                               n->bk[0], n->du[0], c->cr[k ^ 1], c->cr[k], n->bk[1], n->du[1], 0.5 * Jc, n->part);
            long double pt[4];
            if ((rc = sur_partials(n, 4, pt))) return rc;
            ro = pt[1] > 0 ? (double)std::sqrt(pt[0] / pt[1]) : INFINITY;
            rcr = pt[3] > 0 ? (double)std::sqrt(pt[2] / pt[3]) : INFINITY;
            if (evo) {
                NN_LAUNCH(k_sur_eta_trap, dim3(SUR_BLOCKS), dim3(256), 0, n->stream, c->ncell,
                                   c->P.fstride, c->cfg.n_eta, (double)Jc, (const double*)c->eta[k ^ 1],
                                   (const double*)c->eta[k], (const double*)spare, n->part);
                long double pr[2];
                if ((rc = sur_partials(n, 2, pr))) return rc;
                reta = pr[1] > 0 ? (double)std::sqrt(pr[0] / pr[1]) : (pr[0] == 0 ? 0. : INFINITY);
            }
            if (!(ro <= opt->residual_tol && rcr <= opt->residual_tol && reta <= opt->residual_tol)) ok = false;
        }
        rep->last_residual[0] = ro;
        rep->last_residual[1] = rcr;
        rep->last_residual[2] = reta;
        if (trace)
            std::fprintf(stderr, "[sur] left %d jump %d pre %d residual O %.4g Cr %.4g eta %.4g %s\n", left, Jc, pre,
                         ro, rcr, reta, ok ? "accepted" : "rejected");
        if (ok) {
            // the jump's sources by the trapezoid rule over its `jump` steps,
            // into the budgets; with conserve_mass the state after the check
            // step takes the difference to the first-order target above, so
            // the totals stay consistent with the budgets
            DevCtl after;
            if ((rc = sync_ctl(c, &after))) return rc;
            long double src_check[NSLOT + 1];
            sources(saved, after, src_check);
            long double add[NSLOT + 1];
            for (int k = 0; k <= NSLOT; ++k) add[k] = 0.5L * (long double)Jc * (src_last[k] + src_check[k]);
            bool changed = false;
            for (int k = 0; k < NSLOT; ++k) {
                if (add[k] == 0.L) continue;
                const long double v = budget_at(after, k) + add[k];
                after.budget[2 * k] = (double)v;
                after.budget[2 * k + 1] = (double)(v - (long double)after.budget[2 * k]);
                changed = true;
            }
            const long long nclip = (long long)std::llround((double)add[NSLOT]);
            if (nclip) {
                after.clipped += nclip;
                changed = true;
            }
            if (changed) {
                HIPCHK(hipMemcpyAsync(c->ctl, &after, sizeof(after), hipMemcpyHostToDevice, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            if (conserve) {
                double g[2] = {1., 1.};
                for (int sp = 0; sp < 2; ++sp) {
                    const long double now = (long double)target[sp] + species_sum(src_check, sp);
                    const long double want = (long double)m0[sp] + species_sum(add, sp) + species_sum(src_check, sp);
                    if (now > 0.L && want > 0.L && want != now) g[sp] = (double)(want / now);
                }
                if (g[0] != 1. || g[1] != 1.) {
                    const int k = c->cur;
                    NN_LAUNCH(k_scale2, dim3(nblk(V)), dim3(256), 0, n->stream, V, c->o[k], g[0], c->cr[k],
                                       g[1]);
                    HIPCHK(hipStreamSynchronize(n->stream));
                }
            }
            for (int k = 0; k <= NSLOT; ++k) src_last[k] = src_check[k];
            ++rep->jumps_accepted;
            rep->max_residual_o = std::max(rep->max_residual_o, ro);
            rep->max_residual_cr = std::max(rep->max_residual_cr, rcr);
            rep->max_residual_eta = std::max(rep->max_residual_eta, reta);
            rep->plain_steps += 1;
            left -= Jc + 1;
            consecutive = true;
            continue;
        }
        // reject: restore the control block (state index, budgets) and the
        // fields before the jump (static eta and mask are not touched by a
        // jump; a check step wrote only the other buffers), then take the
        // plain steps
        ++rep->jumps_rejected;
        const double rmax = std::max(std::max(ro, rcr), reta);
        if (std::isfinite(rmax))
            rep->min_rejected_residual = rep->jumps_rejected == 1 ? rmax : std::min(rep->min_rejected_residual, rmax);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpyAsync(c->ctl, &saved, sizeof(saved), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->cur = saved.cur;
        const int cb = c->cur;
        HIPCHK(hipMemcpyAsync(c->o[cb], n->bk[0], V * 8, hipMemcpyDeviceToDevice, n->stream));
        HIPCHK(hipMemcpyAsync(c->cr[cb], n->bk[1], V * 8, hipMemcpyDeviceToDevice, n->stream));
        HIPCHK(hipMemcpyAsync(c->p[cb], n->bk[2], V * 8, hipMemcpyDeviceToDevice, n->stream));
        if (evo) {
            std::swap(c->eta[cb], spare);   // eta_0 back into the context
            HIPCHK(hipMemcpyAsync(c->s2[cb], n->bk_mask, V * 8, hipMemcpyDeviceToDevice, n->stream));
            HIPCHK(hipMemcpyAsync(c->pairs[cb], n->bk_mask + V, V * 8, hipMemcpyDeviceToDevice, n->stream));
            HIPCHK(hipMemcpyAsync(c->gb[cb], n->bk_mask + 2 * V, V * 8, hipMemcpyDeviceToDevice, n->stream));
        }
        HIPCHK(hipStreamSynchronize(n->stream));
        c->refresh_valid = false;
        if ((rc = plain_tracked(std::min(fallback, left)))) return rc;
        consecutive = true;
    }
    rep->steps_done = nsteps - left;
    return NICR_OK;
}

// Entry points that launch kernels: the launch checks of the call decide its
// result when the call itself succeeded.
int nicr_unet_create(int s0, int s1, int s2, int max_batch, int train, int max_samples, uint64_t seed, void** net) {
    nicr::nn::g_launch = {};
    const int rc = nicr_unet_create_impl(s0, s1, s2, max_batch, train, max_samples, seed, net);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_unet_set_params(void* h, const float* host) {
    nicr::nn::g_launch = {};
    const int rc = nicr_unet_set_params_impl(h, host);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_unet_forward(void* h, int nb, const float* in, float* out) {
    nicr::nn::g_launch = {};
    const int rc = nicr_unet_forward_impl(h, nb, in, out);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_unet_loss_backward(void* h, int nb, const float* in, const float* target, double* loss) {
    nicr::nn::g_launch = {};
    const int rc = nicr_unet_loss_backward_impl(h, nb, in, target, loss);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_nn_conv_check(int op, int nb, int s0, int s1, int s2, int cin, int cout, int relu, const float* x,
                       const float* w, const float* b, const float* dy, float* out) {
    nicr::nn::g_launch = {};
    const int rc = nicr_nn_conv_check_impl(op, nb, s0, s1, s2, cin, cout, relu, x, w, b, dy, out);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_sur_record(void* h, void* model, int jump, nicr_step_report* rep) {
    nicr::nn::g_launch = {};
    const int rc = nicr_sur_record_impl(h, model, jump, rep);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_sur_fit_norm(void* h) {
    nicr::nn::g_launch = {};
    const int rc = nicr_sur_fit_norm_impl(h);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_sur_train(void* h, int iters, int batch, float lr, uint64_t seed, float* losses) {
    nicr::nn::g_launch = {};
    const int rc = nicr_sur_train_impl(h, iters, batch, lr, seed, losses);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_sur_eval(void* h, int first, int count, double* mse) {
    nicr::nn::g_launch = {};
    const int rc = nicr_sur_eval_impl(h, first, count, mse);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}
int nicr_sur_advance(void* h, void* model, int nsteps, const nicr_sur_options* opt, nicr_sur_report* rep) {
    nicr::nn::g_launch = {};
    const int rc = nicr_sur_advance_impl(h, model, nsteps, opt, rep);
    const int lc = nicr::nn::launch_status();
    return rc ? rc : lc;
}

}  // extern "C"
