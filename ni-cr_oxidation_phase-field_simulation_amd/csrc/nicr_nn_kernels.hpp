// nicr_nn_kernels.hpp -- BF16 / FP32-accumulate kernels of the learned surrogate
// (CDNA4 MFMA), for a 3D U-Net on channels-last tensors.
//
// Layout: a tensor of nb volumes of s0 x s1 x s2 cells (s2 fastest) with C
// channels is [position][channel], channel fastest, position p = ((n*s0 + i0)*s1
// + i1)*s2 + i2 -- the state's own cell order ((x*ny + y)*nz + z). Channel
// counts of convolutions are multiples of 32. Dimension 0 (X) is zero padded;
// dimensions 1 and 2 (Y, Z) are periodic like the model's mesh.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16: M = positions,
// N = output channels, K = taps x input channels. Weights are [tap][cout][cin]
// (the B operand, K contiguous).
//   forward        Y[p][co] = act(b[co] + sum_{tap,ci} X[p + off(tap)][ci] W[tap][co][ci])
//   data gradient  the forward kernel on dY with W'[tap][ci][co] = W[26 - tap][co][ci]
//   weight grad.   dW[tap][co][ci] += sum_p dY[p][co] X[p + off(tap)][ci]  (K = positions)
// MFMA operand lane layout (verified on gfx950): lane l holds A row l % 32 and B
// column l % 32 with k = (l / 32) * 8 .. + 7; accumulator element r is row
// (r / 4) * 8 + (l / 32) * 4 + r % 4, column l % 32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nicr {
namespace nn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct Geo {
    int nb, s0, s1, s2;   // volumes and their cells (s2 fastest)
    int cin, cout;        // multiples of 32
    int taps;             // 27 (3x3x3) or 1
    // k_conv_halo only: input channels 0 .. cu-1 come from a coarse tensor of cu
    // channels (nearest upsampling: the cell's parent at half the sides), the
    // rest from X with cin - cu channels (0: X alone)
    int cu = 0;
    __host__ __device__ int positions() const { return nb * s0 * s1 * s2; }
};

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Neighbour rows of position p for the 27 taps of a 3x3x3 stencil (tap t has
// offsets t/9 - 1, (t/3)%3 - 1, t%3 - 1): the volume base, the three dimension-0
// rows (-1 where outside the volume) and the wrapped dimension-1/2 offsets.
struct Nbr {
    int base;       // n * s0 * s1 * s2
    int r0[3];      // (i0 + d0) * s1 * s2, or -1 outside
    int r1[3];      // wrapped (i1 + d1) * s2
    int r2[3];      // wrapped i2 + d2
    __device__ __forceinline__ void init(const Geo& g, int p) {
        const int vol = g.s0 * g.s1 * g.s2, plane = g.s1 * g.s2;
        const int n = p / vol, rem = p - n * vol;
        const int i0 = rem / plane, rem2 = rem - i0 * plane;
        const int i1 = rem2 / g.s2, i2 = rem2 - i1 * g.s2;
        base = n * vol;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int j0 = i0 + d - 1;
            r0[d] = (j0 >= 0 && j0 < g.s0) ? j0 * plane : -1;
            int j1 = i1 + d - 1;
            j1 = j1 < 0 ? j1 + g.s1 : (j1 >= g.s1 ? j1 - g.s1 : j1);
            r1[d] = j1 * g.s2;
            int j2 = i2 + d - 1;
            j2 = j2 < 0 ? j2 + g.s2 : (j2 >= g.s2 ? j2 - g.s2 : j2);
            r2[d] = j2;
        }
    }
    // neighbour position of tap t, or -1 (zero padding)
    __device__ __forceinline__ int at(int taps, int t) const {
        if (taps == 1) return base + r0[1] + r1[1] + r2[1];
        const int a = t / 9, b = (t / 3) % 3, c = t % 3;
        return r0[a] < 0 ? -1 : base + r0[a] + r1[b] + r2[c];
    }
};

constexpr int CBM = 128;   // positions per block
constexpr int CKC = 32;    // K per LDS stage (two MFMA k-steps)
constexpr int CLD = 40;    // LDS row pitch in bf16 (80 bytes: 16-byte aligned rows)

// Forward convolution, bias and optional ReLU. BN: output channels per block (32
// or 64). Four waves: BN = 64 -> 2 x 2 waves of 64 x 32 outputs; BN = 32 -> 4
// waves of 32 x 32. Grid: ceil(P / 128) * (cout / BN) blocks.
template <int BN, bool RELU>
__global__ __launch_bounds__(256) void k_conv(Geo g, const bf16* __restrict__ X, const bf16* __restrict__ W,
                                              const float* __restrict__ bias, bf16* __restrict__ Y) {
    constexpr int MR = BN == 64 ? 2 : 1;   // 32-row MFMA blocks per wave
    __shared__ __align__(16) bf16 sA[2][CBM * CLD];
    __shared__ __align__(16) bf16 sB[2][BN * CLD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = BN == 64 ? (wave >> 1) : wave, wn = BN == 64 ? (wave & 1) : 0;
    const int P = g.positions();
    const int nco = g.cout / BN;
    const int pb = (blockIdx.x / nco) * CBM;
    const int co0 = (blockIdx.x % nco) * BN;
    // A loads: position pa (two threads per position, 16 channels each)
    const int pa = tid >> 1, ha = tid & 1;
    const int p = pb + pa;
    const bool pvalid = p < P;
    Nbr nb;
    nb.init(g, pvalid ? p : 0);
    // B loads: output channel row tid >> 2, 8 channels of k at (tid & 3) * 8
    const bool bload = tid < BN * 4;
    const int kcs = g.cin / CKC;
    const int nk = g.taps * kcs;
    uint4 a0, a1, b0;
    auto load_stage = [&](int s) {
        const int tap = s / kcs, ci0 = (s - tap * kcs) * CKC;
        const int q = nb.at(g.taps, tap);
        if (pvalid && q >= 0) {
            const uint4* src = reinterpret_cast<const uint4*>(X + (long long)q * g.cin + ci0 + 16 * ha);
            a0 = src[0];
            a1 = src[1];
        } else {
            a0 = a1 = make_uint4(0u, 0u, 0u, 0u);
        }
        if (bload)
            b0 = *reinterpret_cast<const uint4*>(W + ((long long)tap * g.cout + co0 + (tid >> 2)) * g.cin + ci0 +
                                                 8 * (tid & 3));
    };
    auto store_stage = [&](int buf) {
        uint4* da = reinterpret_cast<uint4*>(&sA[buf][pa * CLD + 16 * ha]);
        da[0] = a0;
        da[1] = a1;
        if (bload) *reinterpret_cast<uint4*>(&sB[buf][(tid >> 2) * CLD + 8 * (tid & 3)]) = b0;
    };
    f32x16 acc[MR];
#pragma unroll
    for (int r = 0; r < MR; ++r) acc[r] = f32x16{};
    load_stage(0);
    store_stage(0);
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) load_stage(s + 1);
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int koff = kk * 16 + (lane >> 5) * 8;
            const bf16x8 b = as_bf16x8(*reinterpret_cast<const uint4*>(&sB[buf][(wn * 32 + (lane & 31)) * CLD + koff]));
#pragma unroll
            for (int r = 0; r < MR; ++r) {
                const int row = wm * MR * 32 + r * 32 + (lane & 31);
                const bf16x8 a = as_bf16x8(*reinterpret_cast<const uint4*>(&sA[buf][row * CLD + koff]));
                acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[r], 0, 0, 0);
            }
        }
        if (s + 1 < nk) store_stage(buf ^ 1);
        __syncthreads();
    }
    const int co = co0 + wn * 32 + (lane & 31);
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < MR; ++r)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int q = pb + wm * MR * 32 + r * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
            if (q < P) {
                float v = acc[r][e] + bv;
                if (RELU) v = v > 0.f ? v : 0.f;
                Y[(long long)q * g.cout + co] = (bf16)v;
            }
        }
}

// Halo-tile forward convolution for volumes whose sides are multiples of
// (4, 16, 16). A block of 8 waves owns a run of units (output tile of 4 x 16 x
// 16 positions, 32 output channels); per stage (unit, 32-channel input chunk)
// it holds the 6 x 18 x 18 halo of the tile in LDS, so every tap reads its A
// rows from LDS instead of HBM/L2. The weights (B operand, 1 KB per
// tap and K step, shared by every block) are read from L2 straight into
// registers one 3-tap group ahead, so LDS carries only A: each wave computes four
// 32-position row-blocks per B fragment. LDS rows are 64 bytes (32 channels)
// at a pitch of 80 bytes: 16-byte reads of consecutive rows fall in distinct
// bank groups (5 row mod 16 and mod 8 are both permutations), and within each
// of the 9 (x, y) tap groups every z tap's and row-block's address is the
// group base plus a compile-time offset, so the reads are ds_read_b128 with
// immediate offsets, the A fragments of the next step loading while the
// current step's MFMAs run. Wave w computes 4 consecutive row-blocks (2 y rows
// x 16 z each) of one x plane of the tile.
constexpr int HT_X = 4, HT_Y = 16, HT_Z = 16;
constexpr int HH_X = HT_X + 2, HH_Y = HT_Y + 2, HH_Z = HT_Z + 2;
constexpr int H_ROWS = HH_X * HH_Y * HH_Z;      // 1944 halo positions
constexpr int H_PITCH = 80;                     // bytes per LDS row (64 used)
constexpr int HC_LDS_BYTES = H_ROWS * H_PITCH;  // 155520
constexpr int HC_RB = HT_X * HT_Y / 2;          // 32-position row-blocks per tile

constexpr int HC_THREADS = 512;                 // 8 waves
constexpr int H_LOADS = (H_ROWS * 4 + HC_THREADS - 1) / HC_THREADS;
constexpr int HC_MT = HC_RB / (HC_THREADS / 64);  // row-blocks per wave: 4

// HALFK: the input has one 32-channel chunk whose channels 16..31 are zero
// (the network's input tensor), so only the first K step of each tap is
// computed and only the lower half of each row is loaded; the skipped MFMAs
// would add exact zeros.
template <bool RELU, bool HALFK = false>
__global__ __launch_bounds__(HC_THREADS) void k_conv_halo(Geo g, const bf16* __restrict__ X,
                                                          const bf16* __restrict__ W, const float* __restrict__ bias,
                                                          bf16* __restrict__ Y, int units, int upb,
                                                          const bf16* __restrict__ Xc) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntx = g.s0 / HT_X, nty = g.s1 / HT_Y, ntz = g.s2 / HT_Z;
    const int nco = g.cout / 32, nch = g.cin / 32;
    const int u0 = blockIdx.x * upb, u1 = min(u0 + upb, units);
    if (u0 >= u1) return;
    const int nstage = (u1 - u0) * nch;
    const int plane = g.s1 * g.s2;
    uint4 rh[H_LOADS];
    // global loads of the halo of stage s into registers
    auto load = [&](int s) {
        const int u = u0 + s / nch, ch = s % nch;
        int t = u / nco;
        const int tz = t % ntz;
        t /= ntz;
        const int ty = t % nty;
        t /= nty;
        const int tx = t % ntx, n = t / ntx;
        const int x0 = tx * HT_X - 1, y0 = ty * HT_Y - 1, z0 = tz * HT_Z - 1;
        // the chunk's source: the coarse tensor (upsampled) or X
        const bool coarse = ch * 32 < g.cu;
        const int h1 = g.s1 >> 1, h2 = g.s2 >> 1;
        const long long vbase = coarse ? (long long)n * (g.s0 >> 1) * h1 * h2 : (long long)n * g.s0 * plane;
        const bf16* src = coarse ? Xc + ch * 32 : X + (ch * 32 - g.cu);
        const int stride = coarse ? g.cu : g.cin - g.cu;
#pragma unroll
        for (int i = 0; i < H_LOADS; ++i) {
            const int idx = tid + i * HC_THREADS;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (idx < H_ROWS * 4 && (!HALFK || (idx & 3) < 2)) {
                const int row = idx >> 2, q = idx & 3;
                const int hz = row % HH_Z, hy = (row / HH_Z) % HH_Y, hx = row / (HH_Z * HH_Y);
                const int gx = x0 + hx;
                int gy = y0 + hy, gz = z0 + hz;
                gy = gy < 0 ? gy + g.s1 : (gy >= g.s1 ? gy - g.s1 : gy);
                gz = gz < 0 ? gz + g.s2 : (gz >= g.s2 ? gz - g.s2 : gz);
                if (gx >= 0 && gx < g.s0) {
                    const long long pos = coarse ? vbase + ((long long)(gx >> 1) * h1 + (gy >> 1)) * h2 + (gz >> 1)
                                                 : vbase + (long long)gx * plane + gy * g.s2 + gz;
                    v = *reinterpret_cast<const uint4*>(src + pos * stride + q * 8);
                }
            }
            rh[i] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < H_LOADS; ++i) {
            const int idx = tid + i * HC_THREADS;
            if (idx < H_ROWS * 4) *reinterpret_cast<uint4*>(lds + (idx >> 2) * H_PITCH + (idx & 3) * 16) = rh[i];
        }
    };
    // this lane's rows: position r = lane % 32 of the wave's row-blocks; its
    // LDS base covers row-block 0, tap (0, 0, 0), K quad hq
    const int wx = (wave * HC_MT) / (HT_Y / 2), wy = 2 * ((wave * HC_MT) % (HT_Y / 2));
    const int r = lane & 31, hq = lane >> 5;
    const unsigned char* abase = lds + ((wx * HH_Y + wy + (r >> 4)) * HH_Z + (r & 15)) * H_PITCH + hq * 16;
    const long long wtap = (long long)g.cout * g.cin;   // weight elements per tap
    f32x16 acc[HC_MT];
#pragma unroll
    for (int m = 0; m < HC_MT; ++m) acc[m] = f32x16{};
    load(0);
    store();
    __syncthreads();
    for (int s = 0; s < nstage; ++s) {
        const bool more = s + 1 < nstage;
        const int u = u0 + s / nch, ch = s % nch;
        const int cob = u % nco;
        // B fragment of (tap, k step kk) for this lane: 8 channels of k of output channel cob * 32 + r
        const bf16* wl = W + (long long)(cob * 32 + r) * g.cin + ch * 32 + hq * 8;
        // weights of the 3 taps of a group (c = 0, 1, 2; K steps 0, 1), loaded
        // one group ahead: the next group's fragments are in flight for the 24
        // MFMAs of the current group
        constexpr int KS = HALFK ? 1 : 2;   // K steps of 16 channels per tap
        auto load_group = [&](int ab, uint4 (&w)[3][2]) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
                    w[c][kk] = *reinterpret_cast<const uint4*>(wl + (ab * 3 + c) * wtap + kk * 16);
        };
        uint4 bw[3][2], bn[3][2];
        load_group(0, bw);
        // taps in 9 groups ab = 3a + b of 3 (c = 0, 1, 2); inside a group the A
        // fragment of step (c, kk) and row-block m is at the group base plus a
        // compile-time offset
        auto aoff = [](int c, int kk, int m) { return (2 * m * HH_Z + c) * H_PITCH + kk * 32; };
#pragma unroll 1
        for (int ab = 0; ab < 9; ++ab) {
            if (ab + 1 < 9) load_group(ab + 1, bn);
            const unsigned char* gbase = abase + ((ab / 3) * HH_Y + ab % 3) * HH_Z * H_PITCH;
            uint4 af[2][HC_MT];
#pragma unroll
            for (int m = 0; m < HC_MT; ++m) af[0][m] = *reinterpret_cast<const uint4*>(gbase + aoff(0, 0, m));
#pragma unroll
            for (int st = 0; st < 3 * KS; ++st) {
                const int c = st / KS, kk = st % KS;
                if (st + 1 < 3 * KS) {
#pragma unroll
                    for (int m = 0; m < HC_MT; ++m)
                        af[(st + 1) & 1][m] =
                            *reinterpret_cast<const uint4*>(gbase + aoff((st + 1) / KS, (st + 1) % KS, m));
                }
                const bf16x8 bfr = as_bf16x8(bw[c][kk]);
#pragma unroll
                for (int m = 0; m < HC_MT; ++m)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(af[st & 1][m]), bfr, acc[m], 0, 0, 0);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int kk = 0; kk < KS; ++kk) bw[c][kk] = bn[c][kk];
        }
        if (ch == nch - 1) {
            int t = u / nco;
            const int tz = t % ntz;
            t /= ntz;
            const int ty = t % nty;
            t /= nty;
            const int tx = t % ntx, n = t / ntx;
            const int co = cob * 32 + r;
            const float bv = bias ? bias[co] : 0.f;
            const long long vbase = (long long)n * g.s0 * plane;
#pragma unroll
            for (int m = 0; m < HC_MT; ++m)
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int pr = (e >> 2) * 8 + hq * 4 + (e & 3);   // row within the row-block
                    const int y = ty * HT_Y + wy + 2 * m + (pr >> 4), z = tz * HT_Z + (pr & 15), x = tx * HT_X + wx;
                    float v = acc[m][e] + bv;
                    if (RELU) v = v > 0.f ? v : 0.f;
                    Y[(vbase + (long long)x * plane + y * g.s2 + z) * g.cout + co] = (bf16)v;
                }
#pragma unroll
            for (int m = 0; m < HC_MT; ++m) acc[m] = f32x16{};
        }
        // the next stage's halo: loaded after the compute, so its loads do not
        // share the in-order vmcnt counter with the weight loads the MFMAs wait on
        if (more) load(s + 1);
        __syncthreads();
        if (more) store();
        __syncthreads();
    }
}

// Weight gradient: dW[tap][co][ci] += sum_p dY[p][co] X[p + off(tap)][ci] in FP32
// (atomics). Block: one tap, a 32 x 32 (co, ci) tile and a chunk of positions;
// its 4 waves take interleaved 32-position stages of the chunk. Per stage a wave
// stages 32 rows of dY and of X ([position][channel], channels-last as in HBM)
// in its own LDS slice. The MFMA operands need 8 consecutive positions of one
// channel per lane (K = positions): ds_read_b64_tr_b16 delivers them from the
// channels-last rows. Within a group of 16 lanes, lane i receives element i % 4
// of the 8-byte chunks addressed by lanes 4 j + i / 4 (j = 0 .. 3), so lane m
// of the group addresses row p0 + m / 4, channels cb + 4 (m % 4) .. + 3, and
// lane i then holds channel cb + i at positions p0 .. p0 + 3 (mapping verified
// on gfx950). Two such reads give the 8 K values of an operand.
constexpr int WPC = 32;    // positions per wave stage
constexpr int WLD = 40;    // LDS row pitch (bf16): rows stay 8-byte aligned

typedef short v4i16 __attribute__((ext_vector_type(4)));

// p points into LDS (a __shared__ array)
__device__ __forceinline__ v4i16 lds_tr16(const bf16* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) v4i16*)(p));
}

__global__ __launch_bounds__(256) void k_conv_wgrad(Geo g, const bf16* __restrict__ X, const bf16* __restrict__ dY,
                                                    float* __restrict__ dW, int chunk) {
    __shared__ __align__(16) bf16 sY[4][WPC * WLD];
    __shared__ __align__(16) bf16 sX[4][WPC * WLD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = g.positions();
    const int nco = g.cout / 32, nci = g.cin / 32;
    int b = blockIdx.x;
    const int tile = b % (nco * nci);
    b /= nco * nci;
    const int tap = b % g.taps;
    const int c0 = (b / g.taps) * chunk;
    const int co0 = (tile / nci) * 32, ci0 = (tile % nci) * 32;
    const int d0 = g.taps == 1 ? 0 : tap / 9 - 1, d1 = g.taps == 1 ? 0 : (tap / 3) % 3 - 1,
              d2 = g.taps == 1 ? 0 : tap % 3 - 1;
    f32x16 acc = f32x16{};
    // lane loads: position lane >> 1, 16 channels at (lane & 1) * 16
    const int lp = lane >> 1, lh = lane & 1;
    // transposed operand reads: group gq = lane / 16 -> channel base 16 (gq & 1),
    // positions (gq >> 1) * 8 of the K step; lane m of the group addresses row
    // m / 4, channels 4 (m % 4)
    const int gq = lane >> 4, m = lane & 15;
    const int toff = ((gq >> 1) * 8 + (m >> 2)) * WLD + 16 * (gq & 1) + 4 * (m & 3);
    const int cend = min(c0 + chunk, P);
    // coordinates of this lane's position, advanced by the stage stride (no
    // division inside the loop)
    int pn, i0, i1, i2;
    {
        const int p = min(c0 + wave * WPC + lp, P - 1);
        const int vol = g.s0 * g.s1 * g.s2, plane = g.s1 * g.s2;
        pn = p / vol;
        const int rem = p - pn * vol;
        i0 = rem / plane;
        const int rem2 = rem - i0 * plane;
        i1 = rem2 / g.s2;
        i2 = rem2 - i1 * g.s2;
    }
    for (int p0 = c0 + wave * WPC; p0 < cend; p0 += 4 * WPC) {
        const int p = p0 + lp;
        uint4 y0 = make_uint4(0u, 0u, 0u, 0u), y1 = y0, x0 = y0, x1 = y0;
        if (p < cend) {
            const uint4* sy = reinterpret_cast<const uint4*>(dY + (long long)p * g.cout + co0 + 16 * lh);
            y0 = sy[0];
            y1 = sy[1];
            const int j0 = i0 + d0;
            int j1 = i1 + d1, j2 = i2 + d2;
            j1 = j1 < 0 ? j1 + g.s1 : (j1 >= g.s1 ? j1 - g.s1 : j1);
            j2 = j2 < 0 ? j2 + g.s2 : (j2 >= g.s2 ? j2 - g.s2 : j2);
            const int q = (j0 >= 0 && j0 < g.s0) ? ((pn * g.s0 + j0) * g.s1 + j1) * g.s2 + j2 : -1;
            if (q >= 0) {
                const uint4* sx = reinterpret_cast<const uint4*>(X + (long long)q * g.cin + ci0 + 16 * lh);
                x0 = sx[0];
                x1 = sx[1];
            }
        }
        uint4* wy = reinterpret_cast<uint4*>(&sY[wave][lp * WLD + 16 * lh]);
        wy[0] = y0;
        wy[1] = y1;
        uint4* wx = reinterpret_cast<uint4*>(&sX[wave][lp * WLD + 16 * lh]);
        wx[0] = x0;
        wx[1] = x1;
        __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): this wave's LDS writes have landed
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int o = kk * 16 * WLD + toff;
            const v4i16 a0 = lds_tr16(&sY[wave][o]), a1 = lds_tr16(&sY[wave][o + 4 * WLD]);
            const v4i16 b0 = lds_tr16(&sX[wave][o]), b1 = lds_tr16(&sX[wave][o + 4 * WLD]);
            const bf16x8 a = __builtin_bit_cast(bf16x8, __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7));
            const bf16x8 bb = __builtin_bit_cast(bf16x8, __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
        }
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0xc07f);   // reads done before the next stage overwrites the slice
        __builtin_amdgcn_wave_barrier();
        // advance the coordinates by 4 * WPC positions
        i2 += 4 * WPC;
        while (i2 >= g.s2) {
            i2 -= g.s2;
            if (++i1 == g.s1) {
                i1 = 0;
                if (++i0 == g.s0) {
                    i0 = 0;
                    ++pn;
                }
            }
        }
    }
    const int ci = ci0 + (lane & 31);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int co = co0 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
        atomicAdd(&dW[((long long)tap * g.cout + co) * g.cin + ci], acc[e]);
    }
}

// ---- elementwise and resampling kernels (8 channels per thread) ----

__device__ __forceinline__ uint4 bf8_load(const bf16* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void bf8_store(bf16* p, uint4 v) { *reinterpret_cast<uint4*>(p) = v; }

struct F8 {
    float v[8];
    __device__ __forceinline__ void from(uint4 u) {
        const bf16x8 b = as_bf16x8(u);
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (float)b[i];
    }
    __device__ __forceinline__ uint4 to() const {
        bf16x8 b;
#pragma unroll
        for (int i = 0; i < 8; ++i) b[i] = (bf16)v[i];
        return __builtin_bit_cast(uint4, b);
    }
};

// child position c of a coarse position q: fine volume s0 x s1 x s2 (even)
__device__ __forceinline__ int fine_child(int q, int s0, int s1, int s2, int c) {
    const int h0 = s0 >> 1, h1 = s1 >> 1, h2 = s2 >> 1;
    const int vol = h0 * h1 * h2, n = q / vol, r = q - n * vol;
    const int j0 = r / (h1 * h2), r2 = r - j0 * h1 * h2, j1 = r2 / h2, j2 = r2 - j1 * h2;
    return ((n * s0 + 2 * j0 + (c >> 2)) * s1 + 2 * j1 + ((c >> 1) & 1)) * s2 + 2 * j2 + (c & 1);
}
__device__ __forceinline__ int coarse_parent(int p, int s0, int s1, int s2) {
    const int vol = s0 * s1 * s2, n = p / vol, r = p - n * vol;
    const int i0 = r / (s1 * s2), r2 = r - i0 * s1 * s2, i1 = r2 / s2, i2 = r2 - i1 * s2;
    return ((n * (s0 >> 1) + (i0 >> 1)) * (s1 >> 1) + (i1 >> 1)) * (s2 >> 1) + (i2 >> 1);
}

// 2x2x2 average pooling: Y[q][c] = mean of X over the 8 children of q.
__global__ void k_pool(int nq, int s0, int s1, int s2, int C, const bf16* __restrict__ X, bf16* __restrict__ Y) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c8 = C >> 3;
    if (t >= (long long)nq * c8) return;
    const int q = (int)(t / c8), cc = (int)(t - (long long)q * c8) * 8;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < 8; ++c) {
        F8 f;
        f.from(bf8_load(X + (long long)fine_child(q, s0, s1, s2, c) * C + cc));
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i] += f.v[i];
    }
    F8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = 0.125f * s[i];
    bf8_store(Y + (long long)q * C + cc, o.to());
}

// Nearest upsampling of the coarse tensor L (Cu channels) concatenated with the
// fine skip tensor S (Cs channels): U[p] = (L[parent(p)], S[p]).
__global__ void k_up_cat(int np, int s0, int s1, int s2, int Cu, int Cs, const bf16* __restrict__ L,
                         const bf16* __restrict__ S, bf16* __restrict__ U) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = Cu + Cs, c8 = C >> 3;
    if (t >= (long long)np * c8) return;
    const int p = (int)(t / c8), cc = (int)(t - (long long)p * c8) * 8;
    const uint4 v = cc < Cu ? bf8_load(L + (long long)coarse_parent(p, s0, s1, s2) * Cu + cc)
                            : bf8_load(S + (long long)p * Cs + (cc - Cu));
    bf8_store(U + (long long)p * C + cc, v);
}

// Gradient of the upsampled part of a concatenation, masked by the coarse ReLU
// output: dL[q][c] = (sum over children of dU[child][c]) * (L[q][c] > 0), c < Cu.
__global__ void k_up_bwd(int nq, int s0, int s1, int s2, int Cu, int C, const bf16* __restrict__ dU,
                         const bf16* __restrict__ L, bf16* __restrict__ dL) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c8 = Cu >> 3;
    if (t >= (long long)nq * c8) return;
    const int q = (int)(t / c8), cc = (int)(t - (long long)q * c8) * 8;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < 8; ++c) {
        F8 f;
        f.from(bf8_load(dU + (long long)fine_child(q, s0, s1, s2, c) * C + cc));
#pragma unroll
        for (int i = 0; i < 8; ++i) s[i] += f.v[i];
    }
    F8 l, o;
    l.from(bf8_load(L + (long long)q * Cu + cc));
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = l.v[i] > 0.f ? s[i] : 0.f;
    bf8_store(dL + (long long)q * Cu + cc, o.to());
}

// Gradient of an encoder output E (C channels, ReLU) that feeds both a pooling
// and the skip part of a concatenation dU (channels Cu .. Cu + C of Ct):
// dE[p][c] = (dP[parent(p)][c] / 8 + dU[p][Cu + c]) * (E[p][c] > 0).
__global__ void k_pool_skip_bwd(int np, int s0, int s1, int s2, int C, int Cu, int Ct, const bf16* __restrict__ dP,
                                const bf16* __restrict__ dU, const bf16* __restrict__ E, bf16* __restrict__ dE) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c8 = C >> 3;
    if (t >= (long long)np * c8) return;
    const int p = (int)(t / c8), cc = (int)(t - (long long)p * c8) * 8;
    F8 a, b, e, o;
    a.from(bf8_load(dP + (long long)coarse_parent(p, s0, s1, s2) * C + cc));
    b.from(bf8_load(dU + (long long)p * Ct + Cu + cc));
    e.from(bf8_load(E + (long long)p * C + cc));
#pragma unroll
    for (int i = 0; i < 8; ++i) o.v[i] = e.v[i] > 0.f ? 0.125f * a.v[i] + b.v[i] : 0.f;
    bf8_store(dE + (long long)p * C + cc, o.to());
}

// In-place ReLU gradient: G[i] = Y[i] > 0 ? G[i] : 0 (Y the ReLU output).
__global__ void k_relu_mask(long long n8, const bf16* __restrict__ Y, bf16* __restrict__ G) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n8) return;
    F8 y, gg;
    y.from(bf8_load(Y + t * 8));
    gg.from(bf8_load(G + t * 8));
#pragma unroll
    for (int i = 0; i < 8; ++i) gg.v[i] = y.v[i] > 0.f ? gg.v[i] : 0.f;
    bf8_store(G + t * 8, gg.to());
}

// Bias gradient: dB[c] += sum_p G[p][c]. Block: 256 threads = 8 position lanes
// x 32 channels of a 32-channel group, over a chunk of positions.
__global__ __launch_bounds__(256) void k_bias_grad(int P, int C, int chunk, const bf16* __restrict__ G,
                                                   float* __restrict__ dB) {
    __shared__ float red[8][33];
    const int cg = C / 32;
    const int c = (blockIdx.x % cg) * 32 + (threadIdx.x & 31), pl = threadIdx.x >> 5;
    const int p0 = (blockIdx.x / cg) * chunk, p1 = min(p0 + chunk, P);
    float s = 0.f;
    for (int p = p0 + pl; p < p1; p += 8) s += (float)G[(long long)p * C + c];
    red[pl][threadIdx.x & 31] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.f;
#pragma unroll
        for (int i = 0; i < 8; ++i) t += red[i][threadIdx.x];
        atomicAdd(&dB[c], t);
    }
}

}  // namespace nn
}  // namespace nicr
