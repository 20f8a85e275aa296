// Kernels of the warm-start network: a small 3D convolutional encoder-decoder
// trained and run in BF16 on the matrix cores of gfx950, FP32 accumulation.
//
// Tensors are channels-last (N, D, H, W, C) BF16.  A 3x3x3 convolution
// (stride 1, zero padding 1) is an implicit GEMM:
//   rows    = output positions (N*D*H*W),
//   columns = output channels,
//   K       = 27 taps x input channels,
// computed with v_mfma_f32_16x16x32_bf16.  Operand layout of that
// instruction (checked on the device): lane l holds A[l%16][8*(l/16)+j] and
// B[8*(l/16)+j][l%16] for j = 0..7, and returns D[4*(l/16)+j][l%16], j = 0..3.
//
// The input-gradient (dgrad) of a convolution is the same forward kernel run
// on the output gradient with the weights mirrored in space and transposed
// in channels; the weight gradient (wgrad) has its own MFMA kernel with the
// positions as the K dimension.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "lbm_device.h"

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef uint16_t bf16_t;  // storage

namespace {

__device__ __forceinline__ float bf2f(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}

struct Shape {
    int n, d, h, w;  // batch and spatial extent
};

__device__ __forceinline__ bool tap_pos(const Shape &s, int p, int t, int &q) {
    // p = ((n*D + z)*H + y)*W + x ; tap t = (dz+1)*9 + (dy+1)*3 + (dx+1)
    const int x = p % s.w;
    int r = p / s.w;
    const int y = r % s.h;
    r /= s.h;
    const int z = r % s.d;
    const int n = r / s.d;
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const int X = x + dx, Y = y + dy, Z = z + dz;
    if (X < 0 || X >= s.w || Y < 0 || Y >= s.h || Z < 0 || Z >= s.d) return false;
    q = ((n * s.d + Z) * s.h + Y) * s.w + X;
    return true;
}

union Pack8 {
    uint4 u;
    bf16x8 v;
};

// ---------------------------------------------------------------- forward
// Y[p][co] = act(bias[co] + sum_{t, ci} W[co][t][ci] * X[p + off(t)][ci])
// Workgroup: 4 waves x 32 positions = 128 positions; NB column blocks of 16
// output channels (blockIdx.y selects the 16*NB-channel tile).  With
// part != NULL the 27 taps are split over gridDim.z workgroup slices
// (split-K for the small deep levels): slice z writes its FP32 partial sum to
// part[z][p][co] and k_conv_epilogue finishes the result.
// act: 0 none, 1 ReLU.  Yf (optional) receives the FP32 result.
// PK: the weights come packed (k_conv_pack layout, cout % 32 == 0) in Wp,
// so each B read of a wave is 1 KiB contiguous; otherwise rows of Wt.
template <int NB, bool PK>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16_t *__restrict__ X, const bf16_t *__restrict__ Wt,
                                                  const uint4 *__restrict__ Wp, const float *__restrict__ bias,
                                                  bf16_t *__restrict__ Y, float *__restrict__ Yf,
                                                  float *__restrict__ part, Shape s, int cin, int cout, int act) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = s.n * s.d * s.h * s.w;
    const int p0 = (blockIdx.x * 4 + wave) * 32;
    const int co0 = blockIdx.y * 16 * NB;
    const int r = lane & 15, kq = lane >> 4;  // row/column within a block, k quarter
    const int tbeg = blockIdx.z * 27 / gridDim.z, tend = (blockIdx.z + 1) * 27 / gridDim.z;
    f32x4 acc[2][NB];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    const int prow[2] = {p0 + r, p0 + 16 + r};
    for (int t = tbeg; t < tend; t++) {
        int q[2];
        bool ok[2];
#pragma unroll
        for (int a = 0; a < 2; a++) {
            q[a] = 0;
            ok[a] = prow[a] < P && tap_pos(s, prow[a], t, q[a]);
        }
        for (int c0 = 0; c0 < cin; c0 += 32) {
            Pack8 A[2], B[NB];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                A[a].u = make_uint4(0, 0, 0, 0);
                if (ok[a]) A[a].u = *(const uint4 *)(X + (long long)q[a] * cin + c0 + kq * 8);
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int co = co0 + 16 * b + r;
                B[b].u = make_uint4(0, 0, 0, 0);
                if (co < cout) {
                    if (PK) {
                        const int blk16 = co0 / 16 + b;
                        B[b].u = Wp[(((long long)(blk16 >> 1) * (cin / 32) + c0 / 32) * 27 + t) * 128 +
                                    (blk16 & 1) * 64 + lane];
                    } else {
                        B[b].u = *(const uint4 *)(Wt + ((long long)co * 27 + t) * cin + c0 + kq * 8);
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < NB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a].v, B[b].v, acc[a][b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int co = co0 + 16 * b + r;
            if (co >= cout) continue;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int p = p0 + 16 * a + 4 * kq + j;
                if (p >= P) continue;
                if (part) {
                    part[((long long)blockIdx.z * P + p) * cout + co] = acc[a][b][j];
                    continue;
                }
                float v = acc[a][b][j] + bv;
                if (act == 1) v = fmaxf(v, 0.0f);
                Y[(long long)p * cout + co] = f2bf(v);
                if (Yf) Yf[(long long)p * cout + co] = v;
            }
        }
}

// Brick-tiled forward for extents that divide into 8 x 4 x 4 bricks and
// channel counts that are multiples of 32: one workgroup per (brick, 32
// output channels).  Per 32-channel input chunk the brick's 10 x 6 x 6 input
// halo is staged once in LDS (channels-last rows padded to 40 bf16, so the
// 16 rows a wave reads start on distinct bank groups); wave w then runs the
// 32 positions of brick plane z = w (M = 2 x 16: x fastest, two y rows per
// block) against the 32 output channels (N = 2 x 16) through all 27 taps:
// per tap two A reads from the halo (16 bytes: 8 channels of the tap-shifted
// position) and two B reads of the weights, four MFMAs.  The weights come in
// the packed layout of k_conv_pack: each B read of a wave is 1 KiB
// contiguous (eight whole cache lines) instead of 64 bytes from each of 16
// rows 1.7 KiB apart, which made the weight reads the larger part of the
// kernel's time (level 0: 20.3 us against 14.1 us packed).
// The input is read once per brick plus halo instead of once per tap.
// Small grids split the taps over gridDim.z slices (as k_conv_fwd does):
// slice z writes its FP32 partial sum to part[z][p][co] and k_conv_epilogue
// finishes the result.
constexpr int FB_LD = 40;    // bf16 per staged halo position
constexpr int FB_HALO = 360; // 10 x 6 x 6
constexpr int FB_IT = (FB_HALO * 4 + 255) / 256;  // halo loads per lane and chunk

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_conv_fwd_brick(const bf16_t *__restrict__ X, const uint4 *__restrict__ Wp,
                                                        const float *__restrict__ bias, bf16_t *__restrict__ Y,
                                                        float *__restrict__ Yf, float *__restrict__ part, Shape s,
                                                        int cin, int cout, int act) {
    __shared__ __align__(16) bf16_t halo[FB_HALO * FB_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int co0 = blockIdx.y * 32;
    const int bw = s.w / 8, bh = s.h / 4, bps = bw * bh * (s.d / 4);
    const int sn = blockIdx.x / bps, loc = blockIdx.x % bps;
    const int x0 = (loc % bw) * 8, y0 = ((loc / bw) % bh) * 4, z0 = (loc / (bw * bh)) * 4;
    const long long sbase = (long long)sn * s.d * s.h * s.w;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    // halo of one 32-channel chunk: FB_IT 16-byte loads per lane, all issued
    // before any is written to LDS; the next chunk's are in flight while the
    // current chunk's taps run
    auto fetch = [&](int c0, uint4 (&hv)[FB_IT]) {
#pragma unroll
        for (int k = 0; k < FB_IT; k++) {
            const int e = tid + 256 * k;
            hv[k] = make_uint4(0, 0, 0, 0);
            if (e < FB_HALO * 4) {
                const int hp = e >> 2, cg = e & 3;
                const int X_ = x0 + hp % 10 - 1, Y_ = y0 + (hp / 10) % 6 - 1, Z_ = z0 + hp / 60 - 1;
                if (X_ >= 0 && X_ < s.w && Y_ >= 0 && Y_ < s.h && Z_ >= 0 && Z_ < s.d)
                    hv[k] = *(const uint4 *)(X + (sbase + ((long long)Z_ * s.h + Y_) * s.w + X_) * cin + c0 + 8 * cg);
            }
        }
    };
    uint4 hv[FB_IT];
    fetch(0, hv);
    for (int c0 = 0; c0 < cin; c0 += 32) {
#pragma unroll
        for (int k = 0; k < FB_IT; k++) {
            const int e = tid + 256 * k;
            if (e < FB_HALO * 4) *(uint4 *)(halo + (e >> 2) * FB_LD + 8 * (e & 3)) = hv[k];
        }
        __syncthreads();
        if (c0 + 32 < cin) fetch(c0 + 32, hv);
        // per tap: two A reads from the halo, two B reads of weight rows, four
        // MFMAs; the operands of tap t + 1 are loaded before tap t's MFMAs
        auto load = [&](int t, Pack8 (&a)[2], Pack8 (&b)[2]) {
            const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;  // halo offsets (tap + 1)
#pragma unroll
            for (int mb = 0; mb < 2; mb++) {
                const int hp = ((wave + dz) * 6 + 2 * mb + (r >> 3) + dy) * 10 + (r & 7) + dx;
                a[mb].u = *(const uint4 *)(halo + hp * FB_LD + 8 * kq);
            }
#pragma unroll
            for (int nb = 0; nb < 2; nb++)
                b[nb].u = Wp[(((long long)(blockIdx.y * (cin / 32) + c0 / 32) * 27 + t) * 2 + nb) * 64 + lane];
        };
        auto mma = [&](const Pack8 (&a)[2], const Pack8 (&b)[2]) {
#pragma unroll
            for (int mb = 0; mb < 2; mb++)
#pragma unroll
                for (int nb = 0; nb < 2; nb++)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb].v, b[nb].v, acc[mb][nb], 0, 0, 0);
        };
        const int tbeg = SPLIT ? blockIdx.z * 27 / gridDim.z : 0;
        const int tend = SPLIT ? (blockIdx.z + 1) * 27 / gridDim.z : 27;
        Pack8 a0[2], b0[2], a1[2], b1[2];
        load(tbeg, a0, b0);
        int t = tbeg;
        for (; t + 2 <= tend; t += 2) {
            load(t + 1, a1, b1);
            mma(a0, b0);
            if (t + 2 < tend) load(t + 2, a0, b0);
            mma(a1, b1);
        }
        if (t < tend) mma(a0, b0);
        __syncthreads();
    }
    // D[row = position 4 kq + j][col = channel r]
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
        const int co = co0 + 16 * nb + r;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int mb = 0; mb < 2; mb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int m = 4 * kq + j;
                const long long p = sbase + ((long long)(z0 + wave) * s.h + y0 + 2 * mb + (m >> 3)) * s.w + x0 + (m & 7);
                if (SPLIT) {
                    const long long P = (long long)s.n * s.d * s.h * s.w;
                    part[((long long)blockIdx.z * P + p) * cout + co] = acc[mb][nb][j];
                    continue;
                }
                float v = acc[mb][nb][j] + bv;
                if (act == 1) v = fmaxf(v, 0.0f);
                Y[p * cout + co] = f2bf(v);
                if (Yf) Yf[p * cout + co] = v;
            }
    }
}

// Sum of the split-K partials in slice order (deterministic), bias, activation.
__global__ void k_conv_epilogue(const float *__restrict__ part, int nsplit, const float *__restrict__ bias,
                                bf16_t *__restrict__ Y, float *__restrict__ Yf, long long n, int cout, int act) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // slices added in order; the loads of 3 slices are issued before their adds
    float v = 0.0f;
    int z = 0;
    for (; z + 3 <= nsplit; z += 3) {
        const float a = part[(long long)z * n + i], b = part[(long long)(z + 1) * n + i],
                    c = part[(long long)(z + 2) * n + i];
        v += a;
        v += b;
        v += c;
    }
    for (; z < nsplit; z++) v += part[(long long)z * n + i];
    if (bias) v += bias[i % cout];
    if (act == 1) v = fmaxf(v, 0.0f);
    Y[i] = f2bf(v);
    if (Yf) Yf[i] = v;
}

// Split-K partial sums: one device buffer, grown on demand and kept for the
// life of the process.  Calls from several host threads at once are not
// supported.
float *g_part = nullptr;
size_t g_part_bytes = 0;

float *part_buffer(size_t bytes) {
    if (bytes <= g_part_bytes) return g_part;
    if (g_part) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_part);
        g_part = nullptr;
        g_part_bytes = 0;
    }
    if (hipMalloc((void **)&g_part, bytes) != hipSuccess) {
        g_part = nullptr;
        return nullptr;
    }
    g_part_bytes = bytes;
    return g_part;
}

// Packed forward weights for calls that pass only W[co][t][ci]: one device
// buffer, grown on demand (same threading rule as part_buffer).
uint4 *g_pack = nullptr;
size_t g_pack_bytes = 0;

uint4 *pack_buffer(size_t bytes) {
    if (bytes <= g_pack_bytes) return g_pack;
    if (g_pack) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_pack);
        g_pack = nullptr;
        g_pack_bytes = 0;
    }
    if (hipMalloc((void **)&g_pack, bytes) != hipSuccess) {
        g_pack = nullptr;
        return nullptr;
    }
    g_pack_bytes = bytes;
    return g_pack;
}

// Packed weights of the brick forward (cin, cout multiples of 32): for
// output-channel block cb (32 channels), input chunk ch (32 channels), tap t
// and MFMA column block nb, the 64 uint4 read by the lanes l = 16 kq + r of
// a wave, lane l holding W[32 cb + 16 nb + r][t][32 ch + 8 kq .. + 8]:
//   Wp[(((cb * cin / 32 + ch) * 27 + t) * 2 + nb) * 64 + l].
// k_conv_pack writes, from W[co][t][ci]: the packed forward weights Wp; the
// input-gradient weights Wm[ci][26 - t][co] (as k_conv_mirror); and Wm
// packed (Wmp, the same layout for the convolution with cin and cout
// swapped).  One workgroup of 128 lanes per (layer, tap, 32 x 32 channel
// tile), up to LBMNN_PACK_MAX layers per launch: lane l reads 8 input
// channels of output channel l / 4 (one 16-byte load) and writes them to Wp;
// the tile is transposed through LDS, and lane l then writes 8 consecutive
// output channels of input channel l / 4 to Wm and Wmp, 16 bytes each.
struct PackList {
    const bf16_t *W[LBMNN_PACK_MAX];
    bf16_t *Wm[LBMNN_PACK_MAX];
    uint4 *Wp[LBMNN_PACK_MAX];
    uint4 *Wmp[LBMNN_PACK_MAX];
    int cin[LBMNN_PACK_MAX], cout[LBMNN_PACK_MAX];
    long long tstart[LBMNN_PACK_MAX + 1];  // first tile of each layer
    int count;
};

__global__ __launch_bounds__(128) void k_conv_pack(PackList a) {
    __shared__ bf16_t tile[32][40];  // [co][ci] of the tile, rows padded
    const long long tb = blockIdx.x;
    int k = 0;
    while (k + 1 < a.count && a.tstart[k + 1] <= tb) k++;
    long long rest = tb - a.tstart[k];
    const int cin = a.cin[k], cout = a.cout[k];
    const int cib = (int)(rest % (cin / 32));
    rest /= cin / 32;
    const int cob = (int)(rest % (cout / 32));
    const int t = (int)(rest / (cout / 32));
    const int tid = threadIdx.x;
    const int co = cob * 32 + tid / 4, ci = cib * 32 + 8 * (tid % 4);
    const uint4 v = *(const uint4 *)(a.W[k] + ((long long)co * 27 + t) * cin + ci);
    a.Wp[k][(((long long)cob * (cin / 32) + cib) * 27 + t) * 128 + ((co % 32) / 16) * 64 + 16 * ((ci % 32) / 8) +
            co % 16] = v;
    *(uint4 *)&tile[tid / 4][8 * (tid % 4)] = v;
    __syncthreads();
    // o: an input channel of W (output channel of the mirrored convolution),
    // c .. c + 8: output channels of W (its input channels), tap 26 - t
    const int o = cib * 32 + tid / 4, c = cob * 32 + 8 * (tid % 4), tm = 26 - t;
    uint4 u;
    bf16_t *h = (bf16_t *)&u;
#pragma unroll
    for (int j = 0; j < 8; j++) h[j] = tile[8 * (tid % 4) + j][tid / 4];
    *(uint4 *)(a.Wm[k] + ((long long)o * 27 + tm) * cout + c) = u;
    a.Wmp[k][(((long long)cib * (cout / 32) + cob) * 27 + tm) * 128 + ((o % 32) / 16) * 64 + 16 * ((c % 32) / 8) +
             o % 16] = u;
}

// Wp alone, for callers without packed copies (one lane per 16-byte group of W).
__global__ void k_conv_pack_fwd(const bf16_t *__restrict__ W, uint4 *__restrict__ Wp, int cin, int cout) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)cout * 27 * cin / 8) return;
    const int ci = (int)(i % (cin / 8)) * 8;
    const int t = (int)((i / (cin / 8)) % 27);
    const int co = (int)(i / ((long long)(cin / 8) * 27));
    const int lane = 16 * ((ci % 32) / 8) + co % 16;
    Wp[(((long long)(co / 32) * (cin / 32) + ci / 32) * 27 + t) * 128 + ((co % 32) / 16) * 64 + lane] =
        *(const uint4 *)(W + ((long long)co * 27 + t) * cin + ci);
}

// W[co][t][ci] -> Wm[ci][26 - t][co] (mirror taps, swap channels): the
// weights of the input-gradient convolution.
__global__ void k_conv_mirror(const bf16_t *W, bf16_t *Wm, int cin, int cout) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)cout * 27 * cin;
    if (i >= n) return;
    const int ci = (int)(i % cin);
    const int t = (int)((i / cin) % 27);
    const int co = (int)(i / ((long long)cin * 27));
    Wm[((long long)ci * 27 + (26 - t)) * cout + co] = W[i];
}

// ---------------------------------------------------------------- wgrad
// dW[co][t][ci] += sum_p dZ[p][co] * X[p + off(t)][ci]: a GEMM with the
// positions as K.  A workgroup owns a TCO x TCI block of one tap and a chunk
// of positions; per slab of 32 positions it loads the dZ and shifted X tiles
// with 16-byte loads (one position's 8 consecutive channels per lane),
// stores them transposed into LDS (channel-major, positions contiguous), and
// every wave reads its MFMA operands back as 16-byte LDS loads.  Partial
// sums go to dW with FP32 atomics.  Workgroups of the centre tap with the
// first ci block also accumulate the bias gradient sum_p dZ[p][co] into db.
constexpr int WG_K = 32;       // positions per slab
constexpr int WG_LD = WG_K + 8; // LDS row length (bf16), keeps rows 16-byte aligned

template <int TCO, int TCI>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16_t *__restrict__ dZ, const bf16_t *__restrict__ X,
                                                    float *__restrict__ dW, float *__restrict__ db, Shape s,
                                                    int cin, int cout, int chunk) {
    __shared__ __align__(16) bf16_t Ds[TCO * WG_LD];
    __shared__ __align__(16) bf16_t Xs[TCI * WG_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = s.n * s.d * s.h * s.w;
    const int t = blockIdx.y;
    const int nbci = cin / TCI;
    const int co0 = (blockIdx.z / nbci) * TCO, ci0 = (blockIdx.z % nbci) * TCI;
    const int pbeg = blockIdx.x * chunk, pend = min(P, pbeg + chunk);
    const bool do_bias = db != nullptr && t == 13 && ci0 == 0;
    constexpr int NBLK = (TCO / 16) * (TCI / 16);
    constexpr int PER_WAVE = (NBLK + 3) / 4;
    const int r = lane & 15, kq = lane >> 4;
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int b = 0; b < PER_WAVE; b++) acc[b] = f32x4{0, 0, 0, 0};
    float bsum = 0.0f;
    // one slab: WG_K positions of dZ (TCO channels) and of X at tap t (TCI
    // channels), 16-byte loads of 8 channels; the next slab's loads are in
    // flight while the current slab's MFMAs run
    constexpr int ZN = (WG_K * TCO / 8 + 255) / 256, XN = (WG_K * TCI / 8 + 255) / 256;
    auto fetch = [&](int p0, uint4 (&vz)[ZN], uint4 (&vx)[XN]) {
#pragma unroll
        for (int k = 0; k < ZN; k++) {
            const int e = tid + 256 * k;
            const int pl = e / (TCO / 8), cg = e % (TCO / 8);
            const int p = p0 + pl;
            vz[k] = make_uint4(0, 0, 0, 0);
            if (e < WG_K * TCO / 8 && p < pend) vz[k] = *(const uint4 *)(dZ + (long long)p * cout + co0 + 8 * cg);
        }
#pragma unroll
        for (int k = 0; k < XN; k++) {
            const int e = tid + 256 * k;
            const int pl = e / (TCI / 8), cg = e % (TCI / 8);
            const int p = p0 + pl;
            vx[k] = make_uint4(0, 0, 0, 0);
            int q;
            if (e < WG_K * TCI / 8 && p < pend && tap_pos(s, p, t, q))
                vx[k] = *(const uint4 *)(X + (long long)q * cin + ci0 + 8 * cg);
        }
    };
    uint4 vz[ZN], vx[XN];
    if (pbeg < pend) fetch(pbeg, vz, vx);
    for (int p0 = pbeg; p0 < pend; p0 += WG_K) {
        // transposed into LDS: channel-major rows, positions contiguous
#pragma unroll
        for (int k = 0; k < ZN; k++) {
            const int e = tid + 256 * k;
            if (e < WG_K * TCO / 8) {
                const int pl = e / (TCO / 8), cg = e % (TCO / 8);
                const bf16_t *h = (const bf16_t *)&vz[k];
#pragma unroll
                for (int j = 0; j < 8; j++) Ds[(8 * cg + j) * WG_LD + pl] = h[j];
            }
        }
#pragma unroll
        for (int k = 0; k < XN; k++) {
            const int e = tid + 256 * k;
            if (e < WG_K * TCI / 8) {
                const int pl = e / (TCI / 8), cg = e % (TCI / 8);
                const bf16_t *h = (const bf16_t *)&vx[k];
#pragma unroll
                for (int j = 0; j < 8; j++) Xs[(8 * cg + j) * WG_LD + pl] = h[j];
            }
        }
        __syncthreads();
        if (p0 + WG_K < pend) fetch(p0 + WG_K, vz, vx);
#pragma unroll
        for (int b = 0; b < PER_WAVE; b++) {
            const int blk = wave + 4 * b;
            if (blk < NBLK) {
                const int bco = (blk / (TCI / 16)) * 16, bci = (blk % (TCI / 16)) * 16;
                Pack8 a, bb;
                a.u = *(const uint4 *)(Ds + (bco + r) * WG_LD + 8 * kq);
                bb.u = *(const uint4 *)(Xs + (bci + r) * WG_LD + 8 * kq);
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bb.v, acc[b], 0, 0, 0);
            }
        }
        if (do_bias && tid < TCO) {
#pragma unroll 8
            for (int k = 0; k < WG_K; k++) bsum += bf2f(Ds[tid * WG_LD + k]);
        }
        __syncthreads();
    }
    // D[row = co][col = ci]: lane holds rows 4 kq + j of column r
#pragma unroll
    for (int b = 0; b < PER_WAVE; b++) {
        const int blk = wave + 4 * b;
        if (blk < NBLK) {
            const int bco = (blk / (TCI / 16)) * 16, bci = (blk % (TCI / 16)) * 16;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int co = co0 + bco + 4 * kq + j, ci = ci0 + bci + r;
                atomicAdd(&dW[((long long)co * 27 + t) * cin + ci], acc[b][j]);
            }
        }
    }
    if (do_bias && tid < TCO) atomicAdd(&db[co0 + tid], bsum);
}

// Brick-tiled weight gradient for 32 x 32 (co x ci) blocks, used when the
// spatial extent divides into 8 x 4 x 4 bricks.  A brick of 128 positions is
// four MFMA K-slabs (one per z plane of the brick) for all 27 taps, so each
// input element is fetched once per brick plus a 2.8x halo instead of once
// per tap.  Per brick:
//   * lanes 0..143 fetch the brick's 10 x 6 x 6 X halo (one (hz, hy) row of
//     10 positions x 8 channels each), transpose it in registers and store
//     three copies shifted by dx = -1, 0, +1: Xs[dx+1][ci][hz][hy][bx];
//   * lanes 192..255 fetch the 8 x 4 x 4 dZ tile (one (by, bz) row x 8
//     channels each) and store it transposed, Ds[co][k] with
//     k = bx + 8 (by + 4 bz); they also accumulate the bias gradient;
//   * wave w owns the 16 x 16 block (w / 2, w % 2) of all 27 taps.  In slab
//     bz, MFMA lane group kq holds row by = kq: one A read per slab and, per
//     tap, one aligned 16-byte B read (8 consecutive bx of halo row
//     (bz + dz + 1, kq + dy + 1)) into a 16x16x32 MFMA.
// The next brick's fetch is issued before the current brick's MFMAs.  LDS
// rows are padded so the 16 rows read by a wave start on distinct bank
// groups.  Each workgroup writes its partial sums to
// part[chunk][zblock][27][wave][j][lane] and, for the first ci block, its
// bias partials to bpart[chunk][co]; k_wgrad_reduce adds them in a fixed
// order (no atomics: bias atomics from every workgroup to the same 32
// addresses serialise).
constexpr int BR_ROWS = 36;                  // halo rows (hz, hy), 6 x 6
constexpr int BR_XS_CI = BR_ROWS * 8 + 8;    // bf16 per ci row of one shifted copy (+ pad)
constexpr int BR_DS_LD = 128 + 8;            // bf16 per co row of the dZ tile (+ pad)
constexpr int BR_HALO_LANES = BR_ROWS * 4;   // 144
constexpr int BR_DZ_LANE0 = 192;             // lanes 192..255 fetch dZ

__device__ __forceinline__ uint32_t word_of(const uint4 &u, int k) {
    return k == 0 ? u.x : (k == 1 ? u.y : (k == 2 ? u.z : u.w));
}

// channel j of positions a (low half) and b (high half)
__device__ __forceinline__ uint32_t chan_pair(const uint4 &a, const uint4 &b, int j) {
    const uint32_t wa = word_of(a, j >> 1), wb = word_of(b, j >> 1);
    return (j & 1) ? ((wa >> 16) | (wb & 0xffff0000u)) : ((wa & 0xffffu) | (wb << 16));
}

struct BrickCtx {
    int bw, bh, bps;
};

__device__ __forceinline__ void brick_fetch(const bf16_t *__restrict__ dZ, const bf16_t *__restrict__ X,
                                            const Shape &s, const BrickCtx &c, int b, int tid, int cin, int cout,
                                            int co0, int ci0, uint4 (&v)[10]) {
    const int sn = b / c.bps, loc = b % c.bps;
    const int x0 = (loc % c.bw) * 8, y0 = ((loc / c.bw) % c.bh) * 4, z0 = (loc / (c.bw * c.bh)) * 4;
    const long long sbase = (long long)sn * s.d * s.h * s.w;
    if (tid < BR_HALO_LANES) {
        const int row = tid >> 2, cg = tid & 3;
        const int Y_ = y0 + row % 6 - 1, Z_ = z0 + row / 6 - 1;
        const bool in = Y_ >= 0 && Y_ < s.h && Z_ >= 0 && Z_ < s.d;
        const long long rowbase = sbase + ((long long)Z_ * s.h + Y_) * s.w;
#pragma unroll
        for (int hx = 0; hx < 10; hx++) {
            const int X_ = x0 + hx - 1;
            v[hx] = make_uint4(0, 0, 0, 0);
            if (in && X_ >= 0 && X_ < s.w) v[hx] = *(const uint4 *)(X + (rowbase + X_) * cin + ci0 + 8 * cg);
        }
    } else if (tid >= BR_DZ_LANE0) {
        const int row = (tid - BR_DZ_LANE0) >> 2, cg = tid & 3;  // row = by + 4 bz
        const bf16_t *src =
            dZ + (sbase + ((long long)(z0 + (row >> 2)) * s.h + y0 + (row & 3)) * s.w + x0) * cout + co0 + 8 * cg;
#pragma unroll
        for (int bx = 0; bx < 8; bx++) v[bx] = *(const uint4 *)(src + (long long)bx * cout);
    }
}

__device__ __forceinline__ void brick_store(bf16_t *Ds, bf16_t *Xs, int tid, const uint4 (&v)[10], bool do_bias,
                                            float (&bs)[8]) {
    if (tid < BR_HALO_LANES) {
        const int row = tid >> 2, cg = tid & 3;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t w[9];
#pragma unroll
            for (int p = 0; p < 9; p++) w[p] = chan_pair(v[p], v[p + 1], j);
#pragma unroll
            for (int cx = 0; cx < 3; cx++)
                *(uint4 *)(Xs + (cx * 32 + 8 * cg + j) * BR_XS_CI + row * 8) =
                    make_uint4(w[cx], w[cx + 2], w[cx + 4], w[cx + 6]);
        }
    } else if (tid >= BR_DZ_LANE0) {
        const int row = (tid - BR_DZ_LANE0) >> 2, cg = tid & 3;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t w[4];
#pragma unroll
            for (int m = 0; m < 4; m++) w[m] = chan_pair(v[2 * m], v[2 * m + 1], j);
            *(uint4 *)(Ds + (8 * cg + j) * BR_DS_LD + 8 * row) = make_uint4(w[0], w[1], w[2], w[3]);
            if (do_bias) {
#pragma unroll
                for (int m = 0; m < 4; m++)
                    bs[j] += __uint_as_float(w[m] << 16) + __uint_as_float(w[m] & 0xffff0000u);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_conv_wgrad_brick(const bf16_t *__restrict__ dZ,
                                                          const bf16_t *__restrict__ X, float *__restrict__ part,
                                                          float *__restrict__ bpart, Shape s, int cin, int cout,
                                                          int bricks_per_wg) {
    __shared__ __align__(16) bf16_t Ds[32 * BR_DS_LD];
    __shared__ __align__(16) bf16_t Xs[3 * 32 * BR_XS_CI];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int nbci = cin / 32;
    const int zb = (cout / 32) * nbci;
    const int co0 = (blockIdx.z / nbci) * 32, ci0 = (blockIdx.z % nbci) * 32;
    BrickCtx c;
    c.bw = s.w / 8;
    c.bh = s.h / 4;
    c.bps = c.bw * c.bh * (s.d / 4);
    const int nbricks = s.n * c.bps;
    const int b0 = blockIdx.x * bricks_per_wg, b1 = min(nbricks, b0 + bricks_per_wg);
    const bool do_bias = bpart != nullptr && ci0 == 0;
    const int bco = 16 * (wave >> 1), bci = 16 * (wave & 1);
    f32x4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; t++) acc[t] = f32x4{0, 0, 0, 0};
    float bs[8];
#pragma unroll
    for (int j = 0; j < 8; j++) bs[j] = 0.0f;
    uint4 v[10];
    if (b0 < b1) brick_fetch(dZ, X, s, c, b0, tid, cin, cout, co0, ci0, v);
    for (int b = b0; b < b1; b++) {
        brick_store(Ds, Xs, tid, v, do_bias, bs);
        __syncthreads();
        if (b + 1 < b1) brick_fetch(dZ, X, s, c, b + 1, tid, cin, cout, co0, ci0, v);
#pragma unroll 1
        for (int bz = 0; bz < 4; bz++) {
            Pack8 a;
            a.u = *(const uint4 *)(Ds + (bco + r) * BR_DS_LD + 32 * bz + 8 * kq);
            // per dz plane: the 9 B operands are read back to back, then
            // their 9 MFMAs run (one LDS latency per plane, not per tap)
#pragma unroll
            for (int dz = 0; dz < 3; dz++) {
                Pack8 bb[9];
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    const int row = (bz + dz) * 6 + (kq + k / 3);
                    bb[k].u = *(const uint4 *)(Xs + ((k % 3) * 32 + bci + r) * BR_XS_CI + row * 8);
                }
#pragma unroll
                for (int k = 0; k < 9; k++)
                    acc[9 * dz + k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, bb[k].v, acc[9 * dz + k], 0, 0, 0);
            }
            // keep that order through scheduling: A read, then per plane
            // 9 LDS reads followed by 9 MFMAs
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 9, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 9, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 9, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 9, 0);
        }
        __syncthreads();
    }
    float *pw = part + ((long long)blockIdx.x * zb + blockIdx.z) * 27 * 1024;
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int j = 0; j < 4; j++) pw[((t * 4 + wave) * 4 + j) * 64 + lane] = acc[t][j];
    if (do_bias) {
        // the 16 (by, bz) rows of each channel group meet in LDS, summed in row order
        float *red = (float *)Xs;
        if (tid >= BR_DZ_LANE0) {
#pragma unroll
            for (int j = 0; j < 8; j++) red[(tid - BR_DZ_LANE0) * 8 + j] = bs[j];
        }
        __syncthreads();
        if (tid < 32) {
            const int cg = tid >> 3, j = tid & 7;
            float v = 0.0f;
#pragma unroll
            for (int row = 0; row < 16; row++) v += red[(row * 4 + cg) * 8 + j];
            bpart[(long long)blockIdx.x * cout + co0 + tid] = v;
        }
    }
}

// Sums the brick partials over chunks: thread i handles part element i of
// every (chunk, zblock) in stored order (contiguous loads), i.e. tap t,
// wave w, register j, lane l of 32 x 32 block zb, which is
// dW[co0 + 16 (w / 2) + 4 (l / 16) + j][t][ci0 + 16 (w % 2) + l % 16].
// Chunks are split over RED_SLICES lane groups of a workgroup (8 loads in
// flight each) and the slice sums are added in slice order: a fixed order
// for any chunk count, so results do not depend on timing.  The last
// workgroups do the same for the bias partials.
constexpr int RED_SLICES = 4;

__global__ __launch_bounds__(256) void k_wgrad_reduce(const float *__restrict__ part, const float *__restrict__ bpart,
                                                      int nchunks, int cin, int cout, float *__restrict__ dW,
                                                      float *__restrict__ db) {
    __shared__ float acc[RED_SLICES][64];
    const int o = threadIdx.x & 63, q = threadIdx.x >> 6;
    const long long i = (long long)blockIdx.x * 64 + o;
    const int zb = (cout / 32) * (cin / 32);
    const long long n = (long long)zb * 27 * 1024;
    const bool bias = i >= n && i < n + cout && bpart != nullptr;
    const float *p = nullptr;
    long long stride = 0;
    if (i < n) {
        p = part + i;
        stride = n;
    } else if (bias) {
        p = bpart + (i - n);
        stride = cout;
    }
    float v = 0.0f;
    if (p) {
        int c = q;
        for (; c + 7 * RED_SLICES < nchunks; c += 8 * RED_SLICES) {
            float x[8];
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = p[(long long)(c + k * RED_SLICES) * stride];
#pragma unroll
            for (int k = 0; k < 8; k++) v += x[k];
        }
        for (; c < nchunks; c += RED_SLICES) v += p[(long long)c * stride];
    }
    acc[q][o] = v;
    __syncthreads();
    if (q == 0 && p) {
        float s = acc[0][o];
#pragma unroll
        for (int k = 1; k < RED_SLICES; k++) s += acc[k][o];
        if (i < n) {
            const int zi = (int)(i / (27 * 1024)), e = (int)(i % (27 * 1024));
            const int t = e >> 10, w = (e >> 8) & 3, j = (e >> 6) & 3, l = e & 63;
            const int co = (zi / (cin / 32)) * 32 + 16 * (w >> 1) + 4 * (l >> 4) + j;
            const int ci = (zi % (cin / 32)) * 32 + 16 * (w & 1) + (l & 15);
            dW[((long long)co * 27 + t) * cin + ci] += s;
        } else {
            db[i - n] += s;
        }
    }
}

// Zero-fill of up to LBMNN_ZERO_MAX float buffers in one launch: blockIdx.y
// selects the buffer, the x blocks stride over it.
struct ZeroList {
    float *p[LBMNN_ZERO_MAX];
    long long count[LBMNN_ZERO_MAX];
    int n;
};

__global__ void k_zero_list(ZeroList z) {
    float *p = z.p[blockIdx.y];
    const long long n = z.count[blockIdx.y];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// ------------------------------------------------------------ elementwise
// dZ = dY * (Y > 0)  (ReLU backward on the stored activation), in place on dY
// The elementwise layer kernels below take V = 8 consecutive channels per
// thread (one 16-byte access) when the channel count allows, else V = 1;
// every element is computed with the same operations either way.
template <int V>
__device__ __forceinline__ void ld_bf(const bf16_t *p, float (&f)[V]) {
    if constexpr (V == 8) {
        const uint4 u = *(const uint4 *)p;
        const bf16_t *h = (const bf16_t *)&u;
#pragma unroll
        for (int j = 0; j < 8; j++) f[j] = bf2f(h[j]);
    } else {
        f[0] = bf2f(p[0]);
    }
}
template <int V>
__device__ __forceinline__ void st_bf(bf16_t *p, const float (&f)[V]) {
    if constexpr (V == 8) {
        uint4 u;
        bf16_t *h = (bf16_t *)&u;
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = f2bf(f[j]);
        *(uint4 *)p = u;
    } else {
        p[0] = f2bf(f[0]);
    }
}

// dY[i] = 0 where Y[i] <= 0 (bits of the other elements untouched)
template <int V>
__global__ void k_relu_bwd(bf16_t *dY, const bf16_t *Y, long long n) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (i >= n) return;
    if constexpr (V == 8) {
        uint4 d = *(const uint4 *)(dY + i);
        const uint4 y = *(const uint4 *)(Y + i);
        bf16_t *hd = (bf16_t *)&d;
        const bf16_t *hy = (const bf16_t *)&y;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (!(bf2f(hy[j]) > 0.0f)) hd[j] = 0;
        *(uint4 *)(dY + i) = d;
    } else {
        if (!(bf2f(Y[i]) > 0.0f)) dY[i] = 0;
    }
}

// db[c] += sum over positions of dZ[p][c]
__global__ void k_bias_grad(const bf16_t *dZ, float *db, long long P, int c) {
    const int ch = blockIdx.y * blockDim.x + threadIdx.x;
    if (ch >= c) return;
    float s = 0;
    for (long long p = blockIdx.x; p < P; p += gridDim.x) s += bf2f(dZ[p * c + ch]);
    atomicAdd(&db[ch], s);
}

// 2x2x2 average pooling: Y (D/2, H/2, W/2) from X
template <int V>
__global__ void k_pool_fwd(const bf16_t *X, bf16_t *Y, Shape so, int c) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    const long long n = (long long)so.n * so.d * so.h * so.w * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long long p = i / c;
    const int x = (int)(p % so.w);
    p /= so.w;
    const int y = (int)(p % so.h);
    p /= so.h;
    const int z = (int)(p % so.d);
    const int b = (int)(p / so.d);
    const int W = 2 * so.w, H = 2 * so.h, D = 2 * so.d;
    float s[V];
#pragma unroll
    for (int j = 0; j < V; j++) s[j] = 0;
    for (int k = 0; k < 8; k++) {
        const int Z = 2 * z + (k >> 2), Y2 = 2 * y + ((k >> 1) & 1), X2 = 2 * x + (k & 1);
        float v[V];
        ld_bf<V>(X + ((((long long)b * D + Z) * H + Y2) * W + X2) * c + ch, v);
#pragma unroll
        for (int j = 0; j < V; j++) s[j] += v[j];
    }
#pragma unroll
    for (int j = 0; j < V; j++) s[j] *= 0.125f;
    st_bf<V>(Y + i, s);
}

// average-pool backward: dX (fine) = dY (coarse) / 8
template <int V>
__global__ void k_pool_bwd(const bf16_t *dY, bf16_t *dX, Shape sf, int c) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    const long long n = (long long)sf.n * sf.d * sf.h * sf.w * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long long p = i / c;
    const int x = (int)(p % sf.w);
    p /= sf.w;
    const int y = (int)(p % sf.h);
    p /= sf.h;
    const int z = (int)(p % sf.d);
    const int b = (int)(p / sf.d);
    const long long q = ((((long long)b * (sf.d / 2) + z / 2) * (sf.h / 2) + y / 2) * (sf.w / 2) + x / 2);
    float v[V];
    ld_bf<V>(dY + q * c + ch, v);
#pragma unroll
    for (int j = 0; j < V; j++) v[j] *= 0.125f;
    st_bf<V>(dX + i, v);
}

// nearest x2 upsampling: Y (fine) from X (coarse); optional skip added
template <int V>
__global__ void k_up_fwd(const bf16_t *X, const bf16_t *skip, bf16_t *Y, Shape sf, int c) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    const long long n = (long long)sf.n * sf.d * sf.h * sf.w * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long long p = i / c;
    const int x = (int)(p % sf.w);
    p /= sf.w;
    const int y = (int)(p % sf.h);
    p /= sf.h;
    const int z = (int)(p % sf.d);
    const int b = (int)(p / sf.d);
    const long long q = ((((long long)b * (sf.d / 2) + z / 2) * (sf.h / 2) + y / 2) * (sf.w / 2) + x / 2);
    float v[V];
    ld_bf<V>(X + q * c + ch, v);
    if (skip) {
        float k[V];
        ld_bf<V>(skip + i, k);
#pragma unroll
        for (int j = 0; j < V; j++) v[j] += k[j];
    }
    st_bf<V>(Y + i, v);
}

// upsampling backward: dX (coarse) = sum over the 8 children of dY (fine)
template <int V>
__global__ void k_up_bwd(const bf16_t *dY, bf16_t *dX, Shape so, int c) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    const long long n = (long long)so.n * so.d * so.h * so.w * c;
    if (i >= n) return;
    const int ch = (int)(i % c);
    long long p = i / c;
    const int x = (int)(p % so.w);
    p /= so.w;
    const int y = (int)(p % so.h);
    p /= so.h;
    const int z = (int)(p % so.d);
    const int b = (int)(p / so.d);
    const int W = 2 * so.w, H = 2 * so.h, D = 2 * so.d;
    float s[V];
#pragma unroll
    for (int j = 0; j < V; j++) s[j] = 0;
    for (int k = 0; k < 8; k++) {
        const int Z = 2 * z + (k >> 2), Y2 = 2 * y + ((k >> 1) & 1), X2 = 2 * x + (k & 1);
        float v[V];
        ld_bf<V>(dY + ((((long long)b * D + Z) * H + Y2) * W + X2) * c + ch, v);
#pragma unroll
        for (int j = 0; j < V; j++) s[j] += v[j];
    }
    st_bf<V>(dX + i, s);
}

// a += b (gradient accumulation into a skip branch)
template <int V>
__global__ void k_add_inplace(bf16_t *a, const bf16_t *b, long long n) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (i >= n) return;
    float x[V], y[V];
    ld_bf<V>(a + i, x);
    ld_bf<V>(b + i, y);
#pragma unroll
    for (int j = 0; j < V; j++) x[j] += y[j];
    st_bf<V>(a + i, x);
}

// MSE loss on FP32 predictions: dY = 2 (Y - T) / n (BF16); the loss is
// reduced per workgroup (wave shuffles, then LDS) into one atomic add.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Yf, const float *__restrict__ T,
                                             bf16_t *__restrict__ dY, float *loss, long long n) {
    __shared__ float part[4];
    float l = 0.0f;
    const float inv = 1.0f / (float)n;
    const long long stride = (long long)gridDim.x * 256;
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    // four grid strides at a time: their loads are issued before the first
    // use; the loss still adds the elements of a lane in index order
    for (; i + 3 * stride < n; i += 4 * stride) {
        float y[4], t[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            y[k] = Yf[i + k * stride];
            t[k] = T[i + k * stride];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float d = y[k] - t[k];
            dY[i + k * stride] = f2bf(2.0f * d * inv);
            l += d * d * inv;
        }
    }
    for (; i < n; i += stride) {
        const float d = Yf[i] - T[i];
        dY[i] = f2bf(2.0f * d * inv);
        l += d * d * inv;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// AdamW on FP32 master weights, refreshing the BF16 copy (decoupled decay).
__global__ void k_adamw(float *w, bf16_t *wb, const float *g, float *m, float *v, long long n, float lr, float b1,
                        float b2, float eps, float wd, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i] * (1.0f - lr * wd);
    wi -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// AdamW over up to LBMNN_ADAM_MAX tensors in one launch: element i of the
// concatenation belongs to tensor k with start[k] <= i < start[k + 1]; each
// element is updated exactly as k_adamw updates it.
struct AdamList {
    float *w[LBMNN_ADAM_MAX];
    bf16_t *wb[LBMNN_ADAM_MAX];
    const float *g[LBMNN_ADAM_MAX];
    float *m[LBMNN_ADAM_MAX];
    float *v[LBMNN_ADAM_MAX];
    float wd[LBMNN_ADAM_MAX];
    long long start[LBMNN_ADAM_MAX + 1];
    int count;
    float lr, b1, b2, eps, c1, c2;
};

__global__ void k_adamw_list(AdamList a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.start[a.count]) return;
    int k = 0;
    while (k + 1 < a.count && a.start[k + 1] <= i) k++;
    const long long j = i - a.start[k];
    const float gi = a.g[k][j];
    const float mi = a.b1 * a.m[k][j] + (1.0f - a.b1) * gi;
    const float vi = a.b2 * a.v[k][j] + (1.0f - a.b2) * gi * gi;
    a.m[k][j] = mi;
    a.v[k][j] = vi;
    float wi = a.w[k][j] * (1.0f - a.lr * a.wd[k]);
    wi -= a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
    a.w[k][j] = wi;
    if (a.wb[k]) a.wb[k][j] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float *a, bf16_t *b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

inline dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

// ======================================================================
// C ABI (device pointers; work on the default stream, caller synchronises)
// ======================================================================
extern "C" {

static int ws_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "warm-start %s: %s\n", what, hipGetErrorString(e));
        return 0;
    }
    return 1;
}

int lbmnn_conv_pack_list(const void *const *W, void *const *Wm, void *const *Wp, void *const *Wmp, const int *cin,
                         const int *cout, int count) {
    if (count <= 0 || count > LBMNN_PACK_MAX) return 0;
    PackList a;
    a.count = count;
    a.tstart[0] = 0;
    for (int k = 0; k < count; k++) {
        if (cin[k] % 32 != 0 || cout[k] % 32 != 0) return 0;
        a.W[k] = (const bf16_t *)W[k];
        a.Wm[k] = (bf16_t *)Wm[k];
        a.Wp[k] = (uint4 *)Wp[k];
        a.Wmp[k] = (uint4 *)Wmp[k];
        a.cin[k] = cin[k];
        a.cout[k] = cout[k];
        a.tstart[k + 1] = a.tstart[k] + 27LL * (cout[k] / 32) * (cin[k] / 32);
    }
    hipLaunchKernelGGL(k_conv_pack, dim3((unsigned)a.tstart[count]), dim3(128), 0, 0, a);
    return ws_check("weight pack");
}

int lbmnn_conv_pack(const void *W, void *Wm, void *Wp, void *Wmp, int cin, int cout) {
    return lbmnn_conv_pack_list(&W, &Wm, &Wp, &Wmp, &cin, &cout, 1);
}

// cin % 32 == 0, cout % 16 == 0.  Wp: the packed weights of k_conv_pack
// (read by both paths), or NULL: the brick path then packs W into a scratch
// buffer and the per-tile path reads W.
static int conv_fwd_any(const void *X, const void *W, const void *Wp, const float *bias, void *Y, float *Yf, int n,
                        int d, int h, int w, int cin, int cout, int relu) {
    if (cin % 32 != 0 || cout % 16 != 0) return 0;
    Shape s{n, d, h, w};
    const long long P = (long long)n * d * h * w;
    // brick path; the taps split over 3, 9 or 27 slices while the
    // (brick, 32-channel) workgroups are fewer than 128 (measured at the
    // network's shapes: splitting 128 or more workgroups only adds partial
    // traffic, below that the slices fill the chip)
    if (w % 8 == 0 && h % 4 == 0 && d % 4 == 0 && cout % 32 == 0) {
        const long long wgs = (P / 128) * (cout / 32);
        int nsplit = 1;
        while (nsplit < 27 && wgs * nsplit < 128) nsplit *= 3;
        float *part = nullptr;
        if (nsplit > 1) {
            part = part_buffer(sizeof(float) * (size_t)nsplit * (size_t)P * (size_t)cout);
            if (!part) {
                fprintf(stderr, "warm-start conv forward: split-K buffer allocation failed\n");
                return 0;
            }
        }
        if (!Wp) {  // pack W here (the network keeps packed copies instead)
            const long long n8 = (long long)cout * 27 * cin / 8;
            uint4 *buf = pack_buffer(sizeof(uint4) * (size_t)n8);
            if (!buf) {
                fprintf(stderr, "warm-start conv forward: weight pack buffer allocation failed\n");
                return 0;
            }
            hipLaunchKernelGGL(k_conv_pack_fwd, g1(n8), dim3(256), 0, 0, (const bf16_t *)W, buf, cin, cout);
            Wp = buf;
        }
        dim3 grid((unsigned)(P / 128), (unsigned)(cout / 32), (unsigned)nsplit);
        if (part)
            hipLaunchKernelGGL(k_conv_fwd_brick<true>, grid, dim3(256), 0, 0, (const bf16_t *)X, (const uint4 *)Wp,
                               bias, (bf16_t *)Y, Yf, part, s, cin, cout, relu);
        else
            hipLaunchKernelGGL(k_conv_fwd_brick<false>, grid, dim3(256), 0, 0, (const bf16_t *)X, (const uint4 *)Wp,
                               bias, (bf16_t *)Y, Yf, part, s, cin, cout, relu);
        if (part) {
            const long long m = P * cout;
            hipLaunchKernelGGL(k_conv_epilogue, g1(m), dim3(256), 0, 0, part, nsplit, bias, (bf16_t *)Y, Yf, m, cout,
                               relu);
        }
        return ws_check("conv forward");
    }
    const int nbTotal = cout / 16;
    const int NB = nbTotal >= 4 ? 4 : nbTotal;
    const long long base = ((P + 127) / 128) * ((nbTotal + NB - 1) / NB);
    // split the taps over 3, 9 or 27 slices when the grid is far below the chip
    int nsplit = 1;
    while (nsplit < 27 && base * nsplit < 512) nsplit *= 3;
    float *part = nullptr;
    if (nsplit > 1) {
        part = part_buffer(sizeof(float) * (size_t)nsplit * (size_t)P * (size_t)cout);
        if (!part) {
            fprintf(stderr, "warm-start conv forward: split-K buffer allocation failed\n");
            return 0;
        }
    }
    dim3 grid((unsigned)((P + 127) / 128), (unsigned)((nbTotal + NB - 1) / NB), (unsigned)nsplit);
    const bf16_t *x = (const bf16_t *)X, *wt = (const bf16_t *)W;
    const uint4 *wp = (const uint4 *)Wp;
    bf16_t *y = (bf16_t *)Y;
    if (wp && cout % 32 == 0) {
        switch (NB) {
        case 1: hipLaunchKernelGGL((k_conv_fwd<1, true>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        case 2: hipLaunchKernelGGL((k_conv_fwd<2, true>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        case 3: hipLaunchKernelGGL((k_conv_fwd<3, true>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        default: hipLaunchKernelGGL((k_conv_fwd<4, true>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        }
    } else {
        switch (NB) {
        case 1: hipLaunchKernelGGL((k_conv_fwd<1, false>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        case 2: hipLaunchKernelGGL((k_conv_fwd<2, false>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        case 3: hipLaunchKernelGGL((k_conv_fwd<3, false>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        default: hipLaunchKernelGGL((k_conv_fwd<4, false>), grid, dim3(256), 0, 0, x, wt, wp, bias, y, Yf, part, s, cin, cout, relu); break;
        }
    }
    if (part) {
        const long long m = P * cout;
        hipLaunchKernelGGL(k_conv_epilogue, g1(m), dim3(256), 0, 0, part, nsplit, bias, y, Yf, m, cout, relu);
    }
    return ws_check("conv forward");
}

int lbmnn_conv_fwd(const void *X, const void *W, const float *bias, void *Y, float *Yf, int n, int d, int h, int w,
                   int cin, int cout, int relu) {
    return conv_fwd_any(X, W, nullptr, bias, Y, Yf, n, d, h, w, cin, cout, relu);
}

int lbmnn_conv_fwd_packed(const void *X, const void *W, const void *Wp, const float *bias, void *Y, float *Yf, int n,
                          int d, int h, int w, int cin, int cout, int relu) {
    return conv_fwd_any(X, W, Wp, bias, Y, Yf, n, d, h, w, cin, cout, relu);
}

int lbmnn_conv_mirror(const void *W, void *Wm, int cin, int cout) {
    const long long n = (long long)cout * 27 * cin;
    hipLaunchKernelGGL(k_conv_mirror, g1(n), dim3(256), 0, 0, (const bf16_t *)W, (bf16_t *)Wm, cin, cout);
    return ws_check("weight mirror");
}

// dW (FP32, zeroed by the caller) += weight gradient; db (may be NULL,
// zeroed by the caller) += bias gradient.  cin, cout % 32 == 0.
int lbmnn_conv_wgrad(const void *dZ, const void *X, float *dW, float *db, int n, int d, int h, int w, int cin,
                     int cout) {
    if (cin % 32 != 0 || cout % 32 != 0) return 0;
    Shape s{n, d, h, w};
    const long long P = (long long)n * d * h * w;
    const bf16_t *z = (const bf16_t *)dZ, *x = (const bf16_t *)X;
    if (w % 8 == 0 && h % 4 == 0 && d % 4 == 0) {
        // brick path: about 256 workgroups over (brick chunks) x (32 x 32 blocks)
        const long long nbricks = P / 128;
        const int zb = (cout / 32) * (cin / 32);
        long long per = (nbricks * zb + 255) / 256;
        if (per < 1) per = 1;
        const long long nchunks = (nbricks + per - 1) / per;
        const size_t nw = (size_t)27 * 1024 * (size_t)nchunks * (size_t)zb;
        float *part = part_buffer(sizeof(float) * (nw + (size_t)nchunks * (size_t)cout));
        if (!part) {
            fprintf(stderr, "warm-start conv wgrad: partial-sum buffer allocation failed\n");
            return 0;
        }
        float *bpart = db ? part + nw : nullptr;
        hipLaunchKernelGGL(k_conv_wgrad_brick, dim3((unsigned)nchunks, 1, (unsigned)zb), dim3(256), 0, 0, z, x, part,
                           bpart, s, cin, cout, (int)per);
        const long long m = (long long)zb * 27 * 1024 + (db ? cout : 0);
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((m + 63) / 64)), dim3(256), 0, 0, part, bpart,
                           (int)nchunks, cin, cout, dW, db);
        return ws_check("conv wgrad");
    }
    const int tco = cout % 64 == 0 ? 64 : 32, tci = cin % 64 == 0 ? 64 : 32;
    const int zb = (cout / tco) * (cin / tci);
    // about 512 workgroups (a few slabs each, so the next slab's loads overlap
    // the current MFMAs and fewer partial sums meet in the atomics), chunks a
    // multiple of the 32-position slab
    long long chunks = (512 + 27LL * zb - 1) / (27LL * zb);
    long long chunk = (P + chunks - 1) / chunks;
    chunk = ((chunk + WG_K - 1) / WG_K) * WG_K;
    if (chunk < WG_K) chunk = WG_K;
    chunks = (P + chunk - 1) / chunk;
    dim3 grid((unsigned)chunks, 27, (unsigned)zb);
    if (tco == 64 && tci == 64)
        hipLaunchKernelGGL((k_conv_wgrad<64, 64>), grid, dim3(256), 0, 0, z, x, dW, db, s, cin, cout, (int)chunk);
    else if (tco == 64)
        hipLaunchKernelGGL((k_conv_wgrad<64, 32>), grid, dim3(256), 0, 0, z, x, dW, db, s, cin, cout, (int)chunk);
    else if (tci == 64)
        hipLaunchKernelGGL((k_conv_wgrad<32, 64>), grid, dim3(256), 0, 0, z, x, dW, db, s, cin, cout, (int)chunk);
    else
        hipLaunchKernelGGL((k_conv_wgrad<32, 32>), grid, dim3(256), 0, 0, z, x, dW, db, s, cin, cout, (int)chunk);
    return ws_check("conv wgrad");
}

int lbmnn_relu_bwd(void *dY, const void *Y, long long n) {
    if (n % 8 == 0)
        hipLaunchKernelGGL(k_relu_bwd<8>, g1(n / 8), dim3(256), 0, 0, (bf16_t *)dY, (const bf16_t *)Y, n);
    else
        hipLaunchKernelGGL(k_relu_bwd<1>, g1(n), dim3(256), 0, 0, (bf16_t *)dY, (const bf16_t *)Y, n);
    return ws_check("relu backward");
}

int lbmnn_bias_grad(const void *dZ, float *db, long long P, int c) {
    dim3 grid(256, (unsigned)((c + 63) / 64));
    hipLaunchKernelGGL(k_bias_grad, grid, dim3(64), 0, 0, (const bf16_t *)dZ, db, P, c);
    return ws_check("bias gradient");
}

int lbmnn_pool_fwd(const void *X, void *Y, int n, int dOut, int hOut, int wOut, int c) {
    Shape so{n, dOut, hOut, wOut};
    const long long m = (long long)n * dOut * hOut * wOut * c;
    if (c % 8 == 0)
        hipLaunchKernelGGL(k_pool_fwd<8>, g1(m / 8), dim3(256), 0, 0, (const bf16_t *)X, (bf16_t *)Y, so, c);
    else
        hipLaunchKernelGGL(k_pool_fwd<1>, g1(m), dim3(256), 0, 0, (const bf16_t *)X, (bf16_t *)Y, so, c);
    return ws_check("pool forward");
}

int lbmnn_pool_bwd(const void *dY, void *dX, int n, int dIn, int hIn, int wIn, int c) {
    Shape sf{n, dIn, hIn, wIn};
    const long long m = (long long)n * dIn * hIn * wIn * c;
    if (c % 8 == 0)
        hipLaunchKernelGGL(k_pool_bwd<8>, g1(m / 8), dim3(256), 0, 0, (const bf16_t *)dY, (bf16_t *)dX, sf, c);
    else
        hipLaunchKernelGGL(k_pool_bwd<1>, g1(m), dim3(256), 0, 0, (const bf16_t *)dY, (bf16_t *)dX, sf, c);
    return ws_check("pool backward");
}

int lbmnn_up_fwd(const void *X, const void *skip, void *Y, int n, int dOut, int hOut, int wOut, int c) {
    Shape sf{n, dOut, hOut, wOut};
    const long long m = (long long)n * dOut * hOut * wOut * c;
    if (c % 8 == 0)
        hipLaunchKernelGGL(k_up_fwd<8>, g1(m / 8), dim3(256), 0, 0, (const bf16_t *)X, (const bf16_t *)skip,
                           (bf16_t *)Y, sf, c);
    else
        hipLaunchKernelGGL(k_up_fwd<1>, g1(m), dim3(256), 0, 0, (const bf16_t *)X, (const bf16_t *)skip, (bf16_t *)Y,
                           sf, c);
    return ws_check("upsample forward");
}

int lbmnn_up_bwd(const void *dY, void *dX, int n, int dIn, int hIn, int wIn, int c) {
    Shape so{n, dIn, hIn, wIn};
    const long long m = (long long)n * dIn * hIn * wIn * c;
    if (c % 8 == 0)
        hipLaunchKernelGGL(k_up_bwd<8>, g1(m / 8), dim3(256), 0, 0, (const bf16_t *)dY, (bf16_t *)dX, so, c);
    else
        hipLaunchKernelGGL(k_up_bwd<1>, g1(m), dim3(256), 0, 0, (const bf16_t *)dY, (bf16_t *)dX, so, c);
    return ws_check("upsample backward");
}

int lbmnn_add_inplace(void *a, const void *b, long long n) {
    if (n % 8 == 0)
        hipLaunchKernelGGL(k_add_inplace<8>, g1(n / 8), dim3(256), 0, 0, (bf16_t *)a, (const bf16_t *)b, n);
    else
        hipLaunchKernelGGL(k_add_inplace<1>, g1(n), dim3(256), 0, 0, (bf16_t *)a, (const bf16_t *)b, n);
    return ws_check("add");
}

int lbmnn_mse(const float *Yf, const float *T, void *dY, float *loss, long long n) {
    const long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(256), 0, 0, Yf, T, (bf16_t *)dY, loss, n);
    return ws_check("mse");
}

int lbmnn_adamw(float *w, void *wb, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                float eps, float wd, int t) {
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adamw, g1(n), dim3(256), 0, 0, w, (bf16_t *)wb, g, m, v, n, lr, b1, b2, eps, wd, c1, c2);
    return ws_check("adamw");
}

int lbmnn_adamw_list(float *const *w, void *const *wb, const float *const *g, float *const *m, float *const *v,
                     const long long *n, const float *wd, int count, float lr, float b1, float b2, float eps, int t) {
    if (count <= 0 || count > LBMNN_ADAM_MAX) return 0;
    AdamList a;
    a.count = count;
    a.start[0] = 0;
    for (int k = 0; k < count; k++) {
        a.w[k] = w[k];
        a.wb[k] = (bf16_t *)wb[k];
        a.g[k] = g[k];
        a.m[k] = m[k];
        a.v[k] = v[k];
        a.wd[k] = wd[k];
        a.start[k + 1] = a.start[k] + n[k];
    }
    a.lr = lr;
    a.b1 = b1;
    a.b2 = b2;
    a.eps = eps;
    a.c1 = 1.0f - powf(b1, (float)t);
    a.c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adamw_list, g1(a.start[count]), dim3(256), 0, 0, a);
    return ws_check("adamw");
}

int lbmnn_f32_to_bf16(const float *a, void *b, long long n) {
    hipLaunchKernelGGL(k_f32_to_bf16, g1(n), dim3(256), 0, 0, a, (bf16_t *)b, n);
    return ws_check("convert");
}

int lbmnn_memset(void *p, int value, unsigned long long bytes) {
    return hipMemset(p, value, bytes) == hipSuccess;
}

int lbmnn_zero_list(float *const *ptrs, const long long *counts, int n) {
    if (n <= 0 || n > LBMNN_ZERO_MAX) return 0;
    ZeroList z;
    z.n = n;
    for (int i = 0; i < n; i++) {
        z.p[i] = ptrs[i];
        z.count[i] = counts[i];
    }
    hipLaunchKernelGGL(k_zero_list, dim3(64, (unsigned)n), dim3(256), 0, 0, z);
    return ws_check("zero");
}

void *lbmnn_host_alloc(size_t bytes) {
    void *p = nullptr;
    return hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

void lbmnn_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

int lbmnn_copy_to_host_async(void *host, const void *dev, size_t bytes) {
    return hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, 0) == hipSuccess;
}

int lbmnn_synchronize(void) { return hipDeviceSynchronize() == hipSuccess; }

}  // extern "C"
