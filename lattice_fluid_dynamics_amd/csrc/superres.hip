// Super-resolution upscaler (include/lbm_superres.h): the per-coarse-cell MLP
// of simulation/shaders/superres_upscale.comp as FP32 matrix-core GEMMs.
//
// Workgroup: one 4x4x4 brick of coarse cells (64 cells, the shader's own
// workgroup), 4 waves; wave w owns the 16 cells of brick plane z = w as the
// M rows of v_mfma_f32_16x16x4_f32.
//   1. gather: the 27 neighbours (edge clamped) of every cell, normalised
//      ((v - mean) / std), into LDS rows act[cell][108] (+4 zero columns);
//   2. skip: the shader's trilinear skip of each of the 8 sub-cells from the
//      centre and its +x/+y/+z neighbours, into LDS skip[cell][32];
//   3. four layers 108(112) -> 128 -> 128 -> 128 -> 32: per 16-wide K chunk
//      the workgroup stages the chunk's weight rows in LDS once (the next
//      chunk's loads in flight meanwhile), and a lane reads a
//      float4 of its activation row and a float4 of its weight row; the 4 K
//      steps of the MFMA use the 4 components.  This permutes K inside each
//      chunk identically for A and B, so both operands come from row-major
//      layouts.  The bias seeds the accumulators; ReLU follows layers 0-2;
//      each wave writes its outputs back over its own activation rows (no
//      other wave reads them);
//   4. the last layer's 32 outputs + skip are de-normalised (v * std + mean)
//      and stored to the 2x2x2 fine cells.
// Arithmetic per coarse cell: 2 (108 128 + 2 128 128 + 128 32) = 101,376
// FLOP, FP32 products with FP32 accumulation (in MFMA order, not the
// shader's sequential order).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm_superres.h"
#include "ltws.h"

namespace {

constexpr int H = 128, IN = 108, INP = 112, OUT = 32;
constexpr int LD = 132;  // activation row stride (floats): 16-byte rows, distinct bank groups
constexpr int NW = H + H * IN + H + H * H + H + H * H + OUT + OUT * H;  // 51,104 floats in the file

typedef __attribute__((ext_vector_type(4))) float f32x4;

struct SrArgs {
    const float4 *coarse;
    float4 *fine;
    const float *w0, *w1, *w2, *w3;  // row-major [N][K]; w0 rows padded to 112
    const float *b0, *b1, *b2, *b3;
    float4 mean, stdv;
    int cx, cy, cz;
    int fx, fy;
};

// Weight rows of one 16-wide K chunk, staged in LDS for all four waves:
// Bs[n][k] with rows padded to 20 floats (16-byte aligned, and the 16 rows a
// wave reads start on distinct bank groups).
constexpr int BS_LD = 20;

// Chunk c of W (N rows of K floats) through registers: the N x 4 float4
// grid (row n = e / 4, column quad e % 4) has 512 cells for N = 128 (two
// per thread) and 128 for N = 32 (threads 0..127, one each).
struct ChunkRegs {
    float4 v0, v1;
};

template <int N>
__device__ __forceinline__ void fetch_chunk(const float *__restrict__ W, int K, int c, int tid, ChunkRegs &v) {
    if (N == 128) {
        v.v0 = *(const float4 *)(W + (tid >> 2) * K + 16 * c + 4 * (tid & 3));
        v.v1 = *(const float4 *)(W + ((tid + 256) >> 2) * K + 16 * c + 4 * (tid & 3));
    } else if (tid < N * 4) {
        v.v0 = *(const float4 *)(W + (tid >> 2) * K + 16 * c + 4 * (tid & 3));
    }
}

template <int N>
__device__ __forceinline__ void put_chunk(float *Bs, int tid, const ChunkRegs &v) {
    if (N == 128) {
        *(float4 *)(Bs + (tid >> 2) * BS_LD + 4 * (tid & 3)) = v.v0;
        *(float4 *)(Bs + ((tid + 256) >> 2) * BS_LD + 4 * (tid & 3)) = v.v1;
    } else if (tid < N * 4) {
        *(float4 *)(Bs + (tid >> 2) * BS_LD + 4 * (tid & 3)) = v.v0;
    }
}

// acc[nb] = bias + act(16 rows) x W^T for output columns 16 nb .. 16 nb + 15.
// The chunks of W pass through the buffer Bs[128 * BS_LD]: while the waves
// multiply chunk c, the next chunk's global loads are in flight; it is stored
// after a barrier.  One buffer (52 KB of LDS per workgroup, three workgroups
// per CU) measured faster than two (62 KB, two per CU).  Every wave of the
// workgroup calls this with the same K and NB (barriers).
template <int K, int NB>
__device__ __forceinline__ void mlp_layer(const float *rows, const float *__restrict__ W, const float *__restrict__ b,
                                          float *Bs, int tid, int r, int q, f32x4 (&acc)[NB]) {
    constexpr int N = 16 * NB, NC = K / 16;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const float bv = b[16 * nb + r];
        acc[nb] = f32x4{bv, bv, bv, bv};
    }
    ChunkRegs v;
    v.v0 = v.v1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    fetch_chunk<N>(W, K, 0, tid, v);
    put_chunk<N>(Bs, tid, v);
    __syncthreads();
    for (int c = 0; c < NC; c++) {
        float *cur = Bs;
        if (c + 1 < NC) fetch_chunk<N>(W, K, c + 1, tid, v);
        const float4 a = *(const float4 *)(rows + r * LD + 16 * c + 4 * q);
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const float4 w = *(const float4 *)(cur + (16 * nb + r) * BS_LD + 4 * q);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, w.x, acc[nb], 0, 0, 0);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, w.y, acc[nb], 0, 0, 0);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, w.z, acc[nb], 0, 0, 0);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, w.w, acc[nb], 0, 0, 0);
        }
        __syncthreads();
        if (c + 1 < NC) {
            put_chunk<N>(Bs, tid, v);
            __syncthreads();
        }
    }
}

// D[row 4 q + j][col 16 nb + r] -> rows (ReLU)
template <int NB>
__device__ __forceinline__ void store_relu(float *rows, int r, int q, const f32x4 (&acc)[NB]) {
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int j = 0; j < 4; j++) rows[(4 * q + j) * LD + 16 * nb + r] = fmaxf(acc[nb][j], 0.0f);
}

__device__ __forceinline__ float comp(const float4 &v, int c) {
    return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}

__global__ __launch_bounds__(256) void k_sr_upscale(SrArgs a) {
    __shared__ __align__(16) float act[64 * LD];
    __shared__ float skip[64 * OUT];
    __shared__ __align__(16) float Bs[128 * BS_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * 4, y0 = blockIdx.y * 4, z0 = blockIdx.z * 4;
    // 1. gather the normalised 3x3x3 patches (x fastest, then y, z; 4 channels)
    for (int e = tid; e < 64 * 27; e += 256) {
        const int cell = e / 27, nb = e % 27;
        const int X = min(max(x0 + (cell & 3) + nb % 3 - 1, 0), a.cx - 1);
        const int Y = min(max(y0 + ((cell >> 2) & 3) + (nb / 3) % 3 - 1, 0), a.cy - 1);
        const int Z = min(max(z0 + (cell >> 4) + nb / 9 - 1, 0), a.cz - 1);
        const float4 v = a.coarse[((long long)Z * a.cy + Y) * a.cx + X];
        float4 n;
        n.x = (v.x - a.mean.x) / a.stdv.x;
        n.y = (v.y - a.mean.y) / a.stdv.y;
        n.z = (v.z - a.mean.z) / a.stdv.z;
        n.w = (v.w - a.mean.w) / a.stdv.w;
        *(float4 *)(act + cell * LD + 4 * nb) = n;
    }
    if (tid < 64) *(float4 *)(act + tid * LD + IN) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    // 2. trilinear skip of sub-cell (dz, dy, dx) from patch cells (1 + c)
    for (int e = tid; e < 64 * 8; e += 256) {
        const int cell = e >> 3, sub = e & 7;
        const float fx = 0.25f + 0.5f * (float)(sub & 1);
        const float fy = 0.25f + 0.5f * (float)((sub >> 1) & 1);
        const float fz = 0.25f + 0.5f * (float)(sub >> 2);
        float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int cz = 0; cz < 2; cz++)
            for (int cy = 0; cy < 2; cy++)
                for (int cxx = 0; cxx < 2; cxx++) {
                    const float wz = cz ? fz : 1.0f - fz, wy = cy ? fy : 1.0f - fy, wx = cxx ? fx : 1.0f - fx;
                    const float w = wz * wy * wx;
                    const float *p = act + cell * LD + ((1 + cz) * 9 + (1 + cy) * 3 + (1 + cxx)) * 4;
#pragma unroll
                    for (int c = 0; c < 4; c++) s[c] += w * p[c];
                }
#pragma unroll
        for (int c = 0; c < 4; c++) skip[cell * OUT + sub * 4 + c] = s[c];
    }
    __syncthreads();
    // 3. the MLP on this wave's 16 rows
    const int r = lane & 15, q = lane >> 4;
    float *rows = act + 16 * wave * LD;
    {
        f32x4 acc[8];
        mlp_layer<INP, 8>(rows, a.w0, a.b0, Bs, tid, r, q, acc);
        store_relu<8>(rows, r, q, acc);
        mlp_layer<H, 8>(rows, a.w1, a.b1, Bs, tid, r, q, acc);
        store_relu<8>(rows, r, q, acc);
        mlp_layer<H, 8>(rows, a.w2, a.b2, Bs, tid, r, q, acc);
        store_relu<8>(rows, r, q, acc);
    }
    f32x4 out[2];
    mlp_layer<H, 2>(rows, a.w3, a.b3, Bs, tid, r, q, out);
    // 4. skip + correction, de-normalised, to the fine cells
    const int ch = r & 3;
    const float sd = comp(a.stdv, ch), mu = comp(a.mean, ch);
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
        const int n = 16 * nb + r, sub = n >> 2;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int cell = 16 * wave + 4 * q + j;
            const int X = x0 + (cell & 3), Y = y0 + ((cell >> 2) & 3), Z = z0 + (cell >> 4);
            if (X >= a.cx || Y >= a.cy || Z >= a.cz) continue;
            const float v = out[nb][j] + skip[cell * OUT + n];
            const float t = v * sd;
            const int FX = 2 * X + (sub & 1), FY = 2 * Y + ((sub >> 1) & 1), FZ = 2 * Z + (sub >> 2);
            ((float *)a.fine)[(((long long)FZ * a.fy + FY) * a.fx + FX) * 4 + ch] = t + mu;
        }
    }
}

}  // namespace

struct SuperResUpscaler {
    int coarseX, coarseY, coarseZ;
    int fineX, fineY, fineZ;
    float mean[4], stdv[4];
    float *weights;  // device: w0 (padded), b0, w1, b1, w2, b2, w3, b3
    float *fine;     // device float4 per fine cell
    hipStream_t stream;
};

namespace {

// All floats of an LTWS file, tensors concatenated in file order
// (superres.c:43-101, read by ltws.c); NULL on a missing, foreign or
// truncated file.
float *read_ltws(const char *path, long *count) {
    LBMLtws t;
    const int rc = lbm_ltws_read(path, &t);
    if (rc == LBM_LTWS_CANNOT_OPEN) {
        printf("SR: cannot open weights file: %s\n", path);
        return NULL;
    }
    if (rc == LBM_LTWS_BAD_MAGIC) {
        printf("SR: bad magic in %s\n", path);
        return NULL;
    }
    if (rc != LBM_LTWS_OK) {
        printf("SR: truncated weights file: %s\n", path);
        return NULL;
    }
    float *data = t.data ? t.data : (float *)malloc(sizeof(float));  // a file of no tensors holds 0 floats
    *count = t.total;
    t.data = NULL;
    lbm_ltws_free(&t);
    return data;
}

void read_normalizer(const char *path, float mean[4], float stdv[4]) {
    for (int i = 0; i < 4; i++) {
        mean[i] = 0.0f;
        stdv[i] = 1.0f;
    }
    FILE *f = path ? fopen(path, "rb") : NULL;
    if (!f) {
        printf("SR: cannot open normalizer: %s\n", path ? path : "(null)");
        return;
    }
    float v[8];
    if (fread(v, sizeof(float), 8, f) == 8) {
        memcpy(mean, v, sizeof(float) * 4);
        memcpy(stdv, v + 4, sizeof(float) * 4);
    } else {
        printf("SR: short normalizer file %s: mean 0, std 1\n", path);
    }
    fclose(f);
}

// device layout: w0 [128][112] (zero columns 108..111), b0, w1, b1, w2, b2, w3, b3
constexpr long OFF_W0 = 0, OFF_B0 = OFF_W0 + (long)H * INP, OFF_W1 = OFF_B0 + H, OFF_B1 = OFF_W1 + (long)H * H,
               OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + (long)H * H, OFF_W3 = OFF_B2 + H,
               OFF_B3 = OFF_W3 + (long)OUT * H, DEV_WEIGHTS = OFF_B3 + OUT;

SrArgs make_args(const SuperResUpscaler *sr, const float *coarse) {
    SrArgs a;
    a.coarse = (const float4 *)coarse;
    a.fine = (float4 *)sr->fine;
    a.w0 = sr->weights + OFF_W0;
    a.b0 = sr->weights + OFF_B0;
    a.w1 = sr->weights + OFF_W1;
    a.b1 = sr->weights + OFF_B1;
    a.w2 = sr->weights + OFF_W2;
    a.b2 = sr->weights + OFF_B2;
    a.w3 = sr->weights + OFF_W3;
    a.b3 = sr->weights + OFF_B3;
    a.mean = make_float4(sr->mean[0], sr->mean[1], sr->mean[2], sr->mean[3]);
    a.stdv = make_float4(sr->stdv[0], sr->stdv[1], sr->stdv[2], sr->stdv[3]);
    a.cx = sr->coarseX;
    a.cy = sr->coarseY;
    a.cz = sr->coarseZ;
    a.fx = sr->fineX;
    a.fy = sr->fineY;
    return a;
}

void launch(const SuperResUpscaler *sr, const float *coarse) {
    dim3 grid((unsigned)((sr->coarseX + 3) / 4), (unsigned)((sr->coarseY + 3) / 4), (unsigned)((sr->coarseZ + 3) / 4));
    hipLaunchKernelGGL(k_sr_upscale, grid, dim3(256), 0, sr->stream, make_args(sr, coarse));
}

}  // namespace

extern "C" SuperResUpscaler *SR_Create(int coarseX, int coarseY, int coarseZ, const char *weightsPath,
                                       const char *normPath) {
    if (coarseX <= 0 || coarseY <= 0 || coarseZ <= 0 || !weightsPath) return NULL;
    if ((coarseY + 3) / 4 > 65535 || (coarseZ + 3) / 4 > 65535) {  // grid y/z limits, 4 cells per block
        printf("SR: coarse grid too large\n");
        return NULL;
    }
    long count = 0;
    float *flat = read_ltws(weightsPath, &count);
    if (!flat) return NULL;
    if (count < NW) {
        printf("SR: %s holds %ld floats, the network needs %d\n", weightsPath, count, NW);
        free(flat);
        return NULL;
    }
    // file order b0, W0, b1, W1, b2, W2, b3, W3 (superres.c:124-154)
    float *host = (float *)calloc((size_t)DEV_WEIGHTS, sizeof(float));
    SuperResUpscaler *sr = (SuperResUpscaler *)calloc(1, sizeof(SuperResUpscaler));
    if (!host || !sr) {
        free(flat);
        free(host);
        free(sr);
        return NULL;
    }
    const float *p = flat;
    memcpy(host + OFF_B0, p, sizeof(float) * H);
    p += H;
    for (int n = 0; n < H; n++) memcpy(host + OFF_W0 + (long)n * INP, p + (long)n * IN, sizeof(float) * IN);
    p += (long)H * IN;
    memcpy(host + OFF_B1, p, sizeof(float) * H);
    p += H;
    memcpy(host + OFF_W1, p, sizeof(float) * H * H);
    p += (long)H * H;
    memcpy(host + OFF_B2, p, sizeof(float) * H);
    p += H;
    memcpy(host + OFF_W2, p, sizeof(float) * H * H);
    p += (long)H * H;
    memcpy(host + OFF_B3, p, sizeof(float) * OUT);
    p += OUT;
    memcpy(host + OFF_W3, p, sizeof(float) * OUT * H);
    free(flat);

    sr->coarseX = coarseX;
    sr->coarseY = coarseY;
    sr->coarseZ = coarseZ;
    sr->fineX = 2 * coarseX;
    sr->fineY = 2 * coarseY;
    sr->fineZ = 2 * coarseZ;
    read_normalizer(normPath, sr->mean, sr->stdv);
    const size_t fine_bytes = (size_t)sr->fineX * sr->fineY * sr->fineZ * 4 * sizeof(float);
    hipError_t e = hipStreamCreateWithFlags(&sr->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&sr->weights, sizeof(float) * (size_t)DEV_WEIGHTS);
    if (e == hipSuccess) e = hipMalloc((void **)&sr->fine, fine_bytes);
    if (e == hipSuccess) e = hipMemcpy(sr->weights, host, sizeof(float) * (size_t)DEV_WEIGHTS, hipMemcpyHostToDevice);
    free(host);
    if (e != hipSuccess) {
        printf("SR: device setup failed (%s)\n", hipGetErrorString(e));
        SR_Free(sr);
        return NULL;
    }
    printf("SR: coarse %dx%dx%d -> fine %dx%dx%d\n", coarseX, coarseY, coarseZ, sr->fineX, sr->fineY, sr->fineZ);
    return sr;
}

extern "C" void SR_Free(SuperResUpscaler *sr) {
    if (!sr) return;
    if (sr->stream) (void)hipStreamSynchronize(sr->stream);
    if (sr->fine) (void)hipFree(sr->fine);
    if (sr->weights) (void)hipFree(sr->weights);
    if (sr->stream) (void)hipStreamDestroy(sr->stream);
    free(sr);
}

extern "C" void SR_Upscale(SuperResUpscaler *sr, const float *coarseVel) {
    if (!sr || !coarseVel) return;
    launch(sr, coarseVel);
    hipError_t e = hipStreamSynchronize(sr->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) fprintf(stderr, "SR_Upscale: %s\n", hipGetErrorString(e));
}

extern "C" float *SR_GetFineVelocityBuffer(SuperResUpscaler *sr) { return sr ? sr->fine : NULL; }

extern "C" void SR_GetFineSize(SuperResUpscaler *sr, int *fineX, int *fineY, int *fineZ) {
    if (!sr) return;
    if (fineX) *fineX = sr->fineX;
    if (fineY) *fineY = sr->fineY;
    if (fineZ) *fineZ = sr->fineZ;
}

extern "C" int SR_Profile(SuperResUpscaler *sr, const float *coarseVel, int n, float *msPerCall) {
    if (!sr || !coarseVel || n <= 0) return 0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 0;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 0;
    }
    (void)hipEventRecord(e0, sr->stream);
    for (int i = 0; i < n; i++) launch(sr, coarseVel);
    (void)hipEventRecord(e1, sr->stream);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) e = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        fprintf(stderr, "SR_Profile: %s\n", hipGetErrorString(e));
        return 0;
    }
    if (msPerCall) *msPerCall = ms / (float)n;
    return 1;
}
