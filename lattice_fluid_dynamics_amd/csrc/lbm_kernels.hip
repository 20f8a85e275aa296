// D3Q19 lattice-Boltzmann kernels for gfx950 (MI355X) and the thin C-ABI
// device layer declared in lbm_device.h.
//
// Population storage: 19 direction planes (SoA), plane i at pop + i*stride,
// cell = x + NX*(y + NY*z), updated IN PLACE by a two-phase streaming scheme
// (the "esoteric pull" family of Geier & Schoenherr / Lehmann):
//
//   phase A (even):  cell x reads f_i from plane i at x, collides, writes
//                    f*_i to plane opp(i) at x                (all local)
//   phase B (odd):   cell x reads f_i from plane opp(i) at x - e_i, collides,
//                    writes f*_i to plane i at x + e_i
//
// Each cell reads and writes the same 19 slots within a phase, so the update
// is race-free with one buffer: 76 B/cell of storage for FP32 and exactly
// 152 B/cell/step of HBM traffic (19 loads + 19 stores).  Neighbour indices
// wrap on the full 3-torus, which makes the scheme a pure permutation; the
// reference's non-periodic stream rules (lbm_stream.comp:85-128: Bouzidi
// interpolated bounce-back, clamped Y/Z, the x-edge self-copy of cells whose
// source lies outside the domain) are applied by a prepass over a compacted
// list of boundary cells, which evaluates the rule on the stored
// post-collision populations and hands the main kernel the 19 incoming
// values of those cells.  The x-edge self-copy only matters for body cells on
// the x faces: for every other x-face cell the Zou-He rewrite of
// lbm_collide.comp:336-408 replaces exactly the self-copied directions.
//
// The stored state after a step is the post-collision field f* (the
// reference's f_new); the reference's post-stream f is f = S(f*), which the
// force, export and mass kernels evaluate on demand.  The first step after
// LBM_InitializeFlow collides the initial feq directly, as the reference
// does (phase A without the prepass).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "lbm_device.h"

namespace {

// D3Q19 velocity set in the reference order (lbm_collide.comp:33-52).
constexpr int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

template <typename T>
__host__ __device__ constexpr T wgt(int i) {
    return i == 0 ? T(1) / T(3) : (i < 7 ? T(1) / T(18) : T(1) / T(36));
}

constexpr int BLOCK = 256;
constexpr int OP_BGK = 0, OP_MRT = 1, OP_REG = 2;
constexpr int ST_INIT = 0, ST_A = 1, ST_B = 2;  // layout of the stored populations

// ---------------------------------------------------------------- storage
// FP16S / BF16S store the equilibrium-shifted value f_i - w_i (FP16 scaled
// by 2^15 to keep typical deviations out of the subnormal range); arithmetic
// is FP32 in every format but FP64.
template <int FMT> struct Pop;
template <> struct Pop<LBMDEV_FP32> {
    typedef float S;
    typedef float T;
    static __device__ __forceinline__ T dec(S v, int) { return v; }
    static __device__ __forceinline__ S enc(T v, int) { return v; }
};
template <> struct Pop<LBMDEV_FP64> {
    typedef double S;
    typedef double T;
    static __device__ __forceinline__ T dec(S v, int) { return v; }
    static __device__ __forceinline__ S enc(T v, int) { return v; }
};
template <> struct Pop<LBMDEV_FP16S> {
    typedef uint16_t S;  // IEEE binary16 bits
    typedef float T;
    static __device__ __forceinline__ T dec(S v, int i) {
        return fmaf(__half2float(__ushort_as_half(v)), 1.0f / 32768.0f, wgt<float>(i));
    }
    static __device__ __forceinline__ S enc(T v, int i) {
        return __half_as_ushort(__float2half_rn((v - wgt<float>(i)) * 32768.0f));
    }
};
template <> struct Pop<LBMDEV_BF16S> {
    typedef uint16_t S;
    typedef float T;
    static __device__ __forceinline__ T dec(S v, int i) {
        return __uint_as_float(((uint32_t)v) << 16) + wgt<float>(i);
    }
    static __device__ __forceinline__ S enc(T v, int i) {
        uint32_t u = __float_as_uint(v - wgt<float>(i));
        u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
        return (uint16_t)(u >> 16);
    }
};

struct Geo {
    int nx, ny, nz, ncell;
    long long stride;
    // n / nx and n / ny for 0 <= n < 2^31 as umulhi(n, m) >> s (m = 0: divisor 1);
    // Granlund & Montgomery's round-up method with a 31-bit numerator
    uint32_t mx, my;
    int sx, sy;
};

static void fast_div_magic(uint32_t d, uint32_t &m, int &s) {
    if (d <= 1) {
        m = 0;
        s = 0;
        return;
    }
    int l = 0;
    while ((1ull << l) < d) l++;  // ceil(log2 d)
    m = (uint32_t)((1ull << (31 + l)) / d + 1);
    s = l - 1;
}

static Geo make_geo(int nx, int ny, int nz, long long stride) {
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.ncell = nx * ny * nz;
    g.stride = stride;
    fast_div_magic((uint32_t)nx, g.mx, g.sx);
    fast_div_magic((uint32_t)ny, g.my, g.sy);
    return g;
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, int s) {
    return m == 0 ? n : (__umulhi(n, m) >> s);
}

// Each of the 8 XCDs receives workgroups b, b+8, b+16, ... in dispatch order.
// xcd_remap gives every XCD one contiguous range of cells, so neighbouring
// rows and partially shared cache lines stay in one L2; without it the eight
// XCDs sweep the lattice side by side.  Lattices that partly fit the caches
// gain up to 2 % from the remap; lattices far larger lose 6-9 % (512x256x256
// FP32 and FP16), so the order is a launch parameter that the autotuner
// measures (`remap` in StepArgs / BatchArgs).
__device__ __forceinline__ int xcd_remap(int b, int n) {
    int q = n >> 3, r = n & 7, k = b & 7, j = b >> 3;
    return k * q + (k < r ? k : r) + j;
}

__device__ __forceinline__ void decompose(const Geo &g, int cell, int &x, int &y, int &z) {
    const uint32_t t = fast_div((uint32_t)cell, g.mx, g.sx);
    x = cell - (int)t * g.nx;
    const uint32_t zz = fast_div(t, g.my, g.sy);
    y = (int)t - (int)zz * g.ny;
    z = (int)zz;
}

// Base of direction plane i: uniform across the workgroup, so loads take it
// in scalar registers and only a 32-bit per-lane offset.
template <typename S>
__device__ __forceinline__ S *plane(S *pop, const Geo &g, int i) {
    return pop + (long long)i * g.stride;
}

// Element idx of a plane, addressed as uniform base + 32-bit byte offset
// (the plane is far below 4 GiB), which maps onto the scalar-base form of
// the global load/store instructions.
template <typename S>
__device__ __forceinline__ S &at(S *base, uint32_t idx) {
    return *(S *)((char *)base + idx * (uint32_t)sizeof(S));
}

// Slots of one lattice as (plane, cell).  32- and 16-bit storage go through
// one buffer resource per plane: the plane base is scalar and the cell's byte
// offset a 32-bit per-lane value (buffer_load/store ... offen), a form the
// compiler cannot fold into per-lane 64-bit addresses kept live across the
// collision.  FP64 storage uses ordinary global addressing.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base) {
    // gfx9-family resource: stride 0, no range limit, raw 32-bit data format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, -1, 0x00020000);
}
// CP is the instruction's cache-policy field: 0 = default, CP_NT = the
// non-temporal bit (lines are streamed through L2 and the memory-side cache
// instead of being kept).  It pays only when a cell's 19 loads and 19 stores
// both carry it.  Phase A (own slots, whole-line stores) gains 5-8 % on
// lattices far larger than the caches; phase B gains only with row-aligned
// stores and loses up to 10 % with per-cell (partial-line) stores; lattices
// that partly fit the caches lose.  The choice per phase is a launch
// parameter (LBMDEV_NT_A / LBMDEV_NT_B) that the autotuner measures.
constexpr int CP_NT = 2;
template <int CP>
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, float) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, CP));
}
template <int CP>
__device__ __forceinline__ uint16_t buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, uint16_t) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, CP);
}
template <int CP>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, CP);
}
template <int CP>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t off, uint16_t v) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, off, 0, CP);
}

template <typename S, int CP = 0>
struct SlotIO {
    S *pop;
    long long stride;
    __device__ __forceinline__ SlotIO(S *p, const Geo &g) : pop(p), stride(g.stride) {}
    __device__ __forceinline__ S ld(int i, uint32_t cell) const {
        return buf_ld<CP>(plane_rsrc(pop + (long long)i * stride), cell * (uint32_t)sizeof(S), S());
    }
    __device__ __forceinline__ void st(int i, uint32_t cell, S v) const {
        buf_st<CP>(plane_rsrc(pop + (long long)i * stride), cell * (uint32_t)sizeof(S), v);
    }
};
template <int CP>
struct SlotIO<double, CP> {
    double *pop;
    long long stride;
    __device__ __forceinline__ SlotIO(double *p, const Geo &g) : pop(p), stride(g.stride) {}
    __device__ __forceinline__ double ld(int i, uint32_t cell) const { return at(plane(pop, Geo{0, 0, 0, 0, stride}, i), cell); }
    __device__ __forceinline__ void st(int i, uint32_t cell, double v) const {
        at(pop + (long long)i * stride, cell) = v;
    }
};

__device__ __forceinline__ int cidx(const Geo &g, int x, int y, int z) {
    return x + g.nx * (y + g.ny * z);
}

__device__ __forceinline__ int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// ------------------------------------------------------- stored-field access
// f*(x, i) of the last step, given the layout the last step left behind.
template <int FMT, int STATE>
__device__ __forceinline__ typename Pop<FMT>::T fstar(const typename Pop<FMT>::S *pop, const Geo &g,
                                                      int x, int y, int z, int i) {
    typedef Pop<FMT> P;
    if (STATE == ST_A) {
        int j = OPP[i];
        return P::dec(pop[(long long)j * g.stride + cidx(g, x, y, z)], j);
    } else if (STATE == ST_B) {
        int X = wrapc(x + EX[i], g.nx), Y = wrapc(y + EY[i], g.ny), Z = wrapc(z + EZ[i], g.nz);
        return P::dec(pop[(long long)i * g.stride + cidx(g, X, Y, Z)], i);
    } else {
        return P::dec(pop[(long long)i * g.stride + cidx(g, x, y, z)], i);
    }
}

__device__ __forceinline__ void handle_yz(const Geo &g, int periodic, int &y, int &z) {
    if (periodic) {
        y = wrapc(y, g.ny);
        z = wrapc(z, g.nz);
    } else {
        y = y < 0 ? 0 : (y > g.ny - 1 ? g.ny - 1 : y);
        z = z < 0 ? 0 : (z > g.nz - 1 ? g.nz - 1 : z);
    }
}

// The reference stream rule (lbm_stream.comp:85-128) for cell (x,y,z) and
// direction i, on the stored post-collision field.  qo = q[cell][opp(i)].
template <int FMT, int STATE>
__device__ __forceinline__ typename Pop<FMT>::T stream_value(const typename Pop<FMT>::S *pop,
                                                             const Geo &g, int periodic, int x, int y,
                                                             int z, int i, float qo) {
    typedef typename Pop<FMT>::T T;
    const int io = OPP[i];
    if (qo >= 0.0f) {
        T fos = fstar<FMT, STATE>(pop, g, x, y, z, io);
        T q = (T)qo;
        if (qo >= 0.5f) {
            T fis = fstar<FMT, STATE>(pop, g, x, y, z, i);
            T inv2q = T(1) / (T(2) * q);
            return inv2q * fos + (T(1) - inv2q) * fis;
        }
        int X = x + EX[i], Y = y + EY[i], Z = z + EZ[i];
        handle_yz(g, periodic, Y, Z);
        T ff = (X >= 0 && X < g.nx) ? fstar<FMT, STATE>(pop, g, X, Y, Z, io) : fos;
        T twoq = T(2) * q;
        return twoq * fos + (T(1) - twoq) * ff;
    }
    int X = x - EX[i], Y = y - EY[i], Z = z - EZ[i];
    handle_yz(g, periodic, Y, Z);
    if (X < 0 || X >= g.nx) return fstar<FMT, STATE>(pop, g, x, y, z, i);
    return fstar<FMT, STATE>(pop, g, X, Y, Z, i);
}

// Post-stream population f(x, i) as the reference holds it after a step.
template <int FMT, int STATE>
__device__ __forceinline__ typename Pop<FMT>::T post_stream(const typename Pop<FMT>::S *pop,
                                                            const Geo &g, int periodic, int x, int y,
                                                            int z, int i, float qo) {
    if (STATE == ST_INIT) return fstar<FMT, ST_INIT>(pop, g, x, y, z, i);
    return stream_value<FMT, STATE>(pop, g, periodic, x, y, z, i, qo);
}

// ------------------------------------------------------------ collision
template <typename T>
__device__ __forceinline__ T feq_i(int i, T rho, T ux, T uy, T uz, T c) {
    // w rho (1 + 3 eu + 4.5 eu^2 - 1.5 u^2), c = 1 - 1.5 u^2
    T eu = T(EX[i]) * ux + T(EY[i]) * uy + T(EZ[i]) * uz;
    return wgt<T>(i) * rho * (c + eu * (T(3) + T(4.5) * eu));
}

// Square root and reciprocal of the collision.  FP32: the hardware's
// v_sqrt_f32 / v_rcp_f32 (1 ulp; one instruction each, where a correctly
// rounded sqrtf or division expands to 10-18 instructions, about 14 % of the
// FP32 MRT + Smagorinsky collision).  The reference's shaders run the same
// operations at GLSL precision (division 2.5 ulp, sqrt derived from
// inversesqrt).  FP64: correctly rounded.
template <typename T>
__device__ __forceinline__ T tsqrt(T v);
template <>
__device__ __forceinline__ float tsqrt<float>(float v) { return __builtin_amdgcn_sqrtf(v); }
template <>
__device__ __forceinline__ double tsqrt<double>(double v) { return sqrt(v); }
template <typename T>
__device__ __forceinline__ T trcp(T v);
template <>
__device__ __forceinline__ float trcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <>
__device__ __forceinline__ double trcp<double>(double v) { return 1.0 / v; }

// |u| > 0.4 of the guard.  In FP32, sqrtf(u2) > 0.4f holds exactly when
// u2 > 0.1600000113... (the largest float whose correctly rounded square root
// is <= 0.4f is 0x3E23D70B), so the square root is not needed.
template <typename T>
__device__ __forceinline__ bool speed_exceeds(T u2);
template <>
__device__ __forceinline__ bool speed_exceeds<float>(float u2) { return u2 > __uint_as_float(0x3E23D70Bu); }
template <>
__device__ __forceinline__ bool speed_exceeds<double>(double u2) { return sqrt(u2) > 0.40; }

struct CollideParams {
    float tau, cs2x18, inx, iny, inz;
    int smag;
    unsigned long long *guard;  // running count of stability-guard resets
};

// lbm_collide.comp:313-626 for one non-body cell: Zou-He, moments, guard,
// Smagorinsky, then regularized / MRT / BGK.  f is updated by the Zou-He
// rewrite; fo receives the post-collision populations and rho, u the values
// written to the velocity buffer.  A reset by the stability guard is counted
// in p.guard.
template <int OP, typename T>
__device__ __forceinline__ void collide_fluid(T f[19], T fo[19], const CollideParams &p, int x, int nx,
                                              T &rho, T &ux, T &uy, T &uz) {
    if (x == 0) {  // Zou-He velocity inlet, lbm_collide.comp:336-375
        const T inx = (T)p.inx, iny = (T)p.iny, inz = (T)p.inz;
        T k0 = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
        T kn = f[2] + f[8] + f[10] + f[12] + f[14];
        T rin = (k0 + T(2) * kn) / (T(1) - inx);
        f[1] = f[2] + (T(1) / T(3)) * rin * inx;
        f[7] = f[10] + (T(1) / T(6)) * rin * (inx + iny);
        f[9] = f[8] + (T(1) / T(6)) * rin * (inx - iny);
        f[11] = f[14] + (T(1) / T(6)) * rin * (inx + inz);
        f[13] = f[12] + (T(1) / T(6)) * rin * (inx - inz);
    }
    if (x == nx - 1) {  // Zou-He pressure outlet, rho = 1, lbm_collide.comp:379-408
        T k0 = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
        T kp = f[1] + f[7] + f[9] + f[11] + f[13];
        T uo = T(-1) + (k0 + T(2) * kp);
        f[2] = f[1] - (T(1) / T(3)) * uo;
        f[10] = f[7] - (T(1) / T(6)) * uo;
        f[8] = f[9] - (T(1) / T(6)) * uo;
        f[14] = f[11] - (T(1) / T(6)) * uo;
        f[12] = f[13] - (T(1) / T(6)) * uo;
    }
    // moments (lbm_collide.comp:411-424)
    // MRT works through the sums sp and differences dp of the nine opposite
    // pairs (1,2) (3,4) (5,6) (7,10) (8,9) (11,14) (12,13) (15,18) (16,17):
    // every even moment is a sum of sp, every odd moment one of dp, and the
    // collision's correction splits the same way (below).
    T jx, jy, jz;
    T sp[10], m16 = 0, m17 = 0, m18 = 0;
    if constexpr (OP == OP_MRT) {
        constexpr int PA[10] = {0, 1, 3, 5, 7, 8, 11, 12, 15, 16};
        T dp[10];
#pragma unroll
        for (int q = 1; q < 10; q++) {
            sp[q] = f[PA[q]] + f[OPP[PA[q]]];
            dp[q] = f[PA[q]] - f[OPP[PA[q]]];
        }
        rho = f[0] + ((sp[1] + sp[2]) + sp[3]) + (((sp[4] + sp[5]) + (sp[6] + sp[7])) + (sp[8] + sp[9]));
        const T t45 = dp[4] - dp[5], t67 = dp[6] - dp[7], u45 = dp[4] + dp[5], u67 = dp[6] + dp[7];
        const T t89 = dp[8] - dp[9], u89 = dp[8] + dp[9];
        jx = dp[1] + (t45 + t67);
        jy = dp[2] + (u45 + t89);
        jz = dp[3] + (u67 + u89);
        m16 = t45 - t67;  // f7 - f8 + f9 - f10 - f11 + f12 - f13 + f14
        m17 = t89 - u45;  // -f7 - f8 + f9 + f10 + f15 - f16 + f17 - f18
        m18 = u67 - u89;  // f11 + f12 - f13 - f14 - f15 - f16 + f17 + f18
    } else {
        rho = T(0);
#pragma unroll
        for (int i = 0; i < 19; i++) rho += f[i];
        jx = (f[1] - f[2]) + (f[7] - f[8]) + (f[9] - f[10]) + (f[11] - f[12]) + (f[13] - f[14]);
        jy = (f[3] - f[4]) + (f[7] - f[9]) + (f[8] - f[10]) + (f[15] - f[16]) + (f[17] - f[18]);
        jz = (f[5] - f[6]) + (f[11] - f[13]) + (f[12] - f[14]) + (f[15] - f[17]) + (f[16] - f[18]);
    }
    if (rho > T(0.0001)) {
        T ir = trcp<T>(rho);
        ux = jx * ir;
        uy = jy * ir;
        uz = jz * ir;
    } else {
        ux = jx;
        uy = jy;
        uz = jz;
    }
    // stability guard (lbm_collide.comp:429-440): reset to feq(1, 0) = w
    if (rho < T(0.3) || rho > T(2.0) || speed_exceeds<T>(ux * ux + uy * uy + uz * uz) || isinf(rho) ||
        isnan(rho)) {
#pragma unroll
        for (int i = 0; i < 19; i++) fo[i] = wgt<T>(i);
        rho = T(1);
        ux = uy = uz = T(0);
        atomicAdd(p.guard, 1ull);  // one atomic per reset cell; the guard normally never fires
        return;
    }
    const T u2 = ux * ux + uy * uy + uz * uz;
    const T c = T(1) - T(1.5) * u2;
    T tau_eff = (T)p.tau;
    T Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
    if (OP == OP_MRT && p.smag) {
        // as below, from the pair sums
        const T sxy = sp[4] + sp[5], sxz = sp[6] + sp[7], syz = sp[8] + sp[9];
        const T r3 = rho * (T(1) / T(3));
        Pxx = (sp[1] + sxy + sxz) - (r3 + rho * ux * ux);
        Pyy = (sp[2] + sxy + syz) - (r3 + rho * uy * uy);
        Pzz = (sp[3] + sxz + syz) - (r3 + rho * uz * uz);
        Pxy = (sp[4] - sp[5]) - rho * ux * uy;
        Pxz = (sp[6] - sp[7]) - rho * ux * uz;
        Pyz = (sp[8] - sp[9]) - rho * uy * uz;
    } else if (p.smag || OP == OP_REG) {
        // non-equilibrium stress sum_i e_a e_b (f_i - feq_i), using the exact
        // D3Q19 identity sum_i e_a e_b feq_i = rho (delta_ab / 3 + u_a u_b)
        T sxy = f[7] + f[8] + f[9] + f[10];
        T sxz = f[11] + f[12] + f[13] + f[14];
        T syz = f[15] + f[16] + f[17] + f[18];
        T r3 = rho * (T(1) / T(3));
        Pxx = (f[1] + f[2] + sxy + sxz) - (r3 + rho * ux * ux);
        Pyy = (f[3] + f[4] + sxy + syz) - (r3 + rho * uy * uy);
        Pzz = (f[5] + f[6] + sxz + syz) - (r3 + rho * uz * uz);
        Pxy = ((f[7] + f[10]) - (f[8] + f[9])) - rho * ux * uy;
        Pxz = ((f[11] + f[14]) - (f[12] + f[13])) - rho * ux * uz;
        Pyz = ((f[15] + f[18]) - (f[16] + f[17])) - rho * uy * uz;
    }
    if (p.smag) {  // lbm_collide.comp:448-484
        T pn = tsqrt<T>(Pxx * Pxx + Pyy * Pyy + Pzz * Pzz + T(2) * (Pxy * Pxy + Pxz * Pxz + Pyz * Pyz));
        T tau = (T)p.tau;
        tau_eff = T(0.5) * (tau + tsqrt<T>(tau * tau + (T)p.cs2x18 * pn * trcp<T>(rho + T(1e-10))));
    }
    const T omega = trcp<T>(tau_eff);
    if (OP == OP_REG) {  // lbm_collide.comp:488-543
        const T a = (T(1) - omega) * T(4.5);
        const T tr3 = (Pxx + Pyy + Pzz) * (T(1) / T(3));
        T QP[19];
        QP[0] = -tr3;
        QP[1] = QP[2] = Pxx - tr3;
        QP[3] = QP[4] = Pyy - tr3;
        QP[5] = QP[6] = Pzz - tr3;
        QP[7] = QP[10] = Pxx + Pyy + T(2) * Pxy - tr3;
        QP[8] = QP[9] = Pxx + Pyy - T(2) * Pxy - tr3;
        QP[11] = QP[14] = Pxx + Pzz + T(2) * Pxz - tr3;
        QP[12] = QP[13] = Pxx + Pzz - T(2) * Pxz - tr3;
        QP[15] = QP[18] = Pyy + Pzz + T(2) * Pyz - tr3;
        QP[16] = QP[17] = Pyy + Pzz - T(2) * Pyz - tr3;
#pragma unroll
        for (int i = 0; i < 19; i++)
            fo[i] = feq_i<T>(i, rho, ux, uy, uz, c) + a * wgt<T>(i) * QP[i];
    } else if (OP == OP_MRT) {  // lbm_collide.comp:544-616
        // moments of the d'Humieres basis from the pair sums (m0, m3, m5, m7
        // are rho and j; m16..m18 came with them)
        const T sf = (sp[1] + sp[2]) + sp[3];
        const T sXY = sp[4] + sp[5], sXZ = sp[6] + sp[7], sYZ = sp[8] + sp[9];
        const T se = (sXY + sXZ) + sYZ;
        const T m1 = T(-30) * f[0] - T(11) * sf + T(8) * se;
        const T m2 = T(12) * f[0] - T(4) * sf + se;
        const T m4 = jx - T(5) * (f[1] - f[2]);  // -4 (f1 - f2) + ax
        const T m6 = jy - T(5) * (f[3] - f[4]);
        const T m8 = jz - T(5) * (f[5] - f[6]);
        const T A = T(2) * sp[1] - sp[2] - sp[3], B = (sXY + sXZ) - T(2) * sYZ;
        const T m9 = A + B, m10 = B - T(2) * A;
        const T C = sp[2] - sp[3], D = sXY - sXZ;
        const T m11 = C + D, m12 = D - T(2) * C;
        const T m13 = sp[4] - sp[5], m14 = sp[8] - sp[9], m15 = sp[6] - sp[7];
        // equilibrium moments (lbm_collide.comp:560-580)
        const T e9 = rho * (T(2) * ux * ux - uy * uy - uz * uz), e11 = rho * (uy * uy - uz * uz);
        // Relaxation rates (lbm_collide.comp:585-604) over the squared row
        // norms of M (lbm_collide.comp:188-206): c_k = s_k (m_k - meq_k) / |M_k|^2.
        // f* = M^-1 (m - S (m - meq)) is applied as f* = f - M^T c: the
        // conserved moments (rho, j; s = 0) drop out, so mass and momentum
        // are conserved to unbiased rounding.
        const T c1 = (m1 - (T(-11) * rho + T(19) * rho * u2)) * T(1.19 / 2394.0);
        const T c2 = (m2 - (T(3) * rho - T(5.5) * rho * u2)) * T(1.4 / 252.0);
        const T c4 = (m4 - T(-2) / T(3) * jx) * T(1.2 / 40.0);
        const T c6 = (m6 - T(-2) / T(3) * jy) * T(1.2 / 40.0);
        const T c8 = (m8 - T(-2) / T(3) * jz) * T(1.2 / 40.0);
        const T c9 = omega * (m9 - e9) * T(1.0 / 36.0);
        const T c10 = (m10 - T(-0.5) * e9) * T(1.4 / 72.0);
        const T c11 = omega * (m11 - e11) * T(1.0 / 12.0);
        const T c12 = (m12 - T(-0.5) * e11) * T(1.4 / 24.0);
        const T c13 = omega * (m13 - rho * ux * uy) * T(0.25);
        const T c14 = omega * (m14 - rho * uy * uz) * T(0.25);
        const T c15 = omega * (m15 - rho * ux * uz) * T(0.25);
        const T c16 = m16 * T(1.98 * 0.125), c17 = m17 * T(1.98 * 0.125), c18 = m18 * T(1.98 * 0.125);
        // M^T c per pair: the even part E (shared by both members) and the
        // odd part O (+ for the first member, - for the second)
        const T b0 = T(-11) * c1 - T(4) * c2, b1 = T(8) * c1 + c2;
        const T p910 = c9 + c10, p1112 = c11 + c12;
        const T t0 = b0 - c9 + T(2) * c10, u0 = c11 - T(2) * c12;
        const T E1 = b0 + T(2) * c9 - T(4) * c10, E2 = t0 + u0, E3 = t0 - u0;
        const T q0 = b1 + p910, r0 = q0 + p1112, r1 = q0 - p1112, v0 = b1 - T(2) * p910;
        const T E4 = r0 + c13, E5 = r0 - c13, E6 = r1 + c15, E7 = r1 - c15, E8 = v0 + c14, E9 = v0 - c14;
        const T g0 = c6 - c17, h0 = c4 + c16, k0 = c8 + c18, l0 = c4 - c16, n0 = c8 - c18, o0 = c6 + c17;
        const T O4 = g0 + h0, O5 = g0 - h0, O6 = k0 + l0, O7 = k0 - l0, O8 = n0 + o0, O9 = n0 - o0;
        fo[0] = f[0] - (T(-30) * c1 + T(12) * c2);
        fo[1] = f[1] - (E1 - T(4) * c4);
        fo[2] = f[2] - (E1 + T(4) * c4);
        fo[3] = f[3] - (E2 - T(4) * c6);
        fo[4] = f[4] - (E2 + T(4) * c6);
        fo[5] = f[5] - (E3 - T(4) * c8);
        fo[6] = f[6] - (E3 + T(4) * c8);
        fo[7] = f[7] - (E4 + O4);
        fo[10] = f[10] - (E4 - O4);
        fo[8] = f[8] - (E5 + O5);
        fo[9] = f[9] - (E5 - O5);
        fo[11] = f[11] - (E6 + O6);
        fo[14] = f[14] - (E6 - O6);
        fo[12] = f[12] - (E7 + O7);
        fo[13] = f[13] - (E7 - O7);
        fo[15] = f[15] - (E8 + O8);
        fo[18] = f[18] - (E8 - O8);
        fo[16] = f[16] - (E9 + O9);
        fo[17] = f[17] - (E9 - O9);
    } else {  // BGK, lbm_collide.comp:618-625
#pragma unroll
        for (int i = 0; i < 19; i++) fo[i] = f[i] - omega * (f[i] - feq_i<T>(i, rho, ux, uy, uz, c));
    }
}

template <int FMT>
struct StepArgs {
    typename Pop<FMT>::S *pop;
    const uint8_t *mask;
    const int32_t *ovr_index;
    const uint32_t *ovr_bits;  // per listed cell: directions whose value comes from the list
    const typename Pop<FMT>::T *ovr_vals;
    int ovr_stride;
    int use_ovr;
    float4 *vel;
    int write_vel;
    int remap;  // 1: XCD-contiguous workgroup order (xcd_remap)
    Geo g;
    CollideParams cp;
};

// One cell of the fused stream + collide update.  PHASE 0 = A, 1 = B (see
// file header).
// Directions with ex = +1 and ex = -1, for the row-aligned phase-B stores.
constexpr int XPLUS[5] = {1, 7, 9, 11, 13};
constexpr int XMINUS[5] = {2, 8, 10, 12, 14};

// The value of lane l - 1 (UP) or l + 1 (!UP) of the 64-lane wave, through a
// whole-wave DPP shift (one VALU instruction, no LDS traffic).  The lane with
// no source (0 for UP, 63 otherwise) keeps its own value.
template <bool UP>
__device__ __forceinline__ float wave_shift1(float v) {
    const int i = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(i, i, UP ? 0x138 : 0x130, 0xF, 0xF, false));
}
template <bool UP>
__device__ __forceinline__ double wave_shift1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const int slo = __builtin_amdgcn_update_dpp(lo, lo, UP ? 0x138 : 0x130, 0xF, 0xF, false);
    const int shi = __builtin_amdgcn_update_dpp(hi, hi, UP ? 0x138 : 0x130, 0xF, 0xF, false);
    return __longlong_as_double(((long long)shi << 32) | (unsigned)slo);
}
template <bool UP>
__device__ __forceinline__ uint16_t wave_shift1(uint16_t v) {
    const int i = (int)v;
    return (uint16_t)__builtin_amdgcn_update_dpp(i, i, UP ? 0x138 : 0x130, 0xF, 0xF, false);
}

// The same uniform pointer, opaque to the compiler (an empty asm statement
// that claims to modify it).  Plane resources derived from it are then
// rebuilt where they are used instead of being kept live in scalar registers
// across the collision, which would cost a wave per SIMD of occupancy.
template <typename S>
__device__ __forceinline__ S *opaque_ptr(S *p) {
    const uint64_t v = (uint64_t)p;
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    asm volatile("" : "+s"(lo), "+s"(hi));
    return (S *)(((uint64_t)hi << 32) | lo);
}

// Position of direction i in XPLUS (ex = +1) or XMINUS (ex = -1).
__host__ __device__ constexpr int xdir_slot(int i) {
    return (i == 1 || i == 2) ? 0 : (i == 7 || i == 8) ? 1 : (i == 9 || i == 10) ? 2 : (i == 11 || i == 12) ? 3 : 4;
}

template <int FMT, int OP, int PHASE, int ROW = 0, int NT = 0>
__device__ __forceinline__ void step_cell(const StepArgs<FMT> &a, const int cell) {
    typedef Pop<FMT> P;
    typedef typename P::T T;
    const Geo &g = a.g;
    int x, y, z;
    decompose(g, cell, x, y, z);
    // torus neighbour offsets
    const int X[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
    const int pl = g.nx * g.ny;
    const int Y[3] = {(y == 0 ? g.ny - 1 : y - 1) * g.nx, y * g.nx, (y == g.ny - 1 ? 0 : y + 1) * g.nx};
    const int Z[3] = {(z == 0 ? g.nz - 1 : z - 1) * pl, z * pl, (z == g.nz - 1 ? 0 : z + 1) * pl};

    const uint8_t m = a.mask[cell];
    const bool ovr = a.use_ovr && (m & LBMDEV_MASK_OVR);
    const SlotIO<typename P::S, NT ? CP_NT : 0> io(a.pop, g);
    // Row geometry for ROW = 1: wave w of the workgroup covers row positions
    // [64 rw, 64 rw + 64) of its row; wprev / wnext are the row's neighbouring
    // waves, wrapping at the row ends (x is periodic on the torus).
    // The wave index is uniform: readfirstlane keeps it and its neighbours
    // in scalar registers.
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpr = ROW ? g.nx >> 6 : 1, rw = w % wpr, w0 = w - rw;
    const int wprev = rw == 0 ? w0 + wpr - 1 : w - 1, wnext = rw == wpr - 1 ? w0 : w + 1;
    T f[19];
    if (PHASE == 1 && ROW) {
        // Row-aligned phase-B loads, the mirror of the row-aligned stores
        // below.  Cell x needs slot (opp(i), x - e_i).  Lane l of a wave
        // starting at x0 loads slot (opp(i), x0 + l) at the y/z offset of
        // direction i, so every load covers whole cache lines, and the value
        // moves to the lane that needs it: for ex_i = +1 from lane l - 1, for
        // ex_i = -1 from lane l + 1 (DPP), the wave-edge values through LDS
        // from the neighbouring wave of the row.  Every lane loads, listed
        // cells included, because its neighbours need the values.
        typedef typename P::S S;
        __shared__ S lp[8][5], lm[8][5];
        S raw[19];
#pragma unroll
        for (int i = 0; i < 19; i++)
            raw[i] = io.ld(OPP[i], (uint32_t)(X[1] + Y[1 - EY[i]] + Z[1 - EZ[i]]));
        // One branch per edge lane for the LDS writes; the reads are
        // uniform-address broadcasts taken by every lane and kept by a
        // select, so no per-direction exec-mask branches.
        if (lane == 63) {
#pragma unroll
            for (int k = 0; k < 5; k++) lp[w][k] = raw[XPLUS[k]];
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) lm[w][k] = raw[XMINUS[k]];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 19; i++) {
            S v = raw[i];
            if (EX[i] == 1) {
                const S e = lp[wprev][xdir_slot(i)];
                v = wave_shift1<true>(v);
                v = lane == 0 ? e : v;
            } else if (EX[i] == -1) {
                const S e = lm[wnext][xdir_slot(i)];
                v = wave_shift1<false>(v);
                v = lane == 63 ? e : v;
            }
            f[i] = P::dec(v, OPP[i]);
        }
    } else if (PHASE == 0) {
#pragma unroll
        for (int i = 0; i < 19; i++) f[i] = P::dec(io.ld(i, (uint32_t)cell), i);
    } else {
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const int j = OPP[i];
            const uint32_t src = (uint32_t)(X[1 - EX[i]] + Y[1 - EY[i]] + Z[1 - EZ[i]]);
            f[i] = P::dec(io.ld(j, src), j);
        }
    }
    // A listed cell's loaded value is the plain pull of the torus streaming,
    // which is the reference's stream rule for every direction without a
    // Bouzidi link and with its source inside the domain (or across a
    // periodic face).  The other directions take the prepass's value.
    if (ovr) {
        const int k = a.ovr_index[cell];
        const uint32_t bits = a.ovr_bits[k];
#pragma unroll
        for (int i = 0; i < 19; i++)
            if (bits & (1u << i)) f[i] = a.ovr_vals[(long long)i * a.ovr_stride + k];
    }

    T fo[19];
    T rho, ux, uy, uz;
    if (m & LBMDEV_MASK_BODY) {  // full bounce-back, lbm_collide.comp:322-330
#pragma unroll
        for (int i = 0; i < 19; i++) fo[i] = f[OPP[i]];
        rho = T(1);
        ux = uy = uz = T(0);
    } else {
        collide_fluid<OP, T>(f, fo, a.cp, x, g.nx, rho, ux, uy, uz);
    }

    if (a.write_vel) {
        // a non-temporal store: the 16 B/cell velocity output streams past the
        // caches instead of displacing the populations' lines (0.5 % per step
        // on 512x256x256 writing it every 10th step)
        typedef float v4f __attribute__((ext_vector_type(4)));
        const v4f v = {(float)ux, (float)uy, (float)uz, (float)rho};
        __builtin_nontemporal_store(v, (v4f *)&a.vel[cell]);
    }

    if (PHASE == 0) {
        const SlotIO<typename P::S, NT ? CP_NT : 0> ios(opaque_ptr(a.pop), g);
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const int j = OPP[i];
            ios.st(j, (uint32_t)cell, P::enc(fo[i], j));
        }
    } else if (!ROW) {
        // Offsets recomputed from an opaque copy of the cell index (see the
        // row-aligned branch below).
        int c2 = cell;
        asm volatile("" : "+v"(c2));
        int x2, y2, z2;
        decompose(g, c2, x2, y2, z2);
        const int Xs[3] = {x2 == 0 ? g.nx - 1 : x2 - 1, x2, x2 == g.nx - 1 ? 0 : x2 + 1};
        const int Ys[3] = {(y2 == 0 ? g.ny - 1 : y2 - 1) * g.nx, y2 * g.nx, (y2 == g.ny - 1 ? 0 : y2 + 1) * g.nx};
        const int Zs[3] = {(z2 == 0 ? g.nz - 1 : z2 - 1) * pl, z2 * pl, (z2 == g.nz - 1 ? 0 : z2 + 1) * pl};
        const SlotIO<typename P::S, NT ? CP_NT : 0> ios(opaque_ptr(a.pop), g);
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const uint32_t dst = (uint32_t)(Xs[1 + EX[i]] + Ys[1 + EY[i]] + Zs[1 + EZ[i]]);
            ios.st(i, dst, P::enc(fo[i], i));
        }
    } else {
        // Row-aligned phase-B stores (the workgroup holds whole x-rows, nx a
        // multiple of 64).  Slot (i, x + e_i) is read in this phase only by
        // cell x itself, so after a workgroup barrier any lane of the row may
        // store it.  Lane l of a wave starting at row position x0 stores
        // slot (i, x0 + l): for ex_i = +1 the value of cell x0 + l - 1 (the
        // lane below, or lane 63 of the previous wave of the row, wrapping),
        // for ex_i = -1 that of cell x0 + l + 1; every store then covers
        // whole cache lines.  Values move between lanes by whole-wave DPP
        // shifts; the wave-edge values come through LDS.
        __shared__ T xp[8][5], xm[8][5];
        // The store offsets are recomputed from an opaque copy of the cell
        // index rather than kept live from the loads (whose y/z offsets are
        // the same nine), which would cost a wave per SIMD of occupancy.
        const SlotIO<typename P::S, NT ? CP_NT : 0> ios(opaque_ptr(a.pop), g);
        int c2 = cell;
        asm volatile("" : "+v"(c2));
        int x2, y2, z2;
        decompose(g, c2, x2, y2, z2);
        const int Ys[3] = {(y2 == 0 ? g.ny - 1 : y2 - 1) * g.nx, y2 * g.nx, (y2 == g.ny - 1 ? 0 : y2 + 1) * g.nx};
        const int Zs[3] = {(z2 == 0 ? g.nz - 1 : z2 - 1) * pl, z2 * pl, (z2 == g.nz - 1 ? 0 : z2 + 1) * pl};
        if (lane == 63) {
#pragma unroll
            for (int k = 0; k < 5; k++) xp[w][k] = fo[XPLUS[k]];
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) xm[w][k] = fo[XMINUS[k]];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const uint32_t dst = (uint32_t)(x2 + Ys[1 + EY[i]] + Zs[1 + EZ[i]]);
            T v = fo[i];
            if (EX[i] == 1) {
                const T e = xp[wprev][xdir_slot(i)];
                v = wave_shift1<true>(v);
                v = lane == 0 ? e : v;
            } else if (EX[i] == -1) {
                const T e = xm[wnext][xdir_slot(i)];
                v = wave_shift1<false>(v);
                v = lane == 63 ? e : v;
            }
            ios.st(i, dst, P::enc(v, i));
        }
    }
}

// The fused kernel over one lattice, BS lanes per workgroup.
// ROW = 1 (phase B only): row-aligned loads and stores; the launch
// guarantees that every workgroup holds whole rows and no lane is past the
// lattice.
// NT = 1: non-temporal population loads and stores.
template <int FMT, int OP, int PHASE, int BS, int ROW = 0, int NT = 0>
__global__ __launch_bounds__(BS) void k_step(StepArgs<FMT> a) {
    const int cell = (a.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x) * BS + threadIdx.x;
    if (!ROW && cell >= a.g.ncell) return;
    step_cell<FMT, OP, PHASE, ROW, NT>(a, cell);
}

// ---------------------------------------------------------- batched tunnels
// Many lattices of different sizes and parameters advanced by one launch:
// workgroup b (after the XCD remap) belongs to tunnel block_tunnel[b] and
// covers cells (b - blockStart) * BS + [0, BS) of it.  Every cell runs the
// same step_cell as the single-lattice kernel.
struct TunnelDesc {
    Geo g;
    long long popOff;  // element offset of the tunnel's 19 planes
    int cellOff;       // offset into the concatenated per-cell arrays
    int blockStart;    // first workgroup of the tunnel
    float tau, inx, iny, inz;
};

template <int FMT>
struct BatchArgs {
    typename Pop<FMT>::S *pop;
    const uint8_t *mask;
    const int32_t *ovr_index;
    const uint32_t *ovr_bits;
    const typename Pop<FMT>::T *ovr_vals;
    int ovr_stride;
    int use_ovr;
    float4 *vel;
    int write_vel;
    float cs2x18;
    int smag;
    int remap;  // 1: XCD-contiguous workgroup order (xcd_remap)
    unsigned long long *guard;  // running count of stability-guard resets
    const TunnelDesc *desc;
    const int32_t *block_tunnel;
};

// ROW = 1 (phase B only): every tunnel's rows fill whole workgroups.
// NT = 1: non-temporal population loads and stores.
template <int FMT, int OP, int PHASE, int BS, int ROW = 0, int NT = 0>
__global__ __launch_bounds__(BS) void k_step_batch(BatchArgs<FMT> b) {
    const int blk = b.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const TunnelDesc d = b.desc[b.block_tunnel[blk]];
    const int cell = (blk - d.blockStart) * BS + threadIdx.x;
    if (!ROW && cell >= d.g.ncell) return;
    StepArgs<FMT> a;
    a.pop = b.pop + d.popOff;
    a.mask = b.mask + d.cellOff;
    a.ovr_index = b.ovr_index + d.cellOff;
    a.ovr_bits = b.ovr_bits;
    a.ovr_vals = b.ovr_vals;
    a.ovr_stride = b.ovr_stride;
    a.use_ovr = b.use_ovr;
    a.vel = b.vel + d.cellOff;
    a.write_vel = b.write_vel;
    a.g = d.g;
    a.cp.tau = d.tau;
    a.cp.cs2x18 = b.cs2x18;
    a.cp.inx = d.inx;
    a.cp.iny = d.iny;
    a.cp.inz = d.inz;
    a.cp.smag = b.smag;
    a.cp.guard = b.guard;
    step_cell<FMT, OP, PHASE, ROW, NT>(a, cell);
}

// One prepass work item: the list slot k of the cell (-1 for padding), the
// cell (local to its tunnel), the Bouzidi q of direction opp(i) at it, and
// its tunnel (0 for one lattice).  One 16-byte load gives the lane all it
// needs before the population loads.
struct alignas(16) PrepassItem {
    int32_t k, cell;
    float qo;
    int32_t tunnel;
};
static_assert(sizeof(PrepassItem) == 16, "PrepassItem is one 16-byte load");

// Work item idx as one 16-byte load.  A buffer-load intrinsic: a plain struct
// load is split into its fields, and the loads of cell and q are then sunk
// below the k < 0 test, which costs a second dependent memory round trip.
// The list holds at most 2^28 items, so the byte offset fits in 32 bits.
__device__ __forceinline__ PrepassItem load_item(const PrepassItem *items, int idx) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u v = __builtin_amdgcn_raw_buffer_load_b128(
        __builtin_amdgcn_make_buffer_rsrc(const_cast<PrepassItem *>(items), 0, -1, 0x00020000), (uint32_t)idx * 16u,
        0, 0);
    return PrepassItem{(int32_t)v.x, (int32_t)v.y, __uint_as_float(v.z), (int32_t)v.w};
}

template <int FMT>
struct BoundArgs {
    const typename Pop<FMT>::S *pop;
    typename Pop<FMT>::T *ovr_vals;
    int n_ovr;
    int periodic;
    Geo g;
    // Work list: entries seg[i] .. seg[i+1] - 1 of items are the (cell,
    // direction i) pairs that need the rule (ovr_bits), each direction's
    // segment padded with k = -1 to a multiple of 64.
    const PrepassItem *items;
    int seg[20];
};

// Boundary prepass: the incoming populations of the listed cells under the
// reference stream rule, evaluated on the stored post-collision field, for
// the (cell, direction) pairs whose value the plain torus pull of the main
// kernel does not give (a Bouzidi link, a source across an x face or a
// clamped y/z face).  One lane per pair; the pairs are grouped by direction
// and each group padded to a multiple of 64, so each wavefront evaluates a
// single direction (found from its uniform first index and dispatched to a
// constant-direction copy of the rule).
#define LBM_PREPASS_CASE(I) \
    case I: v = stream_value<FMT, STATE>(pop, g, periodic, x, y, z, I, qo); break;

template <int FMT, int STATE>
__device__ __forceinline__ typename Pop<FMT>::T prepass_value(const typename Pop<FMT>::S *pop, const Geo &g,
                                                              int periodic, int x, int y, int z, int i, float qo) {
    typename Pop<FMT>::T v = 0;
    switch (i) {
        LBM_PREPASS_CASE(0) LBM_PREPASS_CASE(1) LBM_PREPASS_CASE(2) LBM_PREPASS_CASE(3) LBM_PREPASS_CASE(4)
        LBM_PREPASS_CASE(5) LBM_PREPASS_CASE(6) LBM_PREPASS_CASE(7) LBM_PREPASS_CASE(8) LBM_PREPASS_CASE(9)
        LBM_PREPASS_CASE(10) LBM_PREPASS_CASE(11) LBM_PREPASS_CASE(12) LBM_PREPASS_CASE(13)
        LBM_PREPASS_CASE(14) LBM_PREPASS_CASE(15) LBM_PREPASS_CASE(16) LBM_PREPASS_CASE(17)
        LBM_PREPASS_CASE(18)
    }
    return v;
}
#undef LBM_PREPASS_CASE

// The directions of listed cell k (slot k of a [19][qstride] q table) whose
// incoming value the plain torus pull does not give: a Bouzidi link (q of the
// opposite direction set), a source across an x face, or across a clamped
// y/z face.
static uint32_t special_dirs(int nx, int ny, int nz, int periodic, int cell, const float *q, long long qstride,
                             long long k) {
    const int x = cell % nx, t = cell / nx, y = t % ny, z = t / ny;
    uint32_t bits = 0;
    for (int i = 0; i < 19; i++) {
        const int X = x - EX[i], Y = y - EY[i], Z = z - EZ[i];
        const bool link = q[(long long)OPP[i] * qstride + k] >= 0.0f;
        const bool xout = X < 0 || X >= nx;
        const bool yzout = !periodic && (Y < 0 || Y >= ny || Z < 0 || Z >= nz);
        if (link || xout || yzout) bits |= 1u << i;
    }
    return bits;
}

// The prepass work list of n listed cells (cells[k], q planes q[i*qstride +
// k], tunnel tun[k] or 0 when tun is NULL) with direction sets bits[k]: seg
// and a malloc'd item array (NULL with seg[19] = 0 when empty).  Returns 1,
// 0 on allocation failure, or -1 when the list would exceed 2^28 items.
static int build_items(int n, const uint32_t *bits, const int32_t *cells, const float *q, long long qstride,
                       const int32_t *tun, PrepassItem **items, int seg[20]) {
    long long total = 0;
    *items = NULL;
    for (int i = 0; i < 19; i++) {
        long long c = 0;
        for (int k = 0; k < n; k++) c += (bits[k] >> i) & 1;
        seg[i] = (int)total;
        total += (c + 63) & ~63LL;
        if (total > (1LL << 28)) {
            seg[19] = 0;
            return -1;
        }
    }
    seg[19] = (int)total;
    if (total == 0) return 1;
    PrepassItem *p = (PrepassItem *)malloc(sizeof(PrepassItem) * (size_t)total);
    if (!p) return 0;
    for (int i = 0; i < 19; i++) {
        int w = seg[i];
        for (int k = 0; k < n; k++)
            if ((bits[k] >> i) & 1) p[w++] = PrepassItem{k, cells[k], q[OPP[i] * qstride + k], tun ? tun[k] : 0};
        while (w < seg[i + 1]) p[w++] = PrepassItem{-1, 0, 0.0f, 0};
    }
    *items = p;
    return 1;
}

// The direction of the wavefront holding list entry idx (uniform: segments
// and wavefronts are 64-aligned), or 19 past the last segment.
template <int FMT>
__device__ __forceinline__ int prepass_dir(const BoundArgs<FMT> &a, int idx) {
    const int wbase = __builtin_amdgcn_readfirstlane(idx & ~63);
    int i = 0;
#pragma unroll
    for (int j = 1; j < 20; j++) i += wbase >= a.seg[j];
    return i;
}

template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_prepass(BoundArgs<FMT> a) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    const int i = prepass_dir(a, idx);
    if (i >= 19) return;
    const PrepassItem it = load_item(a.items, idx);
    if (it.k < 0) return;
    int x, y, z;
    decompose(a.g, it.cell, x, y, z);
    a.ovr_vals[(long long)i * a.n_ovr + it.k] = prepass_value<FMT, STATE>(a.pop, a.g, a.periodic, x, y, z, i, it.qo);
}

// Batched boundary prepass: each work item names its tunnel.
template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_prepass_batch(BoundArgs<FMT> a, const TunnelDesc *desc) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    const int i = prepass_dir(a, idx);
    if (i >= 19) return;
    const PrepassItem it = load_item(a.items, idx);
    if (it.k < 0) return;
    const TunnelDesc d = desc[it.tunnel];
    int x, y, z;
    decompose(d.g, it.cell, x, y, z);
    a.ovr_vals[(long long)i * a.n_ovr + it.k] =
        prepass_value<FMT, STATE>(a.pop + d.popOff, d.g, a.periodic, x, y, z, i, it.qo);
}

// ---------------------------------------------------------------- force
constexpr int FORCE_BLOCKS = 512;

template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block reduction of NV doubles and NI int64 values into partial slots.
template <int NV, int NI>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], long long (&iv)[NI], double *pd,
                                                   long long *pi) {
    __shared__ double sd[BLOCK / 64][NV > 0 ? NV : 1];
    __shared__ long long si[BLOCK / 64][NI > 0 ? NI : 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
#pragma unroll
    for (int k = 0; k < NI; k++) iv[k] = wave_sum(iv[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) sd[w][k] = v[k];
#pragma unroll
        for (int k = 0; k < NI; k++) si[w][k] = iv[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int j = 0; j < BLOCK / 64; j++) s += sd[j][k];
            pd[(long long)blockIdx.x * NV + k] = s;
        }
        for (int k = 0; k < NI; k++) {
            long long s = 0;
            for (int j = 0; j < BLOCK / 64; j++) s += si[j][k];
            pi[(long long)blockIdx.x * NI + k] = s;
        }
    }
}

template <int FMT>
struct ForceArgs {
    const typename Pop<FMT>::S *pop;
    const uint8_t *mask;
    const int32_t *ovr_cells;
    const float *ovr_q;
    const float4 *vel;
    int n_ovr;
    int q_stride;  // plane stride of ovr_q (n_ovr for one lattice, the batch total otherwise)
    int periodic;
    Geo g;
    double *pd;
    long long *pi;
};

template <typename T>
__device__ __forceinline__ long long fixed_round(T v) {
    return (long long)round((double)(v * T(10000)));
}

// Momentum exchange (lbm_force.comp:94-151) over the boundary list: for every
// fluid cell with links, F = sum e_i (f_new_i + f_opp(i)), P = sum e_i (feq_i
// + feq_opp(i)) with rho, u from the velocity buffer.  Per-cell values are
// accumulated both in FP64 and as the reference's round(x*1e4) integers.
// Block `blk` of `nblk` over the list: the lane's share of the sums.
template <int FMT, int STATE>
__device__ __forceinline__ void force_block(const ForceArgs<FMT> &a, int blk, int nblk, double (&acc)[6],
                                            long long (&fix)[7]) {
    typedef typename Pop<FMT>::T T;
    for (int k = blk * BLOCK + threadIdx.x; k < a.n_ovr; k += nblk * BLOCK) {
        const int cell = a.ovr_cells[k];
        if (a.mask[cell] & (LBMDEV_MASK_BODY | LBMDEV_MASK_GROUND)) continue;
        float q[19];
#pragma unroll
        for (int i = 1; i < 19; i++) q[i] = a.ovr_q[(long long)i * a.q_stride + k];
        bool any = false;
#pragma unroll
        for (int i = 1; i < 19; i++) any |= (q[i] >= 0.0f);
        if (!any) continue;
        int x, y, z;
        decompose(a.g, cell, x, y, z);
        const float4 v = a.vel[cell];
        const T ux = v.x, uy = v.y, uz = v.z, rho = v.w;
        const T c = T(1) - T(1.5) * (ux * ux + uy * uy + uz * uz);
        T F[3] = {0, 0, 0}, Pp[3] = {0, 0, 0};
#pragma unroll
        for (int i = 1; i < 19; i++) {
            if (q[i] >= 0.0f) {
                const int io = OPP[i];
                T fis = fstar<FMT, STATE>(a.pop, a.g, x, y, z, i);
                T fo = post_stream<FMT, STATE>(a.pop, a.g, a.periodic, x, y, z, io, q[i]);
                T ft = fis + fo;
                F[0] += ft * T(EX[i]);
                F[1] += ft * T(EY[i]);
                F[2] += ft * T(EZ[i]);
                T fp = feq_i<T>(i, rho, ux, uy, uz, c) + feq_i<T>(io, rho, ux, uy, uz, c);
                Pp[0] += fp * T(EX[i]);
                Pp[1] += fp * T(EY[i]);
                Pp[2] += fp * T(EZ[i]);
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            acc[d] += (double)F[d];
            acc[3 + d] += (double)Pp[d];
            fix[d] += fixed_round<T>(F[d]);
            fix[4 + d] += fixed_round<T>(Pp[d]);
        }
        fix[3] += 1;
    }
}

template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_force(ForceArgs<FMT> a) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    long long fix[7] = {0, 0, 0, 0, 0, 0, 0};
    force_block<FMT, STATE>(a, blockIdx.x, gridDim.x, acc, fix);
    block_reduce_store<6, 7>(acc, fix, a.pd, a.pi);
}

// Sums per-block partials in a fixed order (deterministic): lane l adds
// blocks l, l+64, ... then lane 0 adds the 64 lane sums in lane order.
template <int NV, int NI>
__device__ __forceinline__ void finalize_body(const double *pd, const long long *pi, int nblocks, double *outd,
                                              long long *outi) {
    __shared__ double sd[64][NV > 0 ? NV : 1];
    __shared__ long long si[64][NI > 0 ? NI : 1];
    const int l = threadIdx.x;
    // lane l adds partials l, l + 64, l + 128, ... in that order; 16 loads are
    // in flight before their adds (the order of the adds is unchanged)
    constexpr int U = 16;
    for (int k = 0; k < NV; k++) {
        double s = 0;
        int b = l;
        for (; b + 64 * (U - 1) < nblocks; b += 64 * U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = pd[(long long)(b + 64 * u) * NV + k];
#pragma unroll
            for (int u = 0; u < U; u++) s += v[u];
        }
        for (; b < nblocks; b += 64) s += pd[(long long)b * NV + k];
        sd[l][k] = s;
    }
    for (int k = 0; k < NI; k++) {
        long long s = 0;
        int b = l;
        for (; b + 64 * (U - 1) < nblocks; b += 64 * U) {
            long long v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = pi[(long long)(b + 64 * u) * NI + k];
#pragma unroll
            for (int u = 0; u < U; u++) s += v[u];
        }
        for (; b < nblocks; b += 64) s += pi[(long long)b * NI + k];
        si[l][k] = s;
    }
    __syncthreads();
    if (l != 0) return;
    for (int k = 0; k < NV; k++) {
        double s = 0;
        for (int j = 0; j < 64; j++) s += sd[j][k];
        outd[k] = s;
    }
    for (int k = 0; k < NI; k++) {
        long long s = 0;
        for (int j = 0; j < 64; j++) s += si[j][k];
        outi[k] = s;
    }
}

template <int NV, int NI>
__global__ __launch_bounds__(64) void k_finalize(const double *pd, const long long *pi, int nblocks,
                                                 double *outd, long long *outi) {
    finalize_body<NV, NI>(pd, pi, nblocks, outd, outi);
}

// ------------------------------------------------- export / diagnostics
template <int FMT>
struct ExportArgs {
    const typename Pop<FMT>::S *pop;
    const uint8_t *mask;
    const int32_t *ovr_index;
    const float *ovr_q;
    int n_ovr;
    int periodic;
    Geo g;
    void *out_stream;     // AoS [cell*19 + i] post-stream (may be null)
    void *out_collision;  // AoS post-collision (may be null)
    int as_double;
    double *pd;  // mass partials (may be null)
    long long *pi;
};

template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_export(ExportArgs<FMT> a) {
    typedef typename Pop<FMT>::T T;
    const int cell = blockIdx.x * BLOCK + threadIdx.x;
    double mass[1] = {0.0};
    long long dummy[1] = {0};
    if (cell < a.g.ncell) {
        int x, y, z;
        decompose(a.g, cell, x, y, z);
        const bool ovr = (a.mask[cell] & LBMDEV_MASK_OVR) != 0;
        const int k = ovr ? a.ovr_index[cell] : 0;
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const float qo = ovr ? a.ovr_q[(long long)OPP[i] * a.n_ovr + k] : -1.0f;
            if (a.out_stream || a.pd) {
                T v = post_stream<FMT, STATE>(a.pop, a.g, a.periodic, x, y, z, i, qo);
                mass[0] += (double)v;
                if (a.out_stream) {
                    if (a.as_double)
                        ((double *)a.out_stream)[(long long)cell * 19 + i] = (double)v;
                    else
                        ((float *)a.out_stream)[(long long)cell * 19 + i] = (float)v;
                }
            }
            if (a.out_collision) {
                T v = fstar<FMT, STATE>(a.pop, a.g, x, y, z, i);
                if (a.as_double)
                    ((double *)a.out_collision)[(long long)cell * 19 + i] = (double)v;
                else
                    ((float *)a.out_collision)[(long long)cell * 19 + i] = (float)v;
            }
        }
    }
    if (a.pd) block_reduce_store<1, 1>(mass, dummy, a.pd, a.pi);
}

template <int FMT>
__global__ __launch_bounds__(BLOCK) void k_init(typename Pop<FMT>::S *pop, Geo g, float4 *vel,
                                                const float *feq, float4 v4) {
    const int cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= g.ncell) return;
#pragma unroll
    for (int i = 0; i < 19; i++) pop[(long long)i * g.stride + cell] = Pop<FMT>::enc((typename Pop<FMT>::T)feq[i], i);
    vel[cell] = v4;
}

// Initial populations at equilibrium with a (ux, uy, uz, rho) field given on
// a coarse cell-centred lattice of cnx x cny x cnz cells, trilinearly
// interpolated: fine cell x sits at coarse coordinate (x + 1/2) cnx/nx - 1/2,
// clamped to the coarse cell centres at the edges.  Used by warm starts.
template <int FMT>
__global__ __launch_bounds__(BLOCK) void k_init_field(typename Pop<FMT>::S *pop, Geo g, float4 *vel,
                                                      const float4 *__restrict__ coarse, int cnx, int cny,
                                                      int cnz) {
    typedef typename Pop<FMT>::T T;
    const int cell = blockIdx.x * BLOCK + threadIdx.x;
    if (cell >= g.ncell) return;
    int x, y, z;
    decompose(g, cell, x, y, z);
    const float gx = fminf(fmaxf(((float)x + 0.5f) * (float)cnx / (float)g.nx - 0.5f, 0.0f), (float)(cnx - 1));
    const float gy = fminf(fmaxf(((float)y + 0.5f) * (float)cny / (float)g.ny - 0.5f, 0.0f), (float)(cny - 1));
    const float gz = fminf(fmaxf(((float)z + 0.5f) * (float)cnz / (float)g.nz - 0.5f, 0.0f), (float)(cnz - 1));
    const int x0 = (int)gx, y0 = (int)gy, z0 = (int)gz;
    const int x1 = min(x0 + 1, cnx - 1), y1 = min(y0 + 1, cny - 1), z1 = min(z0 + 1, cnz - 1);
    const float fx = gx - (float)x0, fy = gy - (float)y0, fz = gz - (float)z0;
    const long long sy = cnx, sz = (long long)cnx * cny;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float w = ((k & 1) ? fx : 1.0f - fx) * ((k & 2) ? fy : 1.0f - fy) * ((k & 4) ? fz : 1.0f - fz);
        const float4 c = coarse[((k & 1) ? x1 : x0) + ((k & 2) ? y1 : y0) * sy + ((k & 4) ? z1 : z0) * sz];
        v.x += w * c.x;
        v.y += w * c.y;
        v.z += w * c.z;
        v.w += w * c.w;
    }
    const T ux = (T)v.x, uy = (T)v.y, uz = (T)v.z, rho = (T)v.w;
    const T c = T(1) - T(1.5) * (ux * ux + uy * uy + uz * uz);
#pragma unroll
    for (int i = 0; i < 19; i++) pop[(long long)i * g.stride + cell] = Pop<FMT>::enc(feq_i<T>(i, rho, ux, uy, uz, c), i);
    vel[cell] = v;
}

// ------------------------------------------------------------ particles
struct Particle {
    float x, y, z, pad, vx, vy, vz, life;
};

__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// Trilinear sampler of particle_lbm.comp:46-90.
__device__ __forceinline__ float3 sample_trilinear(const float4 *vel, int nx, int ny, int nz, float px,
                                                   float py, float pz) {
    float gx = (px + 4.0f) / 8.0f * (float)nx;
    float gy = (py + 2.0f) / 4.0f * (float)ny;
    float gz = (pz + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.0f), (float)nx - 1.0f);
    gy = fminf(fmaxf(gy, 0.0f), (float)ny - 1.0f);
    gz = fminf(fmaxf(gz, 0.0f), (float)nz - 1.0f);
    const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
    const int x0 = (int)flx, y0 = (int)fly, z0 = (int)flz;
    const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
    const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
    const long long sy = nx, sz = (long long)nx * ny;
    const float4 v000 = vel[x0 + y0 * sy + z0 * sz], v100 = vel[x1 + y0 * sy + z0 * sz];
    const float4 v010 = vel[x0 + y1 * sy + z0 * sz], v110 = vel[x1 + y1 * sy + z0 * sz];
    const float4 v001 = vel[x0 + y0 * sy + z1 * sz], v101 = vel[x1 + y0 * sy + z1 * sz];
    const float4 v011 = vel[x0 + y1 * sy + z1 * sz], v111 = vel[x1 + y1 * sy + z1 * sz];
    float3 r;
    r.x = mixf(mixf(mixf(v000.x, v100.x, fx), mixf(v010.x, v110.x, fx), fy),
               mixf(mixf(v001.x, v101.x, fx), mixf(v011.x, v111.x, fx), fy), fz);
    r.y = mixf(mixf(mixf(v000.y, v100.y, fx), mixf(v010.y, v110.y, fx), fy),
               mixf(mixf(v001.y, v101.y, fx), mixf(v011.y, v111.y, fx), fy), fz);
    r.z = mixf(mixf(mixf(v000.z, v100.z, fx), mixf(v010.z, v110.z, fx), fy),
               mixf(mixf(v001.z, v101.z, fx), mixf(v011.z, v111.z, fx), fy), fz);
    return r;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

__device__ __forceinline__ void spawn_yz(uint32_t seed, uint32_t step, uint32_t idx, float &y, float &z) {
    const uint32_t base = mix32(step ^ mix32(seed + 0x9e3779b9u));
    y = -2.0f + 4.0f * u01(mix32(2u * idx + base));
    z = -2.0f + 4.0f * u01(mix32(2u * idx + 1u + base));
}

__global__ __launch_bounds__(BLOCK) void k_particles_seed(Particle *p, int n, uint32_t seed) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    float y, z;
    spawn_yz(seed, 0u, (uint32_t)i, y, z);
    Particle q;
    q.x = -4.0f;
    q.y = y;
    q.z = z;
    q.pad = 0.0f;
    q.vx = q.vy = q.vz = 0.0f;
    q.life = 1.0f;
    p[i] = q;
}

// Midpoint RK2 through the trilinear field; respawn on the inlet plane when a
// particle leaves |y|,|z| <= 2, x <= 4 (particle_lbm.comp:279-280).
__global__ __launch_bounds__(BLOCK) void k_particles_advect(Particle *p, int n, const float4 *vel, int nx,
                                                            int ny, int nz, float dt, float vs,
                                                            uint32_t seed, uint32_t step) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    Particle q = p[i];
    const float3 v1 = sample_trilinear(vel, nx, ny, nz, q.x, q.y, q.z);
    const float h = 0.5f * dt * vs;
    const float3 v2 = sample_trilinear(vel, nx, ny, nz, q.x + h * v1.x, q.y + h * v1.y, q.z + h * v1.z);
    const float s = dt * vs;
    float X = q.x + s * v2.x, Y = q.y + s * v2.y, Z = q.z + s * v2.z;
    q.vx = vs * v2.x;
    q.vy = vs * v2.y;
    q.vz = vs * v2.z;
    q.pad = sqrtf(q.vx * q.vx + q.vy * q.vy + q.vz * q.vz);
    if (X > 4.0f || fabsf(Y) > 2.0f || fabsf(Z) > 2.0f) {
        spawn_yz(seed, step, (uint32_t)i, Y, Z);
        X = -4.0f;
        q.life = 1.0f;
    }
    q.x = X;
    q.y = Y;
    q.z = Z;
    p[i] = q;
}

// ---------------------------------------------------- live particle update
// The particle update of the reference's running application
// (simulation/shaders/particle.comp:218-407, dispatched by main.c:1012-1101):
// nearest-cell velocity sampling, relaxation of the particle velocity towards
// mix(wind, 8 u_lattice, 0.5), a forward bias outside the model's box, four
// collision modes and the kill box.  The reference draws respawn positions
// from a time-seeded sin hash; here the three uniforms come from the counter
// hash of (seed, step, particle), so runs are reproducible.  Floating point
// is evaluated in the order written, without contraction, as the oracle's
// C restatement is.
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3add(V3 a, V3 b) {
#pragma clang fp contract(off)
    return V3{a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ V3 v3sub(V3 a, V3 b) {
#pragma clang fp contract(off)
    return V3{a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ V3 v3scale(V3 a, float s) {
#pragma clang fp contract(off)
    return V3{a.x * s, a.y * s, a.z * s};
}
__device__ __forceinline__ float v3dot(V3 a, V3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float v3len(V3 a) { return sqrtf(v3dot(a, a)); }
__device__ __forceinline__ V3 v3mix(V3 a, V3 b, float t) {
#pragma clang fp contract(off)
    return V3{a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t};
}
__device__ __forceinline__ float gsign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float v3comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ void v3set(V3 &a, int k, float v) {
    if (k == 0) a.x = v;
    else if (k == 1) a.y = v;
    else a.z = v;
}

// particle.comp:66-84 (and 86-103 for the density): the cell index of the
// nearest-cell sampler, grid position clamped to [0.5, N - 1.5].
__device__ __forceinline__ int live_cell_index(V3 p, int nx, int ny, int nz) {
#pragma clang fp contract(off)
    float gx = (p.x + 4.0f) / 8.0f * (float)nx;
    float gy = (p.y + 2.0f) / 4.0f * (float)ny;
    float gz = (p.z + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.5f), (float)nx - 1.5f);
    gy = fminf(fmaxf(gy, 0.5f), (float)ny - 1.5f);
    gz = fminf(fmaxf(gz, 0.5f), (float)nz - 1.5f);
    const int cx = min(max((int)floorf(gx), 0), nx - 1);
    const int cy = min(max((int)floorf(gy), 0), ny - 1);
    const int cz = min(max((int)floorf(gz), 0), nz - 1);
    const long long idx = cx + (long long)cy * nx + (long long)cz * nx * ny;
    const long long maxIdx = (long long)nx * ny * nz - 1;
    return (int)(idx < 0 ? 0 : (idx > maxIdx ? maxIdx : idx));
}

__device__ __forceinline__ bool inside_box(V3 p, V3 lo, V3 hi) {
    return p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z;
}

// particle.comp:112-125: the box face normal nearest to the particle.
__device__ __forceinline__ V3 box_normal(V3 p, V3 lo, V3 hi, V3 c) {
#pragma clang fp contract(off)
    const V3 t = v3sub(p, c);
    const V3 a = v3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
    const V3 h = v3scale(v3sub(hi, lo), 0.5f);
    V3 n = v3(0.0f, 0.0f, 0.0f);
    if (a.x / h.x > a.y / h.y && a.x / h.x > a.z / h.z) n.x = gsign(t.x);
    else if (a.y / h.y > a.z / h.z) n.y = gsign(t.y);
    else n.z = gsign(t.z);
    return n;
}

// particle.comp:127-145: out through the face of least penetration, 0.03 beyond.
__device__ __forceinline__ V3 box_push_out(V3 p, V3 lo, V3 hi) {
#pragma clang fp contract(off)
    const V3 c = v3scale(v3add(lo, hi), 0.5f);
    const V3 h = v3scale(v3sub(hi, lo), 0.5f);
    const V3 t = v3sub(p, c);
    const V3 pen = v3(h.x - fabsf(t.x), h.y - fabsf(t.y), h.z - fabsf(t.z));
    if (pen.x < pen.y && pen.x < pen.z) p.x = c.x + gsign(t.x) * (h.x + 0.03f);
    else if (pen.y < pen.z) p.y = c.y + gsign(t.y) * (h.y + 0.03f);
    else p.z = c.z + gsign(t.z) * (h.z + 0.03f);
    return p;
}

// Distance from p to triangle (v0, v1, v2), its closest point and unit normal:
// the region classification of the minimiser of |v0 + s e0 + t e1 - p|^2
// over the triangle (D. Eberly, "Distance Between Point and Triangle in 3D",
// Geometric Tools), as particle.comp:147-216 evaluates it.
__device__ __forceinline__ float point_triangle(V3 p, V3 v0, V3 v1, V3 v2, V3 &closest, V3 &normal) {
#pragma clang fp contract(off)
    const V3 e0 = v3sub(v1, v0), e1 = v3sub(v2, v0), d0 = v3sub(v0, p);
    const float a = v3dot(e0, e0), b = v3dot(e0, e1), c = v3dot(e1, e1);
    const float d = v3dot(e0, d0), e = v3dot(e1, d0);
    const float det = a * c - b * b;
    float s = b * e - c * d, t = b * d - a * e;
    if (s + t <= det) {
        if (s < 0.0f) {
            if (t < 0.0f) {
                if (d < 0.0f) {
                    t = 0.0f;
                    s = fminf(fmaxf(-d / a, 0.0f), 1.0f);
                } else {
                    s = 0.0f;
                    t = fminf(fmaxf(-e / c, 0.0f), 1.0f);
                }
            } else {
                s = 0.0f;
                t = fminf(fmaxf(-e / c, 0.0f), 1.0f);
            }
        } else if (t < 0.0f) {
            t = 0.0f;
            s = fminf(fmaxf(-d / a, 0.0f), 1.0f);
        } else {
            const float inv = 1.0f / det;
            s *= inv;
            t *= inv;
        }
    } else {
        if (s < 0.0f) {
            const float t0 = b + d, t1 = c + e;
            if (t1 > t0) {
                s = fminf(fmaxf((t1 - t0) / (a - 2.0f * b + c), 0.0f), 1.0f);
                t = 1.0f - s;
            } else {
                s = 0.0f;
                t = fminf(fmaxf(-e / c, 0.0f), 1.0f);
            }
        } else if (t < 0.0f) {
            const float t0 = b + e, t1 = a + d;
            if (t1 > t0) {
                t = fminf(fmaxf((t1 - t0) / (a - 2.0f * b + c), 0.0f), 1.0f);
                s = 1.0f - t;
            } else {
                t = 0.0f;
                s = fminf(fmaxf(-d / a, 0.0f), 1.0f);
            }
        } else {
            s = fminf(fmaxf(((c + e) - (b + d)) / (a - 2.0f * b + c), 0.0f), 1.0f);
            t = 1.0f - s;
        }
    }
    closest = v3add(v3add(v0, v3scale(e0, s)), v3scale(e1, t));
    const V3 cr = v3(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x);
    const float cl = v3len(cr);
    normal = v3(cr.x / cl, cr.y / cl, cr.z / cl);
    return v3len(v3sub(p, closest));
}

struct LiveArgs {
    V3 wind, carMin, carMax, carCenter;
    int mode, useLBM, vizMode;
    const float *tris;
    int ntri;
    const int32_t *cellStart, *cellCount, *triIdx;
    int nIdx;  // entries of triIdx
    V3 gridMin, gridCell;
    int gridRes;
};

__global__ __launch_bounds__(BLOCK) void k_particles_live(Particle *pp, int n, const float4 *vel,
                                                          const int32_t *solid, int nx, int ny, int nz, float dt,
                                                          LiveArgs f, uint32_t seed, uint32_t step) {
#pragma clang fp contract(off)
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= n) return;
    Particle q = pp[idx];
    V3 pos = v3(q.x, q.y, q.z), v = v3(q.vx, q.vy, q.vz);
    float pad = q.pad, life = q.life;
    V3 flow = f.wind;
    if (f.useLBM) {
        const float4 s = vel[live_cell_index(pos, nx, ny, nz)];
        flow = v3mix(f.wind, v3(s.x * 8.0f, s.y * 8.0f, s.z * 8.0f), 0.5f);
    }
    if (f.vizMode == 8 && f.useLBM) {
        const float rho = vel[live_cell_index(pos, nx, ny, nz)].w;
        const float cp = (rho - 1.0f) / (0.5f * 0.05f * 0.05f / (1.0f / 3.0f));
        pad = pad * (1.0f - 0.05f) + cp * 0.05f;
    } else {
        pad = pad * (1.0f - 0.02f) + v3len(flow) * 0.02f;
    }
    v = v3mix(v, flow, 0.05f);
    const V3 pad5 = v3(0.05f, 0.05f, 0.05f);
    const bool near = f.mode > 0 && inside_box(pos, v3sub(f.carMin, pad5), v3add(f.carMax, pad5));
    if (!near) v.x = fmaxf(v.x, f.wind.x * 0.3f);
    pos = v3add(pos, v3scale(v, dt));

    if (f.mode == 1) {
        if (inside_box(pos, f.carMin, f.carMax)) {
            const V3 nrm = box_normal(pos, f.carMin, f.carMax, f.carCenter);
            pos = box_push_out(pos, f.carMin, f.carMax);
            const float vdn = v3dot(v, nrm);
            if (vdn < 0.0f) {
                v = v3sub(v, v3scale(nrm, 1.05f * vdn));
                V3 away = v3sub(pos, f.carCenter);
                away.x = 0.0f;
                const float al = v3len(away);
                if (al > 0.001f) v = v3add(v, v3scale(v3(away.x / al, away.y / al, away.z / al), fabsf(vdn) * 0.4f));
            }
        }
    } else if (f.mode == 2 && f.ntri > 0 && f.gridRes > 0) {
        const float radius = 0.03f;
        const V3 rel = v3(v3sub(pos, f.gridMin).x / f.gridCell.x, v3sub(pos, f.gridMin).y / f.gridCell.y,
                          v3sub(pos, f.gridMin).z / f.gridCell.z);
        const int cx = (int)floorf(rel.x), cy = (int)floorf(rel.y), cz = (int)floorf(rel.z);
        const int R = f.gridRes;
        if (cx >= -1 && cx <= R && cy >= -1 && cy <= R && cz >= -1 && cz <= R) {
            const int lx = max(cx - 1, 0), ly = max(cy - 1, 0), lz = max(cz - 1, 0);
            const int hx = min(cx + 1, R - 1), hy = min(cy + 1, R - 1), hz = min(cz + 1, R - 1);
            bool hit = false;
            for (int z = lz; z <= hz && !hit; z++)
                for (int y = ly; y <= hy && !hit; y++)
                    for (int x = lx; x <= hx && !hit; x++) {
                        const int ci = x + y * R + z * R * R;
                        const int start = f.cellStart[ci], cnt = f.cellCount[ci];
                        // The cell's triangles in rounds of TB: the indices and
                        // vertices of a round are fetched together, then
                        // tested in list order (the first hit wins, as in the
                        // one-at-a-time loop), so a round waits for memory once.
                        constexpr int TB = 4;
                        for (int j0 = 0; j0 < cnt && !hit; j0 += TB) {
                            int tix[TB];
#pragma unroll
                            for (int k = 0; k < TB; k++) {
                                // entries outside the arrays (an inconsistent grid) are skipped
                                const int e = start + j0 + k;
                                tix[k] = j0 + k < cnt && e >= 0 && e < f.nIdx ? f.triIdx[e] : -1;
                                if (tix[k] >= f.ntri) tix[k] = -1;
                            }
                            float4 tv[TB][3];
#pragma unroll
                            for (int k = 0; k < TB; k++) {
                                const float4 *t = (const float4 *)(f.tris + 12 * (long long)(tix[k] < 0 ? 0 : tix[k]));
                                tv[k][0] = t[0];
                                tv[k][1] = t[1];
                                tv[k][2] = t[2];
                            }
#pragma unroll
                            for (int k = 0; k < TB; k++) {
                                if (hit || tix[k] < 0) continue;
                                // A triangle whose bounding box lies more than
                                // radius + 1e-3 from the particle on some axis
                                // cannot give dist < radius (rounding is far
                                // below the margin), so skipping it changes
                                // nothing; it saves the full distance test for
                                // nearly every candidate.
                                const float m = radius + 1e-3f;
                                if (pos.x + m < fminf(fminf(tv[k][0].x, tv[k][1].x), tv[k][2].x) ||
                                    pos.x - m > fmaxf(fmaxf(tv[k][0].x, tv[k][1].x), tv[k][2].x) ||
                                    pos.y + m < fminf(fminf(tv[k][0].y, tv[k][1].y), tv[k][2].y) ||
                                    pos.y - m > fmaxf(fmaxf(tv[k][0].y, tv[k][1].y), tv[k][2].y) ||
                                    pos.z + m < fminf(fminf(tv[k][0].z, tv[k][1].z), tv[k][2].z) ||
                                    pos.z - m > fmaxf(fmaxf(tv[k][0].z, tv[k][1].z), tv[k][2].z))
                                    continue;
                                V3 cp, nrm;
                                const float dist = point_triangle(pos, v3(tv[k][0].x, tv[k][0].y, tv[k][0].z),
                                                                  v3(tv[k][1].x, tv[k][1].y, tv[k][1].z),
                                                                  v3(tv[k][2].x, tv[k][2].y, tv[k][2].z), cp, nrm);
                                if (dist < radius) {
                                    pos = v3add(cp, v3scale(nrm, radius + 0.001f));
                                    const float vdn = v3dot(v, nrm);
                                    if (vdn < 0.0f) v = v3sub(v, v3scale(nrm, 1.05f * vdn));
                                    hit = true;
                                }
                            }
                        }
                    }
        }
    } else if (f.mode == 3) {
        const float gx = (pos.x + 4.0f) / 8.0f * (float)nx, gy = (pos.y + 2.0f) / 4.0f * (float)ny,
                    gz = (pos.z + 2.0f) / 4.0f * (float)nz;
        const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
        if (ix >= 0 && ix < nx && iy >= 0 && iy < ny && iz >= 0 && iz < nz &&
            solid[ix + (long long)iy * nx + (long long)iz * nx * ny] == 1) {
            const V3 cell = v3(8.0f / (float)nx, 4.0f / (float)ny, 4.0f / (float)nz);
            const V3 ctr = v3(((float)ix + 0.5f) / (float)nx * 8.0f - 4.0f, ((float)iy + 0.5f) / (float)ny * 4.0f - 2.0f,
                              ((float)iz + 0.5f) / (float)nz * 4.0f - 2.0f);
            const V3 toP = v3sub(pos, ctr);
            const V3 ab = v3(fabsf(toP.x) / (cell.x * 0.5f), fabsf(toP.y) / (cell.y * 0.5f),
                             fabsf(toP.z) / (cell.z * 0.5f));
            // axes by penetration depth: the largest |offset| (the shallowest) first
            int ax[3] = {0, 1, 2};
            if (v3comp(ab, ax[0]) < v3comp(ab, ax[1])) { const int t = ax[0]; ax[0] = ax[1]; ax[1] = t; }
            if (v3comp(ab, ax[1]) < v3comp(ab, ax[2])) { const int t = ax[1]; ax[1] = ax[2]; ax[2] = t; }
            if (v3comp(ab, ax[0]) < v3comp(ab, ax[1])) { const int t = ax[0]; ax[0] = ax[1]; ax[1] = t; }
            bool pushed = false;
            const int dims[3] = {nx, ny, nz};
            for (int k = 0; k < 3 && !pushed; k++) {
                const int a = ax[k];
                V3 nrm = v3(0.0f, 0.0f, 0.0f);
                v3set(nrm, a, gsign(v3comp(toP, a)));
                const int ni[3] = {ix + (int)nrm.x, iy + (int)nrm.y, iz + (int)nrm.z};
                if (ni[a] >= 0 && ni[a] < dims[a] &&
                    solid[ni[0] + (long long)ni[1] * nx + (long long)ni[2] * nx * ny] == 1)
                    continue;  // the neighbour on that side is solid too: try the next axis
                pos = v3add(ctr, v3(nrm.x * cell.x * 0.52f, nrm.y * cell.y * 0.52f, nrm.z * cell.z * 0.52f));
                const float vdn = v3dot(v, nrm);
                if (vdn < 0.0f) v = v3sub(v, v3scale(nrm, 1.05f * vdn));
                pushed = true;
            }
            if (!pushed) life = -1.0f;
        }
    }
    // kill box (particle.comp:369-405)
    const float killX = f.carMax.x + 2.0f, tunnel = 1.3f;
    pos.y = fminf(fmaxf(pos.y, -tunnel), tunnel);
    pos.z = fminf(fmaxf(pos.z, -tunnel), tunnel);
    const float speed = v3len(v);
    const bool stuck = speed < 0.05f && pos.x > f.carMin.x - 0.5f && pos.x < f.carMax.x + 0.5f;
    if (pos.x > killX || stuck || life < 0.0f) {
        const uint32_t base = mix32(step ^ mix32(seed + 0x5bd1e995u));
        const float ry = u01(mix32(3u * (uint32_t)idx + base)) - 0.5f;
        const float rz = u01(mix32(3u * (uint32_t)idx + 1u + base)) - 0.5f;
        const float rv = u01(mix32(3u * (uint32_t)idx + 2u + base));
        pos = v3(-4.0f, ry * 2.0f * tunnel, rz * 2.0f * tunnel);
        v = v3(0.2f + rv * 0.15f, 0.0f, 0.0f);
        life = 1.0f;
    }
    q.x = pos.x;
    q.y = pos.y;
    q.z = pos.z;
    q.pad = pad;
    q.vx = v.x;
    q.vy = v.y;
    q.vz = v.z;
    q.life = life;
    pp[idx] = q;
}

}  // namespace

// ======================================================================
// Device layer (C ABI)
// ======================================================================

static char g_err[512];

static int set_err(const char *what, hipError_t e) {
    snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
    fprintf(stderr, "lbm: %s\n", g_err);
    return 0;
}

#define HIPCHECK(call, what)                      \
    do {                                          \
        hipError_t _e = (call);                   \
        if (_e != hipSuccess) return set_err(what, _e); \
    } while (0)

struct LbmDev {
    int nx, ny, nz, ncell;
    int fmt;
    long long stride;
    size_t elem;  // bytes per stored population
    size_t telem; // bytes per compute value
    void *pop;
    uint8_t *mask;
    int32_t *solid;
    int32_t *ovr_index;
    float4 *vel;
    int n_ovr, ovr_cap;
    int32_t *ovr_cells;
    float *ovr_q;
    void *ovr_vals;
    uint32_t *ovr_bits;  // per listed cell: directions taken from ovr_vals
    PrepassItem *ovr_items;  // prepass work list (see BoundArgs)
    int item_cap;
    int seg[20];
    int state;  // -1 uninitialised, ST_INIT / ST_A / ST_B
    hipStream_t stream;
    double *pd;
    long long *pi;
    double *resd;
    long long *resi;
    unsigned long long *guard;  // stability-guard resets since the flow was initialised
    float *feq_d;
    void *scratch;  // this device's address of the static staging array `scratch`
    int nparts;  // partial slots allocated
    int bs;      // workgroup size of the step kernel
    int nt;      // LBMDEV_NT_A / LBMDEV_NT_B: non-temporal population access per phase
    int remap;   // XCD-contiguous workgroup order in the step kernel
    // Timed steps not yet collected (lbmdev_step_timed): three events per
    // step (before the prepass, before and after the step kernel), reused
    // across collections; tpre[s] = 1 when step s had a prepass.
    hipEvent_t *tev;
    unsigned char *tpre;
    int tcap, tused;  // steps the pool holds / steps recorded
};

extern "C" const char *lbmdev_last_error(void) { return g_err; }

// ------------------------------------------------------------ streamlines
// The streamline tracer of the application's visualisation mode 9
// (simulation/shaders/streamline_trace.comp:19-107, dispatched by
// main.c:1120-1140): one lane per seed.  Seeds on a gridSide x gridSide
// lattice of the plane x = -2.5, classic RK4 with ds = 8 / traceLen through
// the trilinear sampler clamped to [0.5, N - 1.5], velocities scaled by 8;
// vertex i = (position, 8 |u|) before step i; a step shorter than 1e-6 ends
// the line (the rest repeat the position with speed 0); positions clamped to
// the box.  Every float expression is evaluated in the shader's order
// without contraction, as the oracle's C restatement is.  The clamp keeps
// every sample index inside the grid for any position, NaN included
// (fmaxf/fminf return the non-NaN operand), as long as each dimension is at
// least 2, which the host checks.
__device__ __forceinline__ V3 sample_streamline(const float4 *__restrict__ vel, int nx, int ny, int nz, V3 p) {
#pragma clang fp contract(off)
    float gx = (p.x + 4.0f) / 8.0f * (float)nx;
    float gy = (p.y + 2.0f) / 4.0f * (float)ny;
    float gz = (p.z + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.5f), (float)nx - 1.5f);
    gy = fminf(fmaxf(gy, 0.5f), (float)ny - 1.5f);
    gz = fminf(fmaxf(gz, 0.5f), (float)nz - 1.5f);
    const float flx = floorf(gx), fly = floorf(gy), flz = floorf(gz);
    const int x0 = (int)flx, y0 = (int)fly, z0 = (int)flz;
    const int x1 = min(x0 + 1, nx - 1), y1 = min(y0 + 1, ny - 1), z1 = min(z0 + 1, nz - 1);
    const float fx = gx - flx, fy = gy - fly, fz = gz - flz;
    const long long sy = nx, sz = (long long)nx * ny;
    const float4 v000 = vel[x0 + y0 * sy + z0 * sz], v100 = vel[x1 + y0 * sy + z0 * sz];
    const float4 v010 = vel[x0 + y1 * sy + z0 * sz], v110 = vel[x1 + y1 * sy + z0 * sz];
    const float4 v001 = vel[x0 + y0 * sy + z1 * sz], v101 = vel[x1 + y0 * sy + z1 * sz];
    const float4 v011 = vel[x0 + y1 * sy + z1 * sz], v111 = vel[x1 + y1 * sy + z1 * sz];
    const V3 v00 = v3mix(v3(v000.x, v000.y, v000.z), v3(v100.x, v100.y, v100.z), fx);
    const V3 v10 = v3mix(v3(v010.x, v010.y, v010.z), v3(v110.x, v110.y, v110.z), fx);
    const V3 v01 = v3mix(v3(v001.x, v001.y, v001.z), v3(v101.x, v101.y, v101.z), fx);
    const V3 v11 = v3mix(v3(v011.x, v011.y, v011.z), v3(v111.x, v111.y, v111.z), fx);
    return v3mix(v3mix(v00, v10, fy), v3mix(v01, v11, fy), fz);
}

__global__ __launch_bounds__(BLOCK) void k_streamlines(const float4 *__restrict__ vel, int nx, int ny, int nz,
                                                       int nseeds, int len, float4 *__restrict__ out) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= nseeds) return;
    const int side = (int)ceilf(sqrtf((float)nseeds));
    const int iy = s / side, iz = s % side;
    V3 pos = v3(-2.5f, -1.5f + 3.0f * ((float)iy + 0.5f) / (float)side, -1.5f + 3.0f * ((float)iz + 0.5f) / (float)side);
    const float ds = 8.0f / (float)len, h = 0.5f * ds, c = ds / 6.0f;
    float4 *o = out + (long long)s * len;
    for (int i = 0; i < len; i++) {
        const V3 v = sample_streamline(vel, nx, ny, nz, pos);
        o[i] = make_float4(pos.x, pos.y, pos.z, v3len(v) * 8.0f);
        const V3 k1 = v3scale(v, 8.0f);
        const V3 k2 = v3scale(sample_streamline(vel, nx, ny, nz, v3add(pos, v3scale(k1, h))), 8.0f);
        const V3 k3 = v3scale(sample_streamline(vel, nx, ny, nz, v3add(pos, v3scale(k2, h))), 8.0f);
        const V3 k4 = v3scale(sample_streamline(vel, nx, ny, nz, v3add(pos, v3scale(k3, ds))), 8.0f);
        const V3 st = v3scale(v3add(v3add(v3add(k1, v3scale(k2, 2.0f)), v3scale(k3, 2.0f)), k4), c);
        if (v3len(st) < 1e-6f) {
            for (int j = i + 1; j < len; j++) o[j] = make_float4(pos.x, pos.y, pos.z, 0.0f);
            return;
        }
        pos = v3add(pos, st);
        pos = v3(fminf(fmaxf(pos.x, -4.0f), 4.0f), fminf(fmaxf(pos.y, -2.0f), 2.0f), fminf(fmaxf(pos.z, -2.0f), 2.0f));
    }
}

extern "C" int lbmdev_trace_streamlines(LbmDev *d, void *dev_out, int nseeds, int len) {
    if (d->nx < 2 || d->ny < 2 || d->nz < 2) {
        snprintf(g_err, sizeof(g_err), "LBM_TraceStreamlines: every grid dimension must be at least 2");
        return 0;
    }
    if (nseeds < 1 || len < 1 || !dev_out) {
        snprintf(g_err, sizeof(g_err), "LBM_TraceStreamlines: needs numSeeds >= 1, traceLen >= 1 and an output");
        return 0;
    }
    hipLaunchKernelGGL(k_streamlines, dim3((nseeds + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, d->stream,
                       (const float4 *)d->vel, d->nx, d->ny, d->nz, nseeds, len, (float4 *)dev_out);
    HIPCHECK(hipGetLastError(), "streamline launch");
    return 1;
}

// ------------------------------------------------ staging without allocation
// A file-scope device array: the staging buffer of lbmdev_init_field_noalloc,
// which calls no allocator.  Each GPU the library runs on holds its own copy
// for the life of the process.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU is taken once
// with hipGetSymbolAddress (with that GPU current) and only that address is
// used, as kernel argument and copy target.  Every call on one GPU shares the
// array, so such calls must not overlap on a GPU (they race).
static __device__ __align__(256) unsigned char scratch[LBMDEV_NOALLOC_FIELD_BYTES];

constexpr int kMaxDevices = 64;
static void *g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mutex;

// The current device's address of `scratch`, resolved once per device.
static hipError_t scratch_address(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = NULL;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return hipSuccess;
}

static size_t elem_size(int fmt) {
    switch (fmt) {
    case LBMDEV_FP32: return 4;
    case LBMDEV_FP64: return 8;
    default: return 2;
    }
}

extern "C" LbmDev *lbmdev_create(int nx, int ny, int nz, int storage, char *err, int errlen) {
    LbmDev *d = (LbmDev *)calloc(1, sizeof(LbmDev));
    if (!d) return NULL;
    d->nx = nx;
    d->ny = ny;
    d->nz = nz;
    long long nc = (long long)nx * ny * nz;
    if (nx < 2 || ny < 1 || nz < 1 || nc > 0x7fffffffLL - 256) {
        snprintf(err, errlen, "grid %dx%dx%d is outside the supported range", nx, ny, nz);
        free(d);
        return NULL;
    }
    // 32- and 16-bit slots are addressed by 32-bit byte offsets within a plane
    if (storage != LBMDEV_FP64 && (unsigned long long)nc * elem_size(storage) > (1ULL << 32)) {
        snprintf(err, errlen,
                 "grid %dx%dx%d has %lld cells; FP32 populations address at most 2^30 cells per grid "
                 "(32-bit slot offsets): use FP16/BF16 or FP64 storage",
                 nx, ny, nz, nc);
        free(d);
        return NULL;
    }
    d->ncell = (int)nc;
    d->fmt = storage;
    d->elem = elem_size(storage);
    d->telem = storage == LBMDEV_FP64 ? 8 : 4;
    d->stride = (nc + 63) / 64 * 64;
    d->state = -1;
    d->bs = 256;
    d->remap = 1;
    hipError_t e;
    const size_t popBytes = (size_t)19 * d->stride * d->elem;
    d->nparts = (int)((nc + BLOCK - 1) / BLOCK);
    if (d->nparts < FORCE_BLOCKS) d->nparts = FORCE_BLOCKS;
#define ALLOC(ptr, bytes, name)                                                        \
    if ((e = hipMalloc((void **)&(ptr), (bytes))) != hipSuccess) {                     \
        snprintf(err, errlen, "device allocation of %s (%.1f MB) failed: %s", name,  \
                 (double)(bytes) / 1048576.0, hipGetErrorString(e));                   \
        lbmdev_destroy(d);                                                             \
        return NULL;                                                                   \
    }
    ALLOC(d->pop, popBytes, "populations");
    ALLOC(d->mask, (size_t)nc, "mask");
    ALLOC(d->solid, (size_t)nc * 4, "solid");
    ALLOC(d->ovr_index, (size_t)nc * 4, "boundary index");
    ALLOC(d->vel, (size_t)nc * 16, "velocity");
    ALLOC(d->pd, (size_t)d->nparts * 8 * 8, "reduction partials");
    ALLOC(d->pi, (size_t)d->nparts * 8 * 8, "reduction partials");
    ALLOC(d->resd, 16 * 8, "reduction result");
    ALLOC(d->resi, 16 * 8, "reduction result");
    ALLOC(d->feq_d, 19 * 4, "feq constants");
    ALLOC(d->guard, 8, "guard counter");
#undef ALLOC
    if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) {
        snprintf(err, errlen, "stream creation failed: %s", hipGetErrorString(e));
        lbmdev_destroy(d);
        return NULL;
    }
    if ((e = scratch_address(&d->scratch)) != hipSuccess) {
        snprintf(err, errlen, "address of the staging array: %s", hipGetErrorString(e));
        lbmdev_destroy(d);
        return NULL;
    }
    hipMemsetAsync(d->mask, 0, (size_t)nc, d->stream);
    hipMemsetAsync(d->solid, 0, (size_t)nc * 4, d->stream);
    hipMemsetAsync(d->vel, 0, (size_t)nc * 16, d->stream);
    hipMemsetAsync(d->pop, 0, popBytes, d->stream);
    if ((e = hipStreamSynchronize(d->stream)) != hipSuccess) {
        snprintf(err, errlen, "device initialisation failed: %s", hipGetErrorString(e));
        lbmdev_destroy(d);
        return NULL;
    }
    return d;
}

extern "C" void lbmdev_destroy(LbmDev *d) {
    if (!d) return;
    if (d->stream) hipStreamSynchronize(d->stream);
    hipFree(d->pop);
    hipFree(d->mask);
    hipFree(d->solid);
    hipFree(d->ovr_index);
    hipFree(d->vel);
    hipFree(d->ovr_cells);
    hipFree(d->ovr_q);
    hipFree(d->ovr_vals);
    hipFree(d->ovr_bits);
    hipFree(d->ovr_items);
    hipFree(d->pd);
    hipFree(d->pi);
    hipFree(d->resd);
    hipFree(d->resi);
    hipFree(d->feq_d);
    hipFree(d->guard);
    for (int k = 0; k < 3 * d->tcap; k++) hipEventDestroy(d->tev[k]);
    free(d->tev);
    free(d->tpre);
    if (d->stream) hipStreamDestroy(d->stream);
    free(d);
}

extern "C" int lbmdev_storage(const LbmDev *d) { return d->fmt; }

extern "C" int lbmdev_set_geometry(LbmDev *d, const int32_t *solid, int n_ovr, const int32_t *ovr_cells,
                                   const float *ovr_q, int periodic) {
    const int nc = d->ncell;
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    uint8_t *mask = (uint8_t *)malloc((size_t)nc);
    int32_t *index = (int32_t *)malloc((size_t)nc * 4);
    uint32_t *bits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n_ovr > 0 ? n_ovr : 1));
    PrepassItem *items = NULL;
    int seg[20];
    if (mask && index && bits)
        for (int k = 0; k < n_ovr; k++) bits[k] = special_dirs(d->nx, d->ny, d->nz, periodic, ovr_cells[k], ovr_q, n_ovr, k);
    const int built = mask && index && bits ? build_items(n_ovr, bits, ovr_cells, ovr_q, n_ovr, NULL, &items, seg) : 0;
    if (built <= 0) {
        free(mask);
        free(index);
        free(bits);
        snprintf(g_err, sizeof(g_err), built < 0 ? "boundary work list exceeds 2^28 (cell, direction) pairs"
                                                  : "host allocation failed");
        return 0;
    }
    for (int c = 0; c < nc; c++) {
        mask[c] = (solid[c] == 1 ? LBMDEV_MASK_BODY : 0) | (solid[c] == 2 ? LBMDEV_MASK_GROUND : 0);
        index[c] = -1;
    }
    for (int k = 0; k < n_ovr; k++) {
        mask[ovr_cells[k]] |= LBMDEV_MASK_OVR;
        index[ovr_cells[k]] = k;
    }
    if (n_ovr > d->ovr_cap) {
        hipFree(d->ovr_cells);
        hipFree(d->ovr_q);
        hipFree(d->ovr_vals);
        d->ovr_cells = NULL;
        d->ovr_q = NULL;
        d->ovr_vals = NULL;
        d->ovr_cap = 0;
        int cap = n_ovr + 1024;
        if (hipMalloc((void **)&d->ovr_cells, (size_t)cap * 4) != hipSuccess ||
            hipMalloc((void **)&d->ovr_q, (size_t)cap * 19 * 4) != hipSuccess ||
            hipMalloc((void **)&d->ovr_vals, (size_t)cap * 19 * d->telem) != hipSuccess) {
            free(mask);
            free(index);
            free(bits);
            free(items);
            snprintf(g_err, sizeof(g_err), "device allocation of the boundary list failed");
            return 0;
        }
        d->ovr_cap = cap;
        hipFree(d->ovr_bits);
        d->ovr_bits = NULL;
        if (hipMalloc((void **)&d->ovr_bits, (size_t)cap * 4) != hipSuccess) {
            free(mask);
            free(index);
            free(bits);
            free(items);
            d->ovr_cap = 0;
            snprintf(g_err, sizeof(g_err), "device allocation of the boundary list failed");
            return 0;
        }
    }
    if (seg[19] > d->item_cap) {
        hipFree(d->ovr_items);
        d->ovr_items = NULL;
        d->item_cap = 0;
        const int cap = seg[19] + 19 * 64;
        if (hipMalloc((void **)&d->ovr_items, (size_t)cap * sizeof(PrepassItem)) != hipSuccess) {
            free(mask);
            free(index);
            free(bits);
            free(items);
            snprintf(g_err, sizeof(g_err), "device allocation of the boundary work list failed");
            return 0;
        }
        d->item_cap = cap;
    }
    memcpy(d->seg, seg, sizeof(seg));
    d->n_ovr = n_ovr;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(d->mask, mask, (size_t)nc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->ovr_index, index, (size_t)nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->solid, solid, (size_t)nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_ovr > 0) e = hipMemcpy(d->ovr_cells, ovr_cells, (size_t)n_ovr * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_ovr > 0)
        e = hipMemcpy(d->ovr_q, ovr_q, (size_t)n_ovr * 19 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_ovr > 0) e = hipMemcpy(d->ovr_bits, bits, (size_t)n_ovr * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && seg[19] > 0)
        e = hipMemcpy(d->ovr_items, items, (size_t)seg[19] * sizeof(PrepassItem), hipMemcpyHostToDevice);
    free(mask);
    free(index);
    free(bits);
    free(items);
    if (e != hipSuccess) return set_err("geometry upload", e);
    return 1;
}

template <int FMT>
static int init_t(LbmDev *d, const float4 &v4) {
    const int nb = (d->ncell + BLOCK - 1) / BLOCK;
    Geo g{d->nx, d->ny, d->nz, d->ncell, d->stride};
    hipLaunchKernelGGL(k_init<FMT>, dim3(nb), dim3(BLOCK), 0, d->stream, (typename Pop<FMT>::S *)d->pop, g,
                       d->vel, d->feq_d, v4);
    return 1;
}

extern "C" int lbmdev_init_flow(LbmDev *d, const float feq[19], const float vel4[4]) {
    HIPCHECK(hipMemcpyAsync(d->feq_d, feq, 19 * 4, hipMemcpyHostToDevice, d->stream), "feq upload");
    HIPCHECK(hipMemsetAsync(d->guard, 0, 8, d->stream), "guard counter");
    float4 v4 = make_float4(vel4[0], vel4[1], vel4[2], vel4[3]);
    switch (d->fmt) {
    case LBMDEV_FP32: init_t<LBMDEV_FP32>(d, v4); break;
    case LBMDEV_FP16S: init_t<LBMDEV_FP16S>(d, v4); break;
    case LBMDEV_BF16S: init_t<LBMDEV_BF16S>(d, v4); break;
    case LBMDEV_FP64: init_t<LBMDEV_FP64>(d, v4); break;
    }
    HIPCHECK(hipGetLastError(), "init launch");
    HIPCHECK(hipStreamSynchronize(d->stream), "init");
    d->state = ST_INIT;
    return 1;
}

template <int FMT>
static void init_field_t(typename Pop<FMT>::S *pop, const Geo &g, float4 *vel, const float *coarse, int cnx, int cny,
                         int cnz, hipStream_t s) {
    const int nb = (int)((g.ncell + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_init_field<FMT>, dim3(nb), dim3(BLOCK), 0, s, pop, g, vel, (const float4 *)coarse, cnx,
                       cny, cnz);
}

static void init_field_dispatch(int fmt, void *pop, const Geo &g, float4 *vel, const float *coarse, int cnx, int cny,
                                int cnz, hipStream_t s) {
    switch (fmt) {
    case LBMDEV_FP32: init_field_t<LBMDEV_FP32>((float *)pop, g, vel, coarse, cnx, cny, cnz, s); break;
    case LBMDEV_FP16S: init_field_t<LBMDEV_FP16S>((uint16_t *)pop, g, vel, coarse, cnx, cny, cnz, s); break;
    case LBMDEV_BF16S: init_field_t<LBMDEV_BF16S>((uint16_t *)pop, g, vel, coarse, cnx, cny, cnz, s); break;
    case LBMDEV_FP64: init_field_t<LBMDEV_FP64>((double *)pop, g, vel, coarse, cnx, cny, cnz, s); break;
    }
}

// Equilibrium initial state from a coarse (ux, uy, uz, rho) field on the host
// (cnz x cny x cnx x 4 floats), upsampled on the device.
// As lbmdev_init_field, staging the coarse field in the static array
// `scratch` (through the address resolved when the grid was created) instead
// of an allocation.  Rejects fields larger than the array.
extern "C" int lbmdev_init_field_noalloc(LbmDev *d, const float *coarse_host, int cnx, int cny, int cnz) {
    const size_t bytes = sizeof(float) * 4 * (size_t)cnx * cny * cnz;
    if (bytes > (size_t)LBMDEV_NOALLOC_FIELD_BYTES) {
        snprintf(g_err, sizeof(g_err),
                 "coarse field of %dx%dx%d cells (%zu bytes) exceeds the %d-byte staging array (%d cells)", cnx, cny,
                 cnz, bytes, LBMDEV_NOALLOC_FIELD_BYTES, LBMDEV_NOALLOC_FIELD_BYTES / 16);
        return 0;
    }
    if (!d->scratch) {
        snprintf(g_err, sizeof(g_err), "the staging array has no device address");
        return 0;
    }
    HIPCHECK(hipMemcpyAsync(d->scratch, coarse_host, bytes, hipMemcpyHostToDevice, d->stream), "field staging");
    HIPCHECK(hipMemsetAsync(d->guard, 0, 8, d->stream), "guard counter");
    init_field_dispatch(d->fmt, d->pop, make_geo(d->nx, d->ny, d->nz, d->stride), d->vel, (const float *)d->scratch,
                        cnx, cny, cnz, d->stream);
    HIPCHECK(hipGetLastError(), "field init");
    HIPCHECK(hipStreamSynchronize(d->stream), "field init");
    d->state = ST_INIT;
    return 1;
}

extern "C" int lbmdev_init_field(LbmDev *d, const float *coarse_host, int cnx, int cny, int cnz) {
    const size_t bytes = sizeof(float) * 4 * (size_t)cnx * cny * cnz;
    float *c = NULL;
    HIPCHECK(hipMalloc((void **)&c, bytes), "coarse field");
    hipError_t e = hipMemcpy(c, coarse_host, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(d->guard, 0, 8, d->stream);
    if (e == hipSuccess) {
        init_field_dispatch(d->fmt, d->pop, make_geo(d->nx, d->ny, d->nz, d->stride), d->vel, c, cnx, cny, cnz,
                            d->stream);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(d->stream);
    hipFree(c);
    if (e != hipSuccess) return set_err("field init", e);
    if (e2 != hipSuccess) return set_err("field init", e2);
    d->state = ST_INIT;
    return 1;
}

template <int FMT>
static void launch_prepass(LbmDev *d, int periodic) {
    if (d->n_ovr == 0) return;
    BoundArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)d->pop;
    a.ovr_vals = (typename Pop<FMT>::T *)d->ovr_vals;
    a.n_ovr = d->n_ovr;
    a.periodic = periodic;
    a.g = make_geo(d->nx, d->ny, d->nz, d->stride);
    a.items = d->ovr_items;
    memcpy(a.seg, d->seg, sizeof(a.seg));
    if (a.seg[19] == 0) return;
    const int nb = (a.seg[19] + BLOCK - 1) / BLOCK;
    if (d->state == ST_A)
        hipLaunchKernelGGL((k_prepass<FMT, ST_A>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
    else
        hipLaunchKernelGGL((k_prepass<FMT, ST_B>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
}

template <int FMT, int OP, int BS, int NT>
static void launch_step_phase(int phase, int ncell, hipStream_t st, const StepArgs<FMT> &a) {
    const int nb = (ncell + BS - 1) / BS;
    // whole rows per workgroup, rows of whole wavefronts, no partial workgroup
    const bool row = FMT != LBMDEV_FP64 && a.g.nx % 64 == 0 && BS % a.g.nx == 0 && ncell % BS == 0;
    if (phase == 0)
        hipLaunchKernelGGL((k_step<FMT, OP, 0, BS, 0, NT>), dim3(nb), dim3(BS), 0, st, a);
    else if (row)
        hipLaunchKernelGGL((k_step<FMT, OP, 1, BS, 1, NT>), dim3(nb), dim3(BS), 0, st, a);
    else
        hipLaunchKernelGGL((k_step<FMT, OP, 1, BS, 0, NT>), dim3(nb), dim3(BS), 0, st, a);
}

// Workgroup sizes the autotuner chooses from.
static const int kBlockSizes[3] = {128, 256, 512};

template <int FMT, int OP, int NT>
static void launch_step_bs(int bs, int phase, int ncell, hipStream_t st, const StepArgs<FMT> &a) {
    if (bs == 128)
        launch_step_phase<FMT, OP, 128, NT>(phase, ncell, st, a);
    else if (bs == 512)
        launch_step_phase<FMT, OP, 512, NT>(phase, ncell, st, a);
    else
        launch_step_phase<FMT, OP, 256, NT>(phase, ncell, st, a);
}

// Streaming modes the autotuner chooses from: none, phase A only, both.
static const int kStreamModes[3] = {0, LBMDEV_NT_A, LBMDEV_NT_A | LBMDEV_NT_B};

template <int FMT, int OP>
static void launch_main(LbmDev *d, const LbmStepParams *p, int phase, int use_ovr, int write_vel) {
    StepArgs<FMT> a;
    a.pop = (typename Pop<FMT>::S *)d->pop;
    a.mask = d->mask;
    a.ovr_index = d->ovr_index;
    a.ovr_bits = d->ovr_bits;
    a.ovr_vals = (const typename Pop<FMT>::T *)d->ovr_vals;
    a.ovr_stride = d->n_ovr;
    a.use_ovr = use_ovr && d->n_ovr > 0;
    a.vel = d->vel;
    a.write_vel = write_vel;
    a.g = make_geo(d->nx, d->ny, d->nz, d->stride);
    a.cp.tau = p->tau;
    a.cp.cs2x18 = 18.0f * p->smagorinskyCs * p->smagorinskyCs;
    a.cp.inx = p->inletX;
    a.cp.iny = p->inletY;
    a.cp.inz = p->inletZ;
    a.cp.smag = p->useSmagorinsky == 1;
    a.remap = d->remap;
    a.cp.guard = d->guard;
    // FP64 storage uses global addressing and has no streaming variant
    if (FMT != LBMDEV_FP64 && (d->nt & (phase == 0 ? LBMDEV_NT_A : LBMDEV_NT_B)))
        launch_step_bs<FMT, OP, FMT != LBMDEV_FP64>(d->bs, phase, d->ncell, d->stream, a);
    else
        launch_step_bs<FMT, OP, 0>(d->bs, phase, d->ncell, d->stream, a);
}

template <int FMT>
static void launch_main_op(LbmDev *d, const LbmStepParams *p, int phase, int use_ovr, int write_vel) {
    // Operator precedence of lbm_collide.comp:488-625: regularized, else MRT, else BGK.
    if (p->useRegularized == 1)
        launch_main<FMT, OP_REG>(d, p, phase, use_ovr, write_vel);
    else if (p->useMRT == 1)
        launch_main<FMT, OP_MRT>(d, p, phase, use_ovr, write_vel);
    else
        launch_main<FMT, OP_BGK>(d, p, phase, use_ovr, write_vel);
}

// One LBM step: prepass (if the stored field is post-collision) then the
// fused kernel in the phase opposite to the stored layout.
template <int FMT>
static void step_once(LbmDev *d, const LbmStepParams *p, int write_vel, hipEvent_t *ev) {
    int phase, use_ovr;
    if (d->state == ST_INIT) {
        phase = 0;
        use_ovr = 0;
    } else {
        if (ev) hipEventRecord(ev[0], d->stream);
        launch_prepass<FMT>(d, p->periodicYZ == 1);
        phase = d->state == ST_A ? 1 : 0;
        use_ovr = 1;
    }
    if (ev) hipEventRecord(ev[1], d->stream);
    launch_main_op<FMT>(d, p, phase, use_ovr, write_vel);
    if (ev) hipEventRecord(ev[2], d->stream);
    d->state = phase == 0 ? ST_A : ST_B;
}

static int dispatch_steps(LbmDev *d, const LbmStepParams *p, int nsteps, hipEvent_t *ev) {
    for (int s = 0; s < nsteps; s++) {
        const int wv = (s == nsteps - 1);
        hipEvent_t *e = ev ? ev + 3 * s : NULL;
        switch (d->fmt) {
        case LBMDEV_FP32: step_once<LBMDEV_FP32>(d, p, wv, e); break;
        case LBMDEV_FP16S: step_once<LBMDEV_FP16S>(d, p, wv, e); break;
        case LBMDEV_BF16S: step_once<LBMDEV_BF16S>(d, p, wv, e); break;
        case LBMDEV_FP64: step_once<LBMDEV_FP64>(d, p, wv, e); break;
        }
    }
    HIPCHECK(hipGetLastError(), "step launch");
    return 1;
}

extern "C" int lbmdev_step(LbmDev *d, const LbmStepParams *p, int nsteps) {
    if (d->state < 0) {
        snprintf(g_err, sizeof(g_err), "LBM_Step before LBM_InitializeFlow");
        return 0;
    }
    if (nsteps <= 0) return 1;
    return dispatch_steps(d, p, nsteps, NULL);
}

extern "C" int lbmdev_profile_steps(LbmDev *d, const LbmStepParams *p, int nsteps, float *total_ms,
                                    float *main_ms, float *pre_ms) {
    if (d->state < 0 || nsteps <= 0) return 0;
    hipEvent_t t0, t1;
    HIPCHECK(hipEventCreate(&t0), "event");
    HIPCHECK(hipEventCreate(&t1), "event");
    hipEvent_t *ev = NULL;
    const int per = (main_ms || pre_ms) ? 1 : 0;
    if (per) {
        // Timing-only events: no system-scope fence (cache write-back and
        // invalidation) when they are recorded, so they disturb the launches
        // they bracket as little as possible.  t1 below keeps the fence and is
        // what the host waits on.
        ev = (hipEvent_t *)calloc((size_t)nsteps * 3, sizeof(hipEvent_t));
        for (int k = 0; k < 3 * nsteps; k++) hipEventCreateWithFlags(&ev[k], hipEventDisableSystemFence);
    }
    // No host synchronisation before the region: t0 and the per-launch events
    // are stream-ordered, so work queued earlier is not counted, and the GPU
    // does not idle while the host waits.  Whether the first step had a
    // prepass (only the first step after LBM_InitializeFlow has none).
    const int first_init = d->state == ST_INIT;
    hipEventRecord(t0, d->stream);
    int ok = dispatch_steps(d, p, nsteps, ev);
    hipEventRecord(t1, d->stream);
    HIPCHECK(hipEventSynchronize(t1), "profile");
    float tot = 0;
    hipEventElapsedTime(&tot, t0, t1);
    if (total_ms) *total_ms = tot;
    if (per) {
        double sm = 0, sp = 0;
        for (int s = 0; s < nsteps; s++) {
            float a = 0, b = 0;
            if (!(s == 0 && first_init)) {
                hipEventElapsedTime(&a, ev[3 * s + 0], ev[3 * s + 1]);
                sp += a;
            }
            hipEventElapsedTime(&b, ev[3 * s + 1], ev[3 * s + 2]);
            sm += b;
        }
        if (main_ms) *main_ms = (float)sm;
        if (pre_ms) *pre_ms = (float)sp;
        for (int k = 0; k < 3 * nsteps; k++) hipEventDestroy(ev[k]);
        free(ev);
    }
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return ok;
}

// Steps with per-launch events and no host wait: the events go to the
// grid's pool (grown on the host while earlier work may still run) and are
// read by lbmdev_collect_timings after the caller's own synchronisation, so
// sampling the launches of a long run adds no idle time to it.
extern "C" int lbmdev_step_timed(LbmDev *d, const LbmStepParams *p, int nsteps) {
    if (d->state < 0) {
        snprintf(g_err, sizeof(g_err), "LBM_Step before LBM_InitializeFlow");
        return 0;
    }
    if (nsteps <= 0) return 1;
    const int need = d->tused + nsteps;
    if (need > (1 << 16)) {
        snprintf(g_err, sizeof(g_err), "more than 65536 timed steps without LBM_CollectStepTimings");
        return 0;
    }
    if (need > d->tcap) {
        int cap = d->tcap ? d->tcap : 64;
        while (cap < need) cap *= 2;
        hipEvent_t *ev = (hipEvent_t *)realloc(d->tev, sizeof(hipEvent_t) * 3 * (size_t)cap);
        if (!ev) {
            snprintf(g_err, sizeof(g_err), "host allocation failed");
            return 0;
        }
        d->tev = ev;
        unsigned char *pre = (unsigned char *)realloc(d->tpre, (size_t)cap);
        if (!pre) {
            snprintf(g_err, sizeof(g_err), "host allocation failed");
            return 0;
        }
        d->tpre = pre;
        for (int k = 3 * d->tcap; k < 3 * cap; k++) {
            // timing-only events: no system-scope fence when recorded
            if (hipEventCreateWithFlags(&d->tev[k], hipEventDisableSystemFence) != hipSuccess) {
                for (int j = 3 * d->tcap; j < k; j++) hipEventDestroy(d->tev[j]);
                snprintf(g_err, sizeof(g_err), "event creation failed");
                return 0;
            }
        }
        d->tcap = cap;
    }
    const int s0 = d->tused;
    for (int s = 0; s < nsteps; s++) d->tpre[s0 + s] = s > 0 || d->state != ST_INIT;
    if (!dispatch_steps(d, p, nsteps, d->tev + 3 * s0)) return 0;
    d->tused = need;
    return 1;
}

extern "C" int lbmdev_collect_timings(LbmDev *d, float *main_ms, float *pre_ms, int *launches) {
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    double sm = 0, sp = 0;
    for (int s = 0; s < d->tused; s++) {
        float a = 0, b = 0;
        if (d->tpre[s]) {
            HIPCHECK(hipEventElapsedTime(&a, d->tev[3 * s + 0], d->tev[3 * s + 1]), "event time");
            sp += a;
        }
        HIPCHECK(hipEventElapsedTime(&b, d->tev[3 * s + 1], d->tev[3 * s + 2]), "event time");
        sm += b;
    }
    if (main_ms) *main_ms = (float)sm;
    if (pre_ms) *pre_ms = (float)sp;
    if (launches) *launches = d->tused;
    d->tused = 0;
    return 1;
}

template <int FMT, int STATE>
static void launch_force(LbmDev *d, int periodic, int nb) {
    ForceArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)d->pop;
    a.mask = d->mask;
    a.ovr_cells = d->ovr_cells;
    a.ovr_q = d->ovr_q;
    a.vel = d->vel;
    a.n_ovr = d->n_ovr;
    a.q_stride = d->n_ovr;
    a.periodic = periodic;
    a.g = make_geo(d->nx, d->ny, d->nz, d->stride);
    a.pd = d->pd;
    a.pi = d->pi;
    hipLaunchKernelGGL((k_force<FMT, STATE>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
}

template <int FMT>
static void launch_force_state(LbmDev *d, int periodic, int nb) {
    if (d->state == ST_INIT) launch_force<FMT, ST_INIT>(d, periodic, nb);
    else if (d->state == ST_A) launch_force<FMT, ST_A>(d, periodic, nb);
    else launch_force<FMT, ST_B>(d, periodic, nb);
}

extern "C" int lbmdev_force(LbmDev *d, int periodicYZ, double exact[6], long long fixed[7]) {
    for (int k = 0; k < 6; k++) exact[k] = 0;
    for (int k = 0; k < 7; k++) fixed[k] = 0;
    if (d->state < 0 || d->n_ovr == 0) return 1;
    int nb = (d->n_ovr + BLOCK - 1) / BLOCK;
    if (nb > FORCE_BLOCKS) nb = FORCE_BLOCKS;
    switch (d->fmt) {
    case LBMDEV_FP32: launch_force_state<LBMDEV_FP32>(d, periodicYZ, nb); break;
    case LBMDEV_FP16S: launch_force_state<LBMDEV_FP16S>(d, periodicYZ, nb); break;
    case LBMDEV_BF16S: launch_force_state<LBMDEV_BF16S>(d, periodicYZ, nb); break;
    case LBMDEV_FP64: launch_force_state<LBMDEV_FP64>(d, periodicYZ, nb); break;
    }
    hipLaunchKernelGGL((k_finalize<6, 7>), dim3(1), dim3(64), 0, d->stream, d->pd, d->pi, nb, d->resd, d->resi);
    HIPCHECK(hipGetLastError(), "force launch");
    HIPCHECK(hipMemcpyAsync(exact, d->resd, 6 * 8, hipMemcpyDeviceToHost, d->stream), "force readback");
    HIPCHECK(hipMemcpyAsync(fixed, d->resi, 7 * 8, hipMemcpyDeviceToHost, d->stream), "force readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "force");
    return 1;
}

template <int FMT, int STATE>
static void launch_export_s(LbmDev *d, const ExportArgs<FMT> &a) {
    const int nb = (d->ncell + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((k_export<FMT, STATE>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
}

template <int FMT>
static void launch_export(LbmDev *d, int periodic, void *os, void *oc, int as_double, int mass) {
    ExportArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)d->pop;
    a.mask = d->mask;
    a.ovr_index = d->ovr_index;
    a.ovr_q = d->ovr_q;
    a.n_ovr = d->n_ovr;
    a.periodic = periodic;
    a.g = make_geo(d->nx, d->ny, d->nz, d->stride);
    a.out_stream = os;
    a.out_collision = oc;
    a.as_double = as_double;
    a.pd = mass ? d->pd : NULL;
    a.pi = mass ? d->pi : NULL;
    if (d->state == ST_INIT) launch_export_s<FMT, ST_INIT>(d, a);
    else if (d->state == ST_A) launch_export_s<FMT, ST_A>(d, a);
    else launch_export_s<FMT, ST_B>(d, a);
}

static int export_any(LbmDev *d, int periodic, void *os, void *oc, int as_double, int mass) {
    switch (d->fmt) {
    case LBMDEV_FP32: launch_export<LBMDEV_FP32>(d, periodic, os, oc, as_double, mass); break;
    case LBMDEV_FP16S: launch_export<LBMDEV_FP16S>(d, periodic, os, oc, as_double, mass); break;
    case LBMDEV_BF16S: launch_export<LBMDEV_BF16S>(d, periodic, os, oc, as_double, mass); break;
    case LBMDEV_FP64: launch_export<LBMDEV_FP64>(d, periodic, os, oc, as_double, mass); break;
    }
    HIPCHECK(hipGetLastError(), "export launch");
    return 1;
}

static int export_populations(LbmDev *d, int periodicYZ, void *f_ps, void *f_pc, int as_double) {
    if (d->state < 0) {
        snprintf(g_err, sizeof(g_err), "populations read before LBM_InitializeFlow");
        return 0;
    }
    const size_t es = as_double ? 8 : 4;
    const size_t bytes = (size_t)d->ncell * 19 * es;
    void *dps = NULL, *dpc = NULL;
    if (f_ps) HIPCHECK(hipMalloc(&dps, bytes), "export buffer");
    if (f_pc) {
        hipError_t e = hipMalloc(&dpc, bytes);
        if (e != hipSuccess) {
            hipFree(dps);
            return set_err("export buffer", e);
        }
    }
    int ok = export_any(d, periodicYZ, dps, dpc, as_double, 0);
    hipError_t e = hipSuccess;
    if (ok && f_ps) e = hipMemcpyAsync(f_ps, dps, bytes, hipMemcpyDeviceToHost, d->stream);
    if (ok && e == hipSuccess && f_pc) e = hipMemcpyAsync(f_pc, dpc, bytes, hipMemcpyDeviceToHost, d->stream);
    hipError_t e2 = hipStreamSynchronize(d->stream);
    hipFree(dps);
    hipFree(dpc);
    if (e != hipSuccess) return set_err("export readback", e);
    if (e2 != hipSuccess) return set_err("export", e2);
    return ok;
}

extern "C" int lbmdev_export_populations(LbmDev *d, int periodicYZ, float *f_ps, float *f_pc) {
    return export_populations(d, periodicYZ, f_ps, f_pc, 0);
}

extern "C" int lbmdev_export_populations_f64(LbmDev *d, int periodicYZ, double *f_ps, double *f_pc) {
    return export_populations(d, periodicYZ, f_ps, f_pc, 1);
}

extern "C" int lbmdev_total_mass(LbmDev *d, int periodicYZ, double *mass) {
    *mass = 0;
    if (d->state < 0) return 0;
    if (!export_any(d, periodicYZ, NULL, NULL, 1, 1)) return 0;
    const int nb = (d->ncell + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL((k_finalize<1, 1>), dim3(1), dim3(64), 0, d->stream, d->pd, d->pi, nb, d->resd, d->resi);
    HIPCHECK(hipMemcpyAsync(mass, d->resd, 8, hipMemcpyDeviceToHost, d->stream), "mass readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "mass");
    return 1;
}

// Raw state for checkpoints: the stored populations exactly as the device
// holds them (format, padding and the phase layout) and the velocity buffer.
extern "C" size_t lbmdev_population_bytes(const LbmDev *d) { return (size_t)19 * d->stride * d->elem; }

extern "C" int lbmdev_save_state(LbmDev *d, void *pop, void *vel, int *state) {
    if (d->state < 0) {
        snprintf(g_err, sizeof(g_err), "no flow to save: call LBM_InitializeFlow first");
        return 0;
    }
    HIPCHECK(hipMemcpyAsync(pop, d->pop, lbmdev_population_bytes(d), hipMemcpyDeviceToHost, d->stream), "state readback");
    HIPCHECK(hipMemcpyAsync(vel, d->vel, (size_t)d->ncell * 16, hipMemcpyDeviceToHost, d->stream), "state readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "state readback");
    *state = d->state;
    return 1;
}

extern "C" int lbmdev_load_state(LbmDev *d, const void *pop, const void *vel, int state) {
    if (state != ST_INIT && state != ST_A && state != ST_B) {
        snprintf(g_err, sizeof(g_err), "saved state has an unknown population layout (%d)", state);
        return 0;
    }
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    HIPCHECK(hipMemcpyAsync(d->pop, pop, lbmdev_population_bytes(d), hipMemcpyHostToDevice, d->stream), "state upload");
    HIPCHECK(hipMemcpyAsync(d->vel, vel, (size_t)d->ncell * 16, hipMemcpyHostToDevice, d->stream), "state upload");
    HIPCHECK(hipStreamSynchronize(d->stream), "state upload");
    d->state = state;
    return 1;
}

extern "C" float *lbmdev_velocity_ptr(LbmDev *d) { return (float *)d->vel; }
extern "C" int32_t *lbmdev_solid_ptr(LbmDev *d) { return d->solid; }

extern "C" int lbmdev_copy_velocity(LbmDev *d, float *host) {
    HIPCHECK(hipMemcpyAsync(host, d->vel, (size_t)d->ncell * 16, hipMemcpyDeviceToHost, d->stream), "velocity readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "velocity readback");
    return 1;
}

// Mean of the velocity buffer over factor^3 blocks: coarse cell (X, Y, Z)
// averages fine cells [fX, fX + f) x [fY, fY + f) x [fZ, fZ + f).
__global__ __launch_bounds__(BLOCK) void k_coarsen(const float4 *__restrict__ vel, int nx, int ny, int f,
                                                   float4 *__restrict__ out, int cnx, int cny, int ncoarse) {
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= ncoarse) return;
    const int X = c % cnx, Y = (c / cnx) % cny, Z = c / (cnx * cny);
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int dz = 0; dz < f; dz++)
        for (int dy = 0; dy < f; dy++) {
            const long long row = ((long long)(Z * f + dz) * ny + (Y * f + dy)) * nx + (long long)X * f;
            for (int dx = 0; dx < f; dx++) {
                const float4 v = vel[row + dx];
                s.x += v.x;
                s.y += v.y;
                s.z += v.z;
                s.w += v.w;
            }
        }
    const float inv = 1.0f / (float)(f * f * f);
    out[c] = make_float4(s.x * inv, s.y * inv, s.z * inv, s.w * inv);
}

static int coarsen_to_host(const float4 *vel, int nx, int ny, int nz, int f, float *host, hipStream_t s) {
    const int cnx = nx / f, cny = ny / f, cnz = nz / f, nc = cnx * cny * cnz;
    float4 *c = NULL;
    HIPCHECK(hipMalloc((void **)&c, sizeof(float4) * (size_t)nc), "coarse velocity");
    hipLaunchKernelGGL(k_coarsen, dim3((nc + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, vel, nx, ny, f, c, cnx, cny, nc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, c, sizeof(float4) * (size_t)nc, hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipStreamSynchronize(s);
    hipFree(c);
    if (e != hipSuccess) return set_err("coarse velocity", e);
    if (e2 != hipSuccess) return set_err("coarse velocity", e2);
    return 1;
}

extern "C" int lbmdev_copy_velocity_coarse(LbmDev *d, int f, float *host) {
    return coarsen_to_host(d->vel, d->nx, d->ny, d->nz, f, host, d->stream);
}

extern "C" int lbmdev_copy_solid(LbmDev *d, int32_t *host) {
    HIPCHECK(hipMemcpyAsync(host, d->solid, (size_t)d->ncell * 4, hipMemcpyDeviceToHost, d->stream), "solid readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "solid readback");
    return 1;
}

extern "C" int lbmdev_particles_seed(LbmDev *d, void *dev_particles, int n, uint32_t seed) {
    if (n <= 0) return 1;
    hipLaunchKernelGGL(k_particles_seed, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, d->stream,
                       (Particle *)dev_particles, n, seed);
    HIPCHECK(hipGetLastError(), "particle seed launch");
    return 1;
}

extern "C" int lbmdev_particles_live(LbmDev *d, void *dev_particles, int n, float dt, const LbmParticleFlow *fl,
                                    uint32_t seed, uint32_t step) {
    if (n <= 0) return 1;
    LiveArgs a;
    a.wind = V3{fl->wind[0], fl->wind[1], fl->wind[2]};
    a.carMin = V3{fl->carMin[0], fl->carMin[1], fl->carMin[2]};
    a.carMax = V3{fl->carMax[0], fl->carMax[1], fl->carMax[2]};
    a.carCenter = V3{fl->carCenter[0], fl->carCenter[1], fl->carCenter[2]};
    a.mode = fl->collisionMode;
    a.useLBM = fl->useLBM;
    a.vizMode = fl->vizMode;
    a.tris = fl->triangles;
    a.ntri = fl->numTriangles;
    a.cellStart = fl->cellStart;
    a.cellCount = fl->cellCount;
    a.triIdx = fl->triIndices;
    a.nIdx = fl->numIndices;
    a.gridMin = V3{fl->gridMin[0], fl->gridMin[1], fl->gridMin[2]};
    a.gridCell = V3{fl->gridCellSize[0], fl->gridCellSize[1], fl->gridCellSize[2]};
    a.gridRes = fl->gridRes;
    if (a.mode == 2 && a.ntri > 0 && a.gridRes > 0 && (!a.tris || !a.cellStart || !a.cellCount || !a.triIdx)) {
        snprintf(g_err, sizeof(g_err), "collision mode 2 needs the triangles and the collision grid on the device");
        return 0;
    }
    if (a.mode == 2 && a.ntri > 0 && ((uintptr_t)a.tris & 15)) {
        // the kernel reads each 12-float triangle as three float4
        snprintf(g_err, sizeof(g_err), "collision mode 2 needs a 16-byte aligned triangle array");
        return 0;
    }
    hipLaunchKernelGGL(k_particles_live, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, d->stream,
                       (Particle *)dev_particles, n, (const float4 *)d->vel, (const int32_t *)d->solid, d->nx, d->ny,
                       d->nz, dt, a, seed, step);
    HIPCHECK(hipGetLastError(), "live particle launch");
    return 1;
}

extern "C" int lbmdev_particles_advect(LbmDev *d, void *dev_particles, int n, float dt, float velScale,
                                       uint32_t seed, uint32_t step) {
    if (n <= 0) return 1;
    hipLaunchKernelGGL(k_particles_advect, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, d->stream,
                       (Particle *)dev_particles, n, (const float4 *)d->vel, d->nx, d->ny, d->nz, dt, velScale,
                       seed, step);
    HIPCHECK(hipGetLastError(), "particle advect launch");
    return 1;
}

extern "C" int lbmdev_synchronize(LbmDev *d) {
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    return 1;
}

extern "C" void *lbmdev_stream(LbmDev *d) { return (void *)d->stream; }

extern "C" void *lbmdev_malloc(size_t bytes) {
    void *p = NULL;
    if (hipMalloc(&p, bytes) != hipSuccess) return NULL;
    return p;
}

extern "C" void lbmdev_free(void *p) {
    if (p) hipFree(p);
}

extern "C" int lbmdev_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "copy to device");
    return 1;
}

extern "C" int lbmdev_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copy to host");
    return 1;
}

extern "C" int lbmdev_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "device copy");
    return 1;
}

extern "C" int lbmdev_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ======================================================================
// Batched tunnels (C ABI): many lattices in one allocation, one launch per
// step covering all of them, the step sequence replayed from HIP graphs.
// ======================================================================

struct LbmBatchDev {
    int fmt;
    size_t elem, telem;
    int nt;
    TunnelDesc *desc_h;
    TunnelDesc *desc_d;
    int *dims;  // 3 per tunnel
    long long total_cells, total_pop;
    void *pop;
    uint8_t *mask;
    int32_t *ovr_index;
    float4 *vel;
    int n_ovr;
    int *kstart;  // nt + 1 list offsets per tunnel (host)
    int32_t *ovr_cells;
    float *ovr_q;
    void *ovr_vals;
    uint32_t *ovr_bits;  // per listed cell: directions taken from ovr_vals
    PrepassItem *ovr_items;  // prepass work list (see BoundArgs)
    int seg[20];
    int32_t *block_tunnel;
    int nblocks, bs;
    int smode;  // LBMDEV_NT_A / LBMDEV_NT_B streaming mode of the step kernel
    int remap;  // XCD-contiguous workgroup order in the step kernel
    int state;
    hipStream_t stream;
    double *pd;
    long long *pi;
    double *resd;
    long long *resi;
    unsigned long long *guard;  // stability-guard resets over all tunnels
    int nparts;
    // One launch for every tunnel's force: ftun[t] = (list start, list
    // length, first block, blocks) with the block count the per-tunnel launch
    // would use, fblk_tunnel[block] = its tunnel; fblocks blocks in all.
    int4 *ftun;
    int32_t *fblk_tunnel;
    int fblocks;
    hipGraphExec_t graph[2];
    int graph_ok;
    LbmBatchParams graph_params;
    int graph_bs, graph_smode, graph_remap;
};

extern "C" void lbmbatch_destroy(LbmBatchDev *b);

static void batch_free_graphs(LbmBatchDev *b) {
    for (int k = 0; k < 2; k++)
        if (b->graph[k]) {
            hipGraphExecDestroy(b->graph[k]);
            b->graph[k] = NULL;
        }
    b->graph_ok = 0;
}

// Lays the workgroups of every tunnel end to end for workgroup size bs.
static int batch_layout_blocks(LbmBatchDev *b, int bs) {
    int nb = 0;
    for (int t = 0; t < b->nt; t++) nb += (b->desc_h[t].g.ncell + bs - 1) / bs;
    int32_t *map = (int32_t *)malloc(sizeof(int32_t) * (size_t)nb);
    if (!map) return 0;
    int at = 0;
    for (int t = 0; t < b->nt; t++) {
        b->desc_h[t].blockStart = at;
        const int n = (b->desc_h[t].g.ncell + bs - 1) / bs;
        for (int k = 0; k < n; k++) map[at + k] = t;
        at += n;
    }
    hipFree(b->block_tunnel);
    b->block_tunnel = NULL;
    hipError_t e = hipMalloc((void **)&b->block_tunnel, sizeof(int32_t) * (size_t)nb);
    if (e == hipSuccess) e = hipMemcpy(b->block_tunnel, map, sizeof(int32_t) * (size_t)nb, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->desc_d, b->desc_h, sizeof(TunnelDesc) * (size_t)b->nt, hipMemcpyHostToDevice);
    free(map);
    if (e != hipSuccess) return set_err("batch layout", e);
    b->nblocks = nb;
    b->bs = bs;
    batch_free_graphs(b);
    return 1;
}

extern "C" LbmBatchDev *lbmbatch_create(int storage, int nt, const int *dims, const float *tau,
                                        const float *inlet, char *err, int errlen) {
    LbmBatchDev *b = (LbmBatchDev *)calloc(1, sizeof(LbmBatchDev));
    if (!b) return NULL;
    b->fmt = storage;
    b->elem = elem_size(storage);
    b->telem = storage == LBMDEV_FP64 ? 8 : 4;
    b->nt = nt;
    b->remap = 1;
    b->state = -1;
    b->desc_h = (TunnelDesc *)calloc((size_t)nt, sizeof(TunnelDesc));
    b->dims = (int *)malloc(sizeof(int) * 3 * (size_t)nt);
    b->kstart = (int *)calloc((size_t)nt + 1, sizeof(int));
    if (!b->desc_h || !b->dims || !b->kstart) {
        snprintf(err, errlen, "host allocation failed");
        lbmbatch_destroy(b);
        return NULL;
    }
    long long cells = 0, popel = 0;
    for (int t = 0; t < nt; t++) {
        const int nx = dims[3 * t], ny = dims[3 * t + 1], nz = dims[3 * t + 2];
        const long long nc = (long long)nx * ny * nz;
        if (nx < 2 || ny < 1 || nz < 1 || cells + nc > 0x7fffffffLL - 256 ||
            (storage != LBMDEV_FP64 && (unsigned long long)nc * b->elem > (1ULL << 32))) {
            snprintf(err, errlen, "tunnel %d (%dx%dx%d) is outside the supported range", t, nx, ny, nz);
            lbmbatch_destroy(b);
            return NULL;
        }
        memcpy(b->dims + 3 * t, dims + 3 * t, 3 * sizeof(int));
        TunnelDesc &d = b->desc_h[t];
        d.g = make_geo(nx, ny, nz, (nc + 63) / 64 * 64);
        d.popOff = popel;
        d.cellOff = (int)cells;
        d.tau = tau[t];
        d.inx = inlet[3 * t];
        d.iny = inlet[3 * t + 1];
        d.inz = inlet[3 * t + 2];
        cells += nc;
        popel += 19 * d.g.stride;
    }
    b->total_cells = cells;
    b->total_pop = popel;
    b->nparts = (int)((cells + BLOCK - 1) / BLOCK);
    if (b->nparts < FORCE_BLOCKS) b->nparts = FORCE_BLOCKS;
    hipError_t e;
#define BALLOC(ptr, bytes, name)                                                          \
    if ((e = hipMalloc((void **)&(ptr), (bytes))) != hipSuccess) {                        \
        snprintf(err, errlen, "device allocation of %s (%.1f MB) failed: %s", name,     \
                 (double)(bytes) / 1048576.0, hipGetErrorString(e));                      \
        lbmbatch_destroy(b);                                                              \
        return NULL;                                                                      \
    }
    BALLOC(b->pop, (size_t)popel * b->elem, "batched populations");
    BALLOC(b->mask, (size_t)cells, "mask");
    BALLOC(b->ovr_index, (size_t)cells * 4, "boundary index");
    BALLOC(b->vel, (size_t)cells * 16, "velocity");
    BALLOC(b->desc_d, sizeof(TunnelDesc) * (size_t)nt, "tunnel table");
    BALLOC(b->pd, (size_t)b->nparts * 8 * 8, "reduction partials");
    BALLOC(b->pi, (size_t)b->nparts * 8 * 8, "reduction partials");
    BALLOC(b->resd, (size_t)nt * 8 * 8, "reduction results");
    BALLOC(b->resi, (size_t)nt * 8 * 8, "reduction results");
    BALLOC(b->guard, 8, "guard counter");
#undef BALLOC
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
        snprintf(err, errlen, "stream creation failed: %s", hipGetErrorString(e));
        lbmbatch_destroy(b);
        return NULL;
    }
    if (!batch_layout_blocks(b, 256)) {
        snprintf(err, errlen, "%s", g_err);
        lbmbatch_destroy(b);
        return NULL;
    }
    return b;
}

extern "C" void lbmbatch_destroy(LbmBatchDev *b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    batch_free_graphs(b);
    hipFree(b->pop);
    hipFree(b->mask);
    hipFree(b->ovr_index);
    hipFree(b->vel);
    hipFree(b->desc_d);
    hipFree(b->ovr_cells);
    hipFree(b->ovr_q);
    hipFree(b->ovr_vals);
    hipFree(b->ovr_bits);
    hipFree(b->ovr_items);
    hipFree(b->ftun);
    hipFree(b->fblk_tunnel);
    hipFree(b->block_tunnel);
    hipFree(b->pd);
    hipFree(b->pi);
    hipFree(b->resd);
    hipFree(b->resi);
    hipFree(b->guard);
    if (b->stream) hipStreamDestroy(b->stream);
    free(b->desc_h);
    free(b->dims);
    free(b->kstart);
    free(b);
}

// solid: concatenated per tunnel; n_ovr_per[t] list entries of tunnel t;
// ovr_cells: concatenated local cell indices; ovr_q: planes q[i*Ntot + k].
extern "C" int lbmbatch_set_geometry(LbmBatchDev *b, const int32_t *solid, const int *n_ovr_per,
                                     const int32_t *ovr_cells, const float *ovr_q, int periodic) {
    HIPCHECK(hipStreamSynchronize(b->stream), "synchronize");
    const long long nc = b->total_cells;
    int ntot = 0;
    for (int t = 0; t < b->nt; t++) {
        b->kstart[t] = ntot;
        ntot += n_ovr_per[t];
    }
    b->kstart[b->nt] = ntot;
    uint8_t *mask = (uint8_t *)malloc((size_t)nc);
    int32_t *index = (int32_t *)malloc((size_t)nc * 4);
    int32_t *tun = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntot > 0 ? ntot : 1));
    uint32_t *bits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(ntot > 0 ? ntot : 1));
    PrepassItem *items = NULL;
    int seg[20];
    if (bits && tun)
        for (int t = 0; t < b->nt; t++)
            for (int k = b->kstart[t]; k < b->kstart[t + 1]; k++) {
                bits[k] = special_dirs(b->desc_h[t].g.nx, b->desc_h[t].g.ny, b->desc_h[t].g.nz, periodic, ovr_cells[k],
                                       ovr_q, ntot, k);
                tun[k] = t;
            }
    const int built = mask && index && tun && bits ? build_items(ntot, bits, ovr_cells, ovr_q, ntot, tun, &items, seg) : 0;
    if (built <= 0) {
        free(mask);
        free(index);
        free(tun);
        free(bits);
        snprintf(g_err, sizeof(g_err), built < 0 ? "boundary work list exceeds 2^28 (cell, direction) pairs"
                                                  : "host allocation failed");
        return 0;
    }
    for (long long c = 0; c < nc; c++) {
        mask[c] = (solid[c] == 1 ? LBMDEV_MASK_BODY : 0) | (solid[c] == 2 ? LBMDEV_MASK_GROUND : 0);
        index[c] = -1;
    }
    for (int t = 0; t < b->nt; t++)
        for (int k = b->kstart[t]; k < b->kstart[t + 1]; k++) {
            const long long c = b->desc_h[t].cellOff + (long long)ovr_cells[k];
            mask[c] |= LBMDEV_MASK_OVR;
            index[c] = k;
        }
    hipFree(b->ovr_cells);
    hipFree(b->ovr_q);
    hipFree(b->ovr_vals);
    hipFree(b->ovr_bits);
    hipFree(b->ovr_items);
    b->ovr_cells = NULL;
    b->ovr_q = NULL;
    b->ovr_vals = NULL;
    b->ovr_bits = NULL;
    b->ovr_items = NULL;
    hipError_t e = hipSuccess;
    const size_t n1 = (size_t)(ntot > 0 ? ntot : 1);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ovr_bits, n1 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ovr_items, (size_t)(seg[19] > 0 ? seg[19] : 1) * sizeof(PrepassItem));
    if (e == hipSuccess) e = hipMalloc((void **)&b->ovr_cells, n1 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ovr_q, n1 * 19 * 4);
    if (e == hipSuccess) e = hipMalloc((void **)&b->ovr_vals, n1 * 19 * b->telem);
    if (e == hipSuccess) e = hipMemcpy(b->mask, mask, (size_t)nc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->ovr_index, index, (size_t)nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && ntot > 0) e = hipMemcpy(b->ovr_cells, ovr_cells, (size_t)ntot * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && ntot > 0)
        e = hipMemcpy(b->ovr_q, ovr_q, (size_t)ntot * 19 * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && ntot > 0) e = hipMemcpy(b->ovr_bits, bits, (size_t)ntot * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && seg[19] > 0)
        e = hipMemcpy(b->ovr_items, items, (size_t)seg[19] * sizeof(PrepassItem), hipMemcpyHostToDevice);
    free(mask);
    free(index);
    free(tun);
    free(bits);
    free(items);
    if (e != hipSuccess) return set_err("batch geometry upload", e);
    memcpy(b->seg, seg, sizeof(seg));
    b->n_ovr = ntot;
    batch_free_graphs(b);
    // the one-launch force tables (batch_force_t): each tunnel gets the blocks
    // its own launch would, min(ceil(n / BLOCK), FORCE_BLOCKS), at least one
    hipFree(b->ftun);
    hipFree(b->fblk_tunnel);
    b->ftun = NULL;
    b->fblk_tunnel = NULL;
    b->fblocks = 0;
    int4 *ft = (int4 *)malloc(sizeof(int4) * (size_t)b->nt);
    long long total = 0;
    if (ft) {
        for (int t = 0; t < b->nt; t++) {
            const int n = b->kstart[t + 1] - b->kstart[t];
            int nb = (n + BLOCK - 1) / BLOCK;
            if (nb > FORCE_BLOCKS) nb = FORCE_BLOCKS;
            if (nb < 1) nb = 1;
            ft[t] = make_int4(b->kstart[t], n, (int)total, nb);
            total += nb;
        }
    }
    int32_t *bt = ft && total <= b->nparts ? (int32_t *)malloc(sizeof(int32_t) * (size_t)total) : NULL;
    if (bt) {
        for (int t = 0; t < b->nt; t++)
            for (int k = 0; k < ft[t].w; k++) bt[ft[t].z + k] = t;
        // without these tables the per-tunnel launches are used
        if (hipMalloc((void **)&b->ftun, sizeof(int4) * (size_t)b->nt) == hipSuccess &&
            hipMalloc((void **)&b->fblk_tunnel, sizeof(int32_t) * (size_t)total) == hipSuccess &&
            hipMemcpy(b->ftun, ft, sizeof(int4) * (size_t)b->nt, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->fblk_tunnel, bt, sizeof(int32_t) * (size_t)total, hipMemcpyHostToDevice) == hipSuccess)
            b->fblocks = (int)total;
    }
    free(ft);
    free(bt);
    return 1;
}

template <int FMT>
static void batch_init_t(LbmBatchDev *b, const float *feq_d, const float *vel4) {
    typedef typename Pop<FMT>::S S;
    for (int t = 0; t < b->nt; t++) {
        const TunnelDesc &d = b->desc_h[t];
        const int nb = (d.g.ncell + BLOCK - 1) / BLOCK;
        float4 v4 = make_float4(vel4[4 * t], vel4[4 * t + 1], vel4[4 * t + 2], vel4[4 * t + 3]);
        hipLaunchKernelGGL(k_init<FMT>, dim3(nb), dim3(BLOCK), 0, b->stream, (S *)b->pop + d.popOff, d.g,
                           b->vel + d.cellOff, feq_d + 19 * t, v4);
    }
}

extern "C" int lbmbatch_init_flow(LbmBatchDev *b, const float *feq, const float *vel4) {
    float *feq_d = NULL;
    HIPCHECK(hipMalloc((void **)&feq_d, sizeof(float) * 19 * (size_t)b->nt), "feq table");
    hipError_t e = hipMemcpy(feq_d, feq, sizeof(float) * 19 * (size_t)b->nt, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(b->guard, 0, 8, b->stream);
    if (e == hipSuccess) {
        switch (b->fmt) {
        case LBMDEV_FP32: batch_init_t<LBMDEV_FP32>(b, feq_d, vel4); break;
        case LBMDEV_FP16S: batch_init_t<LBMDEV_FP16S>(b, feq_d, vel4); break;
        case LBMDEV_BF16S: batch_init_t<LBMDEV_BF16S>(b, feq_d, vel4); break;
        case LBMDEV_FP64: batch_init_t<LBMDEV_FP64>(b, feq_d, vel4); break;
        }
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(b->stream);
    hipFree(feq_d);
    if (e != hipSuccess) return set_err("batch init", e);
    if (e2 != hipSuccess) return set_err("batch init", e2);
    b->state = ST_INIT;
    return 1;
}

// Overwrites tunnel t of a freshly initialised batch with the equilibrium of
// a coarse field (see lbmdev_init_field).
extern "C" int lbmbatch_init_tunnel_field(LbmBatchDev *b, int t, const float *coarse_host, int cnx, int cny,
                                          int cnz) {
    if (t < 0 || t >= b->nt) return set_err("field init: tunnel index", hipErrorInvalidValue);
    if (b->state != ST_INIT) return set_err("field init: batch must be freshly initialised", hipErrorInvalidValue);
    const size_t bytes = sizeof(float) * 4 * (size_t)cnx * cny * cnz;
    float *c = NULL;
    HIPCHECK(hipMalloc((void **)&c, bytes), "coarse field");
    hipError_t e = hipMemcpy(c, coarse_host, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const TunnelDesc &d = b->desc_h[t];
        init_field_dispatch(b->fmt, (char *)b->pop + (size_t)d.popOff * b->elem, d.g, b->vel + d.cellOff, c, cnx, cny,
                            cnz, b->stream);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(b->stream);
    hipFree(c);
    if (e != hipSuccess) return set_err("field init", e);
    if (e2 != hipSuccess) return set_err("field init", e2);
    return 1;
}

template <int FMT, int OP, int BS, int NT>
static void batch_launch_main(LbmBatchDev *b, const LbmBatchParams *p, int phase, int use_ovr, int wv) {
    BatchArgs<FMT> a;
    a.pop = (typename Pop<FMT>::S *)b->pop;
    a.mask = b->mask;
    a.ovr_index = b->ovr_index;
    a.ovr_vals = (const typename Pop<FMT>::T *)b->ovr_vals;
    a.ovr_bits = b->ovr_bits;
    a.ovr_stride = b->n_ovr;
    a.use_ovr = use_ovr && b->n_ovr > 0;
    a.vel = b->vel;
    a.write_vel = wv;
    a.cs2x18 = 18.0f * p->smagorinskyCs * p->smagorinskyCs;
    a.smag = p->useSmagorinsky == 1;
    a.desc = b->desc_d;
    a.block_tunnel = b->block_tunnel;
    a.remap = b->remap;
    a.guard = b->guard;
    // row-aligned phase B when every tunnel's rows fill whole workgroups
    bool row = FMT != LBMDEV_FP64;
    for (int t = 0; t < b->nt && row; t++) {
        const Geo &g = b->desc_h[t].g;
        row = g.nx % 64 == 0 && BS % g.nx == 0 && g.ncell % BS == 0;
    }
    if (phase == 0)
        hipLaunchKernelGGL((k_step_batch<FMT, OP, 0, BS, 0, NT>), dim3(b->nblocks), dim3(BS), 0, b->stream, a);
    else if (row)
        hipLaunchKernelGGL((k_step_batch<FMT, OP, 1, BS, 1, NT>), dim3(b->nblocks), dim3(BS), 0, b->stream, a);
    else
        hipLaunchKernelGGL((k_step_batch<FMT, OP, 1, BS, 0, NT>), dim3(b->nblocks), dim3(BS), 0, b->stream, a);
}

template <int FMT, int OP, int NT>
static void batch_launch_main_nt(LbmBatchDev *b, const LbmBatchParams *p, int phase, int use_ovr, int wv) {
    if (b->bs == 128)
        batch_launch_main<FMT, OP, 128, NT>(b, p, phase, use_ovr, wv);
    else if (b->bs == 512)
        batch_launch_main<FMT, OP, 512, NT>(b, p, phase, use_ovr, wv);
    else
        batch_launch_main<FMT, OP, 256, NT>(b, p, phase, use_ovr, wv);
}

template <int FMT, int OP>
static void batch_launch_main_bs(LbmBatchDev *b, const LbmBatchParams *p, int phase, int use_ovr, int wv) {
    // FP64 storage uses global addressing and has no streaming variant
    if (FMT != LBMDEV_FP64 && (b->smode & (phase == 0 ? LBMDEV_NT_A : LBMDEV_NT_B)))
        batch_launch_main_nt<FMT, OP, FMT != LBMDEV_FP64>(b, p, phase, use_ovr, wv);
    else
        batch_launch_main_nt<FMT, OP, 0>(b, p, phase, use_ovr, wv);
}

template <int FMT>
static void batch_step_once(LbmBatchDev *b, const LbmBatchParams *p, int wv) {
    int phase, use_ovr;
    if (b->state == ST_INIT) {
        phase = 0;
        use_ovr = 0;
    } else {
        if (b->n_ovr > 0 && b->seg[19] > 0) {
            BoundArgs<FMT> a;
            a.pop = (const typename Pop<FMT>::S *)b->pop;
            a.ovr_vals = (typename Pop<FMT>::T *)b->ovr_vals;
            a.n_ovr = b->n_ovr;
            a.periodic = p->periodicYZ == 1;
            a.g = make_geo(2, 1, 1, 64);
            a.items = b->ovr_items;
            memcpy(a.seg, b->seg, sizeof(a.seg));
            const int nb = (a.seg[19] + BLOCK - 1) / BLOCK;
            if (b->state == ST_A)
                hipLaunchKernelGGL((k_prepass_batch<FMT, ST_A>), dim3(nb), dim3(BLOCK), 0, b->stream, a,
                                   (const TunnelDesc *)b->desc_d);
            else
                hipLaunchKernelGGL((k_prepass_batch<FMT, ST_B>), dim3(nb), dim3(BLOCK), 0, b->stream, a,
                                   (const TunnelDesc *)b->desc_d);
        }
        phase = b->state == ST_A ? 1 : 0;
        use_ovr = 1;
    }
    if (p->useRegularized == 1)
        batch_launch_main_bs<FMT, OP_REG>(b, p, phase, use_ovr, wv);
    else if (p->useMRT == 1)
        batch_launch_main_bs<FMT, OP_MRT>(b, p, phase, use_ovr, wv);
    else
        batch_launch_main_bs<FMT, OP_BGK>(b, p, phase, use_ovr, wv);
    b->state = phase == 0 ? ST_A : ST_B;
}

static void batch_step_any(LbmBatchDev *b, const LbmBatchParams *p, int wv) {
    switch (b->fmt) {
    case LBMDEV_FP32: batch_step_once<LBMDEV_FP32>(b, p, wv); break;
    case LBMDEV_FP16S: batch_step_once<LBMDEV_FP16S>(b, p, wv); break;
    case LBMDEV_BF16S: batch_step_once<LBMDEV_BF16S>(b, p, wv); break;
    case LBMDEV_FP64: batch_step_once<LBMDEV_FP64>(b, p, wv); break;
    }
}

static int same_params(const LbmBatchParams *a, const LbmBatchParams *c) {
    return a->smagorinskyCs == c->smagorinskyCs && a->useRegularized == c->useRegularized &&
           a->useMRT == c->useMRT && a->useSmagorinsky == c->useSmagorinsky && a->periodicYZ == c->periodicYZ;
}

// Captures the two-step sequence (phase B then phase A, each with its
// prepass) that returns the stored layout to phase-A order; graph[1] writes
// the velocity buffer on its second step, graph[0] does not.
static int batch_capture(LbmBatchDev *b, const LbmBatchParams *p) {
    batch_free_graphs(b);
    for (int k = 0; k < 2; k++) {
        hipGraph_t g = NULL;
        HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        const int saved = b->state;
        batch_step_any(b, p, 0);
        batch_step_any(b, p, k);
        b->state = saved;
        HIPCHECK(hipStreamEndCapture(b->stream, &g), "graph capture");
        hipError_t e = hipGraphInstantiate(&b->graph[k], g, NULL, NULL, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return set_err("graph instantiate", e);
    }
    b->graph_ok = 1;
    b->graph_params = *p;
    b->graph_bs = b->bs;
    b->graph_smode = b->smode;
    b->graph_remap = b->remap;
    return 1;
}

extern "C" int lbmbatch_step(LbmBatchDev *b, const LbmBatchParams *p, int nsteps, int use_graph) {
    if (b->state < 0) {
        snprintf(g_err, sizeof(g_err), "batch stepped before initialisation");
        return 0;
    }
    while (nsteps > 0 && b->state != ST_A) {  // bring the layout to phase-A order
        batch_step_any(b, p, nsteps == 1);
        nsteps--;
    }
    if (use_graph && nsteps >= 2) {
        if (!b->graph_ok || !same_params(&b->graph_params, p) || b->graph_bs != b->bs ||
            b->graph_smode != b->smode || b->graph_remap != b->remap)
            if (!batch_capture(b, p)) return 0;
        const int pairs = nsteps / 2, rem = nsteps % 2;
        for (int k = 0; k < pairs; k++) {
            const int last = (k == pairs - 1) && rem == 0;
            HIPCHECK(hipGraphLaunch(b->graph[last ? 1 : 0], b->stream), "graph launch");
        }
        nsteps = rem;  // state is phase-A order again
    }
    for (int s = 0; s < nsteps; s++) batch_step_any(b, p, s == nsteps - 1);
    HIPCHECK(hipGetLastError(), "batch step launch");
    return 1;
}

extern "C" int lbmbatch_set_block_size(LbmBatchDev *b, int bs) {
    if (bs != 128 && bs != 256 && bs != 512) return 0;
    if (bs == b->bs) return 1;
    HIPCHECK(hipStreamSynchronize(b->stream), "synchronize");
    return batch_layout_blocks(b, bs);
}

extern "C" int lbmbatch_block_size(const LbmBatchDev *b) { return b->bs; }

extern "C" int lbmbatch_set_streaming(LbmBatchDev *b, int mode) {
    if (mode < 0 || mode > (LBMDEV_NT_A | LBMDEV_NT_B)) return 0;
    b->smode = b->fmt == LBMDEV_FP64 ? 0 : mode;
    return 1;
}

extern "C" int lbmbatch_streaming(const LbmBatchDev *b) { return b->smode; }

extern "C" void lbmbatch_set_block_order(LbmBatchDev *b, int xcd_contiguous) { b->remap = xcd_contiguous != 0; }

extern "C" int lbmbatch_block_order(const LbmBatchDev *b) { return b->remap; }

extern "C" int lbmbatch_profile(LbmBatchDev *b, const LbmBatchParams *p, int nsteps, int use_graph, float *ms) {
    hipEvent_t t0, t1;
    HIPCHECK(hipEventCreate(&t0), "event");
    HIPCHECK(hipEventCreate(&t1), "event");
    HIPCHECK(hipStreamSynchronize(b->stream), "synchronize");
    hipEventRecord(t0, b->stream);
    int ok = lbmbatch_step(b, p, nsteps, use_graph);
    hipEventRecord(t1, b->stream);
    hipError_t e = hipEventSynchronize(t1);
    float v = 0;
    hipEventElapsedTime(&v, t0, t1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    if (e != hipSuccess) return set_err("batch profile", e);
    *ms = v;
    return ok;
}

// Every tunnel's force in one launch: block B works as block B - first of its
// tunnel's launch of the per-tunnel path (the same lanes, stride and block
// sums, so the per-tunnel results are bitwise those of that path), and
// k_finalize_batch reduces each tunnel's block sums in block order.
template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_force_batch(ForceArgs<FMT> base, const TunnelDesc *desc,
                                                       const int4 *ftun, const int32_t *fblk_tunnel) {
    const int t = fblk_tunnel[blockIdx.x];
    const int4 tt = ftun[t];
    const TunnelDesc d = desc[t];
    ForceArgs<FMT> a = base;
    a.pop = base.pop + d.popOff;
    a.mask = base.mask + d.cellOff;
    a.ovr_cells = base.ovr_cells + tt.x;
    a.ovr_q = base.ovr_q + tt.x;
    a.vel = base.vel + d.cellOff;
    a.n_ovr = tt.y;
    a.g = d.g;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    long long fix[7] = {0, 0, 0, 0, 0, 0, 0};
    force_block<FMT, STATE>(a, (int)blockIdx.x - tt.z, tt.w, acc, fix);
    block_reduce_store<6, 7>(acc, fix, base.pd, base.pi);
}

__global__ __launch_bounds__(64) void k_finalize_batch(const double *pd, const long long *pi, const int4 *ftun,
                                                       double *outd, long long *outi) {
    const int t = blockIdx.x;
    const int4 tt = ftun[t];
    finalize_body<6, 7>(pd + (long long)tt.z * 6, pi + (long long)tt.z * 7, tt.w, outd + 6 * t, outi + 7 * t);
}

template <int FMT, int STATE>
static void batch_force_t(LbmBatchDev *b, int periodic) {
    if (b->fblocks > 0 && b->fblocks <= b->nparts) {
        ForceArgs<FMT> a;
        a.pop = (const typename Pop<FMT>::S *)b->pop;
        a.mask = b->mask;
        a.ovr_cells = b->ovr_cells;
        a.ovr_q = b->ovr_q;
        a.vel = b->vel;
        a.n_ovr = 0;
        a.q_stride = b->n_ovr;
        a.periodic = periodic;
        a.g = b->desc_h[0].g;
        a.pd = b->pd;
        a.pi = b->pi;
        hipLaunchKernelGGL((k_force_batch<FMT, STATE>), dim3(b->fblocks), dim3(BLOCK), 0, b->stream, a,
                           (const TunnelDesc *)b->desc_d, (const int4 *)b->ftun, (const int32_t *)b->fblk_tunnel);
        hipLaunchKernelGGL(k_finalize_batch, dim3(b->nt), dim3(64), 0, b->stream, b->pd, b->pi,
                           (const int4 *)b->ftun, b->resd, b->resi);
        return;
    }
    for (int t = 0; t < b->nt; t++) {
        const int k0 = b->kstart[t], n = b->kstart[t + 1] - k0;
        const TunnelDesc &d = b->desc_h[t];
        int nb = (n + BLOCK - 1) / BLOCK;
        if (nb > FORCE_BLOCKS) nb = FORCE_BLOCKS;
        if (nb < 1) nb = 1;
        ForceArgs<FMT> a;
        a.pop = (const typename Pop<FMT>::S *)b->pop + d.popOff;
        a.mask = b->mask + d.cellOff;
        a.ovr_cells = b->ovr_cells + k0;
        a.ovr_q = b->ovr_q + k0;
        a.vel = b->vel + d.cellOff;
        a.n_ovr = n;
        a.q_stride = b->n_ovr;
        a.periodic = periodic;
        a.g = d.g;
        a.pd = b->pd;
        a.pi = b->pi;
        hipLaunchKernelGGL((k_force<FMT, STATE>), dim3(nb), dim3(BLOCK), 0, b->stream, a);
        hipLaunchKernelGGL((k_finalize<6, 7>), dim3(1), dim3(64), 0, b->stream, b->pd, b->pi, nb, b->resd + 6 * t,
                           b->resi + 7 * t);
    }
}

template <int FMT>
static void batch_force_state(LbmBatchDev *b, int periodic) {
    if (b->state == ST_INIT) batch_force_t<FMT, ST_INIT>(b, periodic);
    else if (b->state == ST_A) batch_force_t<FMT, ST_A>(b, periodic);
    else batch_force_t<FMT, ST_B>(b, periodic);
}

extern "C" int lbmbatch_force(LbmBatchDev *b, int periodic, double *exact, long long *fixed) {
    if (b->state < 0) return 0;
    switch (b->fmt) {
    case LBMDEV_FP32: batch_force_state<LBMDEV_FP32>(b, periodic); break;
    case LBMDEV_FP16S: batch_force_state<LBMDEV_FP16S>(b, periodic); break;
    case LBMDEV_BF16S: batch_force_state<LBMDEV_BF16S>(b, periodic); break;
    case LBMDEV_FP64: batch_force_state<LBMDEV_FP64>(b, periodic); break;
    }
    HIPCHECK(hipGetLastError(), "batch force launch");
    HIPCHECK(hipMemcpyAsync(exact, b->resd, sizeof(double) * 6 * (size_t)b->nt, hipMemcpyDeviceToHost, b->stream),
             "batch force readback");
    HIPCHECK(hipMemcpyAsync(fixed, b->resi, sizeof(long long) * 7 * (size_t)b->nt, hipMemcpyDeviceToHost,
                            b->stream),
             "batch force readback");
    HIPCHECK(hipStreamSynchronize(b->stream), "batch force");
    return 1;
}

template <int FMT>
static void batch_export_t(LbmBatchDev *b, int t, int periodic, void *os, void *oc) {
    const TunnelDesc &d = b->desc_h[t];
    ExportArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)b->pop + d.popOff;
    a.mask = b->mask + d.cellOff;
    a.ovr_index = b->ovr_index + d.cellOff;
    a.ovr_q = b->ovr_q;
    a.n_ovr = b->n_ovr;
    a.periodic = periodic;
    a.g = d.g;
    a.out_stream = os;
    a.out_collision = oc;
    a.as_double = 1;
    a.pd = NULL;
    a.pi = NULL;
    const int nb = (d.g.ncell + BLOCK - 1) / BLOCK;
    if (b->state == ST_INIT)
        hipLaunchKernelGGL((k_export<FMT, ST_INIT>), dim3(nb), dim3(BLOCK), 0, b->stream, a);
    else if (b->state == ST_A)
        hipLaunchKernelGGL((k_export<FMT, ST_A>), dim3(nb), dim3(BLOCK), 0, b->stream, a);
    else
        hipLaunchKernelGGL((k_export<FMT, ST_B>), dim3(nb), dim3(BLOCK), 0, b->stream, a);
}

extern "C" int lbmbatch_export(LbmBatchDev *b, int t, int periodic, double *f_ps, double *f_pc) {
    if (b->state < 0 || t < 0 || t >= b->nt) return 0;
    const size_t bytes = (size_t)b->desc_h[t].g.ncell * 19 * 8;
    void *dps = NULL, *dpc = NULL;
    HIPCHECK(hipMalloc(&dps, bytes), "export buffer");
    hipError_t e = hipMalloc(&dpc, bytes);
    if (e != hipSuccess) {
        hipFree(dps);
        return set_err("export buffer", e);
    }
    switch (b->fmt) {
    case LBMDEV_FP32: batch_export_t<LBMDEV_FP32>(b, t, periodic, dps, dpc); break;
    case LBMDEV_FP16S: batch_export_t<LBMDEV_FP16S>(b, t, periodic, dps, dpc); break;
    case LBMDEV_BF16S: batch_export_t<LBMDEV_BF16S>(b, t, periodic, dps, dpc); break;
    case LBMDEV_FP64: batch_export_t<LBMDEV_FP64>(b, t, periodic, dps, dpc); break;
    }
    e = hipGetLastError();
    if (e == hipSuccess && f_ps) e = hipMemcpyAsync(f_ps, dps, bytes, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess && f_pc) e = hipMemcpyAsync(f_pc, dpc, bytes, hipMemcpyDeviceToHost, b->stream);
    hipError_t e2 = hipStreamSynchronize(b->stream);
    hipFree(dps);
    hipFree(dpc);
    if (e != hipSuccess) return set_err("batch export", e);
    if (e2 != hipSuccess) return set_err("batch export", e2);
    return 1;
}

extern "C" int lbmbatch_copy_velocity(LbmBatchDev *b, int t, float *host) {
    if (t < 0 || t >= b->nt) return 0;
    const TunnelDesc &d = b->desc_h[t];
    HIPCHECK(hipMemcpyAsync(host, b->vel + d.cellOff, (size_t)d.g.ncell * 16, hipMemcpyDeviceToHost, b->stream),
             "velocity readback");
    HIPCHECK(hipStreamSynchronize(b->stream), "velocity readback");
    return 1;
}

extern "C" int lbmbatch_copy_velocity_coarse(LbmBatchDev *b, int t, int f, float *host) {
    if (t < 0 || t >= b->nt) return 0;
    const TunnelDesc &d = b->desc_h[t];
    return coarsen_to_host(b->vel + d.cellOff, d.g.nx, d.g.ny, d.g.nz, f, host, b->stream);
}

extern "C" int lbmbatch_synchronize(LbmBatchDev *b) {
    HIPCHECK(hipStreamSynchronize(b->stream), "synchronize");
    return 1;
}

// Single-lattice workgroup size (autotuned by the host).
extern "C" int lbmdev_set_block_size(LbmDev *d, int bs) {
    if (bs != 128 && bs != 256 && bs != 512) return 0;
    d->bs = bs;
    return 1;
}

extern "C" int lbmdev_block_size(const LbmDev *d) { return d->bs; }

extern "C" int lbmdev_set_streaming(LbmDev *d, int mode) {
    if (mode < 0 || mode > (LBMDEV_NT_A | LBMDEV_NT_B)) return 0;
    d->nt = d->fmt == LBMDEV_FP64 ? 0 : mode;
    return 1;
}

extern "C" int lbmdev_streaming(const LbmDev *d) { return d->nt; }

extern "C" void lbmdev_set_block_order(LbmDev *d, int xcd_contiguous) { d->remap = xcd_contiguous != 0; }

extern "C" int lbmdev_block_order(const LbmDev *d) { return d->remap; }

// ---------------------------------------------------------------- autotuner
// Candidates are timed on the real lattice; the stored state (populations,
// velocity, layout) is copied aside first and put back after every trial.
struct Snapshot {
    void *pop;
    void *vel;
    size_t popBytes, velBytes;
};

static int snapshot_take(Snapshot &s, const void *pop, size_t pb, const void *vel, size_t vb, hipStream_t st) {
    s.popBytes = pb;
    s.velBytes = vb;
    s.pop = s.vel = NULL;
    HIPCHECK(hipMalloc(&s.pop, pb), "autotune snapshot");
    hipError_t e = hipMalloc(&s.vel, vb);
    if (e != hipSuccess) {
        hipFree(s.pop);
        return set_err("autotune snapshot", e);
    }
    HIPCHECK(hipMemcpyAsync(s.pop, pop, pb, hipMemcpyDeviceToDevice, st), "autotune snapshot");
    HIPCHECK(hipMemcpyAsync(s.vel, vel, vb, hipMemcpyDeviceToDevice, st), "autotune snapshot");
    return 1;
}

static int snapshot_put(const Snapshot &s, void *pop, void *vel, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(pop, s.pop, s.popBytes, hipMemcpyDeviceToDevice, st), "autotune restore");
    HIPCHECK(hipMemcpyAsync(vel, s.vel, s.velBytes, hipMemcpyDeviceToDevice, st), "autotune restore");
    HIPCHECK(hipStreamSynchronize(st), "autotune restore");
    return 1;
}

// The guard counter around trials that do not advance the run.
static int guard_read(const unsigned long long *g, unsigned long long *v, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(v, g, 8, hipMemcpyDeviceToHost, st), "guard counter");
    HIPCHECK(hipStreamSynchronize(st), "guard counter");
    return 1;
}
static int guard_write(unsigned long long *g, unsigned long long v, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(g, &v, 8, hipMemcpyHostToDevice, st), "guard counter");
    HIPCHECK(hipStreamSynchronize(st), "guard counter");
    return 1;
}

extern "C" int lbmdev_guard_resets(LbmDev *d, long long *count) {
    unsigned long long v = 0;
    if (!guard_read(d->guard, &v, d->stream)) return 0;
    *count = (long long)v;
    return 1;
}

extern "C" int lbmbatch_guard_resets(LbmBatchDev *b, long long *count) {
    unsigned long long v = 0;
    if (!guard_read(b->guard, &v, b->stream)) return 0;
    *count = (long long)v;
    return 1;
}

extern "C" int lbmdev_autotune(LbmDev *d, const LbmStepParams *p, int steps, float *ms_per_step) {
    if (d->state < 0 || steps <= 0) return 0;
    unsigned long long guard0 = 0;
    if (!guard_read(d->guard, &guard0, d->stream)) return 0;
    Snapshot s;
    const size_t pb = (size_t)19 * d->stride * d->elem, vb = (size_t)d->ncell * 16;
    if (!snapshot_take(s, d->pop, pb, d->vel, vb, d->stream)) return 0;
    const int state0 = d->state, bs0 = d->bs, nt0 = d->nt, remap0 = d->remap;
    int best = 0, bestNt = 0, bestRemap = 1;
    float bestMs = 0;
    int ok = 1;
    // every workgroup size with each streaming mode (FP64: default access
    // only), in XCD-contiguous and in plain workgroup order
    const int nmode = d->fmt == LBMDEV_FP64 ? 1 : 3;
    for (int c = 0; c < 6 * nmode && ok; c++) {
        d->bs = kBlockSizes[c % 3];
        d->nt = kStreamModes[(c / 3) % nmode];
        d->remap = c < 3 * nmode;
        float tot = 0;
        ok = dispatch_steps(d, p, 2, NULL) &&  // warm-up
             lbmdev_profile_steps(d, p, steps, &tot, NULL, NULL);
        const float ms = tot / steps;
        if (ms_per_step && (c < 3 || ms < ms_per_step[c % 3])) ms_per_step[c % 3] = ms;
        if (ok && (best == 0 || ms < bestMs)) {
            best = d->bs;
            bestNt = d->nt;
            bestRemap = d->remap;
            bestMs = ms;
        }
        ok = ok && snapshot_put(s, d->pop, d->vel, d->stream);
        d->state = state0;
    }
    hipFree(s.pop);
    hipFree(s.vel);
    d->bs = ok ? best : bs0;
    d->nt = ok ? bestNt : nt0;
    d->remap = ok ? bestRemap : remap0;
    if (!guard_write(d->guard, guard0, d->stream)) return 0;
    return ok ? best : 0;
}

extern "C" int lbmbatch_autotune(LbmBatchDev *b, const LbmBatchParams *p, int steps, float *ms_per_step) {
    if (b->state < 0 || steps <= 0) return 0;
    unsigned long long guard0 = 0;
    if (!guard_read(b->guard, &guard0, b->stream)) return 0;
    Snapshot s;
    const size_t pb = (size_t)b->total_pop * b->elem, vb = (size_t)b->total_cells * 16;
    if (!snapshot_take(s, b->pop, pb, b->vel, vb, b->stream)) return 0;
    const int state0 = b->state, bs0 = b->bs, smode0 = b->smode, remap0 = b->remap;
    int best = 0, bestMode = 0, bestRemap = 1;
    float bestMs = 0;
    int ok = 1;
    // every workgroup size with each streaming mode (FP64: default access
    // only), in XCD-contiguous and in plain workgroup order
    const int nmode = b->fmt == LBMDEV_FP64 ? 1 : 3;
    for (int c = 0; c < 6 * nmode && ok; c++) {
        ok = lbmbatch_set_block_size(b, kBlockSizes[c % 3]);
        b->smode = kStreamModes[(c / 3) % nmode];
        b->remap = c < 3 * nmode;
        float ms = 0;
        ok = ok && lbmbatch_step(b, p, 2, 1) && lbmbatch_profile(b, p, steps, 1, &ms);
        ms /= steps;
        if (ms_per_step && (c < 3 || ms < ms_per_step[c % 3])) ms_per_step[c % 3] = ms;
        if (ok && (best == 0 || ms < bestMs)) {
            best = kBlockSizes[c % 3];
            bestMode = b->smode;
            bestRemap = b->remap;
            bestMs = ms;
        }
        ok = ok && snapshot_put(s, b->pop, b->vel, b->stream);
        b->state = state0;
    }
    hipFree(s.pop);
    hipFree(s.vel);
    b->smode = ok ? bestMode : smode0;
    b->remap = ok ? bestRemap : remap0;
    if (!guard_write(b->guard, guard0, b->stream)) return 0;
    if (!lbmbatch_set_block_size(b, ok ? best : bs0)) return 0;
    return ok ? best : 0;
}

// ------------------------------------------------------ copy bandwidth probe
// A plain streaming copy used to measure the attainable HBM bandwidth and
// to calibrate byte counters: one 16-byte vector per lane, one load and one
// store, no grid-stride loop.  Of the copy shapes measured on MI355X this is
// the fastest (6.1-6.2 TB/s on 2.5 GB; four vectors per lane reach 5.55,
// 19 separate planes of 4-byte elements 5.5).
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int lbmdev_copy_bandwidth(size_t bytes, int iters, float *gbps, float *ms_per_copy) {
    const long long n = (long long)(bytes / 16);
    float4 *a = NULL, *b = NULL;
    HIPCHECK(hipMalloc((void **)&a, (size_t)n * 16), "copy probe buffer");
    hipError_t e = hipMalloc((void **)&b, (size_t)n * 16);
    if (e != hipSuccess) {
        hipFree(a);
        return set_err("copy probe buffer", e);
    }
    hipMemset(a, 0, (size_t)n * 16);
    hipMemset(b, 0, (size_t)n * 16);
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipEvent_t t0, t1;
    hipEventCreate(&t0);
    hipEventCreate(&t1);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);  // warm-up
    hipEventRecord(t0, 0);
    for (int k = 0; k < iters; k++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n);
    hipEventRecord(t1, 0);
    e = hipEventSynchronize(t1);
    float ms = 0;
    hipEventElapsedTime(&ms, t0, t1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return set_err("copy probe", e);
    const float per = ms / iters;
    if (ms_per_copy) *ms_per_copy = per;
    if (gbps) *gbps = (float)(2.0 * (double)n * 16 / (per * 1e-3) / 1e9);
    return 1;
}
