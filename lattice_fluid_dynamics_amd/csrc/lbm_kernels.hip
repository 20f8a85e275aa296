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
    typedef __half S;
    typedef float T;
    static __device__ __forceinline__ T dec(S v, int i) {
        return fmaf(__half2float(v), 1.0f / 32768.0f, wgt<float>(i));
    }
    static __device__ __forceinline__ S enc(T v, int i) {
        return __float2half_rn((v - wgt<float>(i)) * 32768.0f);
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
};

// Each of the 8 XCDs receives workgroups b, b+8, b+16, ... in dispatch order;
// this gives every XCD one contiguous range of cells so neighbouring rows and
// partially shared cache lines stay in one L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    int q = n >> 3, r = n & 7, k = b & 7, j = b >> 3;
    return k * q + (k < r ? k : r) + j;
}

__device__ __forceinline__ void decompose(const Geo &g, int cell, int &x, int &y, int &z) {
    x = cell % g.nx;
    int t = cell / g.nx;
    y = t % g.ny;
    z = t / g.ny;
}

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

template <typename T>
__device__ __forceinline__ T tsqrt(T v);
template <>
__device__ __forceinline__ float tsqrt<float>(float v) { return sqrtf(v); }
template <>
__device__ __forceinline__ double tsqrt<double>(double v) { return sqrt(v); }

struct CollideParams {
    float tau, cs2x18, inx, iny, inz;
    int smag;
};

// lbm_collide.comp:313-626 for one non-body cell: Zou-He, moments, guard,
// Smagorinsky, then regularized / MRT / BGK.  f is updated by the Zou-He
// rewrite; fo receives the post-collision populations; returns the vec4
// written to the velocity buffer.
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
    rho = T(0);
#pragma unroll
    for (int i = 0; i < 19; i++) rho += f[i];
    T jx = (f[1] - f[2]) + (f[7] - f[8]) + (f[9] - f[10]) + (f[11] - f[12]) + (f[13] - f[14]);
    T jy = (f[3] - f[4]) + (f[7] - f[9]) + (f[8] - f[10]) + (f[15] - f[16]) + (f[17] - f[18]);
    T jz = (f[5] - f[6]) + (f[11] - f[13]) + (f[12] - f[14]) + (f[15] - f[17]) + (f[16] - f[18]);
    if (rho > T(0.0001)) {
        T ir = T(1) / rho;
        ux = jx * ir;
        uy = jy * ir;
        uz = jz * ir;
    } else {
        ux = jx;
        uy = jy;
        uz = jz;
    }
    // stability guard (lbm_collide.comp:429-440): reset to feq(1, 0) = w
    T umag = tsqrt<T>(ux * ux + uy * uy + uz * uz);
    if (rho < T(0.3) || rho > T(2.0) || umag > T(0.40) || isinf(rho) || isnan(rho)) {
#pragma unroll
        for (int i = 0; i < 19; i++) fo[i] = wgt<T>(i);
        rho = T(1);
        ux = uy = uz = T(0);
        return;
    }
    const T u2 = ux * ux + uy * uy + uz * uz;
    const T c = T(1) - T(1.5) * u2;
    T tau_eff = (T)p.tau;
    T Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
    if (p.smag || OP == OP_REG) {
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
        tau_eff = T(0.5) * (tau + tsqrt<T>(tau * tau + (T)p.cs2x18 * pn / (rho + T(1e-10))));
    }
    const T omega = T(1) / tau_eff;
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
        T m[19];
        T sf = f[1] + f[2] + f[3] + f[4] + f[5] + f[6];
        T se = f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
        m[0] = f[0] + sf + se;
        m[1] = T(-30) * f[0] - T(11) * sf + T(8) * se;
        m[2] = T(12) * f[0] - T(4) * sf + se;
        T ax = (f[7] - f[8]) + (f[9] - f[10]) + (f[11] - f[12]) + (f[13] - f[14]);
        T ay = (f[7] + f[8]) - (f[9] + f[10]) + (f[15] - f[16]) + (f[17] - f[18]);
        T az = (f[11] + f[12]) - (f[13] + f[14]) + (f[15] + f[16]) - (f[17] + f[18]);
        m[3] = (f[1] - f[2]) + ax;
        m[4] = T(-4) * (f[1] - f[2]) + ax;
        m[5] = (f[3] - f[4]) + ay;
        m[6] = T(-4) * (f[3] - f[4]) + ay;
        m[7] = (f[5] - f[6]) + az;
        m[8] = T(-4) * (f[5] - f[6]) + az;
        T sXY = f[7] + f[8] + f[9] + f[10];
        T sXZ = f[11] + f[12] + f[13] + f[14];
        T sYZ = f[15] + f[16] + f[17] + f[18];
        T fx2 = f[1] + f[2], fy2 = f[3] + f[4], fz2 = f[5] + f[6];
        m[9] = T(2) * fx2 - fy2 - fz2 + sXY + sXZ - T(2) * sYZ;
        m[10] = T(-4) * fx2 + T(2) * fy2 + T(2) * fz2 + sXY + sXZ - T(2) * sYZ;
        m[11] = fy2 - fz2 + sXY - sXZ;
        m[12] = T(-2) * fy2 + T(2) * fz2 + sXY - sXZ;
        m[13] = f[7] - f[8] - f[9] + f[10];
        m[14] = f[15] - f[16] - f[17] + f[18];
        m[15] = f[11] - f[12] - f[13] + f[14];
        m[16] = f[7] - f[8] + f[9] - f[10] - f[11] + f[12] - f[13] + f[14];
        m[17] = -f[7] - f[8] + f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
        m[18] = f[11] + f[12] - f[13] - f[14] - f[15] - f[16] + f[17] + f[18];
        // equilibrium moments (lbm_collide.comp:560-580)
        T me[19];
        me[0] = rho;
        me[1] = T(-11) * rho + T(19) * rho * u2;
        me[2] = T(3) * rho - T(5.5) * rho * u2;
        me[3] = rho * ux;
        me[4] = T(-2) / T(3) * rho * ux;
        me[5] = rho * uy;
        me[6] = T(-2) / T(3) * rho * uy;
        me[7] = rho * uz;
        me[8] = T(-2) / T(3) * rho * uz;
        me[9] = rho * (T(2) * ux * ux - uy * uy - uz * uz);
        me[10] = T(-0.5) * rho * (T(2) * ux * ux - uy * uy - uz * uz);
        me[11] = rho * (uy * uy - uz * uz);
        me[12] = T(-0.5) * rho * (uy * uy - uz * uz);
        me[13] = rho * ux * uy;
        me[14] = rho * uy * uz;
        me[15] = rho * ux * uz;
        me[16] = me[17] = me[18] = T(0);
        // relaxation rates (lbm_collide.comp:585-604), then scaling by
        // 1/|row|^2 of the inverse (lbm_collide.comp:188-206)
        const T s[19] = {T(0),   T(1.19), T(1.4), T(0),   T(1.2), T(0),   T(1.2), T(0),    T(1.2),   omega,
                         T(1.4), omega,   T(1.4), omega,  omega,  omega,  T(1.98), T(1.98), T(1.98)};
        const T inrm[19] = {T(1) / T(19), T(1) / T(2394), T(1) / T(252), T(1) / T(10), T(1) / T(40),
                            T(1) / T(10), T(1) / T(40),   T(1) / T(10),  T(1) / T(40), T(1) / T(36),
                            T(1) / T(72), T(1) / T(12),   T(1) / T(24),  T(0.25),      T(0.25),
                            T(0.25),      T(0.125),       T(0.125),      T(0.125)};
        T ms[19];
#pragma unroll
        for (int k = 0; k < 19; k++) ms[k] = (m[k] - s[k] * (m[k] - me[k])) * inrm[k];
        const T b0 = ms[0] - T(11) * ms[1] - T(4) * ms[2];
        const T b1 = ms[0] + T(8) * ms[1] + ms[2];
        fo[0] = ms[0] - T(30) * ms[1] + T(12) * ms[2];
        fo[1] = b0 + ms[3] - T(4) * ms[4] + T(2) * ms[9] - T(4) * ms[10];
        fo[2] = b0 - ms[3] + T(4) * ms[4] + T(2) * ms[9] - T(4) * ms[10];
        fo[3] = b0 + ms[5] - T(4) * ms[6] - ms[9] + T(2) * ms[10] + ms[11] - T(2) * ms[12];
        fo[4] = b0 - ms[5] + T(4) * ms[6] - ms[9] + T(2) * ms[10] + ms[11] - T(2) * ms[12];
        fo[5] = b0 + ms[7] - T(4) * ms[8] - ms[9] + T(2) * ms[10] - ms[11] + T(2) * ms[12];
        fo[6] = b0 - ms[7] + T(4) * ms[8] - ms[9] + T(2) * ms[10] - ms[11] + T(2) * ms[12];
        const T p910 = ms[9] + ms[10], p1112 = ms[11] + ms[12];
        fo[7] = b1 + ms[3] + ms[4] + ms[5] + ms[6] + p910 + p1112 + ms[13] + ms[16] - ms[17];
        fo[8] = b1 - ms[3] - ms[4] + ms[5] + ms[6] + p910 + p1112 - ms[13] - ms[16] - ms[17];
        fo[9] = b1 + ms[3] + ms[4] - ms[5] - ms[6] + p910 + p1112 - ms[13] + ms[16] + ms[17];
        fo[10] = b1 - ms[3] - ms[4] - ms[5] - ms[6] + p910 + p1112 + ms[13] - ms[16] + ms[17];
        fo[11] = b1 + ms[3] + ms[4] + ms[7] + ms[8] + p910 - p1112 + ms[15] - ms[16] + ms[18];
        fo[12] = b1 - ms[3] - ms[4] + ms[7] + ms[8] + p910 - p1112 - ms[15] + ms[16] + ms[18];
        fo[13] = b1 + ms[3] + ms[4] - ms[7] - ms[8] + p910 - p1112 - ms[15] - ms[16] - ms[18];
        fo[14] = b1 - ms[3] - ms[4] - ms[7] - ms[8] + p910 - p1112 + ms[15] + ms[16] - ms[18];
        fo[15] = b1 + ms[5] + ms[6] + ms[7] + ms[8] - T(2) * p910 + ms[14] + ms[17] - ms[18];
        fo[16] = b1 - ms[5] - ms[6] + ms[7] + ms[8] - T(2) * p910 - ms[14] - ms[17] - ms[18];
        fo[17] = b1 + ms[5] + ms[6] - ms[7] - ms[8] - T(2) * p910 - ms[14] + ms[17] + ms[18];
        fo[18] = b1 - ms[5] - ms[6] - ms[7] - ms[8] - T(2) * p910 + ms[14] - ms[17] + ms[18];
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
    const typename Pop<FMT>::T *ovr_vals;
    int ovr_stride;
    int use_ovr;
    float4 *vel;
    int write_vel;
    Geo g;
    CollideParams cp;
};

// The fused stream + collide kernel.  PHASE 0 = A, 1 = B (see file header).
template <int FMT, int OP, int PHASE>
__global__ __launch_bounds__(BLOCK) void k_step(StepArgs<FMT> a) {
    typedef Pop<FMT> P;
    typedef typename P::T T;
    const Geo &g = a.g;
    const int cell = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (cell >= g.ncell) return;
    int x, y, z;
    decompose(g, cell, x, y, z);
    // torus neighbour offsets
    const int X[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
    const int pl = g.nx * g.ny;
    const int Y[3] = {(y == 0 ? g.ny - 1 : y - 1) * g.nx, y * g.nx, (y == g.ny - 1 ? 0 : y + 1) * g.nx};
    const int Z[3] = {(z == 0 ? g.nz - 1 : z - 1) * pl, z * pl, (z == g.nz - 1 ? 0 : z + 1) * pl};

    const uint8_t m = a.mask[cell];
    const bool ovr = a.use_ovr && (m & LBMDEV_MASK_OVR);
    T f[19];
    if (ovr) {
        const int k = a.ovr_index[cell];
#pragma unroll
        for (int i = 0; i < 19; i++) f[i] = a.ovr_vals[(long long)i * a.ovr_stride + k];
    } else if (PHASE == 0) {
#pragma unroll
        for (int i = 0; i < 19; i++) f[i] = P::dec(a.pop[(long long)i * g.stride + cell], i);
    } else {
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const int j = OPP[i];
            const int src = X[1 - EX[i]] + Y[1 - EY[i]] + Z[1 - EZ[i]];
            f[i] = P::dec(a.pop[(long long)j * g.stride + src], j);
        }
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
    if (a.write_vel) a.vel[cell] = make_float4((float)ux, (float)uy, (float)uz, (float)rho);

    if (PHASE == 0) {
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const int j = OPP[i];
            a.pop[(long long)j * g.stride + cell] = P::enc(fo[i], j);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 19; i++) {
            const int dst = X[1 + EX[i]] + Y[1 + EY[i]] + Z[1 + EZ[i]];
            a.pop[(long long)i * g.stride + dst] = P::enc(fo[i], i);
        }
    }
}

template <int FMT>
struct BoundArgs {
    const typename Pop<FMT>::S *pop;
    const int32_t *ovr_cells;
    const float *ovr_q;
    typename Pop<FMT>::T *ovr_vals;
    int n_ovr;
    int periodic;
    Geo g;
};

// Boundary prepass: the incoming populations of every listed cell under the
// reference stream rule, evaluated on the stored post-collision field.
template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_prepass(BoundArgs<FMT> a) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= a.n_ovr) return;
    int x, y, z;
    decompose(a.g, a.ovr_cells[k], x, y, z);
#pragma unroll
    for (int i = 0; i < 19; i++) {
        const float qo = a.ovr_q[(long long)OPP[i] * a.n_ovr + k];
        a.ovr_vals[(long long)i * a.n_ovr + k] =
            stream_value<FMT, STATE>(a.pop, a.g, a.periodic, x, y, z, i, qo);
    }
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
template <int FMT, int STATE>
__global__ __launch_bounds__(BLOCK) void k_force(ForceArgs<FMT> a) {
    typedef typename Pop<FMT>::T T;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    long long fix[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = blockIdx.x * BLOCK + threadIdx.x; k < a.n_ovr; k += gridDim.x * BLOCK) {
        const int cell = a.ovr_cells[k];
        if (a.mask[cell] & (LBMDEV_MASK_BODY | LBMDEV_MASK_GROUND)) continue;
        float q[19];
#pragma unroll
        for (int i = 1; i < 19; i++) q[i] = a.ovr_q[(long long)i * a.n_ovr + k];
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
    block_reduce_store<6, 7>(acc, fix, a.pd, a.pi);
}

// Sums per-block partials in block order (deterministic).
template <int NV, int NI>
__global__ __launch_bounds__(64) void k_finalize(const double *pd, const long long *pi, int nblocks,
                                                 double *outd, long long *outi) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < NV; k++) {
        double s = 0;
        for (int b = 0; b < nblocks; b++) s += pd[(long long)b * NV + k];
        outd[k] = s;
    }
    for (int k = 0; k < NI; k++) {
        long long s = 0;
        for (int b = 0; b < nblocks; b++) s += pi[(long long)b * NI + k];
        outi[k] = s;
    }
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
    int state;  // -1 uninitialised, ST_INIT / ST_A / ST_B
    hipStream_t stream;
    double *pd;
    long long *pi;
    double *resd;
    long long *resi;
    float *feq_d;
    int nparts;  // partial slots allocated
};

extern "C" const char *lbmdev_last_error(void) { return g_err; }

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
    d->ncell = (int)nc;
    d->fmt = storage;
    d->elem = elem_size(storage);
    d->telem = storage == LBMDEV_FP64 ? 8 : 4;
    d->stride = (nc + 63) / 64 * 64;
    d->state = -1;
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
#undef ALLOC
    if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) {
        snprintf(err, errlen, "stream creation failed: %s", hipGetErrorString(e));
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
    hipFree(d->pd);
    hipFree(d->pi);
    hipFree(d->resd);
    hipFree(d->resi);
    hipFree(d->feq_d);
    if (d->stream) hipStreamDestroy(d->stream);
    free(d);
}

extern "C" int lbmdev_storage(const LbmDev *d) { return d->fmt; }

extern "C" int lbmdev_set_geometry(LbmDev *d, const int32_t *solid, int n_ovr, const int32_t *ovr_cells,
                                   const float *ovr_q) {
    const int nc = d->ncell;
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    uint8_t *mask = (uint8_t *)malloc((size_t)nc);
    int32_t *index = (int32_t *)malloc((size_t)nc * 4);
    if (!mask || !index) {
        free(mask);
        free(index);
        snprintf(g_err, sizeof(g_err), "host allocation failed");
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
            snprintf(g_err, sizeof(g_err), "device allocation of the boundary list failed");
            return 0;
        }
        d->ovr_cap = cap;
    }
    d->n_ovr = n_ovr;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(d->mask, mask, (size_t)nc, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->ovr_index, index, (size_t)nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->solid, solid, (size_t)nc * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_ovr > 0) e = hipMemcpy(d->ovr_cells, ovr_cells, (size_t)n_ovr * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_ovr > 0)
        e = hipMemcpy(d->ovr_q, ovr_q, (size_t)n_ovr * 19 * 4, hipMemcpyHostToDevice);
    free(mask);
    free(index);
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
static void launch_prepass(LbmDev *d, int periodic) {
    if (d->n_ovr == 0) return;
    BoundArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)d->pop;
    a.ovr_cells = d->ovr_cells;
    a.ovr_q = d->ovr_q;
    a.ovr_vals = (typename Pop<FMT>::T *)d->ovr_vals;
    a.n_ovr = d->n_ovr;
    a.periodic = periodic;
    a.g = Geo{d->nx, d->ny, d->nz, d->ncell, d->stride};
    const int nb = (d->n_ovr + BLOCK - 1) / BLOCK;
    if (d->state == ST_A)
        hipLaunchKernelGGL((k_prepass<FMT, ST_A>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
    else
        hipLaunchKernelGGL((k_prepass<FMT, ST_B>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
}

template <int FMT, int OP>
static void launch_main(LbmDev *d, const LbmStepParams *p, int phase, int use_ovr, int write_vel) {
    StepArgs<FMT> a;
    a.pop = (typename Pop<FMT>::S *)d->pop;
    a.mask = d->mask;
    a.ovr_index = d->ovr_index;
    a.ovr_vals = (const typename Pop<FMT>::T *)d->ovr_vals;
    a.ovr_stride = d->n_ovr;
    a.use_ovr = use_ovr && d->n_ovr > 0;
    a.vel = d->vel;
    a.write_vel = write_vel;
    a.g = Geo{d->nx, d->ny, d->nz, d->ncell, d->stride};
    a.cp.tau = p->tau;
    a.cp.cs2x18 = 18.0f * p->smagorinskyCs * p->smagorinskyCs;
    a.cp.inx = p->inletX;
    a.cp.iny = p->inletY;
    a.cp.inz = p->inletZ;
    a.cp.smag = p->useSmagorinsky == 1;
    const int nb = (d->ncell + BLOCK - 1) / BLOCK;
    if (phase == 0)
        hipLaunchKernelGGL((k_step<FMT, OP, 0>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
    else
        hipLaunchKernelGGL((k_step<FMT, OP, 1>), dim3(nb), dim3(BLOCK), 0, d->stream, a);
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
        ev = (hipEvent_t *)calloc((size_t)nsteps * 3, sizeof(hipEvent_t));
        for (int k = 0; k < 3 * nsteps; k++) hipEventCreate(&ev[k]);
    }
    HIPCHECK(hipStreamSynchronize(d->stream), "synchronize");
    // Whether the first step had a prepass (only the first step after
    // LBM_InitializeFlow has none).
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

template <int FMT, int STATE>
static void launch_force(LbmDev *d, int periodic, int nb) {
    ForceArgs<FMT> a;
    a.pop = (const typename Pop<FMT>::S *)d->pop;
    a.mask = d->mask;
    a.ovr_cells = d->ovr_cells;
    a.ovr_q = d->ovr_q;
    a.vel = d->vel;
    a.n_ovr = d->n_ovr;
    a.periodic = periodic;
    a.g = Geo{d->nx, d->ny, d->nz, d->ncell, d->stride};
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
    a.g = Geo{d->nx, d->ny, d->nz, d->ncell, d->stride};
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

extern "C" float *lbmdev_velocity_ptr(LbmDev *d) { return (float *)d->vel; }
extern "C" int32_t *lbmdev_solid_ptr(LbmDev *d) { return d->solid; }

extern "C" int lbmdev_copy_velocity(LbmDev *d, float *host) {
    HIPCHECK(hipMemcpyAsync(host, d->vel, (size_t)d->ncell * 16, hipMemcpyDeviceToHost, d->stream), "velocity readback");
    HIPCHECK(hipStreamSynchronize(d->stream), "velocity readback");
    return 1;
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

extern "C" int lbmdev_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
