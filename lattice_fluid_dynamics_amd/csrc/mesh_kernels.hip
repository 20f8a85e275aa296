// Triangle-mesh voxelisation on the GPU: the solid mask and the Bouzidi q
// of LBM_SetSolidMesh (simulation/src/lbm.c:964-1223) with the float
// expressions of lbm_mesh.h, shared with the host, compiled without
// contraction; results equal the host evaluation bit for bit.
//
//   k_mesh_solid    one lane per cell: +X ray parity over the bin's culled
//                   candidates (cells outside the model box stay fluid);
//   k_mesh_linkflag one lane per cell: fluid (solid != 1) with at least one
//                   in-range non-fluid neighbour;
//   compaction      k_flag_count + k_flag_scatter keep those cells in cell
//                   order (the order of the host's q rows): per tile of 4096
//                   cells a count, an exclusive scan of the counts on the
//                   host, then wave ballots place each flagged cell;
//   k_mesh_q        one lane per (listed cell, direction): nearest hit of the
//                   link over the same candidates, or -1 for no link.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_device.h"
#include "lbm_mesh.h"

namespace {

constexpr int MEX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int MEY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int MEZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

struct MeshDev {
    const float *tri, *bounds;
    const int *binStart, *binList;
    int ntri, by, bz;
    float y0, z0, iy, iz;
};

__device__ __forceinline__ int mesh_bin(const MeshDev &m, float wy, float wz) {
    return bin_of(wy, m.y0, m.iy, m.by) + m.by * bin_of(wz, m.z0, m.iz, m.bz);
}

__global__ __launch_bounds__(256) void k_mesh_solid(MeshDev m, Scale s, int nx, int ny, long long ncell, float minX,
                                                    float minY, float minZ, float maxX, float maxY, float maxZ,
                                                    int32_t *solid) {
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((long long)nx * ny));
    float wx, wy, wz;
    cell_centre(s, x, y, z, &wx, &wy, &wz);
    int v = 0;
    if (m.ntri > 0 && !(wx < minX || wx > maxX || wy < minY || wy > maxY || wz < minZ || wz > maxZ)) {
        const int k = mesh_bin(m, wy, wz);
        int hits = 0;
        for (int j = m.binStart[k]; j < m.binStart[k + 1]; j++) {
            const int t = m.binList[j];
            if (!mesh_candidate(m.bounds, t, wx, wy, wz)) continue;
            float d;
            if (ray_tri(m.tri + 12L * t, wx, wy, wz, 1.0f, 0.0f, 0.0f, 1, 0.00001f, &d) && d > 0.00001f) hits++;
        }
        v = hits & 1;
    }
    solid[c] = v;
}

__global__ __launch_bounds__(256) void k_mesh_linkflag(const int32_t *solid, int nx, int ny, int nz, long long ncell,
                                                       uint8_t *flag) {
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((long long)nx * ny));
    uint8_t f = 0;
    if (solid[c] != 1) {
#pragma unroll
        for (int i = 1; i < 19; i++) {
            const int X = x + MEX[i], Y = y + MEY[i], Z = z + MEZ[i];
            if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
            if (solid[((long long)Z * ny + Y) * nx + X] != 0) f = 1;
        }
    }
    flag[c] = f;
}

__global__ __launch_bounds__(256) void k_mesh_q(MeshDev m, Scale s, const int32_t *solid, int nx, int ny, int nz,
                                                const int32_t *cells, int n, float *qrows) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx >= 19LL * n) return;
    const int k = (int)(idx / 19), i = (int)(idx % 19);
    const long long c = cells[k];
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((long long)nx * ny));
    float q = -1.0f;
    const int X = x + MEX[i], Y = y + MEY[i], Z = z + MEZ[i];
    if (i > 0 && X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz &&
        solid[((long long)Z * ny + Y) * nx + X] != 0) {
        float wx, wy, wz;
        cell_centre(s, x, y, z, &wx, &wy, &wz);
        const float dx = (float)MEX[i] / s.sx, dy = (float)MEY[i] / s.sy, dz = (float)MEZ[i] / s.sz;
        float best = 2.0f;
        if (m.ntri > 0) {
            const int b = mesh_bin(m, wy, wz);
            for (int j = m.binStart[b]; j < m.binStart[b + 1]; j++) {
                const int t = m.binList[j];
                if (!mesh_candidate(m.bounds, t, wx, wy, wz)) continue;
                float th;
                if (ray_tri(m.tri + 12L * t, wx, wy, wz, dx, dy, dz, 0, 1e-6f, &th) && th > 0.0f && th < best)
                    best = th;
            }
        }
        q = clamp_q(best <= 1.0f ? best : 0.5f);
    }
    qrows[idx] = q;
}

// Order-preserving compaction of a byte flag array into cell indices.  A
// workgroup owns one tile of CMP_TILE consecutive cells, visited in
// CMP_ROUNDS rounds of 256; within a round, cells are placed by wave ballot
// (lanes in index order) after the earlier waves of the round.
constexpr int CMP_ROUNDS = 16, CMP_TILE = 256 * CMP_ROUNDS;

__global__ __launch_bounds__(256) void k_flag_count(const uint8_t *__restrict__ flag, long long n,
                                                    int *__restrict__ tileCount) {
    __shared__ int total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * CMP_TILE;
    int c = 0;
    for (int r = 0; r < CMP_ROUNDS; r++) {
        const long long i = base + r * 256 + threadIdx.x;
        c += (i < n && flag[i]) ? 1 : 0;
    }
    if (c) atomicAdd(&total, c);
    __syncthreads();
    if (threadIdx.x == 0) tileCount[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_flag_scatter(const uint8_t *__restrict__ flag, long long n,
                                                      const int *__restrict__ tileStart, int32_t *__restrict__ list) {
    __shared__ int waveCount[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long base = (long long)blockIdx.x * CMP_TILE;
    int out = tileStart[blockIdx.x];
    for (int r = 0; r < CMP_ROUNDS; r++) {
        const long long i = base + r * 256 + threadIdx.x;
        const bool f = i < n && flag[i];
        const unsigned long long m = __ballot(f);
        if (lane == 0) waveCount[wave] = __popcll(m);
        __syncthreads();
        int off = out;
        for (int w = 0; w < wave; w++) off += waveCount[w];
        if (f) list[off + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
        out += waveCount[0] + waveCount[1] + waveCount[2] + waveCount[3];
        __syncthreads();
    }
}

struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

inline unsigned grid_of(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int lbmdev_mesh_voxelize(int nx, int ny, int nz, const MeshIndex *m, Scale s, const float box[6],
                                    int *solid_host, int *nlinks, int **link_cells, float **qrows) {
    *nlinks = 0;
    *link_cells = NULL;
    *qrows = NULL;
    const long long ncell = (long long)nx * ny * nz;
    const int nbins = m->by * m->bz;
    const long long nlist = m->ntri > 0 ? m->binStart[nbins] : 0;
    Buf tri, bounds, bstart, blist, solid, flag, list, tiles, q;
#define MV_CHECK(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            fprintf(stderr, "mesh voxelisation: %s (%s)\n", #call, hipGetErrorString(e_));     \
            return 0;                                                                           \
        }                                                                                       \
    } while (0)
    MeshDev md;
    md.ntri = m->ntri;
    md.by = m->by;
    md.bz = m->bz;
    md.y0 = m->y0;
    md.z0 = m->z0;
    md.iy = m->iy;
    md.iz = m->iz;
    md.tri = md.bounds = NULL;
    md.binStart = md.binList = NULL;
    if (m->ntri > 0) {
        MV_CHECK(hipMalloc(&tri.p, sizeof(float) * 12 * (size_t)m->ntri));
        MV_CHECK(hipMalloc(&bounds.p, sizeof(float) * 5 * (size_t)m->ntri));
        MV_CHECK(hipMalloc(&bstart.p, sizeof(int) * (size_t)(nbins + 1)));
        MV_CHECK(hipMalloc(&blist.p, sizeof(int) * (size_t)(nlist > 0 ? nlist : 1)));
        MV_CHECK(hipMemcpy(tri.p, m->tri, sizeof(float) * 12 * (size_t)m->ntri, hipMemcpyHostToDevice));
        MV_CHECK(hipMemcpy(bounds.p, m->bounds, sizeof(float) * 5 * (size_t)m->ntri, hipMemcpyHostToDevice));
        MV_CHECK(hipMemcpy(bstart.p, m->binStart, sizeof(int) * (size_t)(nbins + 1), hipMemcpyHostToDevice));
        if (nlist > 0) MV_CHECK(hipMemcpy(blist.p, m->binList, sizeof(int) * (size_t)nlist, hipMemcpyHostToDevice));
        md.tri = (const float *)tri.p;
        md.bounds = (const float *)bounds.p;
        md.binStart = (const int *)bstart.p;
        md.binList = (const int *)blist.p;
    }
    MV_CHECK(hipMalloc(&solid.p, sizeof(int32_t) * (size_t)ncell));
    MV_CHECK(hipMalloc(&flag.p, (size_t)ncell));
    MV_CHECK(hipMalloc(&list.p, sizeof(int32_t) * (size_t)ncell));
    hipLaunchKernelGGL(k_mesh_solid, dim3(grid_of(ncell)), dim3(256), 0, 0, md, s, nx, ny, ncell, box[0], box[1],
                       box[2], box[3], box[4], box[5], (int32_t *)solid.p);
    MV_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_mesh_linkflag, dim3(grid_of(ncell)), dim3(256), 0, 0, (const int32_t *)solid.p, nx, ny, nz,
                       ncell, (uint8_t *)flag.p);
    MV_CHECK(hipGetLastError());
    // order-preserving compaction of the flagged cell indices
    const long long ntile = (ncell + CMP_TILE - 1) / CMP_TILE;
    MV_CHECK(hipMalloc(&tiles.p, sizeof(int) * (size_t)ntile));
    hipLaunchKernelGGL(k_flag_count, dim3((unsigned)ntile), dim3(256), 0, 0, (const uint8_t *)flag.p, ncell,
                       (int *)tiles.p);
    MV_CHECK(hipGetLastError());
    int *tileStart = (int *)malloc(sizeof(int) * (size_t)ntile);
    if (!tileStart) return 0;
    hipError_t ec = hipMemcpy(tileStart, tiles.p, sizeof(int) * (size_t)ntile, hipMemcpyDeviceToHost);
    long long n64 = 0;
    for (long long t = 0; t < ntile && ec == hipSuccess; t++) {  // exclusive scan (cell indices are int32)
        const int c = tileStart[t];
        tileStart[t] = (int)n64;
        n64 += c;
    }
    if (ec == hipSuccess) ec = hipMemcpy(tiles.p, tileStart, sizeof(int) * (size_t)ntile, hipMemcpyHostToDevice);
    free(tileStart);
    MV_CHECK(ec);
    if (n64 > 0x7fffffffLL) return 0;
    const int n = (int)n64;
    if (n > 0) {
        hipLaunchKernelGGL(k_flag_scatter, dim3((unsigned)ntile), dim3(256), 0, 0, (const uint8_t *)flag.p, ncell,
                           (const int *)tiles.p, (int32_t *)list.p);
        MV_CHECK(hipGetLastError());
    }
    if (n > 0) {
        MV_CHECK(hipMalloc(&q.p, sizeof(float) * 19 * (size_t)n));
        hipLaunchKernelGGL(k_mesh_q, dim3(grid_of(19LL * n)), dim3(256), 0, 0, md, s, (const int32_t *)solid.p, nx,
                           ny, nz, (const int32_t *)list.p, n, (float *)q.p);
        MV_CHECK(hipGetLastError());
    }
    MV_CHECK(hipMemcpy(solid_host, solid.p, sizeof(int) * (size_t)ncell, hipMemcpyDeviceToHost));
    int *hc = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    float *hq = (float *)malloc(sizeof(float) * 19 * (size_t)(n > 0 ? n : 1));
    if (!hc || !hq) {
        free(hc);
        free(hq);
        return 0;
    }
    if (n > 0) {
        hipError_t e1 = hipMemcpy(hc, list.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(hq, q.p, sizeof(float) * 19 * (size_t)n, hipMemcpyDeviceToHost);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            free(hc);
            free(hq);
            fprintf(stderr, "mesh voxelisation: readback failed\n");
            return 0;
        }
    }
#undef MV_CHECK
    *nlinks = n;
    *link_cells = hc;
    *qrows = hq;
    return 1;
}
