/*
 * Triangle-mesh geometry shared by the host library (lbm.c, gcc) and the
 * device voxeliser (mesh_kernels.hip, hipcc): the same float expressions on
 * both sides, compiled without contraction, so host and device give the same
 * solid mask and Bouzidi q bit for bit.
 *
 * Semantics of simulation/src/lbm.c:964-1223 (LBM_SetSolidMesh):
 *   solid  odd number of +X ray crossings from the cell centre (Moller &
 *          Trumbore, |det| < 1e-5 rejected, hit distance > 1e-5), over the
 *          triangles whose (y, z) range contains the centre and whose max x
 *          is not behind it;
 *   q      nearest hit t in (0, 1] along each fluid -> non-fluid link
 *          (|det| < 1e-6 rejected), over the SAME +X-culled candidates
 *          (lbm.c:1150-1157), fallback 0.5, clamped to [0.01, 0.99].
 * A uniform (y, z) bin grid narrows the candidates; the exact cull test is
 * applied per triangle, so every result equals testing all triangles.
 */
#ifndef LBM_MESH_H
#define LBM_MESH_H

#include <math.h>

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

/* World box x in [-4,4], y,z in [-2,2] on the lattice (lbm.c:352-354). */
typedef struct {
    float sx, sy, sz;
} Scale;

/* Cell centre in world coordinates (lbm.c:446-448). */
static inline LBM_HD void cell_centre(Scale s, int x, int y, int z, float *wx, float *wy, float *wz) {
    *wx = (x + 0.5f) / s.sx - 4.0f;
    *wy = (y + 0.5f) / s.sy - 2.0f;
    *wz = (z + 0.5f) / s.sz - 2.0f;
}

static inline LBM_HD float clamp_q(float q) {
    if (q < 0.01f) q = 0.01f;
    if (q > 0.99f) q = 0.99f;
    return q;
}

static inline LBM_HD int bin_of(float v, float o, float inv, int n) {
    int b = (int)floorf((v - o) * inv);
    return b < 0 ? 0 : (b >= n ? n - 1 : b);
}

/* Ray-triangle test (Moller & Trumbore 1997) in the float expression order
 * the reference uses; returns the ray parameter through *t.  The +X solid
 * test specialises h = d x e2 and v = f (d . q) for d = (1,0,0).  Vertices
 * are v[0..2], v[4..6], v[8..10] (12 floats per triangle). */
static inline LBM_HD int ray_tri(const float *v, float ox, float oy, float oz, float dx, float dy, float dz,
                                 int plusX, float eps, float *t) {
    float e1x = v[4] - v[0], e1y = v[5] - v[1], e1z = v[6] - v[2];
    float e2x = v[8] - v[0], e2y = v[9] - v[1], e2z = v[10] - v[2];
    float hx, hy, hz;
    if (plusX) {
        hx = 0;
        hy = -e2z;
        hz = e2y;
    } else {
        hx = dy * e2z - dz * e2y;
        hy = dz * e2x - dx * e2z;
        hz = dx * e2y - dy * e2x;
    }
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (a > -eps && a < eps) return 0;
    float f = 1.0f / a;
    float sx = ox - v[0], sy = oy - v[1], sz = oz - v[2];
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return 0;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float vv = plusX ? f * qx : f * (dx * qx + dy * qy + dz * qz);
    if (vv < 0.0f || u + vv > 1.0f) return 0;
    *t = f * (e2x * qx + e2y * qy + e2z * qz);
    return 1;
}

/* Per-triangle cull bounds (minY maxY minZ maxZ maxX) and a by x bz bin grid
 * over (y, z): bin k = yb + by * zb lists its triangles in
 * binList[binStart[k] .. binStart[k+1]). */
typedef struct {
    const float *tri;
    int ntri;
    float *bounds;
    int by, bz;
    float y0, z0, iy, iz;
    int *binStart, *binList;
} MeshIndex;

/* The +X cull test of lbm.c:1034-1040 for triangle t at centre (wx, wy, wz). */
static inline LBM_HD int mesh_candidate(const float *bounds, int t, float wx, float wy, float wz) {
    const float *b = bounds + 5L * t;
    return !(wy < b[0] || wy > b[1] || wz < b[2] || wz > b[3] || wx > b[4]);
}

#ifdef __cplusplus
extern "C" {
#endif
/* mesh_kernels.hip: solid mask (ncell int32, written to solid_host) and the
 * cells with links in cell order (*link_cells, *nlinks) with their 19 q
 * values each (*qrows; -1 = no link), both malloc'd for the caller.
 * box: the model's world bounding box (minX minY minZ maxX maxY maxZ).
 * 1 on success. */
int lbmdev_mesh_voxelize(int nx, int ny, int nz, const MeshIndex *m, Scale s, const float box[6],
                         int *solid_host, int *nlinks, int **link_cells, float **qrows);
#ifdef __cplusplus
}
#endif

#endif /* LBM_MESH_H */
