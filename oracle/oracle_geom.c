/*
 * CPU oracle -- geometry setup (solid mask, Bouzidi q, projected area,
 * equilibrium initial populations).  TEST INFRASTRUCTURE ONLY: this file is
 * linked into the oracle libraries that tests/, the smoke check and the CPU
 * baseline leg of bench.py load as the checker.  The product never links it.
 *
 * Every function restates the host code of the reference solver in plain C,
 * with the same float arithmetic, so that the solid/q arrays it produces can
 * be compared bit-for-bit with the product's.  Arrays use the reference
 * layout: solid[cell] (int 0 fluid / 1 body / 2 ground) and q[19*cell + i]
 * (-1 = no link), cell = x + nx*(y + ny*z).
 *
 * Citations are to MarcosAsh/Lattice_Fluid_Dynamics.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* D3Q19 tables: simulation/src/lbm.c:9-35 (same order as the shaders). */
const int OR_EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int OR_EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
const int OR_EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
const int OR_OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
const float OR_W[19] = {1.0f / 3.0f,  1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f,
                        1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                        1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                        1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};

/* Host feq in float: simulation/src/lbm.c:37-44. */
float or_feq_f32(int i, float rho, float ux, float uy, float uz) {
    float eu = OR_EX[i] * ux + OR_EY[i] * uy + OR_EZ[i] * uz;
    float u2 = ux * ux + uy * uy + uz * uz;
    float cs2 = 1.0f / 3.0f;
    return OR_W[i] * rho * (1.0f + eu / cs2 + (eu * eu) / (2.0f * cs2 * cs2) - u2 / (2.0f * cs2));
}

/* tau = 0.5 + 3 nu clamped to [0.52, 2]: simulation/src/lbm.c:56-60. */
float or_tau_from_viscosity(float viscosity) {
    float tau = 0.5f + 3.0f * viscosity;
    if (tau < 0.52f) tau = 0.52f;
    if (tau > 2.0f) tau = 2.0f;
    return tau;
}

static void q_clear(long ncell, float *q) {
    for (long k = 0; k < 19 * ncell; k++) q[k] = -1.0f;
}

/* simulation/src/lbm.c:346-435: world box -> grid cells (truncation, inclusive,
 * clamped), solid=1, q=0.5 on every in-range fluid->solid link. */
void or_set_solid_aabb(int nx, int ny, int nz, int *solid, float *q, float minX, float minY,
                       float minZ, float maxX, float maxY, float maxZ) {
    long ncell = (long)nx * ny * nz;
    float scaleX = nx / 8.0f, scaleY = ny / 4.0f, scaleZ = nz / 4.0f;
    int gMinX = (int)((minX + 4.0f) * scaleX);
    int gMaxX = (int)((maxX + 4.0f) * scaleX);
    int gMinY = (int)((minY + 2.0f) * scaleY);
    int gMaxY = (int)((maxY + 2.0f) * scaleY);
    int gMinZ = (int)((minZ + 2.0f) * scaleZ);
    int gMaxZ = (int)((maxZ + 2.0f) * scaleZ);
    if (gMinX < 0) gMinX = 0;
    if (gMaxX >= nx) gMaxX = nx - 1;
    if (gMinY < 0) gMinY = 0;
    if (gMaxY >= ny) gMaxY = ny - 1;
    if (gMinZ < 0) gMinZ = 0;
    if (gMaxZ >= nz) gMaxZ = nz - 1;
    memset(solid, 0, sizeof(int) * ncell);
    for (int z = gMinZ; z <= gMaxZ; z++)
        for (int y = gMinY; y <= gMaxY; y++)
            for (int x = gMinX; x <= gMaxX; x++) solid[x + (long)nx * (y + (long)ny * z)] = 1;
    q_clear(ncell, q);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                long c = x + (long)nx * (y + (long)ny * z);
                if (solid[c] == 1) continue;
                for (int i = 1; i < 19; i++) {
                    int X = x + OR_EX[i], Y = y + OR_EY[i], Z = z + OR_EZ[i];
                    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
                    if (solid[X + (long)nx * (Y + (long)ny * Z)] == 1) q[19 * c + i] = 0.5f;
                }
            }
}

/* simulation/src/lbm.c:437-535: cell-centre-in-sphere solid test and analytic
 * ray-sphere q, fallback 0.5, clamp [0.01, 0.99]. */
void or_set_solid_sphere(int nx, int ny, int nz, int *solid, float *q, float cx, float cy, float cz,
                         float radius) {
    long ncell = (long)nx * ny * nz;
    float scaleX = nx / 8.0f, scaleY = ny / 4.0f, scaleZ = nz / 4.0f;
    memset(solid, 0, sizeof(int) * ncell);
    for (int gz = 0; gz < nz; gz++)
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                float wx = (gx + 0.5f) / scaleX - 4.0f;
                float wy = (gy + 0.5f) / scaleY - 2.0f;
                float wz = (gz + 0.5f) / scaleZ - 2.0f;
                float dx = wx - cx, dy = wy - cy, dz = wz - cz;
                if (dx * dx + dy * dy + dz * dz <= radius * radius)
                    solid[gx + (long)nx * (gy + (long)ny * gz)] = 1;
            }
    q_clear(ncell, q);
    for (int gz = 0; gz < nz; gz++)
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                long c = gx + (long)nx * (gy + (long)ny * gz);
                if (solid[c] == 1) continue;
                float wx = (gx + 0.5f) / scaleX - 4.0f;
                float wy = (gy + 0.5f) / scaleY - 2.0f;
                float wz = (gz + 0.5f) / scaleZ - 2.0f;
                for (int i = 1; i < 19; i++) {
                    int X = gx + OR_EX[i], Y = gy + OR_EY[i], Z = gz + OR_EZ[i];
                    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
                    if (solid[X + (long)nx * (Y + (long)ny * Z)] != 1) continue;
                    float dx = (float)OR_EX[i] / scaleX;
                    float dy = (float)OR_EY[i] / scaleY;
                    float dz = (float)OR_EZ[i] / scaleZ;
                    float ox = wx - cx, oy = wy - cy, oz = wz - cz;
                    float a = dx * dx + dy * dy + dz * dz;
                    float b = 2.0f * (ox * dx + oy * dy + oz * dz);
                    float cc = ox * ox + oy * oy + oz * oz - radius * radius;
                    float disc = b * b - 4.0f * a * cc;
                    float qq = 0.5f;
                    if (disc >= 0.0f) {
                        float t = (-b - sqrtf(disc)) / (2.0f * a);
                        if (t > 0.0f && t <= 1.0f) qq = t;
                    }
                    if (qq < 0.01f) qq = 0.01f;
                    if (qq > 0.99f) qq = 0.99f;
                    q[19 * c + i] = qq;
                }
            }
}

/* simulation/src/lbm.c:537-614: cells with z <= gz that are fluid become 2;
 * q = 0.5 on the ez=-1 links of the first layer above, where the link
 * target is non-fluid and no q is set yet. */
void or_add_ground_plane(int nx, int ny, int nz, int *solid, float *q, float worldZ) {
    float scaleZ = nz / 4.0f;
    int gzGround = (int)((worldZ + 2.0f) * scaleZ);
    if (gzGround < 0) gzGround = 0;
    if (gzGround >= nz) gzGround = nz - 1;
    for (int gz = 0; gz <= gzGround; gz++)
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                long c = gx + (long)nx * (gy + (long)ny * gz);
                if (solid[c] == 0) solid[c] = 2;
            }
    int gzAbove = gzGround + 1;
    if (gzAbove < nz) {
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                long c = gx + (long)nx * (gy + (long)ny * gzAbove);
                if (solid[c] != 0) continue;
                for (int i = 1; i < 19; i++) {
                    if (OR_EZ[i] != -1) continue;
                    int Z = gzAbove + OR_EZ[i];
                    int Y = gy + OR_EY[i], X = gx + OR_EX[i];
                    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0) continue;
                    long n = X + (long)nx * (Y + (long)ny * Z);
                    if (solid[n] != 0 && q[19 * c + i] < 0.0f) q[19 * c + i] = 0.5f;
                }
            }
    }
}

/* simulation/src/lbm.c:616-668: number of columns along `axis` holding any
 * solid==1 cell. */
float or_projected_area(int nx, int ny, int nz, const int *solid, int axis) {
    int count = 0;
    if (axis == 0) {
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++) {
                int found = 0;
                for (int x = 0; x < nx && !found; x++)
                    if (solid[x + (long)nx * (y + (long)ny * z)] == 1) found = 1;
                count += found;
            }
    } else if (axis == 1) {
        for (int z = 0; z < nz; z++)
            for (int x = 0; x < nx; x++) {
                int found = 0;
                for (int y = 0; y < ny && !found; y++)
                    if (solid[x + (long)nx * (y + (long)ny * z)] == 1) found = 1;
                count += found;
            }
    } else {
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int found = 0;
                for (int z = 0; z < nz && !found; z++)
                    if (solid[x + (long)nx * (y + (long)ny * z)] == 1) found = 1;
                count += found;
            }
    }
    return (float)count;
}

/* Moller-Trumbore against one packed triangle (12 floats: v0 xyz pad, v1 xyz
 * pad, v2 xyz pad; simulation/lib/model_bounds.h:14-18).  Returns 1 and sets
 * *t when the ray o + t d hits it; `eps` is the parallel-ray threshold.  The
 * +X inside test and the link q test of simulation/src/lbm.c:1043-1090 and
 * :1160-1190 are this routine with d=(1,0,0), eps=1e-5 and with the scaled
 * link vector, eps=1e-6.  The arithmetic follows those lines term by term
 * (including the +X specialisation h=(0,-e2z,e2y), v=f*qx). */
static int mt_hit_x(const float *tri, float wx, float wy, float wz, float *dist) {
    float v0x = tri[0], v0y = tri[1], v0z = tri[2];
    float v1x = tri[4], v1y = tri[5], v1z = tri[6];
    float v2x = tri[8], v2y = tri[9], v2z = tri[10];
    float e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
    float e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
    float hx = 0, hy = -e2z, hz = e2y;
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (a > -0.00001f && a < 0.00001f) return 0;
    float f = 1.0f / a;
    float sx = wx - v0x, sy = wy - v0y, sz = wz - v0z;
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return 0;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float v = f * qx;
    if (v < 0.0f || u + v > 1.0f) return 0;
    *dist = f * (e2x * qx + e2y * qy + e2z * qz);
    return 1;
}

static int mt_hit_dir(const float *tri, float wx, float wy, float wz, float dx, float dy, float dz,
                      float *tHit) {
    float v0x = tri[0], v0y = tri[1], v0z = tri[2];
    float v1x = tri[4], v1y = tri[5], v1z = tri[6];
    float v2x = tri[8], v2y = tri[9], v2z = tri[10];
    float e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
    float e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
    float hx = dy * e2z - dz * e2y;
    float hy = dz * e2x - dx * e2z;
    float hz = dx * e2y - dy * e2x;
    float a = e1x * hx + e1y * hy + e1z * hz;
    if (a > -1e-6f && a < 1e-6f) return 0;
    float f = 1.0f / a;
    float sx = wx - v0x, sy = wy - v0y, sz = wz - v0z;
    float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return 0;
    float qx = sy * e1z - sz * e1y;
    float qy = sz * e1x - sx * e1z;
    float qz = sx * e1y - sy * e1x;
    float v = f * (dx * qx + dy * qy + dz * qz);
    if (v < 0.0f || u + v > 1.0f) return 0;
    *tHit = f * (e2x * qx + e2y * qy + e2z * qz);
    return 1;
}

/* simulation/src/lbm.c:964-1223.  Solid by +X ray parity over triangles
 * culled by (y, z, maxX) bounds; q by nearest hit along each fluid->solid
 * link with the SAME +X culling test applied for every direction (a quirk of
 * the reference kept here because it changes which links fall back to 0.5). */
void or_set_solid_mesh(int nx, int ny, int nz, int *solid, float *q, const float *tris, int ntri,
                       float minX, float minY, float minZ, float maxX, float maxY, float maxZ) {
    long ncell = (long)nx * ny * nz;
    float *tb = (float *)malloc(sizeof(float) * 5 * (ntri > 0 ? ntri : 1));
    for (int t = 0; t < ntri; t++) {
        const float *tri = tris + 12 * t;
        tb[5 * t + 0] = fminf(tri[1], fminf(tri[5], tri[9]));
        tb[5 * t + 1] = fmaxf(tri[1], fmaxf(tri[5], tri[9]));
        tb[5 * t + 2] = fminf(tri[2], fminf(tri[6], tri[10]));
        tb[5 * t + 3] = fmaxf(tri[2], fmaxf(tri[6], tri[10]));
        tb[5 * t + 4] = fmaxf(tri[0], fmaxf(tri[4], tri[8]));
    }
    float scaleX = nx / 8.0f, scaleY = ny / 4.0f, scaleZ = nz / 4.0f;
    memset(solid, 0, sizeof(int) * ncell);
    for (int gz = 0; gz < nz; gz++)
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                float wx = (gx + 0.5f) / scaleX - 4.0f;
                float wy = (gy + 0.5f) / scaleY - 2.0f;
                float wz = (gz + 0.5f) / scaleZ - 2.0f;
                if (wx < minX || wx > maxX || wy < minY || wy > maxY || wz < minZ || wz > maxZ)
                    continue;
                int hits = 0;
                for (int t = 0; t < ntri; t++) {
                    const float *b = tb + 5 * t;
                    if (wy < b[0] || wy > b[1] || wz < b[2] || wz > b[3] || wx > b[4]) continue;
                    float dist;
                    if (mt_hit_x(tris + 12 * t, wx, wy, wz, &dist) && dist > 0.00001f) hits++;
                }
                if (hits % 2 == 1) solid[gx + (long)nx * (gy + (long)ny * gz)] = 1;
            }
    q_clear(ncell, q);
    for (int gz = 0; gz < nz; gz++)
        for (int gy = 0; gy < ny; gy++)
            for (int gx = 0; gx < nx; gx++) {
                long c = gx + (long)nx * (gy + (long)ny * gz);
                if (solid[c] == 1) continue;
                float wx = (gx + 0.5f) / scaleX - 4.0f;
                float wy = (gy + 0.5f) / scaleY - 2.0f;
                float wz = (gz + 0.5f) / scaleZ - 2.0f;
                for (int i = 1; i < 19; i++) {
                    int X = gx + OR_EX[i], Y = gy + OR_EY[i], Z = gz + OR_EZ[i];
                    if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
                    if (solid[X + (long)nx * (Y + (long)ny * Z)] == 0) continue;
                    float dx = (float)OR_EX[i] / scaleX;
                    float dy = (float)OR_EY[i] / scaleY;
                    float dz = (float)OR_EZ[i] / scaleZ;
                    float bestT = 2.0f;
                    for (int t = 0; t < ntri; t++) {
                        const float *b = tb + 5 * t;
                        if (wy < b[0] || wy > b[1] || wz < b[2] || wz > b[3] || wx > b[4])
                            continue;
                        float th;
                        if (mt_hit_dir(tris + 12 * t, wx, wy, wz, dx, dy, dz, &th) && th > 0.0f &&
                            th < bestT)
                            bestT = th;
                    }
                    float qq = (bestT <= 1.0f) ? bestT : 0.5f;
                    if (qq < 0.01f) qq = 0.01f;
                    if (qq > 0.99f) qq = 0.99f;
                    q[19 * c + i] = qq;
                }
            }
    free(tb);
}
