/*
 * CPU oracle -- particle velocity sampling and RK2 advection.
 * TEST INFRASTRUCTURE ONLY (tests/, smoke check).
 *
 * Sampler: the trilinear velocity sampler of
 * simulation/shaders/particle_lbm.comp:46-90 (world->grid (p+(4,2,2))/(8,4,4)*N,
 * clamp to [0, N-1], i1 = min(i0+1, N-1), GLSL mix(a,b,t) = a*(1-t) + b*t),
 * and the nearest-cell sampler of simulation/shaders/particle.comp:66-84
 * (clamp to [0.5, N-1.5], floor).  The velocity field is the reference's
 * vec4 (ux, uy, uz, rho) per cell, index x + NX*(y + NY*z).
 *
 * Advection: explicit midpoint RK2 with the trilinear sampler, respawn on the
 * inlet plane when a particle leaves the box |y|,|z| <= 2, x <= 4 (the exit
 * test of particle_lbm.comp:279-280).  The reference's live particle kernel
 * relaxes toward a nearest-cell sample with a time-seeded hash, which is not
 * reproducible; the RK2 scheme and the counter hash below are this project's
 * definition and the product kernel follows them exactly.
 */
#include <math.h>
#include <stdint.h>

typedef struct {
    float x, y, z, pad, vx, vy, vz, life;
} OrParticle;

static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

void or_sample_trilinear(const float *vel, int nx, int ny, int nz, float px, float py, float pz,
                         float *out) {
    float gx = (px + 4.0f) / 8.0f * (float)nx;
    float gy = (py + 2.0f) / 4.0f * (float)ny;
    float gz = (pz + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.0f), (float)nx - 1.0f);
    gy = fminf(fmaxf(gy, 0.0f), (float)ny - 1.0f);
    gz = fminf(fmaxf(gz, 0.0f), (float)nz - 1.0f);
    int x0 = (int)floorf(gx), y0 = (int)floorf(gy), z0 = (int)floorf(gz);
    int x1 = x0 + 1 < nx - 1 ? x0 + 1 : nx - 1;
    int y1 = y0 + 1 < ny - 1 ? y0 + 1 : ny - 1;
    int z1 = z0 + 1 < nz - 1 ? z0 + 1 : nz - 1;
    float fx = gx - floorf(gx), fy = gy - floorf(gy), fz = gz - floorf(gz);
#define V(a, b, c, k) vel[4 * ((long)(a) + (long)nx * ((b) + (long)ny * (c))) + (k)]
    for (int k = 0; k < 3; k++) {
        float v00 = mixf(V(x0, y0, z0, k), V(x1, y0, z0, k), fx);
        float v10 = mixf(V(x0, y1, z0, k), V(x1, y1, z0, k), fx);
        float v01 = mixf(V(x0, y0, z1, k), V(x1, y0, z1, k), fx);
        float v11 = mixf(V(x0, y1, z1, k), V(x1, y1, z1, k), fx);
        float v0 = mixf(v00, v10, fy);
        float v1 = mixf(v01, v11, fy);
        out[k] = mixf(v0, v1, fz);
    }
#undef V
}

void or_sample_nearest(const float *vel, int nx, int ny, int nz, float px, float py, float pz,
                       float *out) {
    float gx = (px + 4.0f) / 8.0f * (float)nx;
    float gy = (py + 2.0f) / 4.0f * (float)ny;
    float gz = (pz + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.5f), (float)nx - 1.5f);
    gy = fminf(fmaxf(gy, 0.5f), (float)ny - 1.5f);
    gz = fminf(fmaxf(gz, 0.5f), (float)nz - 1.5f);
    int cx = (int)floorf(gx), cy = (int)floorf(gy), cz = (int)floorf(gz);
    if (cx < 0) cx = 0;
    if (cx > nx - 1) cx = nx - 1;
    if (cy < 0) cy = 0;
    if (cy > ny - 1) cy = ny - 1;
    if (cz < 0) cz = 0;
    if (cz > nz - 1) cz = nz - 1;
    long idx = cx + (long)nx * (cy + (long)ny * cz);
    for (int k = 0; k < 3; k++) out[k] = vel[4 * idx + k];
}

/* 32-bit avalanche mix (the "lowbias32" constants of Wellons' hash-prospector). */
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

void or_particle_spawn_yz(uint32_t seed, uint32_t step, uint32_t idx, float *y, float *z) {
    uint32_t base = mix32(step ^ mix32(seed + 0x9e3779b9u));
    *y = -2.0f + 4.0f * u01(mix32(2u * idx + base));
    *z = -2.0f + 4.0f * u01(mix32(2u * idx + 1u + base));
}

void or_particles_seed(OrParticle *p, int n, uint32_t seed) {
    for (int i = 0; i < n; i++) {
        float y, z;
        or_particle_spawn_yz(seed, 0u, (uint32_t)i, &y, &z);
        p[i].x = -4.0f;
        p[i].y = y;
        p[i].z = z;
        p[i].pad = 0.0f;
        p[i].vx = p[i].vy = p[i].vz = 0.0f;
        p[i].life = 1.0f;
    }
}

void or_particles_advect(const float *vel, int nx, int ny, int nz, OrParticle *p, int n, float dt,
                         float velScale, uint32_t seed, uint32_t step) {
    for (int i = 0; i < n; i++) {
        float v1[3], v2[3];
        or_sample_trilinear(vel, nx, ny, nz, p[i].x, p[i].y, p[i].z, v1);
        float h = 0.5f * dt * velScale;
        float mx = p[i].x + h * v1[0], my = p[i].y + h * v1[1], mz = p[i].z + h * v1[2];
        or_sample_trilinear(vel, nx, ny, nz, mx, my, mz, v2);
        float s = dt * velScale;
        float X = p[i].x + s * v2[0], Y = p[i].y + s * v2[1], Z = p[i].z + s * v2[2];
        p[i].vx = velScale * v2[0];
        p[i].vy = velScale * v2[1];
        p[i].vz = velScale * v2[2];
        p[i].pad = sqrtf(p[i].vx * p[i].vx + p[i].vy * p[i].vy + p[i].vz * p[i].vz);
        if (X > 4.0f || fabsf(Y) > 2.0f || fabsf(Z) > 2.0f) {
            or_particle_spawn_yz(seed, step, (uint32_t)i, &Y, &Z);
            X = -4.0f;
            p[i].life = 1.0f;
        }
        p[i].x = X;
        p[i].y = Y;
        p[i].z = Z;
    }
}

/* ------------------------------------------------------------------------
 * The live particle update of simulation/shaders/particle.comp:218-407 (the
 * kernel main.c:1012-1101 dispatches), restated in C in the shader's order of
 * operations.  GLSL mix(a, b, t) = a (1 - t) + b t; sign(0) = 0; normalize(v)
 * = v / length(v).  The respawn draws its three uniforms from the counter
 * hash below instead of the shader's fract(sin(idx + time)) hash.
 */
typedef struct {
    float wind[3], carMin[3], carMax[3], carCenter[3];
    int collisionMode, useLBM, vizMode;
    const float *triangles;
    int numTriangles;
    const int *cellStart, *cellCount, *triIndices;
    int numIndices;
    float gridMin[3], gridCellSize[3];
    int gridRes;
} OrParticleFlow;

static inline float gsign(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
static inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

static int inside_aabb(const float *p, const float *lo, const float *hi) {
    return p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2];
}

/* particle.comp:66-84: nearest-cell index, grid position clamped to [0.5, N-1.5] */
long or_live_cell(int nx, int ny, int nz, const float *p) {
    float g[3] = {(p[0] + 4.0f) / 8.0f * (float)nx, (p[1] + 2.0f) / 4.0f * (float)ny, (p[2] + 2.0f) / 4.0f * (float)nz};
    const int n[3] = {nx, ny, nz};
    int c[3];
    for (int k = 0; k < 3; k++) {
        g[k] = clampf(g[k], 0.5f, (float)n[k] - 1.5f);
        c[k] = (int)floorf(g[k]);
        if (c[k] < 0) c[k] = 0;
        if (c[k] > n[k] - 1) c[k] = n[k] - 1;
    }
    long idx = c[0] + (long)c[1] * nx + (long)c[2] * nx * ny;
    long maxIdx = (long)nx * ny * nz - 1;
    return idx < 0 ? 0 : (idx > maxIdx ? maxIdx : idx);
}

/* particle.comp:147-216 (Eberly's point-triangle distance) */
float or_point_triangle(const float *p, const float *v0, const float *v1, const float *v2, float *closest,
                        float *normal) {
    float e0[3], e1[3], d0[3];
    for (int k = 0; k < 3; k++) {
        e0[k] = v1[k] - v0[k];
        e1[k] = v2[k] - v0[k];
        d0[k] = v0[k] - p[k];
    }
    float a = dot3(e0, e0), b = dot3(e0, e1), c = dot3(e1, e1), d = dot3(e0, d0), e = dot3(e1, d0);
    float det = a * c - b * b, s = b * e - c * d, t = b * d - a * e;
    if (s + t <= det) {
        if (s < 0.0f) {
            if (t < 0.0f) {
                if (d < 0.0f) {
                    t = 0.0f;
                    s = clampf(-d / a, 0.0f, 1.0f);
                } else {
                    s = 0.0f;
                    t = clampf(-e / c, 0.0f, 1.0f);
                }
            } else {
                s = 0.0f;
                t = clampf(-e / c, 0.0f, 1.0f);
            }
        } else if (t < 0.0f) {
            t = 0.0f;
            s = clampf(-d / a, 0.0f, 1.0f);
        } else {
            float inv = 1.0f / det;
            s *= inv;
            t *= inv;
        }
    } else {
        if (s < 0.0f) {
            float t0 = b + d, t1 = c + e;
            if (t1 > t0) {
                s = clampf((t1 - t0) / (a - 2.0f * b + c), 0.0f, 1.0f);
                t = 1.0f - s;
            } else {
                s = 0.0f;
                t = clampf(-e / c, 0.0f, 1.0f);
            }
        } else if (t < 0.0f) {
            float t0 = b + e, t1 = a + d;
            if (t1 > t0) {
                t = clampf((t1 - t0) / (a - 2.0f * b + c), 0.0f, 1.0f);
                s = 1.0f - t;
            } else {
                t = 0.0f;
                s = clampf(-d / a, 0.0f, 1.0f);
            }
        } else {
            s = clampf(((c + e) - (b + d)) / (a - 2.0f * b + c), 0.0f, 1.0f);
            t = 1.0f - s;
        }
    }
    float cr[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
    float cl = sqrtf(dot3(cr, cr));
    float dist[3];
    for (int k = 0; k < 3; k++) {
        closest[k] = (v0[k] + e0[k] * s) + e1[k] * t;
        normal[k] = cr[k] / cl;
        dist[k] = p[k] - closest[k];
    }
    return sqrtf(dot3(dist, dist));
}

void or_particles_live(const float *vel, const int *solid, int nx, int ny, int nz, OrParticle *pp, int n, float dt,
                       const OrParticleFlow *f, uint32_t seed, uint32_t step) {
    for (int idx = 0; idx < n; idx++) {
        OrParticle *q = &pp[idx];
        float pos[3] = {q->x, q->y, q->z}, v[3] = {q->vx, q->vy, q->vz}, flow[3];
        for (int k = 0; k < 3; k++) flow[k] = f->wind[k];
        if (f->useLBM) {
            long c = or_live_cell(nx, ny, nz, pos);
            for (int k = 0; k < 3; k++) flow[k] = f->wind[k] * (1.0f - 0.5f) + (vel[4 * c + k] * 8.0f) * 0.5f;
        }
        if (f->vizMode == 8 && f->useLBM) {
            float rho = vel[4 * or_live_cell(nx, ny, nz, pos) + 3];
            float cp = (rho - 1.0f) / (0.5f * 0.05f * 0.05f / (1.0f / 3.0f));
            q->pad = q->pad * (1.0f - 0.05f) + cp * 0.05f;
        } else {
            q->pad = q->pad * (1.0f - 0.02f) + sqrtf(dot3(flow, flow)) * 0.02f;
        }
        for (int k = 0; k < 3; k++) v[k] = v[k] * (1.0f - 0.05f) + flow[k] * 0.05f;
        float lo5[3], hi5[3];
        for (int k = 0; k < 3; k++) {
            lo5[k] = f->carMin[k] - 0.05f;
            hi5[k] = f->carMax[k] + 0.05f;
        }
        if (!(f->collisionMode > 0 && inside_aabb(pos, lo5, hi5))) v[0] = fmaxf(v[0], f->wind[0] * 0.3f);
        for (int k = 0; k < 3; k++) pos[k] = pos[k] + v[k] * dt;
        float life = q->life;
        if (f->collisionMode == 1) {
            if (inside_aabb(pos, f->carMin, f->carMax)) {
                /* particle.comp:112-125 and 127-145 */
                float t[3], a[3], h[3], nrm[3] = {0, 0, 0}, c[3], pen[3];
                for (int k = 0; k < 3; k++) {
                    t[k] = pos[k] - f->carCenter[k];
                    a[k] = fabsf(t[k]);
                    h[k] = (f->carMax[k] - f->carMin[k]) * 0.5f;
                }
                if (a[0] / h[0] > a[1] / h[1] && a[0] / h[0] > a[2] / h[2]) nrm[0] = gsign(t[0]);
                else if (a[1] / h[1] > a[2] / h[2]) nrm[1] = gsign(t[1]);
                else nrm[2] = gsign(t[2]);
                for (int k = 0; k < 3; k++) {
                    c[k] = (f->carMin[k] + f->carMax[k]) * 0.5f;
                    t[k] = pos[k] - c[k];
                    pen[k] = h[k] - fabsf(t[k]);
                }
                if (pen[0] < pen[1] && pen[0] < pen[2]) pos[0] = c[0] + gsign(t[0]) * (h[0] + 0.03f);
                else if (pen[1] < pen[2]) pos[1] = c[1] + gsign(t[1]) * (h[1] + 0.03f);
                else pos[2] = c[2] + gsign(t[2]) * (h[2] + 0.03f);
                float vdn = dot3(v, nrm);
                if (vdn < 0.0f) {
                    for (int k = 0; k < 3; k++) v[k] = v[k] - nrm[k] * (1.05f * vdn);
                    float away[3] = {0.0f, pos[1] - f->carCenter[1], pos[2] - f->carCenter[2]};
                    float al = sqrtf(dot3(away, away));
                    if (al > 0.001f)
                        for (int k = 0; k < 3; k++) v[k] = v[k] + (away[k] / al) * (fabsf(vdn) * 0.4f);
                }
            }
        } else if (f->collisionMode == 2 && f->numTriangles > 0 && f->gridRes > 0) {
            const int R = f->gridRes;
            int cc[3], lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                cc[k] = (int)floorf((pos[k] - f->gridMin[k]) / f->gridCellSize[k]);
                lo[k] = cc[k] - 1 > 0 ? cc[k] - 1 : 0;
                hi[k] = cc[k] + 1 < R - 1 ? cc[k] + 1 : R - 1;
            }
            if (cc[0] >= -1 && cc[0] <= R && cc[1] >= -1 && cc[1] <= R && cc[2] >= -1 && cc[2] <= R) {
                int hit = 0;
                for (int z = lo[2]; z <= hi[2] && !hit; z++)
                    for (int y = lo[1]; y <= hi[1] && !hit; y++)
                        for (int x = lo[0]; x <= hi[0] && !hit; x++) {
                            int ci = x + y * R + z * R * R;
                            for (int j = 0; j < f->cellCount[ci] && !hit; j++) {
                                int e = f->cellStart[ci] + j;
                                if (e < 0 || e >= f->numIndices) continue;
                                int ti = f->triIndices[e];
                                if (ti < 0 || ti >= f->numTriangles) continue;
                                const float *tr = f->triangles + 12 * (long)ti;
                                float cp[3], nrm[3];
                                float dist = or_point_triangle(pos, tr, tr + 4, tr + 8, cp, nrm);
                                if (dist < 0.03f) {
                                    for (int k = 0; k < 3; k++) pos[k] = cp[k] + nrm[k] * (0.03f + 0.001f);
                                    float vdn = dot3(v, nrm);
                                    if (vdn < 0.0f)
                                        for (int k = 0; k < 3; k++) v[k] = v[k] - nrm[k] * (1.05f * vdn);
                                    hit = 1;
                                }
                            }
                        }
            }
        } else if (f->collisionMode == 3) {
            const int dims[3] = {nx, ny, nz};
            const float ext[3] = {8.0f, 4.0f, 4.0f}, off[3] = {4.0f, 2.0f, 2.0f};
            int ci[3];
            for (int k = 0; k < 3; k++) ci[k] = (int)floorf((pos[k] + off[k]) / ext[k] * (float)dims[k]);
            if (ci[0] >= 0 && ci[0] < nx && ci[1] >= 0 && ci[1] < ny && ci[2] >= 0 && ci[2] < nz &&
                solid[ci[0] + (long)ci[1] * nx + (long)ci[2] * nx * ny] == 1) {
                float cell[3], ctr[3], toP[3], ab[3];
                for (int k = 0; k < 3; k++) {
                    cell[k] = ext[k] / (float)dims[k];
                    ctr[k] = ((float)ci[k] + 0.5f) / (float)dims[k] * ext[k] - off[k];
                    toP[k] = pos[k] - ctr[k];
                    ab[k] = fabsf(toP[k]) / (cell[k] * 0.5f);
                }
                int ax[3] = {0, 1, 2}, tmp;
                if (ab[ax[0]] < ab[ax[1]]) { tmp = ax[0]; ax[0] = ax[1]; ax[1] = tmp; }
                if (ab[ax[1]] < ab[ax[2]]) { tmp = ax[1]; ax[1] = ax[2]; ax[2] = tmp; }
                if (ab[ax[0]] < ab[ax[1]]) { tmp = ax[0]; ax[0] = ax[1]; ax[1] = tmp; }
                int pushed = 0;
                for (int k = 0; k < 3 && !pushed; k++) {
                    int a = ax[k];
                    float nrm[3] = {0, 0, 0};
                    nrm[a] = gsign(toP[a]);
                    int ni[3] = {ci[0] + (int)nrm[0], ci[1] + (int)nrm[1], ci[2] + (int)nrm[2]};
                    if (ni[a] >= 0 && ni[a] < dims[a] && solid[ni[0] + (long)ni[1] * nx + (long)ni[2] * nx * ny] == 1)
                        continue;
                    for (int m = 0; m < 3; m++) pos[m] = ctr[m] + (nrm[m] * cell[m]) * 0.52f;
                    float vdn = dot3(v, nrm);
                    if (vdn < 0.0f)
                        for (int m = 0; m < 3; m++) v[m] = v[m] - nrm[m] * (1.05f * vdn);
                    pushed = 1;
                }
                if (!pushed) life = -1.0f;
            }
        }
        /* kill box, particle.comp:369-405 */
        float killX = f->carMax[0] + 2.0f, tunnel = 1.3f;
        pos[1] = clampf(pos[1], -tunnel, tunnel);
        pos[2] = clampf(pos[2], -tunnel, tunnel);
        float speed = sqrtf(dot3(v, v));
        int stuck = speed < 0.05f && pos[0] > f->carMin[0] - 0.5f && pos[0] < f->carMax[0] + 0.5f;
        if (pos[0] > killX || stuck || life < 0.0f) {
            uint32_t base = mix32(step ^ mix32(seed + 0x5bd1e995u));
            float ry = u01(mix32(3u * (uint32_t)idx + base)) - 0.5f;
            float rz = u01(mix32(3u * (uint32_t)idx + 1u + base)) - 0.5f;
            float rv = u01(mix32(3u * (uint32_t)idx + 2u + base));
            pos[0] = -4.0f;
            pos[1] = ry * 2.0f * tunnel;
            pos[2] = rz * 2.0f * tunnel;
            v[0] = 0.2f + rv * 0.15f;
            v[1] = v[2] = 0.0f;
            life = 1.0f;
        }
        q->x = pos[0];
        q->y = pos[1];
        q->z = pos[2];
        q->vx = v[0];
        q->vy = v[1];
        q->vz = v[2];
        q->life = life;
    }
}

/*
 * Streamline tracer: simulation/shaders/streamline_trace.comp:19-107, the
 * application's visualisation mode 9 (dispatched by main.c:1120-1140 with 256
 * seeds of 64 vertices).  Seeds sit on a gridSide x gridSide lattice
 * (gridSide = ceil(sqrt(numSeeds))) of the plane x = -2.5 with
 * y, z = -1.5 + 3 (i + 0.5) / gridSide (:71-78); classic RK4 with step
 * ds = 8 / traceLen (:81-95) through the trilinear sampler clamped to
 * [0.5, N - 1.5] (:29-59), velocities scaled by 8; vertex i is
 * (position, 8 |u(position)|) before step i (:84-87); a step shorter than
 * 1e-6 ends the line and the remaining vertices repeat the position with
 * speed 0 (:97-102); positions are clamped to [-4, 4] x [-2, 2] x [-2, 2]
 * (:104-105).  out: numSeeds * traceLen vertices of 4 floats.
 */
static void sl_sample(const float *vel, int nx, int ny, int nz, const float p[3], float out[3]) {
    float gx = (p[0] + 4.0f) / 8.0f * (float)nx;
    float gy = (p[1] + 2.0f) / 4.0f * (float)ny;
    float gz = (p[2] + 2.0f) / 4.0f * (float)nz;
    gx = fminf(fmaxf(gx, 0.5f), (float)nx - 1.5f);
    gy = fminf(fmaxf(gy, 0.5f), (float)ny - 1.5f);
    gz = fminf(fmaxf(gz, 0.5f), (float)nz - 1.5f);
    int x0 = (int)floorf(gx), y0 = (int)floorf(gy), z0 = (int)floorf(gz);
    int x1 = x0 + 1 < nx - 1 ? x0 + 1 : nx - 1;
    int y1 = y0 + 1 < ny - 1 ? y0 + 1 : ny - 1;
    int z1 = z0 + 1 < nz - 1 ? z0 + 1 : nz - 1;
    float fx = gx - floorf(gx), fy = gy - floorf(gy), fz = gz - floorf(gz);
#define V(a, b, c, k) vel[4 * ((long)(a) + (long)nx * ((b) + (long)ny * (c))) + (k)]
    for (int k = 0; k < 3; k++) {
        float v00 = mixf(V(x0, y0, z0, k), V(x1, y0, z0, k), fx);
        float v10 = mixf(V(x0, y1, z0, k), V(x1, y1, z0, k), fx);
        float v01 = mixf(V(x0, y0, z1, k), V(x1, y0, z1, k), fx);
        float v11 = mixf(V(x0, y1, z1, k), V(x1, y1, z1, k), fx);
        float v0 = mixf(v00, v10, fy);
        float v1 = mixf(v01, v11, fy);
        out[k] = mixf(v0, v1, fz);
    }
#undef V
}

void or_trace_streamlines(const float *vel, int nx, int ny, int nz, int numSeeds, int traceLen, float *out) {
    const int gridSide = (int)ceilf(sqrtf((float)numSeeds));
    const float ds = 8.0f / (float)traceLen;
    const float lo[3] = {-4.0f, -2.0f, -2.0f}, hi[3] = {4.0f, 2.0f, 2.0f};
    for (int s = 0; s < numSeeds; s++) {
        const int iy = s / gridSide, iz = s % gridSide;
        float pos[3] = {-2.5f, -1.5f + 3.0f * ((float)iy + 0.5f) / (float)gridSide,
                        -1.5f + 3.0f * ((float)iz + 0.5f) / (float)gridSide};
        float *o = out + 4 * (long)s * traceLen;
        for (int i = 0; i < traceLen; i++) {
            float v[3], k1[3], k2[3], k3[3], k4[3], q[3], st[3];
            sl_sample(vel, nx, ny, nz, pos, v);
            const float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * 8.0f;
            o[4 * i + 0] = pos[0];
            o[4 * i + 1] = pos[1];
            o[4 * i + 2] = pos[2];
            o[4 * i + 3] = speed;
            const float h = 0.5f * ds;
            for (int k = 0; k < 3; k++) k1[k] = v[k] * 8.0f;
            for (int k = 0; k < 3; k++) q[k] = pos[k] + h * k1[k];
            sl_sample(vel, nx, ny, nz, q, k2);
            for (int k = 0; k < 3; k++) k2[k] = k2[k] * 8.0f;
            for (int k = 0; k < 3; k++) q[k] = pos[k] + h * k2[k];
            sl_sample(vel, nx, ny, nz, q, k3);
            for (int k = 0; k < 3; k++) k3[k] = k3[k] * 8.0f;
            for (int k = 0; k < 3; k++) q[k] = pos[k] + ds * k3[k];
            sl_sample(vel, nx, ny, nz, q, k4);
            for (int k = 0; k < 3; k++) k4[k] = k4[k] * 8.0f;
            const float c = ds / 6.0f;
            for (int k = 0; k < 3; k++) st[k] = c * (k1[k] + 2.0f * k2[k] + 2.0f * k3[k] + k4[k]);
            if (sqrtf(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]) < 1e-6f) {
                for (int j = i + 1; j < traceLen; j++) {
                    o[4 * j + 0] = pos[0];
                    o[4 * j + 1] = pos[1];
                    o[4 * j + 2] = pos[2];
                    o[4 * j + 3] = 0.0f;
                }
                break;
            }
            for (int k = 0; k < 3; k++) pos[k] = fminf(fmaxf(pos[k] + st[k], lo[k]), hi[k]);
        }
    }
}
