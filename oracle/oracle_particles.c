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
