/*
 * Batched parameter sweeps: many wind tunnels of different sizes, bodies,
 * relaxation times and inlet speeds advanced together on one GPU.
 *
 * No reference counterpart (the reference runs one tunnel per process).
 * Each tunnel's geometry is taken from an LBMGrid set up with the usual
 * LBM_SetSolid* calls, so every tunnel is exactly what a single grid would
 * run; the device side (lbm_kernels.hip) stores all tunnels in one
 * allocation and advances all of them with one launch per step, replayed
 * from HIP graphs, with a workgroup size picked by the autotuner.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm.h"
#include "lbm_device.h"

/* lbm.c */
int lbm_boundary_count(const LBMGrid *g, int periodic);
void lbm_boundary_fill(const LBMGrid *g, int periodic, int32_t *cells, float *q, long qStride, int k0);
const int32_t *lbm_host_solid(const LBMGrid *g);

typedef struct {
    int nx, ny, nz;
    float tau, inlet[3], area;
    int32_t *solid;
    int nb;
    int32_t *cells;
    float *qrows; /* q[i*nb + j] */
} Tunnel;

struct LBMBatch {
    int storage, useRegularized, useMRT, useSmagorinsky, periodicYZ;
    float smagorinskyCs;
    int nt, cap;
    Tunnel *t;
    LbmBatchDev *dev;
    long long stepCount;
};

static void batch_error(const char *m) { printf("ERROR: %s\n", m); }

LBMBatch *LBM_BatchCreate(int populationStorage, int useRegularized, int useMRT, int useSmagorinsky,
                          float smagorinskyCs, int periodicYZ) {
    LBMBatch *b = (LBMBatch *)calloc(1, sizeof(LBMBatch));
    if (!b) return NULL;
    b->storage = populationStorage;
    b->useRegularized = useRegularized;
    b->useMRT = useMRT;
    b->useSmagorinsky = useSmagorinsky;
    b->smagorinskyCs = smagorinskyCs;
    b->periodicYZ = periodicYZ;
    return b;
}

void LBM_BatchFree(LBMBatch *b) {
    if (!b) return;
    lbmbatch_destroy(b->dev);
    for (int k = 0; k < b->nt; k++) {
        free(b->t[k].solid);
        free(b->t[k].cells);
        free(b->t[k].qrows);
    }
    free(b->t);
    free(b);
}

int LBM_BatchAddTunnel(LBMBatch *b, LBMGrid *geometry, float tau, float inletX, float inletY, float inletZ) {
    if (!b || !geometry || b->dev) return -1;
    if (b->nt == b->cap) {
        int cap = b->cap ? 2 * b->cap : 16;
        Tunnel *p = (Tunnel *)realloc(b->t, sizeof(Tunnel) * cap);
        if (!p) return -1;
        b->t = p;
        b->cap = cap;
    }
    Tunnel *t = &b->t[b->nt];
    memset(t, 0, sizeof(*t));
    t->nx = geometry->sizeX;
    t->ny = geometry->sizeY;
    t->nz = geometry->sizeZ;
    t->tau = tau;
    t->inlet[0] = inletX;
    t->inlet[1] = inletY;
    t->inlet[2] = inletZ;
    t->area = LBM_ComputeProjectedArea(geometry, 0);
    const long nc = geometry->totalCells;
    t->solid = (int32_t *)malloc(sizeof(int32_t) * nc);
    t->nb = lbm_boundary_count(geometry, b->periodicYZ == 1);
    t->cells = (int32_t *)malloc(sizeof(int32_t) * (size_t)(t->nb > 0 ? t->nb : 1));
    t->qrows = (float *)malloc(sizeof(float) * 19 * (size_t)(t->nb > 0 ? t->nb : 1));
    if (!t->solid || !t->cells || !t->qrows) {
        free(t->solid);
        free(t->cells);
        free(t->qrows);
        return -1;
    }
    memcpy(t->solid, lbm_host_solid(geometry), sizeof(int32_t) * nc);
    lbm_boundary_fill(geometry, b->periodicYZ == 1, t->cells, t->qrows, t->nb, 0);
    return b->nt++;
}

static float feq_host(int i, float ux, float uy, float uz) {
    static const int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static const int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static const int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    const float cs2 = 1.0f / 3.0f;
    float w = i == 0 ? 1.0f / 3.0f : (i < 7 ? 1.0f / 18.0f : 1.0f / 36.0f);
    float eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
    float uu = ux * ux + uy * uy + uz * uz;
    return w * 1.0f * (1.0f + eu / cs2 + (eu * eu) / (2.0f * cs2 * cs2) - uu / (2.0f * cs2));
}

int LBM_BatchInitialize(LBMBatch *b) {
    if (!b || b->nt == 0) return 0;
    const int nt = b->nt;
    int *dims = (int *)malloc(sizeof(int) * 3 * nt);
    float *tau = (float *)malloc(sizeof(float) * nt), *inlet = (float *)malloc(sizeof(float) * 3 * nt);
    float *feq = (float *)malloc(sizeof(float) * 19 * nt), *vel4 = (float *)malloc(sizeof(float) * 4 * nt);
    int *nper = (int *)malloc(sizeof(int) * nt);
    long long cells = 0;
    int ntot = 0;
    for (int k = 0; k < nt; k++) {
        cells += (long long)b->t[k].nx * b->t[k].ny * b->t[k].nz;
        ntot += b->t[k].nb;
    }
    int32_t *solid = (int32_t *)malloc(sizeof(int32_t) * (size_t)cells);
    int32_t *lcells = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntot > 0 ? ntot : 1));
    float *q = (float *)malloc(sizeof(float) * 19 * (size_t)(ntot > 0 ? ntot : 1));
    int ok = dims && tau && inlet && feq && vel4 && nper && solid && lcells && q;
    if (ok) {
        long long co = 0;
        int k0 = 0;
        for (int k = 0; k < nt; k++) {
            Tunnel *t = &b->t[k];
            dims[3 * k] = t->nx;
            dims[3 * k + 1] = t->ny;
            dims[3 * k + 2] = t->nz;
            tau[k] = t->tau;
            memcpy(inlet + 3 * k, t->inlet, 3 * sizeof(float));
            for (int i = 0; i < 19; i++) feq[19 * k + i] = feq_host(i, t->inlet[0], t->inlet[1], t->inlet[2]);
            vel4[4 * k] = t->inlet[0];
            vel4[4 * k + 1] = t->inlet[1];
            vel4[4 * k + 2] = t->inlet[2];
            vel4[4 * k + 3] = 1.0f;
            const long long nc = (long long)t->nx * t->ny * t->nz;
            memcpy(solid + co, t->solid, sizeof(int32_t) * (size_t)nc);
            co += nc;
            nper[k] = t->nb;
            for (int j = 0; j < t->nb; j++) {
                lcells[k0 + j] = t->cells[j];
                for (int i = 0; i < 19; i++) q[(long)i * ntot + k0 + j] = t->qrows[(long)i * t->nb + j];
            }
            k0 += t->nb;
        }
        char err[256] = {0};
        b->dev = lbmbatch_create(b->storage, nt, dims, tau, inlet, err, sizeof(err));
        if (!b->dev) {
            char msg[400];
            snprintf(msg, sizeof(msg), "LBM_BatchInitialize: %s", err);
            batch_error(msg);
            ok = 0;
        }
    }
    if (ok) ok = lbmbatch_set_geometry(b->dev, solid, nper, lcells, q, b->periodicYZ == 1) && lbmbatch_init_flow(b->dev, feq, vel4);
    if (!ok && b->dev) batch_error(lbmdev_last_error());
    free(dims);
    free(tau);
    free(inlet);
    free(feq);
    free(vel4);
    free(nper);
    free(solid);
    free(lcells);
    free(q);
    if (ok) {
        b->stepCount = 0;
        printf("LBM batch: %d tunnels, %lld cells, %d boundary cells\n", nt, cells, ntot);
    }
    return ok;
}

int LBM_BatchInitializeTunnelFromField(LBMBatch *b, int t, const float *field, int cnx, int cny, int cnz) {
    if (!b || !b->dev || !field || t < 0 || t >= b->nt || cnx <= 0 || cny <= 0 || cnz <= 0 || cnx > b->t[t].nx ||
        cny > b->t[t].ny || cnz > b->t[t].nz) {
        batch_error("LBM_BatchInitializeTunnelFromField: bad arguments");
        return 0;
    }
    if (b->stepCount != 0) {
        batch_error("LBM_BatchInitializeTunnelFromField: the batch has stepped since LBM_BatchInitialize");
        return 0;
    }
    if (!lbmbatch_init_tunnel_field(b->dev, t, field, cnx, cny, cnz)) {
        batch_error(lbmdev_last_error());
        return 0;
    }
    return 1;
}

static void params_of(const LBMBatch *b, LbmBatchParams *p) {
    p->smagorinskyCs = b->smagorinskyCs;
    p->useRegularized = b->useRegularized;
    p->useMRT = b->useMRT;
    p->useSmagorinsky = b->useSmagorinsky;
    p->periodicYZ = b->periodicYZ;
}

void LBM_BatchStep(LBMBatch *b, int n) {
    if (!b || !b->dev || n <= 0) return;
    LbmBatchParams p;
    params_of(b, &p);
    if (!lbmbatch_step(b->dev, &p, n, 1))
        batch_error(lbmdev_last_error());
    else
        b->stepCount += n;
}

int LBM_BatchProfile(LBMBatch *b, int n, int useGraph, float *ms) {
    if (!b || !b->dev || n <= 0) return 0;
    LbmBatchParams p;
    params_of(b, &p);
    int ok = lbmbatch_profile(b->dev, &p, n, useGraph, ms);
    if (ok) b->stepCount += n;
    return ok;
}

int LBM_BatchAutotune(LBMBatch *b, int steps, float *msPerStep) {
    if (!b || !b->dev) return 0;
    LbmBatchParams p;
    params_of(b, &p);
    return lbmbatch_autotune(b->dev, &p, steps, msPerStep);
}

int LBM_BatchSetBlockSize(LBMBatch *b, int bs) { return b && b->dev ? lbmbatch_set_block_size(b->dev, bs) : 0; }

int LBM_BatchSetStreamingAccess(LBMBatch *b, int mode) {
    return b && b->dev ? lbmbatch_set_streaming(b->dev, mode) : 0;
}

int LBM_BatchSetBlockOrder(LBMBatch *b, int xcdContiguous) {
    if (!b || !b->dev) return 0;
    lbmbatch_set_block_order(b->dev, xcdContiguous);
    return 1;
}

int LBM_BatchGetKernelConfig(const LBMBatch *b, int *blockSize, int *streamingMode, int *xcdContiguous) {
    if (!b || !b->dev) return 0;
    if (blockSize) *blockSize = lbmbatch_block_size(b->dev);
    if (streamingMode) *streamingMode = lbmbatch_streaming(b->dev);
    if (xcdContiguous) *xcdContiguous = lbmbatch_block_order(b->dev);
    return 1;
}

int LBM_BatchTunnelCount(const LBMBatch *b) { return b ? b->nt : 0; }

long long LBM_BatchGuardResetCount(LBMBatch *b) {
    long long n = -1;
    if (!b || !b->dev || !lbmbatch_guard_resets(b->dev, &n)) return -1;
    return n;
}

long long LBM_BatchTotalCells(const LBMBatch *b) {
    long long c = 0;
    for (int k = 0; b && k < b->nt; k++) c += (long long)b->t[k].nx * b->t[k].ny * b->t[k].nz;
    return c;
}

int LBM_BatchComputeForcesF64(LBMBatch *b, double *out, long long *fixedOut) {
    if (!b || !b->dev) return 0;
    long long *fx = (long long *)malloc(sizeof(long long) * 7 * b->nt);
    if (!fx) return 0;
    int ok = lbmbatch_force(b->dev, b->periodicYZ == 1, out, fx);
    if (ok && fixedOut) memcpy(fixedOut, fx, sizeof(long long) * 7 * b->nt);
    free(fx);
    return ok;
}

/* Per tunnel: Cd = |Fx|/(0.5 U^2 A), Cl = Fy/(0.5 U^2 A) (signed, as in
 * main.c:1266), Cd_pressure, Cd_friction = max(Cd - Cd_pressure, 0), from the
 * FP64 force sums with U the tunnel's inlet speed and A its projected area. */
int LBM_BatchComputeCoefficients(LBMBatch *b, float *out) {
    if (!b || !b->dev) return 0;
    double *F = (double *)malloc(sizeof(double) * 6 * b->nt);
    if (!F) return 0;
    int ok = LBM_BatchComputeForcesF64(b, F, NULL);
    for (int k = 0; ok && k < b->nt; k++) {
        const float U = b->t[k].inlet[0];
        const double den = 0.5 * U * U * b->t[k].area;
        float cd = den > 1e-10 ? (float)(fabs(F[6 * k]) / den) : 0.0f;
        float cl = den > 1e-10 ? (float)(F[6 * k + 1] / den) : 0.0f;
        float cp = den > 1e-10 ? (float)(fabs(F[6 * k + 3]) / den) : 0.0f;
        out[4 * k] = cd;
        out[4 * k + 1] = cl;
        out[4 * k + 2] = cp;
        out[4 * k + 3] = cd - cp > 0 ? cd - cp : 0.0f;
    }
    free(F);
    return ok;
}

int LBM_BatchCopyPopulationsToHostF64(LBMBatch *b, int tunnel, double *f, double *fNew) {
    if (!b || !b->dev) return 0;
    return lbmbatch_export(b->dev, tunnel, b->periodicYZ == 1, f, fNew);
}

int LBM_BatchCopyVelocityToHost(LBMBatch *b, int tunnel, float *out) {
    if (!b || !b->dev) return 0;
    return lbmbatch_copy_velocity(b->dev, tunnel, out);
}

int LBM_BatchCopyCoarseVelocityToHost(LBMBatch *b, int tunnel, int factor, float *out) {
    if (!b || !b->dev || !out || tunnel < 0 || tunnel >= b->nt || factor <= 0 || b->t[tunnel].nx % factor ||
        b->t[tunnel].ny % factor || b->t[tunnel].nz % factor) {
        batch_error("LBM_BatchCopyCoarseVelocityToHost: bad arguments");
        return 0;
    }
    return lbmbatch_copy_velocity_coarse(b->dev, tunnel, factor, out);
}
