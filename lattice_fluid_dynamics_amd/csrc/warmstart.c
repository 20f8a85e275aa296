/*
 * Learned warm start: host code of the 3D convolutional encoder-decoder
 * declared in include/lbm_warmstart.h (network layout, LTWS weight I/O,
 * packing, training step, inference and the solver hand-off).  The layer
 * kernels are in warmstart_kernels.hip; every launch goes to the default
 * stream, and the entry points return after their device work has finished.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm_warmstart.h"
#include "lbm_device.h"
#include "ltws.h"

/* lbm.c: the host copy of a grid's solid mask (sizeX * sizeY * sizeZ) */
const int32_t *lbm_host_solid(const LBMGrid *g);

enum { C0A, C0B, C1A, C1B, C2A, C2B, C3A, C3B, R3, D2, R2, D1, R1, D0, HEAD, NCONV };
static const int CIN[NCONV] = {32, 32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32};
static const int COUT[NCONV] = {32, 32, 64, 64, 128, 128, 256, 256, 128, 128, 64, 64, 32, 32, 32};
static const int LEVEL[NCONV] = {0, 0, 1, 1, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0};

/* activations (and their gradients), level and width */
enum {
    A_A0, A_S0, A_P1, A_A1, A_S1, A_P2, A_A2, A_S2, A_P3, A_A3, A_B3,
    A_R3, A_U2, A_D2, A_R2, A_U1, A_D1, A_R1, A_U0, A_D0, A_Y, NACT
};
static const int ACT_LEVEL[NACT] = {0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0};
static const int ACT_C[NACT] = {32, 32, 32, 64, 64, 64, 128, 128, 128, 256, 256,
                                128, 128, 128, 64, 64, 64, 32, 32, 32, 32};

typedef struct {
    float *w, *b;   /* FP32 master weights (cout x 27 x cin) and bias */
    void *wb, *wm;  /* BF16 copy and its tap-mirrored transpose (dgrad) */
    void *wp, *wmp; /* wb and wm in the packed layout of the brick forward */
    float *gw, *gb; /* gradients */
    float *mw, *vw, *mb, *vb;
    long long nw;
} Conv;

struct LBMWarmStart {
    int cnx, cny, cnz, maxBatch;
    int dims[4][3]; /* per level: d, h, w */
    Conv conv[NCONV];
    void *act[NACT], *grad[NACT], *tmp[3];
    float *yf, *loss;
    float *lossHost; /* pinned: the loss of the last step, copied asynchronously */
    void *in1;       /* one packed input (inference) */
    float *out1;    /* one packed output */
    int step;
};

static void ws_err(const char *m) { lbm_note_error(m); }

static long long positions(const LBMWarmStart *ws, int level, int batch) {
    return (long long)batch * ws->dims[level][0] * ws->dims[level][1] * ws->dims[level][2];
}

static uint16_t f2bf_host(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/* splitmix64 + Box-Muller: deterministic He initialisation */
static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return ((double)(sm64(s) >> 11) + 0.5) / 9007199254740992.0; }
static double gauss(uint64_t *s) {
    const double u1 = unif(s), u2 = unif(s);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

#define DALLOC(p, bytes)                                 \
    do {                                                 \
        (p) = lbmdev_malloc(bytes);                      \
        if (!(p)) {                                      \
            ws_err("warm start: device allocation failed"); \
            LBM_WarmStartFree(ws);                       \
            return NULL;                                 \
        }                                                \
    } while (0)

void LBM_WarmStartFree(LBMWarmStart *ws) {
    if (!ws) return;
    for (int k = 0; k < NCONV; k++) {
        Conv *c = &ws->conv[k];
        void *ps[] = {c->w, c->b, c->wb, c->wm, c->wp, c->wmp, c->gw, c->gb, c->mw, c->vw, c->mb, c->vb};
        for (size_t j = 0; j < sizeof(ps) / sizeof(ps[0]); j++)
            if (ps[j]) lbmdev_free(ps[j]);
    }
    for (int a = 0; a < NACT; a++) {
        if (ws->act[a]) lbmdev_free(ws->act[a]);
        if (ws->grad[a]) lbmdev_free(ws->grad[a]);
    }
    for (int l = 0; l < 3; l++)
        if (ws->tmp[l]) lbmdev_free(ws->tmp[l]);
    if (ws->yf) lbmdev_free(ws->yf);
    if (ws->loss) lbmdev_free(ws->loss);
    lbmnn_host_free(ws->lossHost);
    if (ws->in1) lbmdev_free(ws->in1);
    if (ws->out1) lbmdev_free(ws->out1);
    free(ws);
}

/* mirrored and packed copies of every layer's BF16 weights, one launch */
static int pack_all(LBMWarmStart *ws) {
    const void *w[NCONV];
    void *wm[NCONV], *wp[NCONV], *wmp[NCONV];
    for (int k = 0; k < NCONV; k++) {
        w[k] = ws->conv[k].wb;
        wm[k] = ws->conv[k].wm;
        wp[k] = ws->conv[k].wp;
        wmp[k] = ws->conv[k].wmp;
    }
    return lbmnn_conv_pack_list(w, wm, wp, wmp, CIN, COUT, NCONV);
}

static int refresh_bf16(LBMWarmStart *ws, int k) {
    Conv *c = &ws->conv[k];
    return lbmnn_f32_to_bf16(c->w, c->wb, c->nw) && lbmnn_conv_pack(c->wb, c->wm, c->wp, c->wmp, CIN[k], COUT[k]);
}

LBMWarmStart *LBM_WarmStartCreate(int cnx, int cny, int cnz, int maxBatch, unsigned int seed) {
    if (cnx <= 0 || cny <= 0 || cnz <= 0 || cnx % 16 || cny % 16 || cnz % 16 || maxBatch <= 0) {
        ws_err("LBM_WarmStartCreate: the coarse lattice must be positive multiples of 16 and maxBatch > 0");
        return NULL;
    }
    LBMWarmStart *ws = (LBMWarmStart *)calloc(1, sizeof(LBMWarmStart));
    if (!ws) return NULL;
    ws->cnx = cnx;
    ws->cny = cny;
    ws->cnz = cnz;
    ws->maxBatch = maxBatch;
    for (int l = 0; l < 4; l++) {
        ws->dims[l][0] = cnz >> (l + 1);
        ws->dims[l][1] = cny >> (l + 1);
        ws->dims[l][2] = cnx >> (l + 1);
    }
    uint64_t rs = 0x5EED0000ull + seed;
    for (int k = 0; k < NCONV; k++) {
        Conv *c = &ws->conv[k];
        c->nw = (long long)COUT[k] * 27 * CIN[k];
        const size_t wbytes = sizeof(float) * (size_t)c->nw, bbytes = sizeof(float) * (size_t)COUT[k];
        DALLOC(c->w, wbytes);
        DALLOC(c->b, bbytes);
        DALLOC(c->wb, 2 * (size_t)c->nw);
        DALLOC(c->wp, 2 * (size_t)c->nw);
        DALLOC(c->wmp, 2 * (size_t)c->nw);
        DALLOC(c->wm, 2 * (size_t)c->nw);
        DALLOC(c->gw, wbytes);
        DALLOC(c->gb, bbytes);
        DALLOC(c->mw, wbytes);
        DALLOC(c->vw, wbytes);
        DALLOC(c->mb, bbytes);
        DALLOC(c->vb, bbytes);
        float *h = (float *)malloc(wbytes);
        float *hb = (float *)calloc((size_t)COUT[k], sizeof(float));
        if (!h || !hb) {
            free(h);
            free(hb);
            LBM_WarmStartFree(ws);
            return NULL;
        }
        const double sd = sqrt(2.0 / (27.0 * CIN[k])) * (k == HEAD ? 0.1 : 1.0);
        for (long long i = 0; i < c->nw; i++) h[i] = (float)(sd * gauss(&rs));
        int ok = lbmdev_memcpy_h2d(c->w, h, wbytes) && lbmdev_memcpy_h2d(c->b, hb, bbytes) &&
                 lbmnn_memset(c->mw, 0, wbytes) && lbmnn_memset(c->vw, 0, wbytes) &&
                 lbmnn_memset(c->mb, 0, bbytes) && lbmnn_memset(c->vb, 0, bbytes) && refresh_bf16(ws, k);
        free(h);
        free(hb);
        if (!ok) {
            ws_err("warm start: weight upload failed");
            LBM_WarmStartFree(ws);
            return NULL;
        }
    }
    for (int a = 0; a < NACT; a++) {
        const size_t bytes = 2 * (size_t)positions(ws, ACT_LEVEL[a], maxBatch) * ACT_C[a];
        DALLOC(ws->act[a], bytes);
        DALLOC(ws->grad[a], bytes);
    }
    for (int l = 0; l < 3; l++) DALLOC(ws->tmp[l], 2 * (size_t)positions(ws, l, maxBatch) * (32 << l));
    DALLOC(ws->yf, sizeof(float) * (size_t)positions(ws, 0, maxBatch) * 32);
    DALLOC(ws->loss, sizeof(float));
    ws->lossHost = (float *)lbmnn_host_alloc(sizeof(float));
    if (!ws->lossHost) {
        ws_err("warm start: pinned host allocation failed");
        LBM_WarmStartFree(ws);
        return NULL;
    }
    DALLOC(ws->in1, 2 * (size_t)positions(ws, 0, 1) * 32);
    DALLOC(ws->out1, sizeof(float) * (size_t)positions(ws, 0, 1) * 32);
    if (!lbmnn_synchronize()) {
        ws_err("warm start: device error during set-up");
        LBM_WarmStartFree(ws);
        return NULL;
    }
    return ws;
}

int LBM_WarmStartParamTensors(const LBMWarmStart *ws) { return ws ? 2 * NCONV : 0; }

long long LBM_WarmStartParamCount(const LBMWarmStart *ws) {
    if (!ws) return 0;
    long long n = 0;
    for (int k = 0; k < NCONV; k++) n += ws->conv[k].nw + COUT[k];
    return n;
}

/* ------------------------------------------------------------ forward */
static int conv_fwd(LBMWarmStart *ws, int k, const void *in, void *out, float *outf, int B, int relu) {
    const int *d = ws->dims[LEVEL[k]];
    const Conv *c = &ws->conv[k];
    return lbmnn_conv_fwd_packed(in, c->wb, c->wp, c->b, out, outf, B, d[0], d[1], d[2], CIN[k], COUT[k], relu);
}

static int forward(LBMWarmStart *ws, const void *x0, int B, float *yf) {
    void **A = ws->act;
    const int(*D)[3] = (const int(*)[3])ws->dims;
    int ok = 1;
    ok = ok && conv_fwd(ws, C0A, x0, A[A_A0], NULL, B, 1);
    ok = ok && conv_fwd(ws, C0B, A[A_A0], A[A_S0], NULL, B, 1);
    ok = ok && lbmnn_pool_fwd(A[A_S0], A[A_P1], B, D[1][0], D[1][1], D[1][2], 32);
    ok = ok && conv_fwd(ws, C1A, A[A_P1], A[A_A1], NULL, B, 1);
    ok = ok && conv_fwd(ws, C1B, A[A_A1], A[A_S1], NULL, B, 1);
    ok = ok && lbmnn_pool_fwd(A[A_S1], A[A_P2], B, D[2][0], D[2][1], D[2][2], 64);
    ok = ok && conv_fwd(ws, C2A, A[A_P2], A[A_A2], NULL, B, 1);
    ok = ok && conv_fwd(ws, C2B, A[A_A2], A[A_S2], NULL, B, 1);
    ok = ok && lbmnn_pool_fwd(A[A_S2], A[A_P3], B, D[3][0], D[3][1], D[3][2], 128);
    ok = ok && conv_fwd(ws, C3A, A[A_P3], A[A_A3], NULL, B, 1);
    ok = ok && conv_fwd(ws, C3B, A[A_A3], A[A_B3], NULL, B, 1);
    ok = ok && conv_fwd(ws, R3, A[A_B3], A[A_R3], NULL, B, 1);
    ok = ok && lbmnn_up_fwd(A[A_R3], A[A_S2], A[A_U2], B, D[2][0], D[2][1], D[2][2], 128);
    ok = ok && conv_fwd(ws, D2, A[A_U2], A[A_D2], NULL, B, 1);
    ok = ok && conv_fwd(ws, R2, A[A_D2], A[A_R2], NULL, B, 1);
    ok = ok && lbmnn_up_fwd(A[A_R2], A[A_S1], A[A_U1], B, D[1][0], D[1][1], D[1][2], 64);
    ok = ok && conv_fwd(ws, D1, A[A_U1], A[A_D1], NULL, B, 1);
    ok = ok && conv_fwd(ws, R1, A[A_D1], A[A_R1], NULL, B, 1);
    ok = ok && lbmnn_up_fwd(A[A_R1], A[A_S0], A[A_U0], B, D[0][0], D[0][1], D[0][2], 32);
    ok = ok && conv_fwd(ws, D0, A[A_U0], A[A_D0], NULL, B, 1);
    ok = ok && conv_fwd(ws, HEAD, A[A_D0], A[A_Y], yf, B, 0);
    return ok;
}

/* ------------------------------------------------------------ backward */
/* gout: gradient of the layer's output (overwritten by the ReLU mask);
 * in: the layer's input; gin: gradient of the input (NULL: not needed). */
static int conv_bwd(LBMWarmStart *ws, int k, void *gout, const void *in, void *gin, const void *out, int B,
                    int relu) {
    const int *d = ws->dims[LEVEL[k]];
    const long long P = positions(ws, LEVEL[k], B);
    Conv *c = &ws->conv[k];
    int ok = 1;
    if (relu) ok = ok && lbmnn_relu_bwd(gout, out, P * COUT[k]);
    ok = ok && lbmnn_conv_wgrad(gout, in, c->gw, c->gb, B, d[0], d[1], d[2], CIN[k], COUT[k]);
    if (gin)
        ok = ok && lbmnn_conv_fwd_packed(gout, c->wm, c->wmp, NULL, gin, NULL, B, d[0], d[1], d[2], COUT[k], CIN[k], 0);
    return ok;
}

static int backward(LBMWarmStart *ws, const void *x0, int B) {
    void **A = ws->act, **G = ws->grad;
    const int(*D)[3] = (const int(*)[3])ws->dims;
    int ok = 1;
    /* decoder; the skip gradients of s0, s1, s2 accumulate in G[U0..U2] */
    ok = ok && conv_bwd(ws, HEAD, G[A_Y], A[A_D0], G[A_D0], A[A_Y], B, 0);
    ok = ok && conv_bwd(ws, D0, G[A_D0], A[A_U0], G[A_U0], A[A_D0], B, 1);
    ok = ok && lbmnn_up_bwd(G[A_U0], G[A_R1], B, D[1][0], D[1][1], D[1][2], 32);
    ok = ok && conv_bwd(ws, R1, G[A_R1], A[A_D1], G[A_D1], A[A_R1], B, 1);
    ok = ok && conv_bwd(ws, D1, G[A_D1], A[A_U1], G[A_U1], A[A_D1], B, 1);
    ok = ok && lbmnn_up_bwd(G[A_U1], G[A_R2], B, D[2][0], D[2][1], D[2][2], 64);
    ok = ok && conv_bwd(ws, R2, G[A_R2], A[A_D2], G[A_D2], A[A_R2], B, 1);
    ok = ok && conv_bwd(ws, D2, G[A_D2], A[A_U2], G[A_U2], A[A_D2], B, 1);
    ok = ok && lbmnn_up_bwd(G[A_U2], G[A_R3], B, D[3][0], D[3][1], D[3][2], 128);
    ok = ok && conv_bwd(ws, R3, G[A_R3], A[A_B3], G[A_B3], A[A_R3], B, 1);
    /* encoder */
    ok = ok && conv_bwd(ws, C3B, G[A_B3], A[A_A3], G[A_A3], A[A_B3], B, 1);
    ok = ok && conv_bwd(ws, C3A, G[A_A3], A[A_P3], G[A_P3], A[A_A3], B, 1);
    ok = ok && lbmnn_pool_bwd(G[A_P3], ws->tmp[2], B, D[2][0], D[2][1], D[2][2], 128);
    ok = ok && lbmnn_add_inplace(G[A_U2], ws->tmp[2], positions(ws, 2, B) * 128);
    ok = ok && conv_bwd(ws, C2B, G[A_U2], A[A_A2], G[A_A2], A[A_S2], B, 1);
    ok = ok && conv_bwd(ws, C2A, G[A_A2], A[A_P2], G[A_P2], A[A_A2], B, 1);
    ok = ok && lbmnn_pool_bwd(G[A_P2], ws->tmp[1], B, D[1][0], D[1][1], D[1][2], 64);
    ok = ok && lbmnn_add_inplace(G[A_U1], ws->tmp[1], positions(ws, 1, B) * 64);
    ok = ok && conv_bwd(ws, C1B, G[A_U1], A[A_A1], G[A_A1], A[A_S1], B, 1);
    ok = ok && conv_bwd(ws, C1A, G[A_A1], A[A_P1], G[A_P1], A[A_A1], B, 1);
    ok = ok && lbmnn_pool_bwd(G[A_P1], ws->tmp[0], B, D[0][0], D[0][1], D[0][2], 32);
    ok = ok && lbmnn_add_inplace(G[A_U0], ws->tmp[0], positions(ws, 0, B) * 32);
    ok = ok && conv_bwd(ws, C0B, G[A_U0], A[A_A0], G[A_A0], A[A_S0], B, 1);
    ok = ok && conv_bwd(ws, C0A, G[A_A0], x0, NULL, A[A_A0], B, 1);
    return ok;
}

/* Queues gradient zeroing, forward, loss and backward, and the loss copy
 * into pinned host memory; nothing here waits for the device, so the host
 * runs ahead of the kernels (a blocking memset per buffer drained the queue
 * 31 times a step). */
static int queue_gradients(LBMWarmStart *ws, const void *input, const float *target, int batch) {
    float *zp[2 * NCONV + 1];
    long long zn[2 * NCONV + 1];
    for (int k = 0; k < NCONV; k++) {
        zp[2 * k] = ws->conv[k].gw;
        zn[2 * k] = ws->conv[k].nw;
        zp[2 * k + 1] = ws->conv[k].gb;
        zn[2 * k + 1] = COUT[k];
    }
    zp[2 * NCONV] = ws->loss;
    zn[2 * NCONV] = 1;
    const long long n = positions(ws, 0, batch) * 32;
    int ok = lbmnn_zero_list(zp, zn, 2 * NCONV + 1);
    ok = ok && forward(ws, input, batch, ws->yf);
    ok = ok && lbmnn_mse(ws->yf, target, ws->grad[A_Y], ws->loss, n);
    ok = ok && backward(ws, input, batch);
    ok = ok && lbmnn_copy_to_host_async(ws->lossHost, ws->loss, sizeof(float));
    return ok;
}

float LBM_WarmStartGradients(LBMWarmStart *ws, const void *input, const float *target, int batch) {
    if (!ws || !input || !target || batch <= 0 || batch > ws->maxBatch) {
        ws_err("LBM_WarmStartGradients: bad arguments (batch must be 1..maxBatch)");
        return -1.0f;
    }
    if (!queue_gradients(ws, input, target, batch) || !lbmnn_synchronize()) {
        ws_err("LBM_WarmStartGradients: device error");
        return -1.0f;
    }
    return *ws->lossHost;
}

float LBM_WarmStartTrainStep(LBMWarmStart *ws, const void *input, const float *target, int batch, float lr,
                             float weightDecay) {
    if (!ws || !input || !target || batch <= 0 || batch > ws->maxBatch) {
        ws_err("LBM_WarmStartTrainStep: bad arguments (batch must be 1..maxBatch)");
        return -1.0f;
    }
    int ok = queue_gradients(ws, input, target, batch);
    ws->step++;
    /* every weight and bias tensor in one AdamW launch, then the BF16 copies */
    float *aw[2 * NCONV], *am[2 * NCONV], *av[2 * NCONV], awd[2 * NCONV];
    void *awb[2 * NCONV];
    const float *ag[2 * NCONV];
    long long an[2 * NCONV];
    for (int k = 0; k < NCONV; k++) {
        Conv *c = &ws->conv[k];
        aw[2 * k] = c->w, awb[2 * k] = c->wb, ag[2 * k] = c->gw, am[2 * k] = c->mw, av[2 * k] = c->vw;
        an[2 * k] = c->nw, awd[2 * k] = weightDecay;
        aw[2 * k + 1] = c->b, awb[2 * k + 1] = NULL, ag[2 * k + 1] = c->gb, am[2 * k + 1] = c->mb;
        av[2 * k + 1] = c->vb, an[2 * k + 1] = COUT[k], awd[2 * k + 1] = 0.0f;
    }
    ok = ok && lbmnn_adamw_list(aw, awb, ag, am, av, an, awd, 2 * NCONV, lr, 0.9f, 0.999f, 1e-8f, ws->step);
    ok = ok && pack_all(ws);
    if (!ok || !lbmnn_synchronize()) {
        ws_err("LBM_WarmStartTrainStep: device error");
        return -1.0f;
    }
    return *ws->lossHost;
}

int LBM_WarmStartForward(LBMWarmStart *ws, const void *input, int batch, float *output) {
    if (!ws || !input || !output || batch <= 0 || batch > ws->maxBatch) {
        ws_err("LBM_WarmStartForward: bad arguments (batch must be 1..maxBatch)");
        return 0;
    }
    if (!forward(ws, input, batch, output) || !lbmnn_synchronize()) {
        ws_err("LBM_WarmStartForward: device error");
        return 0;
    }
    return 1;
}

/* ------------------------------------------------------------ parameters */
int LBM_WarmStartGetParams(const LBMWarmStart *ws, int which, float *out) {
    if (!ws || !out || which < 0 || which > 1) return 0;
    long long o = 0;
    for (int k = 0; k < NCONV; k++) {
        const Conv *c = &ws->conv[k];
        const float *w = which ? c->gw : c->w, *b = which ? c->gb : c->b;
        if (!lbmdev_memcpy_d2h(out + o, w, sizeof(float) * (size_t)c->nw)) return 0;
        o += c->nw;
        if (!lbmdev_memcpy_d2h(out + o, b, sizeof(float) * (size_t)COUT[k])) return 0;
        o += COUT[k];
    }
    return 1;
}

int LBM_WarmStartSetParams(LBMWarmStart *ws, const float *in) {
    if (!ws || !in) return 0;
    long long o = 0;
    for (int k = 0; k < NCONV; k++) {
        Conv *c = &ws->conv[k];
        if (!lbmdev_memcpy_h2d(c->w, in + o, sizeof(float) * (size_t)c->nw)) return 0;
        o += c->nw;
        if (!lbmdev_memcpy_h2d(c->b, in + o, sizeof(float) * (size_t)COUT[k])) return 0;
        o += COUT[k];
        if (!refresh_bf16(ws, k)) return 0;
    }
    return lbmnn_synchronize();
}

int LBM_WarmStartSave(const LBMWarmStart *ws, const char *path) {
    if (!ws || !path) return 0;
    const long long n = LBM_WarmStartParamCount(ws);
    float *h = (float *)malloc(sizeof(float) * (size_t)n);
    if (!h || !LBM_WarmStartGetParams(ws, 0, h)) {
        free(h);
        ws_err("LBM_WarmStartSave: cannot write the weights");
        return 0;
    }
    /* per convolution its weight (cout, 3, 3, 3, cin), then its bias (cout) */
    uint32_t ndim[2 * NCONV], dims[2 * NCONV][LBM_LTWS_MAX_DIMS];
    memset(dims, 0, sizeof dims);
    for (int k = 0; k < NCONV; k++) {
        const uint32_t wd[5] = {(uint32_t)COUT[k], 3, 3, 3, (uint32_t)CIN[k]};
        ndim[2 * k] = 5;
        memcpy(dims[2 * k], wd, sizeof wd);
        ndim[2 * k + 1] = 1;
        dims[2 * k + 1][0] = (uint32_t)COUT[k];
    }
    const int ok = lbm_ltws_write(path, 2 * NCONV, ndim, (const uint32_t(*)[LBM_LTWS_MAX_DIMS])dims, h) ==
                   LBM_LTWS_OK;
    free(h);
    if (!ok) ws_err("LBM_WarmStartSave: write error");
    return ok;
}

int LBM_WarmStartLoad(LBMWarmStart *ws, const char *path) {
    if (!ws || !path) return 0;
    LBMLtws t;
    const int rc = lbm_ltws_read(path, &t);
    if (rc == LBM_LTWS_CANNOT_OPEN) {
        ws_err("LBM_WarmStartLoad: cannot open the weights file");
        return 0;
    }
    int ok = rc == LBM_LTWS_OK && t.version == LBM_LTWS_VERSION && t.tensors == (uint32_t)(2 * NCONV) &&
             t.total == LBM_WarmStartParamCount(ws);
    for (int k = 0; k < NCONV && ok; k++) {
        const uint32_t want5[5] = {(uint32_t)COUT[k], 3, 3, 3, (uint32_t)CIN[k]};
        ok = t.ndim[2 * k] == 5 && memcmp(t.dims[2 * k], want5, sizeof want5) == 0 && t.ndim[2 * k + 1] == 1 &&
             t.dims[2 * k + 1][0] == (uint32_t)COUT[k];
    }
    ok = ok && LBM_WarmStartSetParams(ws, t.data);
    lbm_ltws_free(&t);
    if (!ok) ws_err("LBM_WarmStartLoad: not an LTWS file of this network");
    return ok;
}

/* ------------------------------------------------------------ packing */
int LBM_WarmStartPackInput(const float *mask, float u, float nu, int cnx, int cny, int cnz, uint16_t *out) {
    if (!mask || !out || cnx % 2 || cny % 2 || cnz % 2) return 0;
    const int hx = cnx / 2, hy = cny / 2, hz = cnz / 2;
    for (int Z = 0; Z < hz; Z++)
        for (int Y = 0; Y < hy; Y++)
            for (int X = 0; X < hx; X++) {
                uint16_t *o = out + (((size_t)Z * hy + Y) * hx + X) * 32;
                float v[32] = {0};
                for (int s = 0; s < 8; s++) {
                    const int z = 2 * Z + (s >> 2), y = 2 * Y + ((s >> 1) & 1), x = 2 * X + (s & 1);
                    v[s] = mask[((size_t)z * cny + y) * cnx + x];
                }
                v[8] = 20.0f * u;
                v[9] = 10.0f * nu;
                v[10] = 2.0f * ((float)X + 0.5f) / (float)hx - 1.0f;
                v[11] = 2.0f * ((float)Y + 0.5f) / (float)hy - 1.0f;
                v[12] = 2.0f * ((float)Z + 0.5f) / (float)hz - 1.0f;
                for (int c = 0; c < 32; c++) o[c] = f2bf_host(v[c]);
            }
    return 1;
}

int LBM_WarmStartPackTarget(const float *field, float u, int cnx, int cny, int cnz, float *out) {
    if (!field || !out || u == 0.0f || cnx % 2 || cny % 2 || cnz % 2) return 0;
    const int hx = cnx / 2, hy = cny / 2, hz = cnz / 2;
    const float iu = 1.0f / u, iu2 = 1.0f / (u * u);
    for (int Z = 0; Z < hz; Z++)
        for (int Y = 0; Y < hy; Y++)
            for (int X = 0; X < hx; X++) {
                float *o = out + (((size_t)Z * hy + Y) * hx + X) * 32;
                for (int s = 0; s < 8; s++) {
                    const int z = 2 * Z + (s >> 2), y = 2 * Y + ((s >> 1) & 1), x = 2 * X + (s & 1);
                    const float *c = field + (((size_t)z * cny + y) * cnx + x) * 4;
                    o[4 * s + 0] = (c[3] - 1.0f) * iu2;
                    o[4 * s + 1] = c[0] * iu - 1.0f;
                    o[4 * s + 2] = c[1] * iu;
                    o[4 * s + 3] = c[2] * iu;
                }
            }
    return 1;
}

int LBM_WarmStartUnpackOutput(const float *packed, float u, int cnx, int cny, int cnz, float *field) {
    if (!packed || !field || cnx % 2 || cny % 2 || cnz % 2) return 0;
    const int hx = cnx / 2, hy = cny / 2, hz = cnz / 2;
    for (int Z = 0; Z < hz; Z++)
        for (int Y = 0; Y < hy; Y++)
            for (int X = 0; X < hx; X++) {
                const float *o = packed + (((size_t)Z * hy + Y) * hx + X) * 32;
                for (int s = 0; s < 8; s++) {
                    const int z = 2 * Z + (s >> 2), y = 2 * Y + ((s >> 1) & 1), x = 2 * X + (s & 1);
                    float *c = field + (((size_t)z * cny + y) * cnx + x) * 4;
                    c[0] = u * (o[4 * s + 1] + 1.0f);
                    c[1] = u * o[4 * s + 2];
                    c[2] = u * o[4 * s + 3];
                    c[3] = 1.0f + u * u * o[4 * s + 0];
                }
            }
    return 1;
}

/* ------------------------------------------------------------ inference */
int LBM_WarmStartPredict(LBMWarmStart *ws, const float *mask, float u, float nu, float *field) {
    if (!ws || !mask || !field || u == 0.0f) return 0;
    const size_t P = (size_t)positions(ws, 0, 1);
    uint16_t *in = (uint16_t *)malloc(2 * P * 32);
    float *out = (float *)malloc(sizeof(float) * P * 32);
    int ok = in && out && LBM_WarmStartPackInput(mask, u, nu, ws->cnx, ws->cny, ws->cnz, in) &&
             lbmdev_memcpy_h2d(ws->in1, in, 2 * P * 32) && LBM_WarmStartForward(ws, ws->in1, 1, ws->out1) &&
             lbmdev_memcpy_d2h(out, ws->out1, sizeof(float) * P * 32) &&
             LBM_WarmStartUnpackOutput(out, u, ws->cnx, ws->cny, ws->cnz, field);
    free(in);
    free(out);
    return ok;
}

int LBM_CoarseSolidFraction(LBMGrid *grid, int factor, float *out) {
    if (!grid || !out || factor <= 0 || grid->sizeX % factor || grid->sizeY % factor || grid->sizeZ % factor) {
        ws_err("LBM_CoarseSolidFraction: the grid must be a multiple of the factor on every axis");
        return 0;
    }
    const int nx = grid->sizeX, ny = grid->sizeY, nz = grid->sizeZ;
    const int cx = nx / factor, cy = ny / factor, cz = nz / factor;
    const int32_t *solid = lbm_host_solid(grid);
    memset(out, 0, sizeof(float) * (size_t)cx * cy * cz);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                if (solid[((size_t)z * ny + y) * nx + x])
                    out[((size_t)(z / factor) * cy + y / factor) * cx + x / factor] += 1.0f;
    const float inv = 1.0f / (float)(factor * factor * factor);
    for (size_t i = 0; i < (size_t)cx * cy * cz; i++) out[i] *= inv;
    return 1;
}

int LBM_WarmStartInitialize(LBMGrid *grid, LBMWarmStart *ws, float u) {
    if (!grid || !ws || u == 0.0f) return 0;
    const int f = grid->sizeX / ws->cnx;
    if (f <= 0 || grid->sizeX != f * ws->cnx || grid->sizeY != f * ws->cny || grid->sizeZ != f * ws->cnz) {
        ws_err("LBM_WarmStartInitialize: the grid must be the network's coarse lattice times one integer factor");
        return 0;
    }
    const size_t nc = (size_t)ws->cnx * ws->cny * ws->cnz;
    float *mask = (float *)malloc(sizeof(float) * nc), *field = (float *)malloc(sizeof(float) * nc * 4);
    const float nu = (grid->tau - 0.5f) / 3.0f;
    int ok = mask && field && LBM_CoarseSolidFraction(grid, f, mask) &&
             LBM_WarmStartPredict(ws, mask, u, nu, field) &&
             LBM_InitializeFlowFromField(grid, field, ws->cnx, ws->cny, ws->cnz);
    free(mask);
    free(field);
    return ok;
}
