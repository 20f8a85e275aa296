/* Learned warm start for the D3Q19 solver: a 3D convolutional
 * encoder-decoder that maps a coarse solid-fraction map plus the inlet speed
 * and viscosity of a tunnel to a coarse (ux, uy, uz, rho) field, used with
 * LBM_InitializeFlowFromField / LBM_BatchInitializeTunnelFromField.
 *
 * No entry point here has a counterpart in the reference solver: it has no
 * warm start.  The weight files use the reference's LTWS container
 * (simulation/src/superres.c:43-101: u32 magic 0x4C545753, u32 version,
 * u32 parameter count, then per parameter u32 ndim, u32 dims[ndim] and the
 * float32 data, little endian).
 *
 * Network (all convolutions 3x3x3, stride 1, zero padding 1, BF16 operands
 * on the matrix cores, FP32 accumulation, ReLU after every convolution but
 * the last; channels-last BF16 activations):
 *   input   coarse lattice C = cnx x cny x cnz (multiples of 16), packed 2x2x2
 *           space-to-depth into 32 channels at C/2: channels 0-7 the solid
 *           fraction of the 8 sub-cells (sub-cell s = 4 dz + 2 dy + dx),
 *           8 = 20 u, 9 = 10 nu, 10-12 = cell-centre coordinates in [-1, 1]
 *           (x, y, z), 13-31 = 0
 *   encoder C/2: 32, 32 | C/4: 64, 64 | C/8: 128, 128 | C/16: 256, 256
 *           (2x2x2 average pooling between levels)
 *   decoder per level: 1 convolution to the finer level's width at the
 *           coarse level, nearest x2 upsampling, + the encoder's skip, 1
 *           convolution; then a head convolution 32 -> 32 without ReLU
 *   output  32 channels at C/2 = 4 targets x 8 sub-cells (channel 4 s + k),
 *           k: 0 = (rho - 1) / u^2, 1 = ux / u - 1, 2 = uy / u, 3 = uz / u.
 * Training: mean squared error on the packed targets, AdamW on FP32 master
 * weights, BF16 copies refreshed after every step. */
#ifndef LBM_WARMSTART_H
#define LBM_WARMSTART_H

#include <stdint.h>

#include "lbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct LBMWarmStart LBMWarmStart;

/* cnx, cny, cnz: coarse lattice (multiples of 16); maxBatch: largest batch
 * of TrainStep / Forward.  Weights He-initialised from `seed` (head scaled
 * by 0.1, biases 0).  NULL on failure (LBM_GetLastError). */
LBMWarmStart *LBM_WarmStartCreate(int cnx, int cny, int cnz, int maxBatch, unsigned int seed);
void LBM_WarmStartFree(LBMWarmStart *ws);
/* number of parameter tensors and of FP32 parameters in total */
int LBM_WarmStartParamTensors(const LBMWarmStart *ws);
long long LBM_WarmStartParamCount(const LBMWarmStart *ws);
/* LTWS files: per convolution its weight (cout, 3, 3, 3, cin) then its bias
 * (cout), in network order. */
int LBM_WarmStartSave(const LBMWarmStart *ws, const char *path);
int LBM_WarmStartLoad(LBMWarmStart *ws, const char *path);
/* All parameters as one FP32 host array in file order (which = 0), or the
 * gradients of the last LBM_WarmStartGradients call (which = 1). */
int LBM_WarmStartGetParams(const LBMWarmStart *ws, int which, float *out);
int LBM_WarmStartSetParams(LBMWarmStart *ws, const float *in);

/* Host-side packing of one sample (see above).  mask: cnz x cny x cnx solid
 * fractions; field: cnz x cny x cnx x 4 (ux, uy, uz, rho). */
int LBM_WarmStartPackInput(const float *mask, float u, float nu, int cnx, int cny, int cnz, uint16_t *out);
int LBM_WarmStartPackTarget(const float *field, float u, int cnx, int cny, int cnz, float *out);
int LBM_WarmStartUnpackOutput(const float *packed, float u, int cnx, int cny, int cnz, float *field);

/* Device-pointer training and inference (LBM_DeviceAlloc buffers):
 * input: batch x (cnz/2 cny/2 cnx/2) x 32 BF16; target and output:
 * batch x (cnz/2 cny/2 cnx/2) x 32 FP32.  Gradients accumulates the
 * parameter gradients of the mean squared error (returned, < 0 on error);
 * TrainStep adds one AdamW update (beta 0.9 / 0.999, eps 1e-8, decoupled
 * weight decay on the convolution weights only). */
float LBM_WarmStartGradients(LBMWarmStart *ws, const void *input, const float *target, int batch);
float LBM_WarmStartTrainStep(LBMWarmStart *ws, const void *input, const float *target, int batch, float lr,
                             float weightDecay);
int LBM_WarmStartForward(LBMWarmStart *ws, const void *input, int batch, float *output);

/* Host-array inference of one tunnel: mask -> field (unpacked, physical). */
int LBM_WarmStartPredict(LBMWarmStart *ws, const float *mask, float u, float nu, float *field);
/* Solid fraction of factor^3 blocks of the grid's solid mask (after
 * LBM_SetSolid*): out has (sizeZ/f) x (sizeY/f) x (sizeX/f) floats. */
int LBM_CoarseSolidFraction(LBMGrid *grid, int factor, float *out);
/* LBM_InitializeFlowFromField with the network's prediction for this grid's
 * geometry, inlet speed u (along x) and viscosity nu = (tau - 1/2) / 3. */
int LBM_WarmStartInitialize(LBMGrid *grid, LBMWarmStart *ws, float u);

/* Layer kernels of the network on device buffers, exported for gradient
 * checks against a CPU oracle (shapes: n x d x h x w positions,
 * channels-last BF16 unless noted).  1 on success. */
int lbmnn_conv_fwd(const void *X, const void *W, const float *bias, void *Y, float *Yf, int n, int d, int h, int w,
                   int cin, int cout, int relu);
/* As lbmnn_conv_fwd, with Wp the packed copy of W from lbmnn_conv_pack
 * (used where the extent divides into 8x4x4 bricks; W elsewhere). */
int lbmnn_conv_fwd_packed(const void *X, const void *W, const void *Wp, const float *bias, void *Y, float *Yf, int n,
                          int d, int h, int w, int cin, int cout, int relu);
int lbmnn_conv_mirror(const void *W, void *Wm, int cin, int cout);
/* From W[co][t][ci] (cin, cout % 32 == 0): Wm = lbmnn_conv_mirror's output,
 * Wp = W packed for lbmnn_conv_fwd_packed, Wmp = Wm packed (its convolution
 * has cin and cout swapped).  Wp and Wmp hold as many elements as W. */
int lbmnn_conv_pack(const void *W, void *Wm, void *Wp, void *Wmp, int cin, int cout);
/* dW += weight gradient and (db != NULL) db += bias gradient; cin, cout % 32 == 0 */
int lbmnn_conv_wgrad(const void *dZ, const void *X, float *dW, float *db, int n, int d, int h, int w, int cin,
                     int cout);
int lbmnn_relu_bwd(void *dY, const void *Y, long long n);
int lbmnn_bias_grad(const void *dZ, float *db, long long P, int c);
int lbmnn_pool_fwd(const void *X, void *Y, int n, int dOut, int hOut, int wOut, int c);
int lbmnn_pool_bwd(const void *dY, void *dX, int n, int dIn, int hIn, int wIn, int c);
int lbmnn_up_fwd(const void *X, const void *skip, void *Y, int n, int dOut, int hOut, int wOut, int c);
int lbmnn_up_bwd(const void *dY, void *dX, int n, int dIn, int hIn, int wIn, int c);
int lbmnn_add_inplace(void *a, const void *b, long long n);
int lbmnn_mse(const float *Yf, const float *T, void *dY, float *loss, long long n);
int lbmnn_adamw(float *w, void *wb, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                float eps, float wd, int t);
int lbmnn_f32_to_bf16(const float *a, void *b, long long n);

#ifdef __cplusplus
}
#endif
#endif /* LBM_WARMSTART_H */
