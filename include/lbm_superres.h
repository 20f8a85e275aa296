/*
 * Super-resolution upscaler of the velocity field: the reference's
 * SuperResUpscaler (simulation/lib/superres.h:1-29, simulation/src/
 * superres.c, simulation/shaders/superres_upscale.comp) on the GPU.
 *
 * For every coarse cell, the 3x3x3 neighbourhood of (ux, uy, uz, rho) (edge
 * clamped, normalised by the file's mean/std) feeds a 108-128-128-128-32 MLP
 * (ReLU between layers).  Its 32 outputs correct a trilinear skip of the
 * 2x2x2 fine cells, which are then de-normalised and written to a fine grid
 * of twice the resolution, float4 (ux, uy, uz, rho) per cell, index
 * x + FX (y + FY z).
 *
 * GL buffer names become device pointers: SR_Upscale takes the coarse field
 * LBM_GetVelocityBuffer returns, and SR_GetFineVelocityBuffer returns a
 * device float4 array.  SR_Upscale returns when the fine field is written.
 */
#ifndef LBM_SUPERRES_H
#define LBM_SUPERRES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SuperResUpscaler SuperResUpscaler;

/* superres.h:10-13.  weightsPath: LTWS file (u32 magic 0x4C545753, u32
 * version, u32 tensor count, per tensor u32 ndim, ndim u32 sizes, float32
 * data) holding, in order, b0 (128), W0 (128 x 108), b1, W1 (128 x 128), b2,
 * W2 (128 x 128), b3 (32), W3 (32 x 128) = 51,104 floats (superres.c:124-154).
 * normPath: 4 float32 means then 4 float32 stds; a missing file means mean 0,
 * std 1.  NULL when the weights cannot be read, have a foreign magic or hold
 * fewer floats than the network needs, or on allocation failure. */
SuperResUpscaler *SR_Create(int coarseX, int coarseY, int coarseZ, const char *weightsPath, const char *normPath);
void SR_Free(SuperResUpscaler *sr);

/* superres.h:19.  coarseVel: device float4 per coarse cell (reference: a GL
 * buffer name).  The caller's producer must have finished writing it (for a
 * grid: LBM_Synchronize). */
void SR_Upscale(SuperResUpscaler *sr, const float *coarseVel);

/* superres.h:22: device float4 per fine cell (reference: a GL buffer name) */
float *SR_GetFineVelocityBuffer(SuperResUpscaler *sr);

/* superres.h:25 */
void SR_GetFineSize(SuperResUpscaler *sr, int *fineX, int *fineY, int *fineZ);

/* No counterpart in the reference: n back-to-back upscales timed with device
 * events on the upscaler's stream; *msPerCall gets the mean.  1 on success. */
int SR_Profile(SuperResUpscaler *sr, const float *coarseVel, int n, float *msPerCall);

#ifdef __cplusplus
}
#endif

#endif /* LBM_SUPERRES_H */
