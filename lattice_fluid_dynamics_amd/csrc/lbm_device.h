/*
 * Thin C-ABI between the C host library (lbm.c) and the HIP kernels
 * (lbm_kernels.hip).  Plain pointers and sizes only; the host side never
 * includes a HIP header.  All calls are ordered on one HIP stream per grid.
 *
 * Population storage (device): one buffer of 19 direction planes,
 * plane i at pop + i*stride, cell = x + NX*(y + NY*z).  Updated in place by
 * the two-phase (esoteric-pull) scheme; see lbm_kernels.hip.
 */
#ifndef LBM_DEVICE_H
#define LBM_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* population storage formats */
enum { LBMDEV_FP32 = 0, LBMDEV_FP16S = 1, LBMDEV_BF16S = 2, LBMDEV_FP64 = 3 };
/* streaming (non-temporal) population access of the step kernel, per phase */
enum { LBMDEV_NT_A = 1, LBMDEV_NT_B = 2 };

/* mask byte bits */
enum { LBMDEV_MASK_BODY = 1, LBMDEV_MASK_GROUND = 2, LBMDEV_MASK_OVR = 4 };

typedef struct LbmDev LbmDev;

typedef struct {
    float tau;
    float smagorinskyCs;
    float inletX, inletY, inletZ;
    int useRegularized, useMRT, useSmagorinsky, periodicYZ;
} LbmStepParams;

LbmDev *lbmdev_create(int nx, int ny, int nz, int storage, char *err, int errlen);
void lbmdev_destroy(LbmDev *d);
int lbmdev_storage(const LbmDev *d);
const char *lbmdev_last_error(void);

/* solid: int per cell (0/1/2).  ovr_cells: sorted cell indices of the
 * boundary cells whose incoming populations the stream-rule prepass computes
 * (for the directions with a link or a source outside the domain); ovr_q:
 * q[i*n_ovr + k] for those cells (-1 = no link); periodic: the y/z mode the
 * list was built for. */
int lbmdev_set_geometry(LbmDev *d, const int32_t *solid, int n_ovr, const int32_t *ovr_cells,
                        const float *ovr_q, int periodic);

/* f = f_new = feq (19 host-computed values) everywhere, vel = vel4. */
int lbmdev_init_flow(LbmDev *d, const float feq[19], const float vel4[4]);
/* equilibrium of a coarse (ux, uy, uz, rho) host field, trilinear upsampling */
int lbmdev_init_field(LbmDev *d, const float *coarse_host, int cnx, int cny, int cnz);
/* Bytes of the static staging array of lbmdev_init_field_noalloc: 4 MiB,
 * 262,144 coarse cells of (ux, uy, uz, rho). */
#define LBMDEV_NOALLOC_FIELD_BYTES (4 << 20)
/* As lbmdev_init_field without any device allocation: the coarse field is
 * staged in a static per-GPU array; fields above LBMDEV_NOALLOC_FIELD_BYTES
 * are rejected.  Calls on one GPU share the array and must not overlap. */
int lbmdev_init_field_noalloc(LbmDev *d, const float *coarse_host, int cnx, int cny, int cnz);

/* n LBM steps; the velocity buffer is written on the last one. */
int lbmdev_step(LbmDev *d, const LbmStepParams *p, int nsteps);

/* Momentum exchange over the boundary links (periodicYZ selects the stream
 * rule).  exact[6] = FP64 sums (F xyz, P xyz); fixed[7] = reference fixed
 * point accumulators (round(F*1e4) xyz, link-cell count, round(P*1e4) xyz). */
int lbmdev_force(LbmDev *d, int periodicYZ, double exact[6], long long fixed[7]);

float *lbmdev_velocity_ptr(LbmDev *d);
int32_t *lbmdev_solid_ptr(LbmDev *d);
int lbmdev_copy_velocity(LbmDev *d, float *host);
/* block means over f^3 cells (dimensions divisible by f) */
int lbmdev_copy_velocity_coarse(LbmDev *d, int f, float *host);
int lbmdev_copy_solid(LbmDev *d, int32_t *host);
/* Reference-layout views (f[i + 19*cell]): post-stream f and post-collision
 * f_new of the last step, as FP32 (either may be NULL). */
int lbmdev_export_populations(LbmDev *d, int periodicYZ, float *f_post_stream, float *f_post_collision);
/* Sum of rho over all cells in FP64 (mass check) from the post-stream state. */
int lbmdev_total_mass(LbmDev *d, int periodicYZ, double *mass);

int lbmdev_particles_seed(LbmDev *d, void *dev_particles, int n, uint32_t seed);
/* Uniforms of the live particle update (simulation/shaders/particle.comp,
 * set by main.c:1036-1096); the pointers are device memory. */
typedef struct {
    float wind[3], carMin[3], carMax[3], carCenter[3];
    int collisionMode, useLBM, vizMode;
    const float *triangles; /* 12 floats per triangle */
    int numTriangles;
    const int32_t *cellStart, *cellCount, *triIndices; /* the collision grid */
    int numIndices;                                     /* entries of triIndices */
    float gridMin[3], gridCellSize[3];
    int gridRes;
} LbmParticleFlow;
int lbmdev_particles_live(LbmDev *d, void *dev_particles, int n, float dt, const LbmParticleFlow *flow,
                          uint32_t seed, uint32_t step);
int lbmdev_particles_advect(LbmDev *d, void *dev_particles, int n, float dt, float velScale,
                            uint32_t seed, uint32_t step);
int lbmdev_trace_streamlines(LbmDev *d, void *dev_out, int nseeds, int len);

int lbmdev_synchronize(LbmDev *d);
/* Checkpoint state: the stored populations as the device holds them
 * (population_bytes) and the velocity buffer (16 B per cell), with the layout
 * tag of the last step; load restores all three. */
size_t lbmdev_population_bytes(const LbmDev *d);
int lbmdev_save_state(LbmDev *d, void *pop, void *vel, int *state);
int lbmdev_load_state(LbmDev *d, const void *pop, const void *vel, int state);
/* Times `nsteps` steps with HIP events on the grid's stream.  total_ms is the
 * whole region; main_ms / pre_ms are sums of per-launch event pairs around
 * the fused step kernel and the boundary prepass (either may be NULL). */
int lbmdev_profile_steps(LbmDev *d, const LbmStepParams *p, int nsteps, float *total_ms,
                         float *main_ms, float *pre_ms);
/* Steps with per-launch events recorded and not waited for; collect_timings
 * waits for the stream and returns the sums over all timed steps since the
 * last collection (launches = their number). */
int lbmdev_step_timed(LbmDev *d, const LbmStepParams *p, int nsteps);
int lbmdev_collect_timings(LbmDev *d, float *main_ms, float *pre_ms, int *launches);

int lbmdev_set_block_size(LbmDev *d, int bs);
/* Times `steps` steps per candidate workgroup size (128, 256, 512), restoring
 * the populations and velocity after each trial; sets and returns the best. */
int lbmdev_autotune(LbmDev *d, const LbmStepParams *p, int steps, float *ms_per_step);
int lbmdev_block_size(const LbmDev *d);
int lbmdev_set_streaming(LbmDev *d, int mode);
int lbmdev_streaming(const LbmDev *d);
void lbmdev_set_block_order(LbmDev *d, int xcd_contiguous);
int lbmdev_block_order(const LbmDev *d);
int lbmdev_export_populations_f64(LbmDev *d, int periodicYZ, double *f_post_stream, double *f_post_collision);
void *lbmdev_stream(LbmDev *d);

/* ---- batched tunnels: one allocation, one launch per step over all ---- */
typedef struct LbmBatchDev LbmBatchDev;
typedef struct {
    float smagorinskyCs;
    int useRegularized, useMRT, useSmagorinsky, periodicYZ;
} LbmBatchParams;

/* dims: 3 ints per tunnel; tau: 1 per tunnel; inlet: 3 per tunnel. */
LbmBatchDev *lbmbatch_create(int storage, int ntunnels, const int *dims, const float *tau, const float *inlet,
                             char *err, int errlen);
void lbmbatch_destroy(LbmBatchDev *b);
/* solid: concatenated per tunnel; n_ovr_per: list length per tunnel; ovr_cells:
 * concatenated local cell indices; ovr_q: q[i*Ntot + k] over the whole list. */
int lbmbatch_set_geometry(LbmBatchDev *b, const int32_t *solid, const int *n_ovr_per, const int32_t *ovr_cells,
                          const float *ovr_q, int periodic);
int lbmbatch_init_flow(LbmBatchDev *b, const float *feq, const float *vel4);
int lbmbatch_init_tunnel_field(LbmBatchDev *b, int t, const float *coarse_host, int cnx, int cny, int cnz);
int lbmbatch_step(LbmBatchDev *b, const LbmBatchParams *p, int nsteps, int use_graph);
int lbmbatch_profile(LbmBatchDev *b, const LbmBatchParams *p, int nsteps, int use_graph, float *ms);
int lbmbatch_set_block_size(LbmBatchDev *b, int bs);
int lbmbatch_autotune(LbmBatchDev *b, const LbmBatchParams *p, int steps, float *ms_per_step);
int lbmbatch_block_size(const LbmBatchDev *b);
int lbmbatch_set_streaming(LbmBatchDev *b, int mode);
int lbmbatch_streaming(const LbmBatchDev *b);
void lbmbatch_set_block_order(LbmBatchDev *b, int xcd_contiguous);
int lbmbatch_block_order(const LbmBatchDev *b);
int lbmbatch_force(LbmBatchDev *b, int periodic, double *exact, long long *fixed);
int lbmbatch_export(LbmBatchDev *b, int t, int periodic, double *f_post_stream, double *f_post_collision);
int lbmbatch_copy_velocity(LbmBatchDev *b, int t, float *host);
int lbmbatch_copy_velocity_coarse(LbmBatchDev *b, int t, int f, float *host);
int lbmbatch_synchronize(LbmBatchDev *b);
/* Cells reset by the stability guard since the flow was initialised (all
 * tunnels of a batch together); 1 on success. */
int lbmdev_guard_resets(LbmDev *d, long long *count);
int lbmbatch_guard_resets(LbmBatchDev *b, long long *count);

/* Read + write bandwidth of a plain copy of `bytes` bytes (mean of iters). */
int lbmdev_copy_bandwidth(size_t bytes, int iters, float *gbps, float *ms_per_copy);

void *lbmdev_malloc(size_t bytes);
void lbmdev_free(void *p);
int lbmdev_memcpy_h2d(void *dst, const void *src, size_t bytes);
int lbmdev_memcpy_d2h(void *dst, const void *src, size_t bytes);
int lbmdev_memcpy_d2d(void *dst, const void *src, size_t bytes);
int lbmdev_device_count(void);

/* lbm.c: records the message returned by LBM_GetLastError */
void lbm_note_error(const char *msg);

/* warmstart_kernels.hip: device helpers of the warm-start network */
int lbmnn_memset(void *p, int value, unsigned long long bytes);
int lbmnn_synchronize(void);
/* Zeroes n float buffers (n <= LBMNN_ZERO_MAX) with one asynchronous launch. */
#define LBMNN_ZERO_MAX 40
int lbmnn_zero_list(float *const *ptrs, const long long *counts, int n);
/* AdamW step of count <= LBMNN_ADAM_MAX tensors in one launch, each element
 * as lbmnn_adamw updates it (wb[k] may be NULL). */
#define LBMNN_ADAM_MAX 32
/* lbmnn_conv_pack of count <= LBMNN_PACK_MAX layers in one launch. */
#define LBMNN_PACK_MAX 16
int lbmnn_conv_pack_list(const void *const *W, void *const *Wm, void *const *Wp, void *const *Wmp, const int *cin,
                         const int *cout, int count);
int lbmnn_adamw_list(float *const *w, void *const *wb, const float *const *g, float *const *m, float *const *v,
                     const long long *n, const float *wd, int count, float lr, float b1, float b2, float eps, int t);
/* Pinned host memory (hipHostMalloc) and an asynchronous device-to-host
 * copy ordered after the work already queued by the network. */
void *lbmnn_host_alloc(size_t bytes);
void lbmnn_host_free(void *p);
int lbmnn_copy_to_host_async(void *host, const void *dev, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
