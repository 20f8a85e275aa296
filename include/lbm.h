/*
 * lbm.h -- D3Q19 lattice-Boltzmann wind tunnel for AMD Instinct MI355X
 * (gfx950), C ABI.
 *
 * Drop-in for the solver API of MarcosAsh/Lattice_Fluid_Dynamics
 * (simulation/lib/lbm.h:6-129): the same function names, argument order and
 * meaning, the same public LBMGrid field names for the flags callers poke
 * (tau, useRegularized, useMRT, useSmagorinsky, smagorinskyCs, periodicYZ),
 * the same D3Q19 direction order and the same world box (x in [-4,4],
 * y,z in [-2,2]).  The OpenGL buffer objects become HIP device pointers.
 *
 * Error behaviour follows the reference: LBM_Create returns NULL (and prints
 * a message) when allocation fails; LBM_InitializeFlow returns 0 on failure;
 * other calls return void.  LBM_GetLastError() additionally reports the last
 * device error text.  One HIP stream per grid orders all work; every call
 * that returns host-visible data synchronises that stream.  A grid is not
 * thread-safe; separate grids may be driven from separate threads.
 */
#ifndef LBM_H
#define LBM_H

#ifdef __cplusplus
extern "C" {
#endif

/* Population storage formats (LBMGrid.populationStorage). */
#define LBM_STORAGE_FP32 0  /* 76 B/cell, reference arithmetic */
#define LBM_STORAGE_FP16 1  /* 38 B/cell: f_i - w_i in FP16 (x 2^15), FP32 arithmetic */
#define LBM_STORAGE_BF16 2  /* 38 B/cell: f_i - w_i in BF16, FP32 arithmetic */
#define LBM_STORAGE_FP64 3  /* 152 B/cell: FP64 storage and arithmetic (oracle runs) */

/* Particle record of simulation/lib/particle_system.h:12-17 (32 bytes). */
typedef struct {
    float x, y, z;    /* position (world units) */
    float padding1;   /* |velocity| after the last advection */
    float vx, vy, vz; /* velocity (world units per time unit) */
    float life;
} Particle;

/* Replaces simulation/lib/lbm.h:6-56.  GL handles are gone; the velocity and
 * solid buffers are device pointers with the reference layouts. */
typedef struct {
    int sizeX, sizeY, sizeZ;
    int totalCells;

    float tau;

    float *velocityBuffer; /* device, float4 (ux, uy, uz, rho) per cell, x + NX*(y + NY*z) */
    int *solidBuffer;      /* device, int per cell: 0 fluid, 1 body, 2 ground */

    int useRegularized;  /* 0 = BGK, 1 = regularized */
    int useMRT;          /* 0 = off, 1 = MRT collision operator */
    int useSmagorinsky;  /* 0 = off, 1 = Smagorinsky SGS */
    float smagorinskyCs; /* Smagorinsky constant */
    int periodicYZ;      /* 0 = clamp, 1 = periodic y/z */

    /* Z-slab fields of the reference (GL SSBO size limit); always one slab. */
    int numChunks;
    int slabZ;
    int haloZ;

    /* Additions (no reference counterpart). */
    int forceFixedPoint;   /* 1 (default): accumulate the force as the reference does,
                              round(F*1e4) per boundary cell in integers; 0: FP64 sum of
                              the per-cell forces (LBM_ComputeDragForceF64 gives both) */
    int populationStorage; /* LBM_STORAGE_*; takes effect at LBM_InitializeFlow */
    long long stepCount;   /* LBM_Step calls since LBM_InitializeFlow */
    void *impl;            /* opaque */
} LBMGrid;

/* simulation/lib/lbm.h:59  tau = 0.5 + 3*viscosity clamped to [0.52, 2]. */
LBMGrid *LBM_Create(int sizeX, int sizeY, int sizeZ, float viscosity);

/* simulation/lib/lbm.h:62  NULL-safe. */
void LBM_Free(LBMGrid *grid);

/* simulation/lib/lbm.h:65-71  world AABB -> solid=1, q=0.5 on every fluid->solid link. */
void LBM_SetSolidAABB(LBMGrid *grid, float minX, float minY, float minZ, float maxX, float maxY,
                      float maxZ);

/* simulation/lib/lbm.h:74  f = f_new = feq(1, u), velocity = (u, 1). 1 on success, 0 on failure. */
int LBM_InitializeFlow(LBMGrid *grid, float ux, float uy, float uz);

/* No reference counterpart (warm starts): like LBM_InitializeFlow, but f =
 * f_new = feq(rho, u) of a (ux, uy, uz, rho) field given on a coarse
 * cell-centred lattice (host array, cnz x cny x cnx x 4 floats), trilinearly
 * interpolated to every cell on the device; velocity = the interpolated
 * field.  The inlet velocity passed to LBM_Step is unaffected. */
int LBM_InitializeFlowFromField(LBMGrid *grid, const float *field, int cnx, int cny, int cnz);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops that re-initialise tunnels from predicted fields,
 * harnesses that forbid hipMalloc): LBM_InitializeFlowFromField with the
 * same arguments, results and errors, except that it calls no device
 * allocator -- the coarse field is staged in a static per-GPU array of
 * 4 MiB, so fields above 262,144 coarse cells are rejected (0, with an
 * error).  It also refuses (0) what would allocate: a populationStorage
 * different from the grid's current device format, or geometry / periodicYZ
 * changed since the last LBM_InitializeFlow*; initialise once with an
 * allocating call first.  Calls on one GPU share the staging array: calls
 * that overlap on one GPU (grids driven from several threads) race. */
int LBM_InitializeFlowFromFieldNoAlloc(LBMGrid *grid, const float *field, int cnx, int cny, int cnz);

/* simulation/lib/lbm.h:77  one collide + stream step with the given inlet velocity. */
void LBM_Step(LBMGrid *grid, float inletVelX, float inletVelY, float inletVelZ);

/* simulation/lib/lbm.h:80  (GLuint -> device pointer) float4 per cell: the
 * pre-collision moments of the last step.  Work is queued on the grid's
 * stream: call LBM_Synchronize before reading it from another stream. */
float *LBM_GetVelocityBuffer(LBMGrid *grid);

/* simulation/lib/lbm.h:83-86  momentum-exchange force on the body. */
void LBM_ComputeDragForce(LBMGrid *grid, float *forceX, float *forceY, float *forceZ);

/* simulation/lib/lbm.h:90-96  force with the equilibrium ("pressure") part;
 * pressure outputs may be NULL.  Outputs are zeroed first. */
void LBM_ComputeDragForceDecomposed(LBMGrid *grid, float *forceX, float *forceY, float *forceZ,
                                    float *pressureX, float *pressureY, float *pressureZ);

/* simulation/lib/lbm.h:99-101  |Fx| / (0.5 U^2 A). */
float LBM_ComputeDragCoefficient(LBMGrid *grid, float inletVelocity, float refArea);

/* simulation/lib/lbm.h:104-106  |Fy| / (0.5 U^2 A). */
float LBM_ComputeLiftCoefficient(LBMGrid *grid, float inletVelocity, float refArea);

/* simulation/lib/lbm.h:109-110  sphere, analytic Bouzidi q. */
void LBM_SetSolidSphere(LBMGrid *grid, float cx, float cy, float cz, float radius);

/* simulation/lib/lbm.h:114  ground cells below worldZ become solid=2. */
void LBM_AddGroundPlane(LBMGrid *grid, float worldZ);

/* simulation/lib/lbm.h:118  columns holding a solid==1 cell along axis (0=x, 1=y, 2=z). */
float LBM_ComputeProjectedArea(LBMGrid *grid, int axis);

/* simulation/lib/lbm.h:121-129  triangles: 12 floats each (v0 xyz pad, v1 xyz pad,
 * v2 xyz pad; simulation/lib/model_bounds.h:14-18); solid by +X ray parity,
 * q by ray casting along the links. */
void LBM_SetSolidMesh(LBMGrid *grid, float *triangles, int numTriangles, float minX, float minY,
                      float minZ, float maxX, float maxY, float maxZ);

/* ---- additions replacing direct GL access by callers --------------------- */

/* n LBM_Step calls with the same inlet velocity.  Observable state afterwards
 * is identical to the n single calls (the velocity buffer, which every step
 * overwrites, is written on the last one only). */
void LBM_StepN(LBMGrid *grid, int n, float inletVelX, float inletVelY, float inletVelZ);

/* Replaces glGetBufferSubData on velocityBuffer (test_lbm.c:503-505,
 * vti_export.c:30-34): 4 floats per cell.  Returns 1 on success. */
int LBM_CopyVelocityToHost(LBMGrid *grid, float *out);
/* No reference counterpart (warm-start data): block means of the velocity
 * buffer over factor^3 cells, (sizeZ/f) x (sizeY/f) x (sizeX/f) x 4 floats. */
int LBM_CopyCoarseVelocityToHost(LBMGrid *grid, int factor, float *out);

/* Replaces glGetBufferSubData on solidBuffer (test_lbm.c:258-260). */
int LBM_CopySolidToHost(LBMGrid *grid, int *out);

/* Replaces glGetBufferSubData on the Bouzidi q buffer (simulation/src/lbm.c:
 * 233-250 layout): q[i + 19*cell], -1 where cell has no link in direction i. */
int LBM_CopyQToHost(LBMGrid *grid, float *out);

/* The reference's two population buffers in their layout f[i + 19*cell]:
 * post-stream f and post-collision f_new of the last step (either may be NULL). */
int LBM_CopyPopulationsToHost(LBMGrid *grid, float *f, float *fNew);
/* The same in double precision (exact for every storage format). */
int LBM_CopyPopulationsToHostF64(LBMGrid *grid, double *f, double *fNew);

/* Sum of all post-stream populations (total mass) accumulated in FP64. */
double LBM_ComputeTotalMass(LBMGrid *grid);

/* FP64 force sums: out[0..2] total force, out[3..5] equilibrium part. */
void LBM_ComputeDragForceF64(LBMGrid *grid, double out[6]);

/* Particles (the reference's particle dispatch is inline GL in main.c:1012-1101).
 * `particles` is device memory (LBM_DeviceAlloc).  Seeding puts n particles on
 * the inlet plane x=-4 with y,z uniform in (-2,2) from `seed`.  Advection is
 * midpoint RK2 through the trilinear velocity sampler of
 * particle_lbm.comp:46-90, displacement = dt * velScale * u_lattice; particles
 * leaving the box respawn on the inlet plane from (seed, step). */
void LBM_SeedParticles(LBMGrid *grid, Particle *particles, int n, unsigned int seed);
void LBM_AdvectParticles(LBMGrid *grid, Particle *particles, int n, float dt, float velScale,
                         unsigned int seed, unsigned int step);

/* The particle update of the reference's running application, which is
 * inline GL in main.c:1012-1101 (no C function; simulation/shaders/
 * particle.comp:218-407): per particle, the velocity of the nearest cell
 * (grid position clamped to [0.5, N - 1.5]) scaled by 8 is blended with the
 * wind, flow = mix(wind, 8 u, 0.5); the particle velocity relaxes towards it,
 * v = mix(v, flow, 0.05), with v.x >= 0.3 wind.x outside the model's box
 * grown by 0.05; x += v dt; then the collision mode; then the kill box (y, z
 * clamped to +-1.3; respawn on x = -4 past carMax.x + 2, when slower than
 * 0.05 beside the model, or when killed by a collision).  padding1 tracks
 * the flow speed (or Cp with vizMode 8).  Respawn positions and speeds come
 * from the counter hash of (seed, step, index) instead of the reference's
 * time-seeded hash.  The uniforms are those main.c:1036-1096 sets. */
typedef struct {
    float windX, windY, windZ; /* main.c:1043: (windSpeed * 0.5, 0, 0) */
    float carMinX, carMinY, carMinZ, carMaxX, carMaxY, carMaxZ, carCenterX, carCenterY, carCenterZ;
    int collisionMode; /* 0 none, 1 the model's box, 2 its triangles, 3 the grid's body cells */
    int useLBM;        /* 1: sample the grid's velocity buffer; 0: the wind only */
    int vizMode;       /* 8: padding1 follows Cp from the density instead of the speed */
    /* collisionMode 2: device copies of the triangles (12 floats each, as
     * LBM_SetSolidMesh takes them, 16-byte aligned as LBM_DeviceAlloc returns
     * them) and of buildCollisionGrid's arrays
     * (include/lbm_model.h), its origin, cell size and resolution */
    const float *triangles;
    int numTriangles;
    const int *cellStart, *cellCount, *triIndices;
    int numIndices;
    float gridMinX, gridMinY, gridMinZ, gridCellSizeX, gridCellSizeY, gridCellSizeZ;
    int gridRes;
} LBMParticleFlow;
void LBM_AdvectParticlesLive(LBMGrid *grid, Particle *particles, int n, float dt, const LBMParticleFlow *flow,
                             unsigned int seed, unsigned int step);

/* Streamlines of the velocity buffer, the application's visualisation mode 9
 * (inline GL in main.c:1120-1140 dispatching simulation/shaders/
 * streamline_trace.comp with STREAMLINE_SEEDS 256 and STREAMLINE_LEN 64,
 * main.c:775-776).  Seed s starts at x = -2.5 on a gridSide x gridSide
 * lattice of the y-z cross-section (gridSide = ceil(sqrt(numSeeds)),
 * y, z = -1.5 + 3 (i + 0.5) / gridSide); each of its traceLen vertices is
 * (x, y, z, 8 |u|) before one classic RK4 step of ds = 8 / traceLen through
 * the trilinear sampler clamped to [0.5, N - 1.5] with velocities scaled by
 * 8; a step shorter than 1e-6 ends the line (the remaining vertices repeat
 * the position with speed 0); positions stay in the box [-4, 4] x [-2, 2]^2.
 * `streamlines` is device memory (LBM_DeviceAlloc) of numSeeds * traceLen
 * float4, the GL buffer the reference binds at binding 11.  Returns 1, or 0
 * with LBM_GetLastError set (no grid or output, numSeeds or traceLen below
 * 1, a grid dimension below 2). */
int LBM_TraceStreamlines(LBMGrid *grid, float *streamlines, int numSeeds, int traceLen);

void *LBM_DeviceAlloc(unsigned long long bytes);
void LBM_DeviceFree(void *p);
int LBM_CopyToDevice(void *dst, const void *src, unsigned long long bytes);
int LBM_CopyToHost(void *dst, const void *src, unsigned long long bytes);
int LBM_CopyDeviceToDevice(void *dst, const void *src, unsigned long long bytes);

/* Waits for all queued work of the grid. */
void LBM_Synchronize(LBMGrid *grid);

/* Runs n steps and reports HIP-event times on the grid's stream: total_ms for
 * the whole region, main_ms and boundary_ms summed over the fused step kernel
 * and the boundary prepass launches (NULL to skip per-launch events).  Work
 * queued before the call is not waited for (the events are stream-ordered);
 * the call returns once its last launch has finished. */
int LBM_ProfileSteps(LBMGrid *grid, int n, float inletVelX, float inletVelY, float inletVelZ,
                     float *total_ms, float *main_ms, float *boundary_ms);

/* No reference counterpart (timing a sample of the steps of a long run without
 * host waits inside it): LBM_StepN with HIP events recorded on the grid's
 * stream around every boundary prepass and fused step kernel of the n steps,
 * not waited for.  LBM_CollectStepTimings waits for the grid's queued work and
 * returns main_ms and boundary_ms summed over the launches of every timed step
 * since the last collection, and their number in launches (any output may be
 * NULL).  At most 65536 timed steps may wait for a collection.  Both return 1,
 * or 0 with LBM_GetLastError set. */
int LBM_StepNTimed(LBMGrid *grid, int n, float inletVelX, float inletVelY, float inletVelZ);
int LBM_CollectStepTimings(LBMGrid *grid, float *main_ms, float *boundary_ms, int *launches);

/* No reference counterpart (checkpoint and resume of long runs; the reference
 * keeps no solver state across runs).  LBM_SaveState writes into a caller
 * buffer of LBM_GetStateSize bytes a 64-byte header (magic "LBMSTAT1",
 * dimensions, storage format, population layout of the last step,
 * periodicYZ, stepCount, a hash of the geometry) followed by the stored
 * populations exactly as the device holds them and the velocity buffer.
 * LBM_LoadState restores it into a grid of the same dimensions,
 * populationStorage, periodicYZ and geometry (set by the same LBM_SetSolid* /
 * LBM_AddGroundPlane calls), replacing LBM_InitializeFlow; stepping then
 * continues bit for bit as the saved grid would.  Both return 1, or 0 with
 * LBM_GetLastError set (no flow, a short buffer, a mismatch); the collision
 * flags and tau are the caller's, as for any step. */
long long LBM_GetStateSize(LBMGrid *grid);
int LBM_SaveState(LBMGrid *grid, void *buffer, long long bytes);
int LBM_LoadState(LBMGrid *grid, const void *buffer, long long bytes);

/* Number of cells in the compacted boundary list used by the prepass. */
int LBM_GetBoundaryCellCount(LBMGrid *grid);

/* No reference counterpart (diagnostics for long runs): the number of cell
 * updates the stability guard of lbm_collide.comp:429-440 reset to rest
 * (rho < 0.3 or > 2, |u| > 0.4, or a non-finite rho) since the last
 * LBM_InitializeFlow*, counted on the device; LBM_Autotune's trials do not
 * count.  Synchronises the grid's stream; -1 on error. */
long long LBM_GetGuardResetCount(LBMGrid *grid);

/* Text of the last device or argument error ("" if none). */
const char *LBM_GetLastError(void);

/* ---- caller-side drag bookkeeping (simulation/src/main.c:1250-1391) ------- */

typedef struct LBMCdMonitor LBMCdMonitor;

/* U: reference velocity; refArea: LBM_ComputeProjectedArea(grid, 0).  The
 * blockage factor (1 - A/(NY NZ))^2 of main.c:549-552 is taken from the grid.
 * verbose != 0 prints the reference's "Drag Force: ... Cd=..." lines. */
LBMCdMonitor *LBM_CdMonitorCreate(LBMGrid *grid, float U, float refArea, int verbose);
void LBM_CdMonitorFree(LBMCdMonitor *m);
/* Computes the force and out[6] = {Cd, Cd_corr, Cl (signed), Cd_pressure,
 * Cd_friction, EMA of Cd_corr}; returns 1 when this sample first meets the
 * convergence test (relative std of the last 100 Cd samples < 1 %). */
int LBM_CdMonitorSample(LBMCdMonitor *m, LBMGrid *grid, float out[6]);
/* main.c:439-552 and 1273-1306: the model's world bounding box and Reynolds
 * number; the next verbose sample prints the reference's one-time "Cd
 * calculation breakdown" (Re, body size in cells, projected and bbox area,
 * blockage and its correction, resolution and confinement warnings).  A
 * projected area below 1 falls back to the bbox area, as main.c:548-549. */
void LBM_CdMonitorSetBody(LBMCdMonitor *m, LBMGrid *grid, float reynolds, float minX, float minY, float minZ,
                          float maxX, float maxY, float maxZ);
/* main.c:524-537: Re = U L / nu with L the body length in cells and
 * nu = (tau - 1/2)/3; print != 0 writes the "Effective Re = ..." line. */
float LBM_EffectiveReynolds(LBMGrid *grid, float U, float minX, float maxX, int print);
int LBM_CdMonitorConverged(const LBMCdMonitor *m, float *mean, float *relStd);
/* Strouhal number of the monitor's Cl series (simulation/src/drag_metrics.c:8-68),
 * one sample per sampleInterval lattice steps. */
float LBM_CdMonitorStrouhal(const LBMCdMonitor *m, int sampleInterval, float charLength);
float LBM_ComputeStrouhal(const float *cl, int n, int sampleInterval, float charLength, float U);

/* VTK ImageData dump of velocity, rho and solid to <dir>/field_<step>.vti
 * (simulation/src/vti_export.c:10-116).  Returns 1 on success. */
int LBM_WriteVTI(LBMGrid *grid, const char *dir, int step);

/* ---- workgroup-size autotuner (no reference counterpart) ------------------ */

/* Times `steps` steps of the grid for each candidate configuration of the
 * step kernel: workgroup size 128, 256 or 512 lanes, each with streaming
 * mode 0, 1 and 3 (see LBM_SetStreamingAccess; FP64 storage: mode 0 only)
 * and each in both workgroup orders (see LBM_SetBlockOrder).  Uses the given
 * inlet velocity, restores the simulation state after each trial, keeps the
 * fastest configuration and returns its workgroup size (0 on failure).
 * msPerStep (3 floats, may be NULL) receives the best time per workgroup
 * size. */
int LBM_Autotune(LBMGrid *grid, int steps, float inletVelX, float inletVelY, float inletVelZ,
                 float *msPerStep);

/* Sets the step kernel's workgroup size (128, 256 or 512) instead of
 * tuning it; the results do not depend on it.  1 on success.  This and the
 * streaming mode below belong to the device state of the current storage
 * format: an LBM_InitializeFlow that switches populationStorage starts again
 * from the defaults (256 lanes, mode 0), so set them after it. */
int LBM_SetBlockSize(LBMGrid *grid, int blockSize);

/* Selects the cache policy of the step kernel's population loads and stores
 * instead of tuning it: mode bit 0 makes them non-temporal in phase A (own
 * slots), bit 1 in phase B (neighbour slots); 0 is the default policy.  FP64
 * storage keeps the default.  The results do not depend on it.  1 on success,
 * 0 for a mode outside 0-3. */
int LBM_SetStreamingAccess(LBMGrid *grid, int mode);

/* Workgroup order of the step kernel instead of tuning it: 1 gives each of
 * the 8 XCDs one contiguous range of cells, 0 lets consecutive workgroups
 * go to different XCDs (the dispatch order).  The results do not depend on
 * it.  1 on success. */
int LBM_SetBlockOrder(LBMGrid *grid, int xcdContiguous);

/* The step kernel's current workgroup size, streaming mode and workgroup
 * order (any pointer may be NULL).  1 on success. */
int LBM_GetKernelConfig(const LBMGrid *grid, int *blockSize, int *streamingMode, int *xcdContiguous);

/* Attainable HBM bandwidth in GB/s: read + write bytes of a plain device copy
 * of `bytes` bytes over its mean launch time (iters launches). */
float LBM_MeasureCopyBandwidth(unsigned long long bytes, int iters);

/* ---- batched parameter sweeps (no reference counterpart) ------------------ */

typedef struct LBMBatch LBMBatch;

/* All tunnels of a batch share the collision operator, Smagorinsky constant,
 * Y/Z boundary mode and population storage format. */
LBMBatch *LBM_BatchCreate(int populationStorage, int useRegularized, int useMRT, int useSmagorinsky,
                          float smagorinskyCs, int periodicYZ);
void LBM_BatchFree(LBMBatch *batch);
/* Copies the solid mask and Bouzidi links set on `geometry` (LBM_SetSolid*,
 * LBM_AddGroundPlane); the grid may be freed afterwards.  Each tunnel has its
 * own size, tau and inlet velocity.  Returns the tunnel index (-1 on error,
 * or after LBM_BatchInitialize). */
int LBM_BatchAddTunnel(LBMBatch *batch, LBMGrid *geometry, float tau, float inletX, float inletY,
                       float inletZ);
/* Allocates every tunnel in one device buffer and sets f = feq(1, u_inlet). */
int LBM_BatchInitialize(LBMBatch *batch);
/* Replaces tunnel t's uniform initial state with the equilibrium of a coarse
 * field, as LBM_InitializeFlowFromField; only between LBM_BatchInitialize
 * and the first LBM_BatchStep. */
int LBM_BatchInitializeTunnelFromField(LBMBatch *batch, int tunnel, const float *field, int cnx, int cny,
                                       int cnz);
/* n steps of every tunnel: one launch per step over all tunnels (plus the
 * boundary prepass), replayed from HIP graphs two steps at a time.  Each
 * tunnel evolves exactly as a single grid with the same settings would. */
void LBM_BatchStep(LBMBatch *batch, int n);
int LBM_BatchProfile(LBMBatch *batch, int n, int useGraph, float *ms);
int LBM_BatchAutotune(LBMBatch *batch, int steps, float *msPerStep);
int LBM_BatchSetBlockSize(LBMBatch *batch, int blockSize);
/* As LBM_SetStreamingAccess, LBM_SetBlockOrder and LBM_GetKernelConfig, for
 * the batch step kernel (LBM_BatchAutotune tunes all three).  Available
 * after LBM_BatchInitialize. */
int LBM_BatchSetStreamingAccess(LBMBatch *batch, int mode);
int LBM_BatchSetBlockOrder(LBMBatch *batch, int xcdContiguous);
int LBM_BatchGetKernelConfig(const LBMBatch *batch, int *blockSize, int *streamingMode, int *xcdContiguous);
int LBM_BatchTunnelCount(const LBMBatch *batch);
/* As LBM_GetGuardResetCount, summed over all tunnels of the batch. */
long long LBM_BatchGuardResetCount(LBMBatch *batch);
long long LBM_BatchTotalCells(const LBMBatch *batch);
/* out: 6 doubles per tunnel (F xyz, pressure part xyz, FP64 sums); fixedOut
 * (may be NULL): the 7 reference fixed-point accumulators per tunnel. */
int LBM_BatchComputeForcesF64(LBMBatch *batch, double *out, long long *fixedOut);
/* out: 4 floats per tunnel: Cd, Cl (signed), Cd_pressure, Cd_friction. */
int LBM_BatchComputeCoefficients(LBMBatch *batch, float *out);
int LBM_BatchCopyPopulationsToHostF64(LBMBatch *batch, int tunnel, double *f, double *fNew);
int LBM_BatchCopyVelocityToHost(LBMBatch *batch, int tunnel, float *out);
/* Block means of the velocity buffer (ux, uy, uz, rho) over factor^3 cells:
 * (nz/f) x (ny/f) x (nx/f) x 4 floats; the dimensions must divide by f. */
int LBM_BatchCopyCoarseVelocityToHost(LBMBatch *batch, int tunnel, int factor, float *out);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
