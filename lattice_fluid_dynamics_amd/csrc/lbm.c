/*
 * Host side of the MI355X D3Q19 solver: the C API of include/lbm.h.
 *
 * This file plays the role of the reference's simulation/src/lbm.c: grid
 * creation, solid/Bouzidi setup on the host, equilibrium initialisation,
 * step dispatch and the drag/lift post-processing.  All device work goes
 * through the thin C ABI of lbm_device.h (HIP kernels in lbm_kernels.hip).
 *
 * Host geometry state: solid[cell] (0 fluid, 1 body, 2 ground) and a sparse
 * q table (qRow[cell] -> 19 floats, -1 = no link) replacing the reference's
 * dense 19-float-per-cell q buffer.  Before a step the host compacts the
 * boundary cells (cells with links, plus the Y/Z faces when the Y/Z
 * boundaries clamp, plus body cells on the x faces) into the list that the
 * device prepass evaluates the reference stream rule on.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm.h"
#include "lbm_device.h"
#include "lbm_mesh.h"

/* D3Q19 velocity set, reference order (simulation/src/lbm.c:9-35). */
static const int kEx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int kEy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int kEz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

typedef struct {
    LbmDev *dev;
    int32_t *solid;
    int32_t *qRow;
    float *qRows;
    int nRows, capRows;
    int geomDirty;
    int builtPeriodic;
    int nBoundary;
    int flowReady;
} Impl;

static char g_lastError[512];

static void note_error(const char *msg) {
    snprintf(g_lastError, sizeof(g_lastError), "%s", msg);
    printf("ERROR: %s\n", msg);
}

const char *LBM_GetLastError(void) { return g_lastError; }

/* For the other host-library files (warm start, batches). */
void lbm_note_error(const char *msg) { note_error(msg); }

static Impl *impl_of(LBMGrid *g) { return (Impl *)g->impl; }

static long cell_index(const LBMGrid *g, int x, int y, int z) {
    return x + (long)g->sizeX * (y + (long)g->sizeY * z);
}

/* Host equilibrium in float, the expression of simulation/src/lbm.c:37-44:
 * w rho (1 + eu/cs2 + eu^2/(2 cs2^2) - u2/(2 cs2)). */
static float host_weight(int i) { return i == 0 ? 1.0f / 3.0f : (i < 7 ? 1.0f / 18.0f : 1.0f / 36.0f); }

static float host_feq(int i, float rho, float ux, float uy, float uz) {
    const float cs2 = 1.0f / 3.0f;
    float eu = kEx[i] * ux + kEy[i] * uy + kEz[i] * uz;
    float uu = ux * ux + uy * uy + uz * uz;
    return host_weight(i) * rho * (1.0f + eu / cs2 + (eu * eu) / (2.0f * cs2 * cs2) - uu / (2.0f * cs2));
}

/* ------------------------------------------------------------ q table */

static void q_reset(Impl *im, long ncell) {
    for (long c = 0; c < ncell; c++) im->qRow[c] = -1;
    im->nRows = 0;
}

static float *q_row_for(Impl *im, long cell, int create) {
    int r = im->qRow[cell];
    if (r >= 0) return im->qRows + 19L * r;
    if (!create) return NULL;
    if (im->nRows == im->capRows) {
        int cap = im->capRows ? 2 * im->capRows : 4096;
        float *p = (float *)realloc(im->qRows, sizeof(float) * 19L * cap);
        if (!p) return NULL;
        im->qRows = p;
        im->capRows = cap;
    }
    r = im->nRows++;
    im->qRow[cell] = r;
    float *row = im->qRows + 19L * r;
    for (int i = 0; i < 19; i++) row[i] = -1.0f;
    return row;
}

static float q_get(Impl *im, long cell, int i) {
    int r = im->qRow[cell];
    return r >= 0 ? im->qRows[19L * r + i] : -1.0f;
}

static void q_set(Impl *im, long cell, int i, float q) {
    float *row = q_row_for(im, cell, 1);
    if (row) row[i] = q;
}

/* ------------------------------------------------------------ create / free */

LBMGrid *LBM_Create(int sizeX, int sizeY, int sizeZ, float viscosity) {
    if (sizeX < 2 || sizeY < 1 || sizeZ < 1) {
        note_error("LBM_Create: grid dimensions must be at least 2x1x1");
        return NULL;
    }
    long long cells = (long long)sizeX * sizeY * sizeZ;
    if (cells > 0x7fffff00LL) {
        note_error("LBM_Create: grid exceeds 2^31 cells");
        return NULL;
    }
    LBMGrid *g = (LBMGrid *)calloc(1, sizeof(LBMGrid));
    Impl *im = (Impl *)calloc(1, sizeof(Impl));
    if (!g || !im) {
        free(g);
        free(im);
        note_error("LBM_Create: host allocation failed");
        return NULL;
    }
    g->impl = im;
    g->sizeX = sizeX;
    g->sizeY = sizeY;
    g->sizeZ = sizeZ;
    g->totalCells = (int)cells;
    /* simulation/src/lbm.c:56-60 */
    g->tau = 0.5f + 3.0f * viscosity;
    if (g->tau < 0.52f) g->tau = 0.52f;
    if (g->tau > 2.0f) g->tau = 2.0f;
    g->useRegularized = 0;
    g->useMRT = 0;
    g->useSmagorinsky = 0;
    g->smagorinskyCs = 0.1f;
    g->periodicYZ = 0;
    g->numChunks = 1;
    g->slabZ = sizeZ;
    g->haloZ = 1;
    g->forceFixedPoint = 1;
    g->populationStorage = LBM_STORAGE_FP32;
    printf("LBM Grid: %dx%dx%d (%d cells), tau=%.3f\n", sizeX, sizeY, sizeZ, g->totalCells, g->tau);

    im->solid = (int32_t *)calloc((size_t)cells, sizeof(int32_t));
    im->qRow = (int32_t *)malloc((size_t)cells * sizeof(int32_t));
    if (!im->solid || !im->qRow) {
        note_error("LBM_Create: host allocation failed");
        LBM_Free(g);
        return NULL;
    }
    q_reset(im, (long)cells);
    im->builtPeriodic = -1;
    im->geomDirty = 1;

    char err[256] = {0};
    /* Grids beyond 2^30 cells cannot hold FP32 populations (32-bit slot
     * offsets): their buffers start in FP16S, and LBM_InitializeFlow fails
     * unless populationStorage is set to a 16-bit or FP64 format first. */
    const int initial = cells > (1LL << 30) ? LBM_STORAGE_FP16 : g->populationStorage;
    im->dev = lbmdev_create(sizeX, sizeY, sizeZ, initial, err, sizeof(err));
    if (!im->dev) {
        char msg[400];
        snprintf(msg, sizeof(msg), "LBM_Create: %s", err);
        note_error(msg);
        LBM_Free(g);
        return NULL;
    }
    g->velocityBuffer = lbmdev_velocity_ptr(im->dev);
    g->solidBuffer = lbmdev_solid_ptr(im->dev);
    printf("GPU memory: populations %.1f MB (in place), velocity %.1f MB\n",
           19.0 * 4.0 * (double)cells / 1048576.0, 16.0 * (double)cells / 1048576.0);
    return g;
}

void LBM_Free(LBMGrid *grid) {
    if (!grid) return;
    Impl *im = impl_of(grid);
    if (im) {
        lbmdev_destroy(im->dev);
        free(im->solid);
        free(im->qRow);
        free(im->qRows);
        free(im);
    }
    free(grid);
}

/* ------------------------------------------------------------ geometry */

/* World box x in [-4,4], y,z in [-2,2] mapped onto the lattice
 * (simulation/src/lbm.c:352-354); Scale and cell_centre are in lbm_mesh.h. */
static Scale world_scale(const LBMGrid *g) {
    Scale s = {g->sizeX / 8.0f, g->sizeY / 4.0f, g->sizeZ / 4.0f};
    return s;
}

/* Visits every in-range link from a cell with solid != skipValue toward a
 * neighbour that satisfies `isTarget`, in cell order then direction order. */
typedef float (*LinkQ)(void *ctx, int x, int y, int z, int i);

static void mark_links(LBMGrid *g, int (*isSource)(int32_t), int (*isTarget)(int32_t), LinkQ fq,
                       void *ctx) {
    Impl *im = impl_of(g);
    for (int z = 0; z < g->sizeZ; z++)
        for (int y = 0; y < g->sizeY; y++)
            for (int x = 0; x < g->sizeX; x++) {
                long c = cell_index(g, x, y, z);
                if (!isSource(im->solid[c])) continue;
                for (int i = 1; i < 19; i++) {
                    int X = x + kEx[i], Y = y + kEy[i], Z = z + kEz[i];
                    if (X < 0 || X >= g->sizeX || Y < 0 || Y >= g->sizeY || Z < 0 || Z >= g->sizeZ) continue;
                    if (!isTarget(im->solid[cell_index(g, X, Y, Z)])) continue;
                    q_set(im, c, i, fq(ctx, x, y, z, i));
                }
            }
}

static int not_body(int32_t s) { return s != 1; }
static int is_body(int32_t s) { return s == 1; }
static int not_fluid(int32_t s) { return s != 0; }
static float q_half(void *ctx, int x, int y, int z, int i) {
    (void)ctx, (void)x, (void)y, (void)z, (void)i;
    return 0.5f;
}

static void geometry_changed(LBMGrid *g) {
    impl_of(g)->geomDirty = 1;
}

/* simulation/src/lbm.c:346-435 */
void LBM_SetSolidAABB(LBMGrid *g, float minX, float minY, float minZ, float maxX, float maxY, float maxZ) {
    if (!g) return;
    Impl *im = impl_of(g);
    Scale s = world_scale(g);
    int lo[3] = {(int)((minX + 4.0f) * s.sx), (int)((minY + 2.0f) * s.sy), (int)((minZ + 2.0f) * s.sz)};
    int hi[3] = {(int)((maxX + 4.0f) * s.sx), (int)((maxY + 2.0f) * s.sy), (int)((maxZ + 2.0f) * s.sz)};
    int n[3] = {g->sizeX, g->sizeY, g->sizeZ};
    for (int a = 0; a < 3; a++) {
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] >= n[a]) hi[a] = n[a] - 1;
    }
    printf("LBM solid AABB: grid [%d-%d, %d-%d, %d-%d]\n", lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
    memset(im->solid, 0, sizeof(int32_t) * (size_t)g->totalCells);
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) im->solid[cell_index(g, x, y, z)] = 1;
    q_reset(im, g->totalCells);
    mark_links(g, not_body, is_body, q_half, NULL);
    geometry_changed(g);
}

typedef struct {
    Scale s;
    float cx, cy, cz, r;
} SphereCtx;


/* Nearest root of |o + t d|^2 = r^2 with t in (0,1], else 0.5
 * (simulation/src/lbm.c:492-515). */
static float sphere_q(void *vctx, int x, int y, int z, int i) {
    SphereCtx *c = (SphereCtx *)vctx;
    float wx, wy, wz;
    cell_centre(c->s, x, y, z, &wx, &wy, &wz);
    float dx = (float)kEx[i] / c->s.sx, dy = (float)kEy[i] / c->s.sy, dz = (float)kEz[i] / c->s.sz;
    float ox = wx - c->cx, oy = wy - c->cy, oz = wz - c->cz;
    float a = dx * dx + dy * dy + dz * dz;
    float b = 2.0f * (ox * dx + oy * dy + oz * dz);
    float cc = ox * ox + oy * oy + oz * oz - c->r * c->r;
    float disc = b * b - 4.0f * a * cc;
    float q = 0.5f;
    if (disc >= 0.0f) {
        float t = (-b - sqrtf(disc)) / (2.0f * a);
        if (t > 0.0f && t <= 1.0f) q = t;
    }
    return clamp_q(q);
}

/* simulation/src/lbm.c:437-535 */
void LBM_SetSolidSphere(LBMGrid *g, float cx, float cy, float cz, float radius) {
    if (!g) return;
    Impl *im = impl_of(g);
    SphereCtx ctx = {world_scale(g), cx, cy, cz, radius};
    int count = 0;
    for (int z = 0; z < g->sizeZ; z++)
        for (int y = 0; y < g->sizeY; y++)
            for (int x = 0; x < g->sizeX; x++) {
                float wx, wy, wz;
                cell_centre(ctx.s, x, y, z, &wx, &wy, &wz);
                float dx = wx - cx, dy = wy - cy, dz = wz - cz;
                int in = dx * dx + dy * dy + dz * dz <= radius * radius;
                im->solid[cell_index(g, x, y, z)] = in;
                count += in;
            }
    printf("LBM solid sphere: %d cells (R=%.2f at (%.2f,%.2f,%.2f))\n", count, radius, cx, cy, cz);
    q_reset(im, g->totalCells);
    mark_links(g, not_body, is_body, sphere_q, &ctx);
    geometry_changed(g);
}

/* simulation/src/lbm.c:537-614 */
void LBM_AddGroundPlane(LBMGrid *g, float worldZ) {
    if (!g) return;
    Impl *im = impl_of(g);
    int gz = (int)((worldZ + 2.0f) * (g->sizeZ / 4.0f));
    if (gz < 0) gz = 0;
    if (gz >= g->sizeZ) gz = g->sizeZ - 1;
    int groundCells = 0, links = 0;
    for (int z = 0; z <= gz; z++)
        for (int y = 0; y < g->sizeY; y++)
            for (int x = 0; x < g->sizeX; x++) {
                long c = cell_index(g, x, y, z);
                if (im->solid[c] == 0) {
                    im->solid[c] = 2;
                    groundCells++;
                }
            }
    const int za = gz + 1;
    if (za < g->sizeZ) {
        for (int y = 0; y < g->sizeY; y++)
            for (int x = 0; x < g->sizeX; x++) {
                long c = cell_index(g, x, y, za);
                if (im->solid[c] != 0) continue;
                for (int i = 1; i < 19; i++) {
                    if (kEz[i] != -1) continue;
                    int X = x + kEx[i], Y = y + kEy[i];
                    if (X < 0 || X >= g->sizeX || Y < 0 || Y >= g->sizeY) continue;
                    if (im->solid[cell_index(g, X, Y, gz)] != 0 && q_get(im, c, i) < 0.0f) {
                        q_set(im, c, i, 0.5f);
                        links++;
                    }
                }
            }
    }
    printf("Ground plane: z_lattice=%d, %d cells, %d boundary links\n", gz, groundCells, links);
    geometry_changed(g);
}

/* simulation/src/lbm.c:616-668 */
float LBM_ComputeProjectedArea(LBMGrid *g, int axis) {
    if (!g) return 0.0f;
    Impl *im = impl_of(g);
    int n[3] = {g->sizeX, g->sizeY, g->sizeZ};
    int a = axis == 0 ? 0 : (axis == 1 ? 1 : 2);
    int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    int count = 0;
    for (int j = 0; j < n[c]; j++)
        for (int k = 0; k < n[b]; k++) {
            for (int t = 0; t < n[a]; t++) {
                int p[3];
                p[a] = t;
                p[b] = k;
                p[c] = j;
                if (im->solid[cell_index(g, p[0], p[1], p[2])] == 1) {
                    count++;
                    break;
                }
            }
        }
    return (float)count;
}

/* ---- triangle mesh ---------------------------------------------------------
 * Semantics in lbm_mesh.h (simulation/src/lbm.c:964-1223).  The host builds
 * the (y, z) bin index of the triangles; the GPU marks the solid cells,
 * finds the cells with links and evaluates their q (mesh_kernels.hip), with
 * the float expressions the host would use. */
static int mesh_index_build(MeshIndex *m, const float *tri, int ntri) {
    memset(m, 0, sizeof(*m));
    m->tri = tri;
    m->ntri = ntri;
    m->bounds = (float *)malloc(sizeof(float) * 5 * (size_t)(ntri > 0 ? ntri : 1));
    if (!m->bounds) return 0;
    float ymin = INFINITY, ymax = -INFINITY, zmin = INFINITY, zmax = -INFINITY;
    for (int t = 0; t < ntri; t++) {
        const float *v = tri + 12L * t;
        float *b = m->bounds + 5L * t;
        b[0] = fminf(v[1], fminf(v[5], v[9]));
        b[1] = fmaxf(v[1], fmaxf(v[5], v[9]));
        b[2] = fminf(v[2], fminf(v[6], v[10]));
        b[3] = fmaxf(v[2], fmaxf(v[6], v[10]));
        b[4] = fmaxf(v[0], fmaxf(v[4], v[8]));
        ymin = fminf(ymin, b[0]);
        ymax = fmaxf(ymax, b[1]);
        zmin = fminf(zmin, b[2]);
        zmax = fmaxf(zmax, b[3]);
    }
    int nb = 1;
    while (nb * nb < ntri / 4 && nb < 256) nb *= 2;
    m->by = m->bz = nb;
    m->y0 = ntri > 0 ? ymin : 0.0f;
    m->z0 = ntri > 0 ? zmin : 0.0f;
    float ey = ntri > 0 ? ymax - ymin : 1.0f, ez = ntri > 0 ? zmax - zmin : 1.0f;
    m->iy = ey > 0 ? nb / ey : 1.0f;
    m->iz = ez > 0 ? nb / ez : 1.0f;
    int nbins = nb * nb;
    m->binStart = (int *)calloc((size_t)nbins + 1, sizeof(int));
    if (!m->binStart) return 0;
    for (int pass = 0; pass < 2; pass++) {
        int *fill = NULL;
        if (pass == 1) {
            for (int k = 0; k < nbins; k++) m->binStart[k + 1] += m->binStart[k];
            m->binList = (int *)malloc(sizeof(int) * (size_t)(m->binStart[nbins] > 0 ? m->binStart[nbins] : 1));
            fill = (int *)calloc((size_t)nbins, sizeof(int));
            if (!m->binList || !fill) {
                free(fill);
                return 0;
            }
        }
        for (int t = 0; t < ntri; t++) {
            const float *b = m->bounds + 5L * t;
            int ya = bin_of(b[0], m->y0, m->iy, nb), yb = bin_of(b[1], m->y0, m->iy, nb);
            int za = bin_of(b[2], m->z0, m->iz, nb), zb = bin_of(b[3], m->z0, m->iz, nb);
            for (int zz = za; zz <= zb; zz++)
                for (int yy = ya; yy <= yb; yy++) {
                    int k = yy + nb * zz;
                    if (pass == 0)
                        m->binStart[k + 1]++;
                    else
                        m->binList[m->binStart[k] + fill[k]++] = t;
                }
        }
        free(fill);
    }
    return 1;
}

static void mesh_index_free(MeshIndex *m) {
    free(m->bounds);
    free(m->binStart);
    free(m->binList);
}

/* simulation/src/lbm.c:964-1223 */
void LBM_SetSolidMesh(LBMGrid *g, float *triangles, int numTriangles, float minX, float minY, float minZ,
                      float maxX, float maxY, float maxZ) {
    if (!g) return;
    Impl *im = impl_of(g);
    MeshIndex m;
    if (!mesh_index_build(&m, triangles, numTriangles < 0 ? 0 : numTriangles)) {
        mesh_index_free(&m);
        note_error("LBM_SetSolidMesh: host allocation failed");
        return;
    }
    const float box[6] = {minX, minY, minZ, maxX, maxY, maxZ};
    int n = 0, *cells = NULL;
    float *qr = NULL;
    q_reset(im, g->totalCells);
    if (!lbmdev_mesh_voxelize(g->sizeX, g->sizeY, g->sizeZ, &m, world_scale(g), box, im->solid, &n, &cells, &qr)) {
        mesh_index_free(&m);
        memset(im->solid, 0, sizeof(int32_t) * (size_t)g->totalCells);
        note_error("LBM_SetSolidMesh: GPU voxelisation failed");
        geometry_changed(g);
        return;
    }
    long count = 0;
    for (long c = 0; c < g->totalCells; c++) count += im->solid[c] == 1;
    printf("LBM mesh solid: %ld cells marked as solid\n", count);
    for (int k = 0; k < n; k++) { /* rows in cell order, as the host link walk creates them */
        float *row = q_row_for(im, cells[k], 1);
        if (!row) {
            note_error("LBM_SetSolidMesh: host allocation failed");
            break;
        }
        memcpy(row, qr + 19L * k, sizeof(float) * 19);
    }
    free(cells);
    free(qr);
    printf("Bouzidi: %d boundary cells with links\n", im->nRows);
    mesh_index_free(&m);
    geometry_changed(g);
}

/* ------------------------------------------------------------ device sync */

/* Whether cell c belongs to the boundary list: it has Bouzidi links, or it
 * lies on a clamped Y/Z face, or it is a body cell on an x face. */
static int in_boundary_list(const LBMGrid *g, const Impl *im, long c, int periodic) {
    int x = (int)(c % g->sizeX);
    long t = c / g->sizeX;
    int y = (int)(t % g->sizeY), z = (int)(t / g->sizeY);
    int face = !periodic && (y == 0 || y == g->sizeY - 1 || z == 0 || z == g->sizeZ - 1);
    int xbody = im->solid[c] == 1 && (x == 0 || x == g->sizeX - 1);
    return im->qRow[c] >= 0 || face || xbody;
}

int lbm_boundary_count(const LBMGrid *g, int periodic) {
    const Impl *im = (const Impl *)g->impl;
    int n = 0;
    for (long c = 0; c < g->totalCells; c++) n += in_boundary_list(g, im, c, periodic);
    return n;
}

/* Writes the list in cell order: cells[k0 + j] and q[i*qStride + k0 + j]. */
void lbm_boundary_fill(const LBMGrid *g, int periodic, int32_t *cells, float *q, long qStride, int k0) {
    const Impl *im = (const Impl *)g->impl;
    int k = k0;
    for (long c = 0; c < g->totalCells; c++) {
        if (!in_boundary_list(g, im, c, periodic)) continue;
        cells[k] = (int32_t)c;
        for (int i = 0; i < 19; i++) q[(long)i * qStride + k] = q_get((Impl *)im, c, i);
        k++;
    }
}

const int32_t *lbm_host_solid(const LBMGrid *g) { return ((const Impl *)g->impl)->solid; }

LbmDev *lbm_device_of(LBMGrid *g) { return impl_of(g)->dev; }

/* Builds the boundary list for the current periodicYZ and uploads geometry. */
/* The grid's device state, or NULL (with an error noted) when an
 * LBM_InitializeFlow that switched the storage format could allocate neither
 * the new nor the previous format.  Every call that reaches the device goes
 * through this check. */
static LbmDev *live_dev(LBMGrid *g) {
    LbmDev *d = impl_of(g)->dev;
    if (!d) note_error("the grid has no device state: an earlier LBM_InitializeFlow failed to allocate it");
    return d;
}

static int ensure_geometry(LBMGrid *g) {
    Impl *im = impl_of(g);
    if (!live_dev(g)) return 0;
    const int periodic = g->periodicYZ == 1;
    if (!im->geomDirty && im->builtPeriodic == periodic) return 1;
    const int n = lbm_boundary_count(g, periodic);
    int32_t *cells = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    float *q = (float *)malloc(sizeof(float) * 19 * (size_t)(n > 0 ? n : 1));
    if (!cells || !q) {
        free(cells);
        free(q);
        note_error("boundary list: host allocation failed");
        return 0;
    }
    lbm_boundary_fill(g, periodic, cells, q, n, 0);
    int ok = lbmdev_set_geometry(im->dev, im->solid, n, cells, q, periodic);
    free(cells);
    free(q);
    if (!ok) {
        note_error(lbmdev_last_error());
        return 0;
    }
    im->nBoundary = n;
    im->geomDirty = 0;
    im->builtPeriodic = periodic;
    return 1;
}

int LBM_GetBoundaryCellCount(LBMGrid *g) {
    if (!g || !ensure_geometry(g)) return -1;
    return impl_of(g)->nBoundary;
}

long long LBM_GetGuardResetCount(LBMGrid *g) {
    long long n = -1;
    if (!g || !live_dev(g)) return -1;
    if (!lbmdev_guard_resets(impl_of(g)->dev, &n)) {
        note_error(lbmdev_last_error());
        return -1;
    }
    return n;
}

/* ------------------------------------------------------------ flow */

/* Device buffers in the requested storage format and the geometry uploaded. */
static int prepare_flow(LBMGrid *g) {
    Impl *im = impl_of(g);
    if (!live_dev(g)) return 0;
    if (g->populationStorage != lbmdev_storage(im->dev)) {
        char err[256] = {0};
        const int previous = lbmdev_storage(im->dev);
        lbmdev_destroy(im->dev);
        im->dev = lbmdev_create(g->sizeX, g->sizeY, g->sizeZ, g->populationStorage, err, sizeof(err));
        if (!im->dev) {
            char msg[400];
            snprintf(msg, sizeof(msg), "LBM_InitializeFlow: %s", err);
            note_error(msg);
            /* fall back to the format the grid had, so later calls stay valid */
            char err2[256] = {0};
            g->populationStorage = previous;
            im->dev = lbmdev_create(g->sizeX, g->sizeY, g->sizeZ, previous, err2, sizeof(err2));
            g->velocityBuffer = im->dev ? lbmdev_velocity_ptr(im->dev) : NULL;
            g->solidBuffer = im->dev ? lbmdev_solid_ptr(im->dev) : NULL;
            im->flowReady = 0; /* the populations are gone either way */
            im->geomDirty = 1;
            return 0;
        }
        g->velocityBuffer = lbmdev_velocity_ptr(im->dev);
        g->solidBuffer = lbmdev_solid_ptr(im->dev);
        im->geomDirty = 1;
    }
    return ensure_geometry(g);
}

int LBM_InitializeFlow(LBMGrid *g, float ux, float uy, float uz) {
    if (!g) return 0;
    Impl *im = impl_of(g);
    if (!prepare_flow(g)) return 0;
    float feq[19], v4[4] = {ux, uy, uz, 1.0f};
    for (int i = 0; i < 19; i++) feq[i] = host_feq(i, 1.0f, ux, uy, uz);
    if (!lbmdev_init_flow(im->dev, feq, v4)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    im->flowReady = 1;
    g->stepCount = 0;
    printf("LBM flow initialized: u=(%.3f, %.3f, %.3f)\n", ux, uy, uz);
    return 1;
}

int LBM_InitializeFlowFromField(LBMGrid *g, const float *field, int cnx, int cny, int cnz) {
    if (!g) return 0;
    if (!field || cnx <= 0 || cny <= 0 || cnz <= 0 || cnx > g->sizeX || cny > g->sizeY || cnz > g->sizeZ) {
        note_error("LBM_InitializeFlowFromField: the field must be a non-empty lattice no finer than the grid");
        return 0;
    }
    Impl *im = impl_of(g);
    if (!prepare_flow(g)) return 0;
    if (!lbmdev_init_field(im->dev, field, cnx, cny, cnz)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    im->flowReady = 1;
    g->stepCount = 0;
    return 1;
}

int LBM_InitializeFlowFromFieldNoAlloc(LBMGrid *g, const float *field, int cnx, int cny, int cnz) {
    if (!g) return 0;
    if (!field || cnx <= 0 || cny <= 0 || cnz <= 0 || cnx > g->sizeX || cny > g->sizeY || cnz > g->sizeZ) {
        note_error("LBM_InitializeFlowFromFieldNoAlloc: the field must be a non-empty lattice no finer than the grid");
        return 0;
    }
    Impl *im = impl_of(g);
    if (!live_dev(g)) return 0;
    /* Both of these would allocate device memory (new population buffers, a
     * larger boundary list): the allocating entry points do them. */
    if (g->populationStorage != lbmdev_storage(im->dev)) {
        note_error("LBM_InitializeFlowFromFieldNoAlloc: switching the storage format allocates; initialise once "
                   "with LBM_InitializeFlow or LBM_InitializeFlowFromField first");
        return 0;
    }
    if (im->geomDirty || im->builtPeriodic != (g->periodicYZ == 1)) {
        note_error("LBM_InitializeFlowFromFieldNoAlloc: the geometry or y/z boundary mode changed since the last "
                   "upload, which may allocate; initialise once with LBM_InitializeFlow or "
                   "LBM_InitializeFlowFromField first");
        return 0;
    }
    if (!lbmdev_init_field_noalloc(im->dev, field, cnx, cny, cnz)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    im->flowReady = 1;
    g->stepCount = 0;
    return 1;
}

static void step_params(const LBMGrid *g, float ux, float uy, float uz, LbmStepParams *p) {
    p->tau = g->tau;
    p->smagorinskyCs = g->smagorinskyCs;
    p->inletX = ux;
    p->inletY = uy;
    p->inletZ = uz;
    p->useRegularized = g->useRegularized;
    p->useMRT = g->useMRT;
    p->useSmagorinsky = g->useSmagorinsky;
    p->periodicYZ = g->periodicYZ;
}

void LBM_StepN(LBMGrid *g, int n, float ux, float uy, float uz) {
    if (!g || n <= 0) return;
    Impl *im = impl_of(g);
    if (!im->flowReady) {
        note_error("LBM_Step called before LBM_InitializeFlow");
        return;
    }
    if (!ensure_geometry(g)) return;
    LbmStepParams p;
    step_params(g, ux, uy, uz, &p);
    if (!lbmdev_step(im->dev, &p, n)) {
        note_error(lbmdev_last_error());
        return;
    }
    g->stepCount += n;
}

void LBM_Step(LBMGrid *g, float ux, float uy, float uz) { LBM_StepN(g, 1, ux, uy, uz); }

int LBM_ProfileSteps(LBMGrid *g, int n, float ux, float uy, float uz, float *total_ms, float *main_ms,
                     float *boundary_ms) {
    if (!g || n <= 0) return 0;
    Impl *im = impl_of(g);
    if (!im->flowReady || !ensure_geometry(g)) return 0;
    LbmStepParams p;
    step_params(g, ux, uy, uz, &p);
    int ok = lbmdev_profile_steps(im->dev, &p, n, total_ms, main_ms, boundary_ms);
    if (ok) g->stepCount += n;
    return ok;
}

int LBM_StepNTimed(LBMGrid *g, int n, float ux, float uy, float uz) {
    if (!g || n <= 0) return 0;
    Impl *im = impl_of(g);
    if (!im->flowReady) {
        note_error("LBM_Step called before LBM_InitializeFlow");
        return 0;
    }
    if (!ensure_geometry(g)) return 0;
    LbmStepParams p;
    step_params(g, ux, uy, uz, &p);
    if (!lbmdev_step_timed(im->dev, &p, n)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    g->stepCount += n;
    return 1;
}

int LBM_CollectStepTimings(LBMGrid *g, float *main_ms, float *boundary_ms, int *launches) {
    if (!g || !live_dev(g)) return 0;
    if (!lbmdev_collect_timings(impl_of(g)->dev, main_ms, boundary_ms, launches)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    return 1;
}

float *LBM_GetVelocityBuffer(LBMGrid *g) { return g ? g->velocityBuffer : NULL; }

void LBM_Synchronize(LBMGrid *g) {
    if (g && live_dev(g)) lbmdev_synchronize(impl_of(g)->dev);
}

/* ------------------------------------------------------------ force */

static int force_sums(LBMGrid *g, double exact[6], long long fixed[7]) {
    for (int k = 0; k < 6; k++) exact[k] = 0;
    for (int k = 0; k < 7; k++) fixed[k] = 0;
    Impl *im = impl_of(g);
    if (!im->flowReady || !ensure_geometry(g)) return 0;
    if (!lbmdev_force(im->dev, g->periodicYZ == 1, exact, fixed)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    return 1;
}

/* simulation/src/lbm.c:835-918 */
void LBM_ComputeDragForceDecomposed(LBMGrid *g, float *fx, float *fy, float *fz, float *px, float *py,
                                    float *pz) {
    *fx = *fy = *fz = 0.0f;
    if (px) *px = 0.0f;
    if (py) *py = 0.0f;
    if (pz) *pz = 0.0f;
    if (!g) return;
    double ex[6];
    long long fi[7];
    if (!force_sums(g, ex, fi)) return;
    float out[6];
    if (g->forceFixedPoint) {
        /* results[k] / 10000.0f on the int32 accumulators, lbm.c:909-917 */
        const int src[6] = {0, 1, 2, 4, 5, 6};
        for (int k = 0; k < 6; k++) out[k] = (float)(int32_t)fi[src[k]] / 10000.0f;
    } else {
        for (int k = 0; k < 6; k++) out[k] = (float)ex[k];
    }
    *fx = out[0];
    *fy = out[1];
    *fz = out[2];
    if (px) *px = out[3];
    if (py) *py = out[4];
    if (pz) *pz = out[5];
}

void LBM_ComputeDragForce(LBMGrid *g, float *fx, float *fy, float *fz) {
    LBM_ComputeDragForceDecomposed(g, fx, fy, fz, NULL, NULL, NULL);
}

void LBM_ComputeDragForceF64(LBMGrid *g, double out[6]) {
    long long fi[7];
    for (int k = 0; k < 6; k++) out[k] = 0;
    if (g) force_sums(g, out, fi);
}

/* |F| / (0.5 rho U^2 A) with rho = 1; 0 when the denominator is below 1e-10
 * (simulation/src/lbm.c:920-962). */
static float coefficient(float force, float U, float A) {
    float dyn = 0.5f * 1.0f * U * U;
    if (dyn * A < 1e-10f) return 0.0f;
    return fabsf(force) / (dyn * A);
}

float LBM_ComputeDragCoefficient(LBMGrid *g, float U, float A) {
    float fx, fy, fz;
    LBM_ComputeDragForce(g, &fx, &fy, &fz);
    return coefficient(fx, U, A);
}

float LBM_ComputeLiftCoefficient(LBMGrid *g, float U, float A) {
    float fx, fy, fz;
    LBM_ComputeDragForce(g, &fx, &fy, &fz);
    return coefficient(fy, U, A);
}

/* ------------------------------------------------------------ readback */

int LBM_CopyVelocityToHost(LBMGrid *g, float *out) {
    if (!g || !out || !live_dev(g)) return 0;
    return lbmdev_copy_velocity(impl_of(g)->dev, out);
}

int LBM_CopyCoarseVelocityToHost(LBMGrid *g, int factor, float *out) {
    if (!g || !out || factor <= 0 || g->sizeX % factor || g->sizeY % factor || g->sizeZ % factor) {
        note_error("LBM_CopyCoarseVelocityToHost: the grid must be a multiple of the factor on every axis");
        return 0;
    }
    if (!live_dev(g) || !lbmdev_copy_velocity_coarse(impl_of(g)->dev, factor, out)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    return 1;
}

int LBM_CopySolidToHost(LBMGrid *g, int *out) {
    if (!g || !out || !ensure_geometry(g)) return 0;
    return lbmdev_copy_solid(impl_of(g)->dev, (int32_t *)out);
}

int LBM_CopyQToHost(LBMGrid *g, float *out) {
    if (!g || !out) return 0;
    Impl *im = impl_of(g);
    for (long c = 0; c < g->totalCells; c++)
        for (int i = 0; i < 19; i++) out[19 * c + i] = q_get(im, c, i);
    return 1;
}

/* ------------------------------------------------------------ checkpoints */

/* Checkpoint container: this 64-byte header, the stored populations exactly
 * as the device holds them, the velocity buffer. */
typedef struct {
    char magic[8];
    int32_t nx, ny, nz, storage;
    int32_t layout, periodic;
    int64_t stepCount;
    uint64_t geometryHash;
    uint64_t populationBytes;
    uint64_t velocityBytes;
} StateHeader;
_Static_assert(sizeof(StateHeader) == 64, "checkpoint header is 64 bytes");
static const char kStateMagic[8] = {'L', 'B', 'M', 'S', 'T', 'A', 'T', '1'};

static uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 1099511628211ULL;
    }
    return h;
}

/* FNV-1a of the dimensions, the y/z mode, the solid array and every cell's q
 * row: a checkpoint only continues on the geometry it was taken from. */
static uint64_t geometry_hash(LBMGrid *g) {
    Impl *im = impl_of(g);
    const int32_t head[4] = {g->sizeX, g->sizeY, g->sizeZ, g->periodicYZ == 1};
    uint64_t h = fnv1a(1469598103934665603ULL, head, sizeof(head));
    const long n = (long)g->sizeX * g->sizeY * g->sizeZ;
    h = fnv1a(h, im->solid, sizeof(int32_t) * (size_t)n);
    for (long c = 0; c < n; c++) {
        const int r = im->qRow[c];
        if (r < 0) continue;
        const int32_t cc = (int32_t)c;
        h = fnv1a(h, &cc, sizeof(cc));
        h = fnv1a(h, im->qRows + 19L * r, sizeof(float) * 19);
    }
    return h;
}

long long LBM_GetStateSize(LBMGrid *g) {
    if (!g || !live_dev(g)) return 0;
    return (long long)(sizeof(StateHeader) + lbmdev_population_bytes(impl_of(g)->dev) +
                       (size_t)g->totalCells * 16);
}

int LBM_SaveState(LBMGrid *g, void *buf, long long bytes) {
    if (!g || !buf || !live_dev(g)) {
        note_error("LBM_SaveState: no grid or no buffer");
        return 0;
    }
    Impl *im = impl_of(g);
    if (!im->flowReady || !ensure_geometry(g)) {
        note_error("LBM_SaveState: no flow to save (call LBM_InitializeFlow first)");
        return 0;
    }
    if (bytes < LBM_GetStateSize(g)) {
        note_error("LBM_SaveState: the buffer is smaller than LBM_GetStateSize");
        return 0;
    }
    StateHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kStateMagic, sizeof(h.magic));
    h.nx = g->sizeX;
    h.ny = g->sizeY;
    h.nz = g->sizeZ;
    h.storage = lbmdev_storage(im->dev);
    h.periodic = g->periodicYZ == 1;
    h.stepCount = g->stepCount;
    h.geometryHash = geometry_hash(g);
    h.populationBytes = lbmdev_population_bytes(im->dev);
    h.velocityBytes = (uint64_t)g->totalCells * 16;
    char *p = (char *)buf;
    int layout = 0;
    if (!lbmdev_save_state(im->dev, p + sizeof(h), p + sizeof(h) + h.populationBytes, &layout)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    h.layout = layout;
    memcpy(p, &h, sizeof(h));
    return 1;
}

int LBM_LoadState(LBMGrid *g, const void *buf, long long bytes) {
    if (!g || !buf || !live_dev(g)) {
        note_error("LBM_LoadState: no grid or no buffer");
        return 0;
    }
    StateHeader h;
    if (bytes < (long long)sizeof(h)) {
        note_error("LBM_LoadState: truncated state");
        return 0;
    }
    memcpy(&h, buf, sizeof(h));
    if (memcmp(h.magic, kStateMagic, sizeof(h.magic)) != 0) {
        note_error("LBM_LoadState: not a saved LBM state");
        return 0;
    }
    if (h.nx != g->sizeX || h.ny != g->sizeY || h.nz != g->sizeZ) {
        note_error("LBM_LoadState: the state was saved from a grid of other dimensions");
        return 0;
    }
    if (h.storage != g->populationStorage) {
        note_error("LBM_LoadState: populationStorage differs from the saved state's");
        return 0;
    }
    if (h.periodic != (g->periodicYZ == 1)) {
        note_error("LBM_LoadState: periodicYZ differs from the saved state's");
        return 0;
    }
    Impl *im = impl_of(g);
    if (!prepare_flow(g)) return 0;
    if (h.populationBytes != lbmdev_population_bytes(im->dev) || h.velocityBytes != (uint64_t)g->totalCells * 16 ||
        (uint64_t)bytes < sizeof(h) + h.populationBytes + h.velocityBytes) {
        note_error("LBM_LoadState: truncated state");
        return 0;
    }
    if (h.geometryHash != geometry_hash(g)) {
        note_error("LBM_LoadState: the grid's geometry (solid cells, q, y/z mode) differs from the saved state's");
        return 0;
    }
    const char *p = (const char *)buf;
    if (!lbmdev_load_state(im->dev, p + sizeof(h), p + sizeof(h) + h.populationBytes, h.layout)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    im->flowReady = 1;
    g->stepCount = h.stepCount;
    return 1;
}

int LBM_CopyPopulationsToHost(LBMGrid *g, float *f, float *fNew) {
    if (!g || !impl_of(g)->flowReady || !ensure_geometry(g)) return 0;
    return lbmdev_export_populations(impl_of(g)->dev, g->periodicYZ == 1, f, fNew);
}

int lbmdev_export_populations_f64(LbmDev *d, int periodicYZ, double *f_ps, double *f_pc);

int LBM_CopyPopulationsToHostF64(LBMGrid *g, double *f, double *fNew) {
    if (!g || !impl_of(g)->flowReady || !ensure_geometry(g)) return 0;
    return lbmdev_export_populations_f64(impl_of(g)->dev, g->periodicYZ == 1, f, fNew);
}

double LBM_ComputeTotalMass(LBMGrid *g) {
    double m = 0.0;
    if (!g || !impl_of(g)->flowReady || !ensure_geometry(g)) return 0.0;
    lbmdev_total_mass(impl_of(g)->dev, g->periodicYZ == 1, &m);
    return m;
}

/* ------------------------------------------------------------ particles */

void LBM_SeedParticles(LBMGrid *g, Particle *p, int n, unsigned int seed) {
    if (!g || !p || !live_dev(g)) return;
    if (!lbmdev_particles_seed(impl_of(g)->dev, p, n, seed)) note_error(lbmdev_last_error());
}

void LBM_AdvectParticles(LBMGrid *g, Particle *p, int n, float dt, float velScale, unsigned int seed,
                         unsigned int step) {
    if (!g || !p || !live_dev(g)) return;
    if (!lbmdev_particles_advect(impl_of(g)->dev, p, n, dt, velScale, seed, step))
        note_error(lbmdev_last_error());
}

int LBM_TraceStreamlines(LBMGrid *g, float *out, int numSeeds, int traceLen) {
    if (!g || !out || !live_dev(g)) {
        note_error("LBM_TraceStreamlines: no grid or no output buffer");
        return 0;
    }
    if (!lbmdev_trace_streamlines(impl_of(g)->dev, out, numSeeds, traceLen)) {
        note_error(lbmdev_last_error());
        return 0;
    }
    return 1;
}

void LBM_AdvectParticlesLive(LBMGrid *g, Particle *p, int n, float dt, const LBMParticleFlow *f, unsigned int seed,
                             unsigned int step) {
    if (!g || !p || !f || !live_dev(g)) return;
    LbmParticleFlow d;
    memset(&d, 0, sizeof(d));
    const float wind[3] = {f->windX, f->windY, f->windZ};
    const float lo[3] = {f->carMinX, f->carMinY, f->carMinZ}, hi[3] = {f->carMaxX, f->carMaxY, f->carMaxZ};
    const float c[3] = {f->carCenterX, f->carCenterY, f->carCenterZ};
    const float gm[3] = {f->gridMinX, f->gridMinY, f->gridMinZ};
    const float gs[3] = {f->gridCellSizeX, f->gridCellSizeY, f->gridCellSizeZ};
    memcpy(d.wind, wind, sizeof(wind));
    memcpy(d.carMin, lo, sizeof(lo));
    memcpy(d.carMax, hi, sizeof(hi));
    memcpy(d.carCenter, c, sizeof(c));
    memcpy(d.gridMin, gm, sizeof(gm));
    memcpy(d.gridCellSize, gs, sizeof(gs));
    d.collisionMode = f->collisionMode;
    d.useLBM = f->useLBM;
    d.vizMode = f->vizMode;
    d.triangles = f->triangles;
    d.numTriangles = f->numTriangles;
    d.cellStart = (const int32_t *)f->cellStart;
    d.cellCount = (const int32_t *)f->cellCount;
    d.triIndices = (const int32_t *)f->triIndices;
    d.numIndices = f->numIndices;
    d.gridRes = f->gridRes;
    if (f->collisionMode == 3 && !ensure_geometry(g)) return;  /* the solid buffer must be current */
    if (!lbmdev_particles_live(impl_of(g)->dev, p, n, dt, &d, seed, step)) note_error(lbmdev_last_error());
}

void *LBM_DeviceAlloc(unsigned long long bytes) { return lbmdev_malloc((size_t)bytes); }
void LBM_DeviceFree(void *p) { lbmdev_free(p); }
int LBM_CopyToDevice(void *dst, const void *src, unsigned long long bytes) {
    return lbmdev_memcpy_h2d(dst, src, (size_t)bytes);
}
int LBM_CopyToHost(void *dst, const void *src, unsigned long long bytes) {
    return lbmdev_memcpy_d2h(dst, src, (size_t)bytes);
}
int LBM_CopyDeviceToDevice(void *dst, const void *src, unsigned long long bytes) {
    return lbmdev_memcpy_d2d(dst, src, (size_t)bytes);
}

/* Workgroup-size autotuner for the step kernel: times `steps` steps with each
 * candidate (128, 256, 512 lanes) on the grid's own lattice and keeps the
 * fastest.  The populations and velocity are saved first and restored after
 * every trial, so tuning leaves the simulation state unchanged. */
int LBM_Autotune(LBMGrid *g, int steps, float ux, float uy, float uz, float *msPerStep) {
    if (!g || steps <= 0) return 0;
    Impl *im = impl_of(g);
    if (!im->flowReady || !ensure_geometry(g)) return 0;
    LbmStepParams p;
    step_params(g, ux, uy, uz, &p);
    int best = lbmdev_autotune(im->dev, &p, steps, msPerStep);
    if (!best) note_error(lbmdev_last_error());
    return best;
}

int LBM_SetBlockSize(LBMGrid *g, int bs) {
    if (!g || !impl_of(g)->dev) return 0;
    if (!lbmdev_set_block_size(impl_of(g)->dev, bs)) {
        note_error("LBM_SetBlockSize: the workgroup size must be 128, 256 or 512");
        return 0;
    }
    return 1;
}

int LBM_SetStreamingAccess(LBMGrid *g, int mode) {
    if (!g || !impl_of(g)->dev) return 0;
    if (!lbmdev_set_streaming(impl_of(g)->dev, mode)) {
        note_error("LBM_SetStreamingAccess: the mode must be 0 to 3");
        return 0;
    }
    return 1;
}

int LBM_SetBlockOrder(LBMGrid *g, int xcdContiguous) {
    if (!g || !impl_of(g)->dev) return 0;
    lbmdev_set_block_order(impl_of(g)->dev, xcdContiguous);
    return 1;
}

int LBM_GetKernelConfig(const LBMGrid *g, int *blockSize, int *streaming, int *xcdContiguous) {
    if (!g || !impl_of((LBMGrid *)g)->dev) return 0;
    const LbmDev *d = impl_of((LBMGrid *)g)->dev;
    if (blockSize) *blockSize = lbmdev_block_size(d);
    if (streaming) *streaming = lbmdev_streaming(d);
    if (xcdContiguous) *xcdContiguous = lbmdev_block_order(d);
    return 1;
}

/* Attainable HBM bandwidth: a plain device copy of `bytes` bytes, read +
 * write bytes over the mean launch time. */
float LBM_MeasureCopyBandwidth(unsigned long long bytes, int iters) {
    float gbps = 0.0f;
    if (!lbmdev_copy_bandwidth((size_t)bytes, iters > 0 ? iters : 10, &gbps, NULL)) note_error(lbmdev_last_error());
    return gbps;
}
