/*
 * Host-side model helpers: OBJ reading, placement, the triangle buffer for
 * LBM_SetSolidMesh and the occupancy voxeliser (declared in lbm_model.h).
 *
 * Results equal the reference helpers bit for bit: every float expression
 * keeps the reference's operand order and precision (single-precision
 * arithmetic, the degree-to-radian factor in double where the reference
 * computes it in double), and the C library does the number parsing.
 *   OBJ records   simulation/obj-file-loader/lib/model_loader.c:7-142
 *   bounds        simulation/src/model_bounds.c:146-212
 *   triangles     simulation/src/model_bounds.c:214-287
 *   placement     simulation/src/main.c:250-349
 *   voxeliser     simulation/src/voxelize.c:17-386
 * The structure is this file's own: the OBJ file is scanned once per pass
 * with one record parser, and the voxeliser gathers the +X ray hits of all
 * (y, z) columns into one counted array before sorting each column.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm_model.h"

#define PI_D 3.14159265358979323846

/* ------------------------------------------------------------- OBJ file */

/* The reference counts records with a 256-byte line reader and parses them
 * with a 512-byte one; both readers are kept so that over-long lines split
 * the same way. */
enum { COUNT_LINE = 256, PARSE_LINE = 512, TOKEN_MAX = 63 };

static void count_records(FILE *fp, int *nv, int *nf) {
    char line[COUNT_LINE];
    *nv = *nf = 0;
    while (fgets(line, sizeof line, fp)) {
        if (line[1] != ' ') continue;
        if (line[0] == 'v') ++*nv;
        if (line[0] == 'f') ++*nf;
    }
}

/* Copies the next blank-delimited token (at most TOKEN_MAX characters; a
 * longer token continues in the next call) and returns the first integer
 * it starts with, 0 when there is none. */
static int face_ref(const char **cursor) {
    const char *p = *cursor;
    char tok[TOKEN_MAX + 1];
    int n = 0, v = 0;
    while (*p == ' ' || *p == '\t') p++;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && n < TOKEN_MAX) tok[n++] = *p++;
    tok[n] = '\0';
    *cursor = p;
    sscanf(tok, "%d", &v);
    return v;
}

Model loadOBJ(const char *filePath) {
    Model m;
    memset(&m, 0, sizeof m);
    FILE *fp = fopen(filePath, "r");
    if (!fp) {
        printf("Error: Could not open file %s\n", filePath);
        return m;
    }
    int nv, nf;
    count_records(fp, &nv, &nf);
    printf("File contains %d vertices and %d faces.\n", nv, nf);
    m.vertices = (Vertex *)calloc(nv > 0 ? (size_t)nv : 1, sizeof(Vertex));
    m.faces = (Face *)calloc(nf > 0 ? (size_t)nf : 1, sizeof(Face));
    if (!m.vertices || !m.faces) {
        printf("Error: Out of memory!\n");
        free(m.vertices);
        free(m.faces);
        memset(&m, 0, sizeof m);
        fclose(fp);
        return m;
    }
    rewind(fp);
    char line[PARSE_LINE];
    int iv = 0, jf = 0;
    while (fgets(line, sizeof line, fp)) {
        if (line[1] != ' ') continue;
        if (line[0] == 'v' && iv < nv) {
            Vertex *v = &m.vertices[iv++];
            sscanf(line, "v %f %f %f", &v->x, &v->y, &v->z);
        } else if (line[0] == 'f' && jf < nf) {
            const char *c = line + 2;
            Face *f = &m.faces[jf++];
            f->v1 = face_ref(&c);
            f->v2 = face_ref(&c);
            f->v3 = face_ref(&c);
        }
    }
    fclose(fp);
    m.vertexCount = nv;
    m.faceCount = jf;
    printf("Successfully loaded model with %d vertices and %d faces.\n", nv, jf);
    return m;
}

void freeModel(Model *model) {
    if (!model) return;
    free(model->vertices);
    free(model->faces);
    memset(model, 0, sizeof *model);
}

/* --------------------------------------------------- placed coordinates */

typedef struct {
    float scale, ox, oy, oz, c, s;
} Placement;

static Placement placement(float scale, float ox, float oy, float oz, float rotationY) {
    Placement p;
    const float rad = (float)(rotationY * PI_D / 180.0f);
    p.scale = scale;
    p.ox = ox;
    p.oy = oy;
    p.oz = oz;
    p.c = cosf(rad);
    p.s = sinf(rad);
    return p;
}

/* rotY(scale * v + offset) */
static void place(const Placement *p, Vertex v, float *X, float *Y, float *Z) {
    const float x = v.x * p->scale + p->ox;
    const float y = v.y * p->scale + p->oy;
    const float z = v.z * p->scale + p->oz;
    *X = x * p->c - z * p->s;
    *Y = y;
    *Z = x * p->s + z * p->c;
}

CarBounds computeModelBounds(Model *model, float scale, float offsetX, float offsetY, float offsetZ,
                             float rotationY) {
    CarBounds b;
    if (!model || model->vertexCount == 0) {
        b.minX = b.minY = b.minZ = -0.5f;
        b.maxX = b.maxY = b.maxZ = 0.5f;
        b.centerX = b.centerY = b.centerZ = 0.0f;
        return b;
    }
    const Placement p = placement(scale, offsetX, offsetY, offsetZ, rotationY);
    b.minX = b.minY = b.minZ = FLT_MAX;
    b.maxX = b.maxY = b.maxZ = -FLT_MAX;
    for (int i = 0; i < model->vertexCount; i++) {
        float X, Y, Z;
        place(&p, model->vertices[i], &X, &Y, &Z);
        if (X < b.minX) b.minX = X;
        if (Y < b.minY) b.minY = Y;
        if (Z < b.minZ) b.minZ = Z;
        if (X > b.maxX) b.maxX = X;
        if (Y > b.maxY) b.maxY = Y;
        if (Z > b.maxZ) b.maxZ = Z;
    }
    b.centerX = (b.minX + b.maxX) * 0.5f;
    b.centerY = (b.minY + b.maxY) * 0.5f;
    b.centerZ = (b.minZ + b.maxZ) * 0.5f;
    return b;
}

static int face_in_range(const Face *f, int nv) {
    return f->v1 >= 1 && f->v1 <= nv && f->v2 >= 1 && f->v2 <= nv && f->v3 >= 1 && f->v3 <= nv;
}

GPUTriangle *createTriangleBuffer(Model *model, float scale, float offsetX, float offsetY, float offsetZ,
                                  float rotationY, int *outCount) {
    *outCount = 0;
    if (!model || model->faceCount == 0 || model->vertexCount == 0) return NULL;
    GPUTriangle *t = (GPUTriangle *)malloc((size_t)model->faceCount * sizeof(GPUTriangle));
    if (!t) return NULL;
    const Placement p = placement(scale, offsetX, offsetY, offsetZ, rotationY);
    int n = 0;
    for (int i = 0; i < model->faceCount; i++) {
        const Face *f = &model->faces[i];
        if (!face_in_range(f, model->vertexCount)) continue;
        float *o = (float *)&t[n++];
        const int idx[3] = {f->v1 - 1, f->v2 - 1, f->v3 - 1};
        for (int k = 0; k < 3; k++) {
            place(&p, model->vertices[idx[k]], &o[4 * k], &o[4 * k + 1], &o[4 * k + 2]);
            o[4 * k + 3] = 0.0f;
        }
    }
    *outCount = n;
    return t;
}

/* The cell range of triangle t along one axis: floor of its extent relative
 * to the grid origin over the cell size, clamped to [0, R - 1] (the
 * expressions of model_bounds.c:45-73). */
static void tri_range(const float *t, int axis, float origin, float cell, int R, int *lo, int *hi) {
    const float a = t[axis], b = t[4 + axis], c = t[8 + axis];
    const float mn = fminf(a, fminf(b, c)), mx = fmaxf(a, fmaxf(b, c));
    *lo = (int)floorf((mn - origin) / cell);
    *hi = (int)floorf((mx - origin) / cell);
    if (*lo < 0) *lo = 0;
    if (*hi >= R) *hi = R - 1;
}

CollisionGrid buildCollisionGrid(GPUTriangle *tris, int numTris, float bminX, float bminY, float bminZ,
                                 float bmaxX, float bmaxY, float bmaxZ) {
    CollisionGrid g;
    memset(&g, 0, sizeof(g));
    const int R = COLL_GRID_RES;
    const float pad = 0.05f;
    g.totalCells = R * R * R;
    g.minX = bminX - pad;
    g.minY = bminY - pad;
    g.minZ = bminZ - pad;
    g.cellSizeX = ((bmaxX + pad) - g.minX) / R;
    g.cellSizeY = ((bmaxY + pad) - g.minY) / R;
    g.cellSizeZ = ((bmaxZ + pad) - g.minZ) / R;
    const float origin[3] = {g.minX, g.minY, g.minZ}, cell[3] = {g.cellSizeX, g.cellSizeY, g.cellSizeZ};
    /* each triangle's clamped cell box, once */
    int *box = (int *)malloc(sizeof(int) * 6 * (size_t)(numTris > 0 ? numTris : 1));
    g.cellCount = (int *)calloc((size_t)g.totalCells, sizeof(int));
    g.cellStart = (int *)malloc(sizeof(int) * (size_t)g.totalCells);
    int *fill = (int *)calloc((size_t)g.totalCells, sizeof(int));
    if (!box || !g.cellCount || !g.cellStart || !fill) goto fail;
    long long total = 0;
    for (int t = 0; t < numTris; t++) {
        int *b = box + 6 * t;
        for (int a = 0; a < 3; a++) tri_range((const float *)&tris[t], a, origin[a], cell[a], R, &b[2 * a], &b[2 * a + 1]);
        for (int z = b[4]; z <= b[5]; z++)
            for (int y = b[2]; y <= b[3]; y++)
                for (int x = b[0]; x <= b[1]; x++) {
                    g.cellCount[x + y * R + z * R * R]++;
                    total++;
                }
    }
    if (total > 0x7fffffff) goto fail;
    g.totalIndices = (int)total;
    g.cellStart[0] = 0;
    for (int c = 1; c < g.totalCells; c++) g.cellStart[c] = g.cellStart[c - 1] + g.cellCount[c - 1];
    /* triangle order within each cell, as the reference's second pass */
    g.triIndices = (int *)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    if (!g.triIndices) goto fail;
    for (int t = 0; t < numTris; t++) {
        const int *b = box + 6 * t;
        for (int z = b[4]; z <= b[5]; z++)
            for (int y = b[2]; y <= b[3]; y++)
                for (int x = b[0]; x <= b[1]; x++) {
                    const int c = x + y * R + z * R * R;
                    g.triIndices[g.cellStart[c] + fill[c]++] = t;
                }
    }
    free(box);
    free(fill);
    return g;
fail:
    free(box);
    free(fill);
    free(g.cellCount);
    free(g.cellStart);
    free(g.triIndices);
    memset(&g, 0, sizeof(g));
    return g;
}

ModelPlacement LBM_AutoPlaceModel(const Model *model, float scale, int scaleGiven) {
    ModelPlacement mp;
    mp.scale = scaleGiven ? scale : 1.0f;
    mp.offsetX = 0.0f;
    mp.offsetY = -0.1f;
    mp.offsetZ = -0.9f;
    mp.rotationY = 0.0f;
    const int nv = model ? model->vertexCount : 0;
    if (nv > 0) {
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int i = 0; i < nv; i++) {
            const float c[3] = {model->vertices[i].x, model->vertices[i].y, model->vertices[i].z};
            for (int k = 0; k < 3; k++) {
                if (c[k] < lo[k]) lo[k] = c[k];
                if (c[k] > hi[k]) hi[k] = c[k];
            }
        }
        float centre[3], extent[3];
        for (int k = 0; k < 3; k++) {
            centre[k] = (lo[k] + hi[k]) * 0.5f;
            extent[k] = hi[k] - lo[k];
        }
        if (!scaleGiven) mp.scale = 2.4f / fmaxf(extent[0], fmaxf(extent[1], extent[2]));
        mp.offsetX = -centre[0] * mp.scale;
        mp.offsetY = -centre[1] * mp.scale;
        mp.offsetZ = -centre[2] * mp.scale;
    }
    /* longest of the placed x and z extents goes streamwise */
    float ex = 0.0f, ez = 0.0f;
    if (nv > 0) {
        float x0 = FLT_MAX, x1 = -FLT_MAX, z0 = FLT_MAX, z1 = -FLT_MAX;
        for (int i = 0; i < nv; i++) {
            const float x = model->vertices[i].x * mp.scale + mp.offsetX;
            const float z = model->vertices[i].z * mp.scale + mp.offsetZ;
            if (x < x0) x0 = x;
            if (x > x1) x1 = x;
            if (z < z0) z0 = z;
            if (z > z1) z1 = z;
        }
        ex = x1 - x0;
        ez = z1 - z0;
    }
    mp.rotationY = ez > ex ? 90.0f : 0.0f;
    /* the offset whose image under the rotation is (-2, 0), in float */
    const float rad = mp.rotationY * (float)PI_D / 180.0f;
    mp.offsetX += -2.0f * cosf(rad);
    mp.offsetZ += 2.0f * sinf(rad);
    return mp;
}

/* ------------------------------------------------------ occupancy grid */

#define VOX_EPS 1e-8f

typedef struct {
    float x, y, z;
} V3;

/* Centred, optionally axis-permuted and scaled copy of the vertices. */
static V3 *normalised_vertices(const Model *m, float padding, int align) {
    const int nv = m->vertexCount;
    V3 *v = (V3 *)malloc((size_t)nv * sizeof(V3));
    if (!v) return NULL;
    float lo[3] = {m->vertices[0].x, m->vertices[0].y, m->vertices[0].z};
    float hi[3] = {lo[0], lo[1], lo[2]};
    for (int i = 0; i < nv; i++) {
        const float c[3] = {m->vertices[i].x, m->vertices[i].y, m->vertices[i].z};
        for (int k = 0; k < 3; k++) {
            if (c[k] < lo[k]) lo[k] = c[k];
            if (c[k] > hi[k]) hi[k] = c[k];
        }
    }
    float ctr[3], ext[3];
    for (int k = 0; k < 3; k++) {
        ctr[k] = 0.5f * (lo[k] + hi[k]);
        ext[k] = hi[k] - lo[k];
    }
    for (int i = 0; i < nv; i++) {
        v[i].x = m->vertices[i].x - ctr[0];
        v[i].y = m->vertices[i].y - ctr[1];
        v[i].z = m->vertices[i].z - ctr[2];
    }
    if (align) {
        int swap = 0; /* 1: x <-> y, 2: x <-> z */
        if (ext[1] > ext[0] && ext[1] >= ext[2])
            swap = 1;
        else if (ext[2] > ext[0] && ext[2] > ext[1])
            swap = 2;
        if (swap) {
            for (int i = 0; i < nv; i++) {
                float *other = swap == 1 ? &v[i].y : &v[i].z;
                const float t = v[i].x;
                v[i].x = *other;
                *other = t;
            }
            const float t = ext[0];
            ext[0] = ext[swap];
            ext[swap] = t;
        }
    }
    float dmax = ext[0];
    if (ext[1] > dmax) dmax = ext[1];
    if (ext[2] > dmax) dmax = ext[2];
    if (dmax < VOX_EPS) {
        free(v);
        return NULL;
    }
    const float s = (2.0f - 2.0f * padding) / dmax;
    for (int i = 0; i < nv; i++) {
        v[i].x *= s;
        v[i].y *= s;
        v[i].z *= s;
    }
    return v;
}

/* Column range [lo, hi) of a face's (y, z) extent, as the reference bins it. */
static int face_columns(float y_min, float y_max, float z_min, float z_max, int R, int *gy0, int *gy1, int *gz0,
                        int *gz1) {
    if (y_max < -1.0f || y_min > 1.0f || z_max < -1.0f || z_min > 1.0f) return 0;
    *gy0 = (int)floorf((y_min + 1.0f) * 0.5f * R);
    *gy1 = (int)ceilf((y_max + 1.0f) * 0.5f * R) + 1;
    *gz0 = (int)floorf((z_min + 1.0f) * 0.5f * R);
    *gz1 = (int)ceilf((z_max + 1.0f) * 0.5f * R) + 1;
    if (*gy0 < 0) *gy0 = 0;
    if (*gz0 < 0) *gz0 = 0;
    if (*gy1 > R) *gy1 = R;
    if (*gz1 > R) *gz1 = R;
    return *gy0 < *gy1 && *gz0 < *gz1;
}

/* Visits every +X ray hit x = t of the face over its columns: the
 * Moller-Trumbore test with direction (1, 0, 0) from (0, wy, wz).  With
 * out == NULL the hits are counted per column, else appended at cursor[]. */
static void face_hits(V3 a, V3 b, V3 c, int R, int *count, int *cursor, float *out) {
    const float y_min = fminf(a.y, fminf(b.y, c.y)), y_max = fmaxf(a.y, fmaxf(b.y, c.y));
    const float z_min = fminf(a.z, fminf(b.z, c.z)), z_max = fmaxf(a.z, fmaxf(b.z, c.z));
    int gy0, gy1, gz0, gz1;
    if (!face_columns(y_min, y_max, z_min, z_max, R, &gy0, &gy1, &gz0, &gz1)) return;
    const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    const float det = -e1y * e2z + e1z * e2y;
    if (fabsf(det) < VOX_EPS) return;
    const float inv = 1.0f / det;
    const float sx = -a.x;
    for (int gy = gy0; gy < gy1; gy++) {
        const float sy = ((float)gy + 0.5f) / (float)R * 2.0f - 1.0f - a.y;
        for (int gz = gz0; gz < gz1; gz++) {
            const float sz = ((float)gz + 0.5f) / (float)R * 2.0f - 1.0f - a.z;
            const float u = inv * (sy * -e2z + sz * e2y);
            if (u < 0.0f || u > 1.0f) continue;
            const float qx = sy * e1z - sz * e1y;
            const float qy = sz * e1x - sx * e1z;
            const float qz = sx * e1y - sy * e1x;
            const float v = inv * qx;
            if (v < 0.0f || u + v > 1.0f) continue;
            const int col = gy * R + gz;
            if (out)
                out[cursor[col]++] = inv * (e2x * qx + e2y * qy + e2z * qz);
            else
                count[col]++;
        }
    }
}

static int cmp_float(const void *pa, const void *pb) {
    const float a = *(const float *)pa, b = *(const float *)pb;
    return (a > b) - (a < b);
}

static int voxelize_solid(const V3 *v, const Face *faces, int nf, int R, uint8_t *grid) {
    const size_t ncol = (size_t)R * R;
    int *start = (int *)calloc(ncol + 1, sizeof(int));
    int *cursor = (int *)malloc(ncol * sizeof(int));
    if (!start || !cursor) {
        free(start);
        free(cursor);
        return 0;
    }
    for (int f = 0; f < nf; f++)
        face_hits(v[faces[f].v1 - 1], v[faces[f].v2 - 1], v[faces[f].v3 - 1], R, start + 1, NULL, NULL);
    for (size_t k = 0; k < ncol; k++) start[k + 1] += start[k];
    float *hits = (float *)malloc((size_t)(start[ncol] > 0 ? start[ncol] : 1) * sizeof(float));
    if (!hits) {
        free(start);
        free(cursor);
        return 0;
    }
    memcpy(cursor, start, ncol * sizeof(int));
    for (int f = 0; f < nf; f++)
        face_hits(v[faces[f].v1 - 1], v[faces[f].v2 - 1], v[faces[f].v3 - 1], R, NULL, cursor, hits);
    for (int gy = 0; gy < R; gy++)
        for (int gz = 0; gz < R; gz++) {
            const int col = gy * R + gz;
            float *x = hits + start[col];
            const int cnt = start[col + 1] - start[col];
            if (cnt < 2) continue;
            qsort(x, (size_t)cnt, sizeof(float), cmp_float);
            /* hits closer than 1e-6 to the last kept one are the same crossing */
            int n = 1;
            for (int i = 1; i < cnt; i++)
                if (fabsf(x[i] - x[n - 1]) > 1e-6f) x[n++] = x[i];
            /* inside between consecutive pairs: voxel centres in [x0, x1] */
            for (int i = 0; i + 1 < n; i += 2) {
                int gx0 = (int)ceilf((x[i] + 1.0f) * 0.5f * R - 0.5f);
                int gx1 = (int)floorf((x[i + 1] + 1.0f) * 0.5f * R - 0.5f) + 1;
                if (gx0 < 0) gx0 = 0;
                if (gx1 > R) gx1 = R;
                for (int gx = gx0; gx < gx1; gx++) grid[gx + (size_t)R * (gy + (size_t)R * gz)] = 1;
            }
        }
    free(hits);
    free(start);
    free(cursor);
    return 1;
}

static void voxelize_surface(const V3 *v, const Face *faces, int nf, int R, uint8_t *grid) {
    for (int f = 0; f < nf; f++) {
        const V3 a = v[faces[f].v1 - 1], b = v[faces[f].v2 - 1], c = v[faces[f].v3 - 1];
        const float lo[3] = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
        const float hi[3] = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
        int g0[3], g1[3];
        for (int k = 0; k < 3; k++) {
            g0[k] = (int)floorf((lo[k] + 1.0f) * 0.5f * R);
            g1[k] = (int)ceilf((hi[k] + 1.0f) * 0.5f * R);
            if (g0[k] < 0) g0[k] = 0;
            if (g1[k] > R) g1[k] = R;
        }
        for (int z = g0[2]; z < g1[2]; z++)
            for (int y = g0[1]; y < g1[1]; y++)
                for (int x = g0[0]; x < g1[0]; x++) grid[x + (size_t)R * (y + (size_t)R * z)] = 1;
    }
}

VoxelGrid *Voxelize_FromModel(const Model *model, int resolution, float padding, int align_longest_x,
                              VoxelizeMode mode) {
    if (!model || resolution <= 0 || model->vertexCount <= 0 || model->faceCount <= 0) return NULL;
    if (padding < 0.0f) padding = 0.0f;
    if (padding >= 0.5f) padding = 0.499f;
    for (int i = 0; i < model->faceCount; i++)
        if (!face_in_range(&model->faces[i], model->vertexCount)) return NULL;
    V3 *v = normalised_vertices(model, padding, align_longest_x);
    if (!v) return NULL;
    const size_t n = (size_t)resolution * resolution * resolution;
    VoxelGrid *g = (VoxelGrid *)malloc(sizeof(VoxelGrid));
    uint8_t *data = (uint8_t *)calloc(n, 1);
    if (!g || !data) {
        free(v);
        free(g);
        free(data);
        return NULL;
    }
    g->resolution = resolution;
    g->data = data;
    if (mode == VOXELIZE_MODE_SURFACE) {
        voxelize_surface(v, model->faces, model->faceCount, resolution, data);
    } else {
        voxelize_solid(v, model->faces, model->faceCount, resolution, data);
        if (mode == VOXELIZE_MODE_AUTO) {
            const float frac = Voxelize_SolidFraction(g);
            if (frac < 0.001f || frac > 0.99f) {
                memset(data, 0, n);
                voxelize_surface(v, model->faces, model->faceCount, resolution, data);
            }
        }
    }
    free(v);
    return g;
}

float Voxelize_SolidFraction(const VoxelGrid *grid) {
    if (!grid || !grid->data || grid->resolution <= 0) return 0.0f;
    const size_t n = (size_t)grid->resolution * grid->resolution * grid->resolution;
    size_t on = 0;
    for (size_t i = 0; i < n; i++) on += grid->data[i] != 0;
    return (float)on / (float)n;
}

int Voxelize_WriteBinary(const VoxelGrid *grid, const char *path) {
    if (!grid || !grid->data || !path) return -1;
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    const uint32_t hdr[3] = {0x42584F56u /* "VOXB" little-endian */, 1u, (uint32_t)grid->resolution};
    const size_t n = (size_t)grid->resolution * grid->resolution * grid->resolution;
    const int ok = fwrite(hdr, sizeof hdr, 1, fp) == 1 && fwrite(grid->data, 1, n, fp) == n;
    fclose(fp);
    return ok ? 0 : -1;
}

void Voxelize_Free(VoxelGrid *grid) {
    if (!grid) return;
    free(grid->data);
    free(grid);
}
