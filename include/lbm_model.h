/*
 * Model helpers on the host: the OBJ reader, the placed triangle buffer that
 * LBM_SetSolidMesh takes, the model bounds, the caller's automatic placement,
 * and the fixed-resolution occupancy voxeliser with its binary file.
 *
 * Names, struct layouts and argument meanings are those of the reference
 * helpers, so a caller of the reference links against this library instead:
 *   simulation/obj-file-loader/lib/model_loader.h:4-20  (Vertex, Face, Model,
 *                                                        loadOBJ, freeModel)
 *   simulation/lib/model_bounds.h:8-54                   (CarBounds,
 *                                                        GPUTriangle,
 *                                                        CollisionGrid,
 *                                                        computeModelBounds,
 *                                                        createTriangleBuffer,
 *                                                        buildCollisionGrid)
 *   simulation/lib/voxelize.h:11-62                      (VoxelizeMode,
 *                                                        VoxelGrid,
 *                                                        Voxelize_*)
 * LBM_AutoPlaceModel has no function in the reference: it is the placement
 * main.c performs inline (simulation/src/main.c:250-349).
 */
#ifndef LBM_MODEL_H
#define LBM_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float x, y, z;
} Vertex;

/* 1-based vertex indices, as written in the file */
typedef struct {
    int v1, v2, v3;
} Face;

typedef struct {
    Vertex *vertices;
    Face *faces;
    int vertexCount;
    int faceCount;
} Model;

/* Reads "v x y z" records and the first three vertex references of each
 * "f ..." record (a quad keeps its first triangle only; "v/vt/vn" keeps v).
 * On a missing file: zero counts and NULL arrays. */
Model loadOBJ(const char *filePath);
void freeModel(Model *model);

typedef struct {
    float minX, minY, minZ;
    float maxX, maxY, maxZ;
    float centerX, centerY, centerZ;
} CarBounds;

/* 48 bytes; the 12-float triangle record LBM_SetSolidMesh reads */
typedef struct {
    float v0x, v0y, v0z, pad0;
    float v1x, v1y, v1z, pad1;
    float v2x, v2y, v2z, pad2;
} GPUTriangle;

/* Bounds of rotY(scale * v + offset) over all vertices; rotation about +y
 * by rotationY degrees.  An empty model gives the box [-0.5, 0.5]^3. */
CarBounds computeModelBounds(Model *model, float scale, float offsetX, float offsetY, float offsetZ,
                             float rotationY);

/* One placed triangle per face whose three indices are in range (others are
 * skipped); malloc'd, *outCount triangles; NULL for an empty model. */
GPUTriangle *createTriangleBuffer(Model *model, float scale, float offsetX, float offsetY, float offsetZ,
                                  float rotationY, int *outCount);

#define COLL_GRID_RES 8

/* Uniform R^3 grid (R = COLL_GRID_RES) over the model box grown by 0.05: the
 * triangles whose bounding box touches each cell, for the per-triangle
 * particle collisions (LBM_AdvectParticlesLive, collisionMode 2).  Arrays
 * are malloc'd; the caller frees cellStart, cellCount and triIndices. */
typedef struct {
    int *cellStart;  /* offset into triIndices per cell */
    int *cellCount;  /* triangle count per cell */
    int *triIndices; /* triangle indices, cell by cell */
    int totalIndices;
    int totalCells;
    float minX, minY, minZ;
    float cellSizeX, cellSizeY, cellSizeZ;
} CollisionGrid;

/* simulation/lib/model_bounds.h:47-54.  All fields zero on allocation failure. */
CollisionGrid buildCollisionGrid(GPUTriangle *tris, int numTris, float bminX, float bminY, float bminZ, float bmaxX,
                                 float bmaxY, float bmaxZ);

typedef struct {
    float scale;
    float offsetX, offsetY, offsetZ;
    float rotationY; /* degrees */
} ModelPlacement;

/* The caller's placement (main.c:250-349): centre the model, scale its
 * largest extent to 2.4 world units (unless scaleGiven, then `scale` is
 * kept), turn it 90 degrees about y when it is longer in z than in x, and
 * move its centre to x = -2.  Offsets start at main.c's defaults
 * (0, -0.1, -0.9) for an empty model. */
ModelPlacement LBM_AutoPlaceModel(const Model *model, float scale, int scaleGiven);

typedef enum {
    VOXELIZE_MODE_SOLID = 0,   /* +X ray parity fill per (y, z) column */
    VOXELIZE_MODE_SURFACE = 1, /* voxels covered by a triangle's bounding box */
    VOXELIZE_MODE_AUTO = 2     /* solid; surface when the solid fraction is < 0.1 % or > 99 % */
} VoxelizeMode;

typedef struct {
    int resolution; /* R: the grid is R x R x R */
    uint8_t *data;  /* R^3 bytes, index x + R (y + R z) */
} VoxelGrid;

/* Occupancy grid of the model centred at the origin, optionally with its
 * longest axis permuted onto x, scaled so its largest extent spans
 * 2 - 2 padding inside [-1, 1]^3 (padding clamped to [0, 0.499]).  NULL for
 * an empty model, a face index out of range, a degenerate extent or an
 * allocation failure. */
VoxelGrid *Voxelize_FromModel(const Model *model, int resolution, float padding, int align_longest_x,
                              VoxelizeMode mode);

/* "VOXB", u32 version 1, u32 R, R^3 occupancy bytes; 0 on success. */
int Voxelize_WriteBinary(const VoxelGrid *grid, const char *path);
float Voxelize_SolidFraction(const VoxelGrid *grid);
void Voxelize_Free(VoxelGrid *grid);

#ifdef __cplusplus
}
#endif

#endif /* LBM_MODEL_H */
