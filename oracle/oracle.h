/*
 * CPU oracle for the D3Q19 lattice-Boltzmann path -- TEST INFRASTRUCTURE.
 * Only tests/, the smoke check and bench.py's CPU-baseline leg load it.
 * It restates the reference's GLSL collide/stream/force passes and the host
 * setup of simulation/src/lbm.c in the reference's AoS two-buffer order
 * (f[i + 19*cell]); it is the checker, never the product.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

extern const int OR_EX[19], OR_EY[19], OR_EZ[19], OR_OPP[19];
extern const float OR_W[19];

float or_feq_f32(int i, float rho, float ux, float uy, float uz);
float or_tau_from_viscosity(float viscosity);
void or_set_solid_aabb(int nx, int ny, int nz, int *solid, float *q, float minX, float minY,
                       float minZ, float maxX, float maxY, float maxZ);
void or_set_solid_sphere(int nx, int ny, int nz, int *solid, float *q, float cx, float cy, float cz,
                         float radius);
void or_add_ground_plane(int nx, int ny, int nz, int *solid, float *q, float worldZ);
float or_projected_area(int nx, int ny, int nz, const int *solid, int axis);
void or_set_solid_mesh(int nx, int ny, int nz, int *solid, float *q, const float *tris, int ntri,
                       float minX, float minY, float minZ, float maxX, float maxY, float maxZ);

#ifdef __cplusplus
}
#endif
#endif
