/*
 * CPU oracle -- D3Q19 dynamics (collide pass, stream pass, force pass).
 * TEST INFRASTRUCTURE ONLY (tests/, smoke check, bench.py CPU baseline).
 *
 * Compiled twice: REAL=float (-> or32_*) restates the reference shaders in
 * their own FP32 arithmetic; REAL=double (-> or64_*) is the FP64 oracle that
 * every result's error is stated against.  Layout and pass order are the
 * reference's: AoS f[i + 19*cell] and f_new[...] two-buffer, one collide pass
 * (simulation/shaders/lbm_collide.comp) then one stream pass
 * (simulation/shaders/lbm_stream.comp) per LBM_Step (simulation/src/lbm.c:
 * 723-821), the force pass of simulation/shaders/lbm_force.comp.
 * Cells are visited in parallel with OpenMP; every cell writes only its own
 * slots, as in the shaders, so the result does not depend on thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifdef REAL_IS_DOUBLE
#define REAL double
#define PFX(name) or64_##name
#else
#define REAL float
#define PFX(name) or32_##name
#endif

#define R(x) ((REAL)(x))
#ifdef REAL_IS_DOUBLE
#define SQRT(x) sqrt(x)
#else
#define SQRT(x) sqrtf(x)
#endif

typedef struct {
    int nx, ny, nz;
    long ncell;
    REAL tau;
    int useRegularized, useMRT, useSmagorinsky, periodicYZ;
    REAL smagorinskyCs;
    REAL *f, *f_new, *vel;
    int *solid;
    float *q;
} OrGrid;

static inline REAL wgt(int i) { return i == 0 ? R(1.0) / R(3.0) : (i < 7 ? R(1.0) / R(18.0) : R(1.0) / R(36.0)); }

/* feq of lbm_collide.comp:94-101 / lbm_force.comp:87-92 in REAL. */
static inline REAL feq(int i, REAL rho, REAL ux, REAL uy, REAL uz) {
    REAL eu = (REAL)OR_EX[i] * ux + (REAL)OR_EY[i] * uy + (REAL)OR_EZ[i] * uz;
    REAL u2 = ux * ux + uy * uy + uz * uz;
    REAL cs2 = R(1.0) / R(3.0);
    return wgt(i) * rho * (R(1.0) + eu / cs2 + (eu * eu) / (R(2.0) * cs2 * cs2) - u2 / (R(2.0) * cs2));
}

REAL PFX(feq)(int i, REAL rho, REAL ux, REAL uy, REAL uz) { return feq(i, rho, ux, uy, uz); }

/* forwardMRT: lbm_collide.comp:110-181 (d'Humieres 2002 basis, sparse rows). */
void PFX(forward_mrt)(const REAL *fi, REAL *m) {
    REAL sf = fi[1] + fi[2] + fi[3] + fi[4] + fi[5] + fi[6];
    REAL se = fi[7] + fi[8] + fi[9] + fi[10] + fi[11] + fi[12] + fi[13] + fi[14] + fi[15] + fi[16] +
              fi[17] + fi[18];
    m[0] = fi[0] + sf + se;
    m[1] = R(-30.0) * fi[0] - R(11.0) * sf + R(8.0) * se;
    m[2] = R(12.0) * fi[0] - R(4.0) * sf + se;
    m[3] = fi[1] - fi[2] + fi[7] - fi[8] + fi[9] - fi[10] + fi[11] - fi[12] + fi[13] - fi[14];
    m[4] = R(-4.0) * (fi[1] - fi[2]) + fi[7] - fi[8] + fi[9] - fi[10] + fi[11] - fi[12] + fi[13] -
           fi[14];
    m[5] = fi[3] - fi[4] + fi[7] + fi[8] - fi[9] - fi[10] + fi[15] - fi[16] + fi[17] - fi[18];
    m[6] = R(-4.0) * (fi[3] - fi[4]) + fi[7] + fi[8] - fi[9] - fi[10] + fi[15] - fi[16] + fi[17] -
           fi[18];
    m[7] = fi[5] - fi[6] + fi[11] + fi[12] - fi[13] - fi[14] + fi[15] + fi[16] - fi[17] - fi[18];
    m[8] = R(-4.0) * (fi[5] - fi[6]) + fi[11] + fi[12] - fi[13] - fi[14] + fi[15] + fi[16] - fi[17] -
           fi[18];
    REAL sXY = fi[7] + fi[8] + fi[9] + fi[10];
    REAL sXZ = fi[11] + fi[12] + fi[13] + fi[14];
    REAL sYZ = fi[15] + fi[16] + fi[17] + fi[18];
    m[9] = R(2.0) * (fi[1] + fi[2]) - (fi[3] + fi[4]) - (fi[5] + fi[6]) + sXY + sXZ - R(2.0) * sYZ;
    m[10] = R(-4.0) * (fi[1] + fi[2]) + R(2.0) * (fi[3] + fi[4]) + R(2.0) * (fi[5] + fi[6]) + sXY +
            sXZ - R(2.0) * sYZ;
    m[11] = (fi[3] + fi[4]) - (fi[5] + fi[6]) + sXY - sXZ;
    m[12] = R(-2.0) * (fi[3] + fi[4]) + R(2.0) * (fi[5] + fi[6]) + sXY - sXZ;
    m[13] = fi[7] - fi[8] - fi[9] + fi[10];
    m[14] = fi[15] - fi[16] - fi[17] + fi[18];
    m[15] = fi[11] - fi[12] - fi[13] + fi[14];
    m[16] = fi[7] - fi[8] + fi[9] - fi[10] - fi[11] + fi[12] - fi[13] + fi[14];
    m[17] = -fi[7] - fi[8] + fi[9] + fi[10] + fi[15] - fi[16] + fi[17] - fi[18];
    m[18] = fi[11] + fi[12] - fi[13] - fi[14] - fi[15] - fi[16] + fi[17] + fi[18];
}

/* inverseMRT: lbm_collide.comp:185-311 (M^-1 = M^T diag(1/|row|^2)). */
void PFX(inverse_mrt)(const REAL *m, REAL *fi) {
    static const double norm[19] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36,
                                    72, 12,   24,  4,  4,  4,  8,  8,  8};
    REAL ms[19];
    for (int k = 0; k < 19; k++) ms[k] = m[k] / (REAL)norm[k];
    fi[0] = ms[0] - R(30.0) * ms[1] + R(12.0) * ms[2];
    fi[1] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] + ms[3] - R(4.0) * ms[4] + R(2.0) * ms[9] -
            R(4.0) * ms[10];
    fi[2] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] - ms[3] + R(4.0) * ms[4] + R(2.0) * ms[9] -
            R(4.0) * ms[10];
    fi[3] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] + ms[5] - R(4.0) * ms[6] - ms[9] +
            R(2.0) * ms[10] + ms[11] - R(2.0) * ms[12];
    fi[4] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] - ms[5] + R(4.0) * ms[6] - ms[9] +
            R(2.0) * ms[10] + ms[11] - R(2.0) * ms[12];
    fi[5] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] + ms[7] - R(4.0) * ms[8] - ms[9] +
            R(2.0) * ms[10] - ms[11] + R(2.0) * ms[12];
    fi[6] = ms[0] - R(11.0) * ms[1] - R(4.0) * ms[2] - ms[7] + R(4.0) * ms[8] - ms[9] +
            R(2.0) * ms[10] - ms[11] + R(2.0) * ms[12];
    fi[7] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[3] + ms[4] + ms[5] + ms[6] + ms[9] + ms[10] + ms[11] +
            ms[12] + ms[13] + ms[16] - ms[17];
    fi[8] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[3] - ms[4] + ms[5] + ms[6] + ms[9] + ms[10] + ms[11] +
            ms[12] - ms[13] - ms[16] - ms[17];
    fi[9] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[3] + ms[4] - ms[5] - ms[6] + ms[9] + ms[10] + ms[11] +
            ms[12] - ms[13] + ms[16] + ms[17];
    fi[10] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[3] - ms[4] - ms[5] - ms[6] + ms[9] + ms[10] +
             ms[11] + ms[12] + ms[13] - ms[16] + ms[17];
    fi[11] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[3] + ms[4] + ms[7] + ms[8] + ms[9] + ms[10] -
             ms[11] - ms[12] + ms[15] - ms[16] + ms[18];
    fi[12] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[3] - ms[4] + ms[7] + ms[8] + ms[9] + ms[10] -
             ms[11] - ms[12] - ms[15] + ms[16] + ms[18];
    fi[13] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[3] + ms[4] - ms[7] - ms[8] + ms[9] + ms[10] -
             ms[11] - ms[12] - ms[15] - ms[16] - ms[18];
    fi[14] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[3] - ms[4] - ms[7] - ms[8] + ms[9] + ms[10] -
             ms[11] - ms[12] + ms[15] + ms[16] - ms[18];
    fi[15] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[5] + ms[6] + ms[7] + ms[8] - R(2.0) * ms[9] -
             R(2.0) * ms[10] + ms[14] + ms[17] - ms[18];
    fi[16] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[5] - ms[6] + ms[7] + ms[8] - R(2.0) * ms[9] -
             R(2.0) * ms[10] - ms[14] - ms[17] - ms[18];
    fi[17] = ms[0] + R(8.0) * ms[1] + ms[2] + ms[5] + ms[6] - ms[7] - ms[8] - R(2.0) * ms[9] -
             R(2.0) * ms[10] - ms[14] + ms[17] + ms[18];
    fi[18] = ms[0] + R(8.0) * ms[1] + ms[2] - ms[5] - ms[6] - ms[7] - ms[8] - R(2.0) * ms[9] -
             R(2.0) * ms[10] + ms[14] - ms[17] + ms[18];
}

/* Equilibrium moments: lbm_collide.comp:560-580. */
void PFX(meq)(REAL rho, REAL ux, REAL uy, REAL uz, REAL *meq) {
    REAL u2 = ux * ux + uy * uy + uz * uz;
    meq[0] = rho;
    meq[1] = R(-11.0) * rho + R(19.0) * rho * u2;
    meq[2] = R(3.0) * rho - R(5.5) * rho * u2;
    meq[3] = rho * ux;
    meq[4] = R(-2.0) / R(3.0) * rho * ux;
    meq[5] = rho * uy;
    meq[6] = R(-2.0) / R(3.0) * rho * uy;
    meq[7] = rho * uz;
    meq[8] = R(-2.0) / R(3.0) * rho * uz;
    meq[9] = rho * (R(2.0) * ux * ux - uy * uy - uz * uz);
    meq[10] = R(-0.5) * rho * (R(2.0) * ux * ux - uy * uy - uz * uz);
    meq[11] = rho * (uy * uy - uz * uz);
    meq[12] = R(-0.5) * rho * (uy * uy - uz * uz);
    meq[13] = rho * ux * uy;
    meq[14] = rho * uy * uz;
    meq[15] = rho * ux * uz;
    meq[16] = 0;
    meq[17] = 0;
    meq[18] = 0;
}

/* One cell of the collide pass, lbm_collide.comp:313-626.  `fin` holds the
 * cell's 19 post-stream populations and is updated in place by the Zou-He
 * rewrites (the shader writes them back into f); `fout` receives f_new and
 * `vel` the vec4 (u, rho).  params: tau, Cs, flags, inlet u, x position. */
void PFX(collide_cell)(REAL *fin, REAL *fout, REAL *vel, int solid, int x, int nx, REAL tau,
                       int useRegularized, int useMRT, int useSmagorinsky, REAL Cs, REAL inx,
                       REAL iny, REAL inz) {
    if (solid == 1) {
        for (int i = 0; i < 19; i++) fout[i] = fin[OR_OPP[i]];
        vel[0] = 0;
        vel[1] = 0;
        vel[2] = 0;
        vel[3] = 1;
        return;
    }
    if (x == 0) { /* Zou-He velocity inlet, :336-375 */
        REAL knowns_0 =
            fin[0] + fin[3] + fin[4] + fin[5] + fin[6] + fin[15] + fin[16] + fin[17] + fin[18];
        REAL knowns_neg = fin[2] + fin[8] + fin[10] + fin[12] + fin[14];
        REAL rho_in = (knowns_0 + R(2.0) * knowns_neg) / (R(1.0) - inx);
        REAL f2 = fin[2], f8 = fin[8], f10 = fin[10], f12 = fin[12], f14 = fin[14];
        fin[1] = f2 + (R(1.0) / R(3.0)) * rho_in * inx;
        fin[7] = f10 + (R(1.0) / R(6.0)) * rho_in * (inx + iny);
        fin[9] = f8 + (R(1.0) / R(6.0)) * rho_in * (inx - iny);
        fin[11] = f14 + (R(1.0) / R(6.0)) * rho_in * (inx + inz);
        fin[13] = f12 + (R(1.0) / R(6.0)) * rho_in * (inx - inz);
    }
    if (x == nx - 1) { /* Zou-He pressure outlet rho=1, :379-408 */
        REAL knowns_0 =
            fin[0] + fin[3] + fin[4] + fin[5] + fin[6] + fin[15] + fin[16] + fin[17] + fin[18];
        REAL knowns_pos = fin[1] + fin[7] + fin[9] + fin[11] + fin[13];
        REAL rho_out = R(1.0);
        REAL ux = R(-1.0) + (knowns_0 + R(2.0) * knowns_pos) / rho_out;
        REAL f1 = fin[1], f7 = fin[7], f9 = fin[9], f11 = fin[11], f13 = fin[13];
        fin[2] = f1 - (R(1.0) / R(3.0)) * rho_out * ux;
        fin[10] = f7 - (R(1.0) / R(6.0)) * rho_out * ux;
        fin[8] = f9 - (R(1.0) / R(6.0)) * rho_out * ux;
        fin[14] = f11 - (R(1.0) / R(6.0)) * rho_out * ux;
        fin[12] = f13 - (R(1.0) / R(6.0)) * rho_out * ux;
    }
    /* moments, :411-424 (summed in direction order) */
    REAL rho = 0, ux = 0, uy = 0, uz = 0;
    for (int i = 0; i < 19; i++) {
        REAL fi = fin[i];
        rho += fi;
        ux += (REAL)OR_EX[i] * fi;
        uy += (REAL)OR_EY[i] * fi;
        uz += (REAL)OR_EZ[i] * fi;
    }
    if (rho > R(0.0001)) {
        ux /= rho;
        uy /= rho;
        uz /= rho;
    }
    /* stability guard, :429-440 */
    REAL umag = SQRT(ux * ux + uy * uy + uz * uz);
    if (rho < R(0.3) || rho > R(2.0) || umag > R(0.40) || isinf(rho) || isnan(rho)) {
        for (int i = 0; i < 19; i++) {
            REAL e = feq(i, R(1.0), 0, 0, 0);
            fin[i] = e;
            fout[i] = e;
        }
        vel[0] = 0;
        vel[1] = 0;
        vel[2] = 0;
        vel[3] = 1;
        return;
    }
    vel[0] = ux;
    vel[1] = uy;
    vel[2] = uz;
    vel[3] = rho;
    REAL tau_eff = tau;
    if (useSmagorinsky == 1) { /* :448-484 */
        REAL Sxx = 0, Syy = 0, Szz = 0, Sxy = 0, Sxz = 0, Syz = 0;
        for (int i = 0; i < 19; i++) {
            REAL fneq = fin[i] - feq(i, rho, ux, uy, uz);
            REAL ex = (REAL)OR_EX[i], ey = (REAL)OR_EY[i], ez = (REAL)OR_EZ[i];
            Sxx += ex * ex * fneq;
            Syy += ey * ey * fneq;
            Szz += ez * ez * fneq;
            Sxy += ex * ey * fneq;
            Sxz += ex * ez * fneq;
            Syz += ey * ez * fneq;
        }
        REAL sq = Sxx * Sxx + Syy * Syy + Szz * Szz + R(2.0) * (Sxy * Sxy + Sxz * Sxz + Syz * Syz);
        REAL Pi_norm = SQRT(sq);
        REAL Cs2 = Cs * Cs;
        REAL inner = tau * tau + R(18.0) * Cs2 * Pi_norm / (rho + R(1e-10));
        REAL root = SQRT(inner);
        tau_eff = R(0.5) * (tau + root);
    }
    REAL omega = R(1.0) / tau_eff;
    if (useRegularized == 1) { /* :488-543 */
        REAL Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
        for (int i = 0; i < 19; i++) {
            REAL fneq = fin[i] - feq(i, rho, ux, uy, uz);
            REAL ex = (REAL)OR_EX[i], ey = (REAL)OR_EY[i], ez = (REAL)OR_EZ[i];
            Pxx += ex * ex * fneq;
            Pyy += ey * ey * fneq;
            Pzz += ez * ez * fneq;
            Pxy += ex * ey * fneq;
            Pxz += ex * ez * fneq;
            Pyz += ey * ez * fneq;
        }
        REAL inv_2cs4 = R(4.5);
        for (int i = 0; i < 19; i++) {
            REAL fe = feq(i, rho, ux, uy, uz);
            REAL ex = (REAL)OR_EX[i], ey = (REAL)OR_EY[i], ez = (REAL)OR_EZ[i];
            REAL cs2 = R(1.0) / R(3.0);
            REAL QP = (ex * ex - cs2) * Pxx + (ey * ey - cs2) * Pyy + (ez * ez - cs2) * Pzz +
                      R(2.0) * ex * ey * Pxy + R(2.0) * ex * ez * Pxz + R(2.0) * ey * ez * Pyz;
            REAL fneq_reg = wgt(i) * inv_2cs4 * QP;
            fout[i] = fe + (R(1.0) - omega) * fneq_reg;
        }
    } else if (useMRT == 1) { /* :544-616 */
        REAL m[19], me[19];
        PFX(forward_mrt)(fin, m);
        PFX(meq)(rho, ux, uy, uz, me);
        REAL s[19] = {0,    R(1.19), R(1.4), 0,     R(1.2), 0,      R(1.2), 0,      R(1.2), omega,
                      R(1.4), omega,   R(1.4), omega, omega,  omega,  R(1.98), R(1.98), R(1.98)};
        for (int k = 0; k < 19; k++) m[k] = m[k] - s[k] * (m[k] - me[k]);
        PFX(inverse_mrt)(m, fout);
    } else { /* BGK, :618-625 */
        for (int i = 0; i < 19; i++) {
            REAL fi = fin[i];
            fout[i] = fi - omega * (fi - feq(i, rho, ux, uy, uz));
        }
    }
}

/* ---- grid object ---------------------------------------------------------------- */

OrGrid *PFX(create)(int nx, int ny, int nz, float viscosity) {
    OrGrid *g = (OrGrid *)calloc(1, sizeof(OrGrid));
    if (!g) return NULL;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->ncell = (long)nx * ny * nz;
    g->tau = (REAL)or_tau_from_viscosity(viscosity);
    g->smagorinskyCs = (REAL)0.1f;
    g->f = (REAL *)calloc(19 * g->ncell, sizeof(REAL));
    g->f_new = (REAL *)calloc(19 * g->ncell, sizeof(REAL));
    g->vel = (REAL *)calloc(4 * g->ncell, sizeof(REAL));
    g->solid = (int *)calloc(g->ncell, sizeof(int));
    g->q = (float *)malloc(19 * g->ncell * sizeof(float));
    if (!g->f || !g->f_new || !g->vel || !g->solid || !g->q) {
        free(g->f);
        free(g->f_new);
        free(g->vel);
        free(g->solid);
        free(g->q);
        free(g);
        return NULL;
    }
    for (long k = 0; k < 19 * g->ncell; k++) g->q[k] = -1.0f;
    return g;
}

void PFX(free)(OrGrid *g) {
    if (!g) return;
    free(g->f);
    free(g->f_new);
    free(g->vel);
    free(g->solid);
    free(g->q);
    free(g);
}

void PFX(set_params)(OrGrid *g, double tau, int useRegularized, int useMRT, int useSmagorinsky,
                     double Cs, int periodicYZ) {
    g->tau = (REAL)tau;
    g->useRegularized = useRegularized;
    g->useMRT = useMRT;
    g->useSmagorinsky = useSmagorinsky;
    g->smagorinskyCs = (REAL)Cs;
    g->periodicYZ = periodicYZ;
}

REAL *PFX(f)(OrGrid *g) { return g->f; }
REAL *PFX(f_new)(OrGrid *g) { return g->f_new; }
REAL *PFX(vel)(OrGrid *g) { return g->vel; }
int *PFX(solid)(OrGrid *g) { return g->solid; }
float *PFX(q)(OrGrid *g) { return g->q; }
double PFX(tau)(OrGrid *g) { return (double)g->tau; }

/* LBM_InitializeFlow, simulation/src/lbm.c:670-721: f = f_new = feq(1,u) of
 * the host float feq, vel = (u, 1).  Values are produced in float as the
 * host does, then widened for the FP64 oracle. */
int PFX(initialize_flow)(OrGrid *g, float ux, float uy, float uz) {
    float e[19];
    for (int i = 0; i < 19; i++) e[i] = or_feq_f32(i, 1.0f, ux, uy, uz);
    for (long c = 0; c < g->ncell; c++) {
        for (int i = 0; i < 19; i++) {
            g->f[19 * c + i] = (REAL)e[i];
            g->f_new[19 * c + i] = (REAL)e[i];
        }
        g->vel[4 * c + 0] = (REAL)ux;
        g->vel[4 * c + 1] = (REAL)uy;
        g->vel[4 * c + 2] = (REAL)uz;
        g->vel[4 * c + 3] = R(1.0);
    }
    return 1;
}

/* handleYZ, lbm_stream.comp:63-74 */
static inline void handle_yz(const OrGrid *g, int *y, int *z) {
    if (g->periodicYZ == 1) {
        if (*y < 0) *y += g->ny;
        if (*y >= g->ny) *y -= g->ny;
        if (*z < 0) *z += g->nz;
        if (*z >= g->nz) *z -= g->nz;
    } else {
        if (*y < 0) *y = 0;
        if (*y > g->ny - 1) *y = g->ny - 1;
        if (*z < 0) *z = 0;
        if (*z > g->nz - 1) *z = g->nz - 1;
    }
}

static void collide_pass(OrGrid *g, float inx, float iny, float inz) {
    const int nx = g->nx;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < g->ncell; c++) {
        int x = (int)(c % nx);
        PFX(collide_cell)(g->f + 19 * c, g->f_new + 19 * c, g->vel + 4 * c, g->solid[c], x, nx,
                          g->tau, g->useRegularized, g->useMRT, g->useSmagorinsky, g->smagorinskyCs,
                          (REAL)inx, (REAL)iny, (REAL)inz);
    }
}

/* The stream rule of lbm_stream.comp:85-128 for one cell and direction,
 * reading post-collision populations from fn. */
static inline REAL stream_value(const OrGrid *g, const REAL *fn, int x, int y, int z, int i) {
    const int nx = g->nx, ny = g->ny;
    long c = x + (long)nx * (y + (long)ny * z);
    int io = OR_OPP[i];
    float q = g->q[19 * c + io];
    if (q >= 0.0f) {
        REAL fos = fn[19 * c + io];
        REAL qq = (REAL)q;
        if (q >= 0.5f) {
            REAL fis = fn[19 * c + i];
            REAL inv2q = R(1.0) / (R(2.0) * qq);
            return inv2q * fos + (R(1.0) - inv2q) * fis;
        } else {
            int X = x + OR_EX[i], Y = y + OR_EY[i], Z = z + OR_EZ[i];
            handle_yz(g, &Y, &Z);
            REAL ff;
            if (X >= 0 && X < nx)
                ff = fn[19 * (X + (long)nx * (Y + (long)ny * Z)) + io];
            else
                ff = fos;
            REAL twoq = R(2.0) * qq;
            return twoq * fos + (R(1.0) - twoq) * ff;
        }
    }
    int X = x - OR_EX[i], Y = y - OR_EY[i], Z = z - OR_EZ[i];
    handle_yz(g, &Y, &Z);
    if (X < 0 || X >= nx) return fn[19 * c + i];
    return fn[19 * (X + (long)nx * (Y + (long)ny * Z)) + i];
}

static void stream_pass(OrGrid *g) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                long c = x + (long)nx * (y + (long)ny * z);
                for (int i = 0; i < 19; i++) g->f[19 * c + i] = stream_value(g, g->f_new, x, y, z, i);
            }
}

/* LBM_Step, simulation/src/lbm.c:723-821 (single slab). */
void PFX(step)(OrGrid *g, float inx, float iny, float inz) {
    collide_pass(g, inx, iny, inz);
    stream_pass(g);
}

void PFX(run)(OrGrid *g, int nsteps, float inx, float iny, float inz) {
    for (int s = 0; s < nsteps; s++) PFX(step)(g, inx, iny, inz);
}

/* Force pass, lbm_force.comp:94-151, over fluid cells (solid==0).  Outputs:
 *   exact[0..5]  = double sums of the per-cell REAL force (F xyz, P xyz),
 *                  summed in cell order;
 *   fixedp[0..6] = the reference's integer accumulators: round(F*1e4) per
 *                  cell, the boundary-cell count, round(P*1e4) per cell.
 * Host conversion /1e4 is simulation/src/lbm.c:908-917. */
void PFX(force)(const OrGrid *g, double *exact, long long *fixedp) {
    for (int k = 0; k < 6; k++) exact[k] = 0;
    for (int k = 0; k < 7; k++) fixedp[k] = 0;
    for (long c = 0; c < g->ncell; c++) {
        if (g->solid[c] != 0) continue;
        REAL ux = g->vel[4 * c + 0], uy = g->vel[4 * c + 1], uz = g->vel[4 * c + 2];
        REAL rho = g->vel[4 * c + 3];
        REAL F[3] = {0, 0, 0}, P[3] = {0, 0, 0};
        int has = 0;
        for (int i = 1; i < 19; i++) {
            if (g->q[19 * c + i] >= 0.0f) {
                has = 1;
                int io = OR_OPP[i];
                REAL ft = g->f_new[19 * c + i] + g->f[19 * c + io];
                F[0] += ft * (REAL)OR_EX[i];
                F[1] += ft * (REAL)OR_EY[i];
                F[2] += ft * (REAL)OR_EZ[i];
                REAL fp = feq(i, rho, ux, uy, uz) + feq(io, rho, ux, uy, uz);
                P[0] += fp * (REAL)OR_EX[i];
                P[1] += fp * (REAL)OR_EY[i];
                P[2] += fp * (REAL)OR_EZ[i];
            }
        }
        if (has) {
            for (int k = 0; k < 3; k++) {
                exact[k] += (double)F[k];
                exact[3 + k] += (double)P[k];
                fixedp[k] += llround((double)(F[k] * R(10000.0)));
                fixedp[4 + k] += llround((double)(P[k] * R(10000.0)));
            }
            fixedp[3] += 1;
        }
    }
}
