/*
 * Post-processing around the solver path, host C:
 *   - the caller-side drag/lift bookkeeping of simulation/src/main.c:1250-1391
 *     (Cd, Cl, pressure/friction split, blockage correction, EMA, convergence
 *     over the last 100 samples) and its stdout line protocol;
 *   - the Strouhal estimate of simulation/src/drag_metrics.c:8-68;
 *   - the VTK ImageData field dump of simulation/src/vti_export.c:10-116.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbm.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* CD_HISTORY_SIZE / CD_SAMPLE_INTERVAL of simulation/lib/drag_metrics.h:4-5 */
#define HISTORY 100

struct LBMCdMonitor {
    float U, refArea, blockageFactor;
    float history[HISTORY];
    int count;
    int converged;
    float ema;
    float convergedMean, convergedRelStd;
    float last[6]; /* Cd, CdCorr, Cl, CdPressure, CdFriction, ema */
    float *cl;
    int clCount, clCap;
    int verbose;
    /* body description for the one-time breakdown (LBM_CdMonitorSetBody) */
    int hasBody, breakdownPrinted;
    float reynolds, charLength, bodyY, bodyZ, bboxArea, epsilon;
};

LBMCdMonitor *LBM_CdMonitorCreate(LBMGrid *grid, float U, float refArea, int verbose) {
    LBMCdMonitor *m = (LBMCdMonitor *)calloc(1, sizeof(LBMCdMonitor));
    if (!m) return NULL;
    m->U = U;
    m->refArea = refArea;
    /* blockage epsilon = A / (NY NZ), factor (1 - eps)^2 (main.c:549-552) */
    float eps = grid ? refArea / ((float)grid->sizeY * grid->sizeZ) : 0.0f;
    m->blockageFactor = (1.0f - eps) * (1.0f - eps);
    m->verbose = verbose;
    return m;
}

/* main.c keeps the model's world bounding box (carBounds) and the Reynolds
 * number; they feed the one-time "Cd calculation breakdown" of
 * main.c:1273-1306, printed by the first verbose LBM_CdMonitorSample.
 * Lattice extents use the world-to-grid scales of main.c:439 and 545-546
 * (x: NX/8, y: NY/4, z: NZ/4). */
void LBM_CdMonitorSetBody(LBMCdMonitor *m, LBMGrid *grid, float reynolds, float minX, float minY, float minZ,
                          float maxX, float maxY, float maxZ) {
    if (!m || !grid) return;
    m->hasBody = 1;
    m->reynolds = reynolds;
    m->charLength = (maxX - minX) * (grid->sizeX / 8.0f);
    m->bodyY = (maxY - minY) * (grid->sizeY / 4.0f);
    m->bodyZ = (maxZ - minZ) * (grid->sizeZ / 4.0f);
    m->bboxArea = m->bodyY * m->bodyZ;
    /* main.c:547-552: projected area (bbox if none), blockage, correction */
    if (m->refArea < 1.0f) m->refArea = m->bboxArea;
    m->epsilon = m->refArea / ((float)grid->sizeY * grid->sizeZ);
    m->blockageFactor = (1.0f - m->epsilon) * (1.0f - m->epsilon);
}

/* Reynolds number from the lattice inlet speed, the body length along x and
 * nu = (tau - 1/2) / 3, with the "Effective Re" line of main.c:524-537. */
float LBM_EffectiveReynolds(LBMGrid *grid, float U, float minX, float maxX, int print) {
    if (!grid) return 0.0f;
    const float charL = (maxX - minX) * (grid->sizeX / 8.0f);
    const float nu = (grid->tau - 0.5f) / 3.0f;
    const float re = (nu > 1e-10f) ? U * charL / nu : 0.0f;
    if (print) printf("Effective Re = %.0f (U=%.3f, L=%.1f, nu=%.4f)\n", re, U, charL, nu);
    return re;
}

static void print_breakdown(LBMCdMonitor *m) {
    printf("  Cd calculation breakdown:\n");
    printf("    Re = %.0f, U_lattice = %.4f, rho_0 = %.1f\n", m->reynolds, m->U, 1.0f);
    printf("    Body (lattice): %.1f x %.1f x %.1f cells\n", m->charLength, m->bodyY, m->bodyZ);
    printf("    Projected area = %.1f cells^2 (bbox = %.1f)\n", m->refArea, m->bboxArea);
    printf("    Blockage = %.1f%%, correction = %.3f (Pope & Harper 1966)\n", m->epsilon * 100.0f,
           m->blockageFactor);
    if (fminf(m->bodyY, m->bodyZ) < 10.0f)
        printf("    WARNING: body < 10 cells across -- Cd is unreliable at this resolution\n");
    if (m->epsilon > 0.05f) printf("    WARNING: blockage > 5%% -- confinement inflates Cd\n");
    printf("    NOTE: published Ahmed body Cd ~0.3 is at Re > 500k. Low-Re simulations give higher Cd.\n");
}

void LBM_CdMonitorFree(LBMCdMonitor *m) {
    if (!m) return;
    free(m->cl);
    free(m);
}

/* One sample: force readback and the coefficient algebra of main.c:1260-1271,
 * EMA (alpha 0.1) of the corrected Cd (1309-1314), convergence when the
 * relative standard deviation of the last 100 raw Cd samples drops below 1 %
 * (1357-1391).  Returns 1 the first time convergence is reached. */
int LBM_CdMonitorSample(LBMCdMonitor *m, LBMGrid *grid, float out[6]) {
    float fx, fy, fz, px, py, pz;
    LBM_ComputeDragForceDecomposed(grid, &fx, &fy, &fz, &px, &py, &pz);
    float dynP = 0.5f * 1.0f * m->U * m->U;
    float denom = dynP * m->refArea;
    float Cd = (denom > 1e-10f) ? fabsf(fx) / denom : 0.0f;
    float Cl = (denom > 1e-10f) ? fy / denom : 0.0f;
    float CdP = (denom > 1e-10f) ? fabsf(px) / denom : 0.0f;
    float CdF = Cd - CdP;
    if (CdF < 0.0f) CdF = 0.0f;
    float CdCorr = Cd * m->blockageFactor;
    if (m->verbose && m->hasBody && !m->breakdownPrinted) {
        print_breakdown(m);
        m->breakdownPrinted = 1;
    }
    /* main.c:1309-1314: the EMA restarts from the sample itself while no Cd
     * has entered the convergence history yet (only 0 < Cd < 1000 does). */
    if (m->count == 0)
        m->ema = CdCorr;
    else
        m->ema = 0.1f * CdCorr + 0.9f * m->ema;
    if (m->clCount == m->clCap) {
        int cap = m->clCap ? 2 * m->clCap : 256;
        float *p = (float *)realloc(m->cl, sizeof(float) * cap);
        if (p) {
            m->cl = p;
            m->clCap = cap;
        }
    }
    if (m->clCount < m->clCap) m->cl[m->clCount++] = Cl;
    if (m->verbose) {
        printf("  Drag Force: (%.4f, %.4f, %.4f), Cd=%.3f Cd_corr=%.3f Cl=%.3f (avg=%.3f)\n", fx, fy, fz,
               Cd, CdCorr, Cl, m->ema);
        printf("  Cd_pressure=%.3f Cd_friction=%.3f\n", CdP, CdF);
    }
    int newly = 0;
    if (Cd > 0 && Cd < 1000) {
        m->history[m->count % HISTORY] = Cd;
        m->count++;
        if (m->count >= HISTORY && !m->converged) {
            float mean = 0;
            for (int j = 0; j < HISTORY; j++) mean += m->history[j];
            mean /= HISTORY;
            float var = 0;
            for (int j = 0; j < HISTORY; j++) {
                float d = m->history[j] - mean;
                var += d * d;
            }
            float rel = sqrtf(var / HISTORY) / (mean + 1e-10f);
            if (rel < 0.01f) {
                m->converged = 1;
                m->convergedMean = mean;
                m->convergedRelStd = rel;
                newly = 1;
                if (m->verbose) printf("  Cd converged (mean=%.3f, relStd=%.4f)\n", mean, rel);
            }
        }
    }
    m->last[0] = Cd;
    m->last[1] = CdCorr;
    m->last[2] = Cl;
    m->last[3] = CdP;
    m->last[4] = CdF;
    m->last[5] = m->ema;
    if (out) memcpy(out, m->last, sizeof(m->last));
    return newly;
}

int LBM_CdMonitorConverged(const LBMCdMonitor *m, float *mean, float *relStd) {
    if (mean) *mean = m->convergedMean;
    if (relStd) *relStd = m->convergedRelStd;
    return m->converged;
}

/* Strouhal number from a Cl series (drag_metrics.c:8-68): drop the first 35 %,
 * remove the mean, Hann window, one-sided DFT without DC, St = f_peak L / U
 * with f in cycles per lattice step (one sample per sampleInterval steps). */
float LBM_ComputeStrouhal(const float *cl, int n, int sampleInterval, float charLength, float U) {
    if (!cl || n < 12 || sampleInterval <= 0) return 0.0f;
    int start = n * 35 / 100;
    int m = n - start;
    const float *sig = cl + start;
    double mean = 0;
    for (int i = 0; i < m; i++) mean += sig[i];
    mean /= m;
    double *win = (double *)malloc(sizeof(double) * m);
    if (!win) return 0.0f;
    for (int i = 0; i < m; i++) win[i] = (sig[i] - mean) * 0.5 * (1.0 - cos(2.0 * M_PI * i / (m - 1)));
    double best = 0, fpk = 0, fs = 1.0 / sampleInterval;
    for (int k = 1; k < m / 2; k++) {
        double re = 0, im = 0;
        for (int i = 0; i < m; i++) {
            double a = 2.0 * M_PI * (double)k * i / m;
            re += win[i] * cos(a);
            im -= win[i] * sin(a);
        }
        double p = re * re + im * im;
        if (p > best) {
            best = p;
            fpk = k * fs / m;
        }
    }
    free(win);
    return U > 1e-10f ? (float)(fpk * charLength / U) : 0.0f;
}

float LBM_CdMonitorStrouhal(const LBMCdMonitor *m, int sampleInterval, float charLength) {
    return LBM_ComputeStrouhal(m->cl, m->clCount, sampleInterval, charLength, m->U);
}

/* VTK ImageData, raw appended, UInt64 block headers: velocity (3 x Float32),
 * rho (Float32), solid (Int32), file <dir>/field_<step %06d>.vti
 * (simulation/src/vti_export.c:10-116). */
int LBM_WriteVTI(LBMGrid *grid, const char *dir, int step) {
    char name[1024];
    snprintf(name, sizeof(name), "%s/field_%06d.vti", dir, step);
    const int nx = grid->sizeX, ny = grid->sizeY, nz = grid->sizeZ;
    const size_t n = (size_t)nx * ny * nz;
    float *vel = (float *)malloc(n * 16);
    int *solid = (int *)malloc(n * 4);
    float *buf = (float *)malloc(n * 12);
    FILE *f = NULL;
    int ok = vel && solid && buf && LBM_CopyVelocityToHost(grid, vel) && LBM_CopySolidToHost(grid, solid);
    if (ok) ok = (f = fopen(name, "wb")) != NULL;
    if (ok) {
        uint64_t velBytes = (uint64_t)n * 12, rhoBytes = (uint64_t)n * 4, solidBytes = (uint64_t)n * 4;
        unsigned long rhoOff = (unsigned long)(8 + velBytes);
        unsigned long solidOff = (unsigned long)(rhoOff + 8 + rhoBytes);
        fprintf(f,
                "<?xml version=\"1.0\"?>\n"
                "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
                "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
                "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n"
                "      <PointData Vectors=\"velocity\" Scalars=\"rho\">\n"
                "        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" "
                "format=\"appended\" offset=\"0\"/>\n"
                "        <DataArray type=\"Float32\" Name=\"rho\" format=\"appended\" offset=\"%lu\"/>\n"
                "        <DataArray type=\"Int32\" Name=\"solid\" format=\"appended\" offset=\"%lu\"/>\n"
                "      </PointData>\n"
                "    </Piece>\n"
                "  </ImageData>\n"
                "  <AppendedData encoding=\"raw\">\n_",
                nx, ny, nz, nx, ny, nz, rhoOff, solidOff);
        for (size_t c = 0; c < n; c++) memcpy(buf + 3 * c, vel + 4 * c, 12);
        ok = fwrite(&velBytes, 8, 1, f) == 1 && fwrite(buf, 12, n, f) == n;
        for (size_t c = 0; c < n; c++) buf[c] = vel[4 * c + 3];
        ok = ok && fwrite(&rhoBytes, 8, 1, f) == 1 && fwrite(buf, 4, n, f) == n;
        ok = ok && fwrite(&solidBytes, 8, 1, f) == 1 && fwrite(solid, 4, n, f) == n;
        ok = ok && fprintf(f, "\n  </AppendedData>\n</VTKFile>\n") > 0;
    }
    if (f) fclose(f);
    free(vel);
    free(solid);
    free(buf);
    if (ok) printf("VTK: wrote %s\n", name);
    return ok;
}
