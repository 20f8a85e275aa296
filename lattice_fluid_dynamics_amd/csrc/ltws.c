/*
 * LTWS reader and writer (container of simulation/src/superres.c:43-101).
 *
 * The reference reads the sizes twice and trusts them (its element counts are
 * int products of u32 sizes, and short reads go unnoticed).  Here one pass
 * reads everything, every size is checked before it is multiplied, the
 * tensor count and the float total are bounded, and a short read is an error.
 */
#include "ltws.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void lbm_ltws_free(LBMLtws *t) {
    if (!t) return;
    free(t->ndim);
    free(t->dims);
    free(t->offset);
    free(t->data);
    memset(t, 0, sizeof *t);
}

static int read_u32(FILE *f, uint32_t *v) { return fread(v, 4, 1, f) == 1; }

int lbm_ltws_read(const char *path, LBMLtws *out) {
    if (!out) return LBM_LTWS_NO_MEMORY;
    memset(out, 0, sizeof *out);
    FILE *f = path ? fopen(path, "rb") : NULL;
    if (!f) return LBM_LTWS_CANNOT_OPEN;
    uint32_t magic = 0, version = 0, tensors = 0;
    if (!read_u32(f, &magic) || magic != LBM_LTWS_MAGIC) {
        fclose(f);
        return LBM_LTWS_BAD_MAGIC;
    }
    if (!read_u32(f, &version) || !read_u32(f, &tensors) || tensors > LBM_LTWS_MAX_TENSORS) {
        fclose(f);
        return LBM_LTWS_TRUNCATED;
    }
    LBMLtws t;
    memset(&t, 0, sizeof t);
    t.version = version;
    t.tensors = tensors;
    const size_t nt = tensors > 0 ? tensors : 1;
    t.ndim = (uint32_t *)calloc(nt, sizeof *t.ndim);
    t.dims = (uint32_t(*)[LBM_LTWS_MAX_DIMS])calloc(nt, sizeof *t.dims);
    t.offset = (long *)calloc(nt, sizeof *t.offset);
    int rc = (t.ndim && t.dims && t.offset) ? LBM_LTWS_OK : LBM_LTWS_NO_MEMORY;
    long cap = 0;
    for (uint32_t p = 0; p < tensors && rc == LBM_LTWS_OK; p++) {
        uint32_t nd = 0;
        if (!read_u32(f, &nd) || nd > LBM_LTWS_MAX_DIMS) {
            rc = LBM_LTWS_TRUNCATED;
            break;
        }
        long numel = 1;
        for (uint32_t d = 0; d < nd; d++) {
            uint32_t s = 0;
            if (!read_u32(f, &s) || (s > 0 && numel > LBM_LTWS_MAX_FLOATS / (long)s)) {
                rc = LBM_LTWS_TRUNCATED;
                break;
            }
            t.dims[p][d] = s;
            numel *= (long)s;
        }
        if (rc != LBM_LTWS_OK || numel > LBM_LTWS_MAX_FLOATS - t.total) {
            rc = LBM_LTWS_TRUNCATED;
            break;
        }
        t.ndim[p] = nd;
        t.offset[p] = t.total;
        if (t.total + numel > cap) {
            /* grow geometrically, never past the limit */
            long want = 2 * (t.total + numel);
            if (want > LBM_LTWS_MAX_FLOATS) want = LBM_LTWS_MAX_FLOATS;
            float *grown = (float *)realloc(t.data, (size_t)(want > 0 ? want : 1) * sizeof(float));
            if (!grown) {
                rc = LBM_LTWS_NO_MEMORY;
                break;
            }
            t.data = grown;
            cap = want;
        }
        if (numel > 0 && fread(t.data + t.total, sizeof(float), (size_t)numel, f) != (size_t)numel) {
            rc = LBM_LTWS_TRUNCATED;
            break;
        }
        t.total += numel;
    }
    fclose(f);
    if (rc != LBM_LTWS_OK) {
        lbm_ltws_free(&t);
        return rc;
    }
    *out = t;
    return LBM_LTWS_OK;
}

int lbm_ltws_write(const char *path, uint32_t tensors, const uint32_t *ndim,
                   const uint32_t (*dims)[LBM_LTWS_MAX_DIMS], const float *data) {
    FILE *f = path ? fopen(path, "wb") : NULL;
    if (!f) return LBM_LTWS_CANNOT_OPEN;
    const uint32_t hdr[3] = {LBM_LTWS_MAGIC, LBM_LTWS_VERSION, tensors};
    int ok = fwrite(hdr, 4, 3, f) == 3;
    long o = 0;
    for (uint32_t p = 0; p < tensors && ok; p++) {
        ok = ndim[p] <= LBM_LTWS_MAX_DIMS && fwrite(&ndim[p], 4, 1, f) == 1 &&
             fwrite(dims[p], 4, ndim[p], f) == ndim[p];
        long numel = 1;
        for (uint32_t d = 0; d < ndim[p] && ok; d++) numel *= (long)dims[p][d];
        ok = ok && fwrite(data + o, sizeof(float), (size_t)numel, f) == (size_t)numel;
        o += numel;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? LBM_LTWS_OK : LBM_LTWS_WRITE_ERROR;
}
