/*
 * LTWS weight container, shared by the warm-start network and the
 * super-resolution upscaler (format of simulation/src/superres.c:43-101):
 *   u32 magic 0x4C545753 ("LTWS"), u32 version, u32 tensor count, then per
 *   tensor u32 ndim, ndim u32 sizes and the float32 data, little endian.
 * Host code only; the reader bounds every count it takes from the file.
 */
#ifndef LBM_LTWS_H
#define LBM_LTWS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_LTWS_MAGIC 0x4C545753u
#define LBM_LTWS_VERSION 1u
#define LBM_LTWS_MAX_DIMS 8
#define LBM_LTWS_MAX_TENSORS 65536u
#define LBM_LTWS_MAX_FLOATS (1L << 28) /* 1 GiB of weights per file */

typedef struct {
    uint32_t version;
    uint32_t tensors;
    uint32_t *ndim;                      /* per tensor */
    uint32_t (*dims)[LBM_LTWS_MAX_DIMS]; /* per tensor, ndim entries used */
    long *offset;                        /* per tensor, first float in data */
    long total;                          /* floats in data */
    float *data;                         /* tensors concatenated in file order */
} LBMLtws;

enum {
    LBM_LTWS_OK = 0,
    LBM_LTWS_CANNOT_OPEN = -1,
    LBM_LTWS_BAD_MAGIC = -2,
    LBM_LTWS_TRUNCATED = -3, /* short read, or a count beyond the limits above */
    LBM_LTWS_NO_MEMORY = -4,
    LBM_LTWS_WRITE_ERROR = -5
};

/* Reads a whole file; on success *out owns its arrays (lbm_ltws_free).  On
 * failure *out is left empty.  Returns one of the codes above. */
int lbm_ltws_read(const char *path, LBMLtws *out);
void lbm_ltws_free(LBMLtws *t);

/* Writes `tensors` tensors whose shapes are ndim[t], dims[t][0..ndim[t]) and
 * whose data follow each other in `data`. */
int lbm_ltws_write(const char *path, uint32_t tensors, const uint32_t *ndim,
                   const uint32_t (*dims)[LBM_LTWS_MAX_DIMS], const float *data);

#ifdef __cplusplus
}
#endif

#endif
