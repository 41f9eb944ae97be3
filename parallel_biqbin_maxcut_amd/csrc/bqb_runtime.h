/* bqb_runtime.h -- internal helpers of the host runtime (C). */
#ifndef BQB_RUNTIME_H
#define BQB_RUNTIME_H
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bqb_buf { void *ptr; size_t cap; } bqb_buf;

void bqb_set_error(const char *fmt, ...);
int bqb_check(hipError_t e, const char *what);
int bqb_ensure_init(void);
int bqb_buf_reserve(bqb_buf *b, size_t bytes);
void bqb_buf_free(bqb_buf *b);
int bqb_hbuf_reserve(bqb_buf *b, size_t bytes);
void bqb_hbuf_free(bqb_buf *b);
int bqb_ipm_ld_for(int nmax);
/* device address of the no-allocation scratch array on the current device */
int bqb_noalloc_setup(void);

#include "../../include/biqbin_gpu.h"

#ifdef __cplusplus
}
#endif
#endif
