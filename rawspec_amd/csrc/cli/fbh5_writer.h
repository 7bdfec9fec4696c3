/*
 * fbh5_writer.h -- FBH5 output plugin of the rawspec program
 * (librawspec_fbh5.so, loaded with dlopen; see fbh5_writer.c).
 */
#ifndef RAWSPEC_FBH5_WRITER_H_
#define RAWSPEC_FBH5_WRITER_H_

#include <stddef.h>

#include "rawspec_fbutils.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fbh5w_ctx fbh5w_ctx;

/* Creates (truncates) path and writes the file and data set attributes for
 * header h; nds spectra per dump set the chunking.  NULL on failure. */
fbh5w_ctx *fbh5w_open(const fb_hdr_t *h, unsigned int nds, const char *path,
                      const char *version_rawspec, const char *version_lib,
                      const char *version_fft);
/* Appends bytes / (nifs * nchans * nbits/8) whole integrations from buf. */
int fbh5w_write(fbh5w_ctx *c, const void *buf, size_t bytes);
/* Sets the dimension labels, closes and frees; non-zero if a close failed. */
int fbh5w_close(fbh5w_ctx *c);

/* Function table the program resolves from the plugin. */
typedef fbh5w_ctx *(*fbh5w_open_fn)(const fb_hdr_t *, unsigned int, const char *, const char *,
                                    const char *, const char *);
typedef int (*fbh5w_write_fn)(fbh5w_ctx *, const void *, size_t);
typedef int (*fbh5w_close_fn)(fbh5w_ctx *);

#ifdef __cplusplus
}
#endif

#endif
