/* zap_tex.h — drop-in replacement for WatchDogStudios/zap's zap_tex.h (1.3.0) whose encoders, decoders and
 * mip downsampler run on an AMD Instinct MI355X (gfx950) through libzaptex (link with -lzaptex).
 *
 * Same types, names and signatures as the reference header (zap_tex.h:4-23, README.md:256-268), so callers
 * recompile against this include path with no source edits; outputs are byte-identical to the reference
 * encoder, rate-distortion mode (rdo > 0) included.  The functions stay `void` like the reference's.  They never fall
 * back to the CPU: if the GPU call fails (no device, bad arguments) the handler zap_tex_on_error points to is
 * called with the status and message; the default handler prints them and aborts.
 */
#ifndef ZAP_TEX_H
#define ZAP_TEX_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "zaptex.h"

typedef enum { ZAP_BC1, ZAP_BC3, ZAP_BC4, ZAP_BC5, ZAP_BC7 } zap_bc_format; /* zap_tex.h:32 */

static inline size_t zap_bc_block_bytes(zap_bc_format f) { return f == ZAP_BC1 || f == ZAP_BC4 ? 8 : 16; } /* :34 */
static inline size_t zap_bc_size(int w, int h, zap_bc_format f) {                                        /* :35-37 */
    return (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4) * zap_bc_block_bytes(f);
}

typedef void (*zap_tex_error_fn)(const char *what, int status, const char *message);
static inline void zap_tex__abort(const char *what, int status, const char *message) {
    fprintf(stderr, "zap_tex: %s failed (status %d): %s\n", what, status, message);
    abort();
}
/* one handler per translation unit; set it before the first call to handle errors without aborting */
static zap_tex_error_fn zap_tex_on_error = zap_tex__abort;
static inline void zap_tex__check(const char *what, int status) {
    if (status != ZAPTEX_OK) zap_tex_on_error(what, status, zaptex_last_error());
}

/* zap_tex.h:1474 */
static inline void zap_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, zap_bc_format f, float rdo, void *out) {
    zap_tex__check("zap_bc_encode", zaptex_bc_encode(rgba, w, h, stride, (int)f, rdo, out));
}
/* zap_tex.h:1526 */
static inline void zap_bc_decode(const void *bc, int w, int h, zap_bc_format f, uint8_t *rgba, size_t stride) {
    zap_tex__check("zap_bc_decode", zaptex_bc_decode(bc, w, h, (int)f, rgba, stride));
}
/* zap_tex.h:1570-1571 */
static inline void zap_bc_split(const void *bc, size_t n, zap_bc_format f, void *out) {
    zap_tex__check("zap_bc_split", zaptex_bc_split(bc, n, (int)f, out));
}
static inline void zap_bc_merge(const void *planes, size_t n, zap_bc_format f, void *out) {
    zap_tex__check("zap_bc_merge", zaptex_bc_merge(planes, n, (int)f, out));
}
/* zap_tex.h:946 / :963 */
static inline void zap_bc6h_encode(const float *rgba, int w, int h, size_t stride, void *out) {
    zap_tex__check("zap_bc6h_encode", zaptex_bc6h_encode(rgba, w, h, stride, out));
}
static inline void zap_bc6h_decode(const void *bc, int w, int h, uint16_t *rgba, size_t stride) {
    zap_tex__check("zap_bc6h_decode", zaptex_bc6h_decode(bc, w, h, rgba, stride));
}
/* zap_tex.h:1291 / :1344 */
static inline void zap_astc_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out) {
    zap_tex__check("zap_astc_encode", zaptex_astc_encode(rgba, w, h, stride, out));
}
static inline void zap_astc_decode(const void *bc, int w, int h, uint8_t *rgba, size_t stride) {
    zap_tex__check("zap_astc_decode", zaptex_astc_decode(bc, w, h, rgba, stride));
}
/* zap_tex.h:1435 */
static inline int zap_mip_levels(int w, int h) {
    int n = 1;
    while (w > 1 || h > 1) { w = w > 1 ? w / 2 : 1; h = h > 1 ? h / 2 : 1; n++; }
    return n;
}
/* zap_tex.h:1442 / :1460 */
static inline void zap_mip_next(const uint8_t *src, int w, int h, size_t stride, int srgb, uint8_t *dst, size_t dst_stride) {
    zap_tex__check("zap_mip_next", zaptex_mip_next(src, w, h, stride, srgb, dst, dst_stride));
}
static inline void zap_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t dst_stride) {
    zap_tex__check("zap_mip_next_f", zaptex_mip_next_f(src, w, h, stride, dst, dst_stride));
}

#endif
