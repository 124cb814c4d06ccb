/* ref_shim.c — exports the reference zap_tex.h (static inline, header-only) as linkable symbols so the
 * tests can run the reference encoder itself.  TEST INFRASTRUCTURE ONLY: compiled by oracle/Makefile
 * from the header where it lies (-I pointing at the reference checkout); nothing is copied. */
#include "zap_tex.h"

void zref_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, int f, float rdo, void *out) { zap_bc_encode(rgba, w, h, stride, (zap_bc_format)f, rdo, out); }
void zref_bc_decode(const void *bc, int w, int h, int f, uint8_t *rgba, size_t stride) { zap_bc_decode(bc, w, h, (zap_bc_format)f, rgba, stride); }
void zref_bc6h_encode(const float *rgba, int w, int h, size_t stride, void *out) { zap_bc6h_encode(rgba, w, h, stride, out); }
void zref_bc6h_decode(const void *bc, int w, int h, uint16_t *rgba, size_t stride) { zap_bc6h_decode(bc, w, h, rgba, stride); }
void zref_astc_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out) { zap_astc_encode(rgba, w, h, stride, out); }
void zref_astc_decode(const void *bc, int w, int h, uint8_t *rgba, size_t stride) { zap_astc_decode(bc, w, h, rgba, stride); }
int zref_mip_levels(int w, int h) { return zap_mip_levels(w, h); }
void zref_mip_next(const uint8_t *src, int w, int h, size_t stride, int srgb, uint8_t *dst, size_t ds) { zap_mip_next(src, w, h, stride, srgb, dst, ds); }
void zref_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t ds) { zap_mip_next_f(src, w, h, stride, dst, ds); }
void zref_bc_split(const void *bc, size_t n, int f, void *out) { zap_bc_split(bc, n, (zap_bc_format)f, out); }
void zref_bc_merge(const void *bc, size_t n, int f, void *out) { zap_bc_merge(bc, n, (zap_bc_format)f, out); }
uint16_t zref_f2h(float f) { return zap__f2h(f); }
float zref_h2f(uint16_t h) { return zap__h2f(h); }
int zref_astc_part(int seed, int x, int y, int count) { return zap__astc_part(seed, x, y, count); }
const float *zref_srgb_lin(void) { return ZAP__SRGB_LIN; }
const float *zref_srgb_thr(void) { return ZAP__SRGB_THR; }
/* one BC7 block through the reference's own search, for per-block probes */
int zref_bc7_block(const uint8_t px[64], uint8_t out[16]) { return zap__bc7_encode((const uint8_t(*)[4])px, out); }
float zref_bc7_est1(const uint8_t px64[64], int part) {
    const uint8_t(*px)[4] = (const uint8_t(*)[4])px64;
    float mo[16][9], tot[9] = { 0 };
    for (int i = 0; i < 16; i++) {
        float r = px[i][0], g = px[i][1], b = px[i][2], v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
        for (int j = 0; j < 9; j++) { mo[i][j] = v[j]; tot[j] += v[j]; }
    }
    return zap__bc7_est1((const float(*)[9])mo, tot, part);
}
