/* A caller of the reference API, compiled unchanged against the drop-in include/zap_tex.h and linked with
 * -lzaptex: it only uses zap_tex.h's own names.  Test infrastructure for tests/test_gpu_dropin.py.
 *
 *   dropin_caller chain fmt srgb mips w h in.rgba out.bin
 *       zap.c:331-343's loop: encode a level with zap_bc_encode, make the next with zap_mip_next, repeat; the
 *       blocks of every level go to out.bin back to back (the payload of the DDS file `zap tex` writes)
 *   dropin_caller small w h in.rgba in.f32 prefix
 *       zap_astc_encode, zap_bc6h_encode (float RGBA input), zap_bc_encode + zap_bc_decode (BC7) and zap_bc_split /
 *       zap_bc_merge (BC7) of one image, each written to prefix.<name>
 *   dropin_caller onerror
 *       a caller-installed zap_tex_on_error receives a failed call (a 0-pixel-wide image) instead of the abort
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zap_tex.h"

static unsigned char *load(const char *path, size_t need) {
    FILE *f = fopen(path, "rb");
    unsigned char *p = (unsigned char *)malloc(need ? need : 1);
    if (!f || !p || fread(p, 1, need, f) != need) { perror(path); exit(1); }
    fclose(f);
    return p;
}

static void save(const char *path, const void *p, size_t n) {
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(p, 1, n, f) != n || fclose(f)) { perror(path); exit(1); }
}

static int chain(int f, int srgb, int mips, int w, int h, const char *in, const char *out) {
    int levels = mips ? zap_mip_levels(w, h) : 1;
    size_t total = 0;
    for (int l = 0, lw = w, lh = h; l < levels; l++, lw = lw > 1 ? lw / 2 : 1, lh = lh > 1 ? lh / 2 : 1)
        total += zap_bc_size(lw, lh, (zap_bc_format)f);
    unsigned char *img = load(in, (size_t)w * h * 4), *bc = (unsigned char *)malloc(total), *p = bc;
    unsigned char *next = (unsigned char *)malloc((size_t)(w > 1 ? w / 2 : 1) * (size_t)(h > 1 ? h / 2 : 1) * 4 + 4);
    for (int l = 0, lw = w, lh = h; l < levels; l++) {
        zap_bc_encode(img, lw, lh, (size_t)lw * 4, (zap_bc_format)f, 0.0f, p);
        p += zap_bc_size(lw, lh, (zap_bc_format)f);
        if (l + 1 < levels) {
            zap_mip_next(img, lw, lh, (size_t)lw * 4, srgb, next, (size_t)(lw > 1 ? lw / 2 : 1) * 4);
            lw = lw > 1 ? lw / 2 : 1;
            lh = lh > 1 ? lh / 2 : 1;
            memcpy(img, next, (size_t)lw * lh * 4);
        }
    }
    save(out, bc, total);
    printf("%d level%s, %zu bytes\n", levels, levels > 1 ? "s" : "", total);
    free(img); free(bc); free(next);
    return 0;
}

static int small(int w, int h, const char *in, const char *inf, const char *prefix) {
    char path[4096];
    size_t n7 = zap_bc_size(w, h, ZAP_BC7);
    unsigned char *img = load(in, (size_t)w * h * 4), *bc = (unsigned char *)malloc(n7), *planes = (unsigned char *)malloc(n7),
                  *back = (unsigned char *)malloc(n7);
    unsigned char *dec = (unsigned char *)malloc((size_t)w * h * 4);
    float *hdr = (float *)load(inf, (size_t)w * h * 16);
    zap_astc_encode(img, w, h, (size_t)w * 4, bc);
    snprintf(path, sizeof path, "%s.astc", prefix); save(path, bc, n7);
    zap_bc6h_encode(hdr, w, h, (size_t)w * 4, bc);
    snprintf(path, sizeof path, "%s.bc6h", prefix); save(path, bc, n7);
    zap_bc_encode(img, w, h, (size_t)w * 4, ZAP_BC7, 0.0f, bc);
    snprintf(path, sizeof path, "%s.bc7", prefix); save(path, bc, n7);
    zap_bc_decode(bc, w, h, ZAP_BC7, dec, (size_t)w * 4);
    snprintf(path, sizeof path, "%s.dec", prefix); save(path, dec, (size_t)w * h * 4);
    zap_bc_split(bc, n7, ZAP_BC7, planes);
    snprintf(path, sizeof path, "%s.split", prefix); save(path, planes, n7);
    zap_bc_merge(planes, n7, ZAP_BC7, back);
    printf("merge %s\n", memcmp(back, bc, n7) ? "differs" : "restores the blocks");
    free(img); free(bc); free(planes); free(back); free(dec); free(hdr);
    return 0;
}

static int handled = 0;
static void on_error(const char *what, int status, const char *message) {
    printf("handled %s status %d: %s\n", what, status, message);
    handled = status;
}

int main(int argc, char **argv) {
    if (argc == 9 && !strcmp(argv[1], "chain"))
        return chain(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]), argv[7], argv[8]);
    if (argc == 7 && !strcmp(argv[1], "small")) return small(atoi(argv[2]), atoi(argv[3]), argv[4], argv[5], argv[6]);
    if (argc == 2 && !strcmp(argv[1], "onerror")) {
        unsigned char px[64] = { 0 }, out[16];
        zap_tex_on_error = on_error;
        zap_bc_encode(px, 0, 4, 16, ZAP_BC7, 0.0f, out);
        return handled == ZAPTEX_E_ARG ? 0 : 1;
    }
    fprintf(stderr, "usage: dropin_caller chain fmt srgb mips w h in out | small w h in.rgba in.f32 prefix | onerror\n");
    return 2;
}
