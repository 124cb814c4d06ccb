/* zaptex-tex — the `zap tex` command (zap.c:309-355, argument syntax of zap.c:355-393) on libzaptex:
 *
 *   zaptex-tex [-f bc1|bc3|bc4|bc5|bc7|astc] [-r rdo] [-m] [-S] w h in.rgba out.dds|.astc
 *
 * reads w*h RGBA8 pixels, encodes them (with -m the whole box mip chain, -S sRGB filtering and format) on the GPU
 * and writes the DDS or .astc file byte for byte as the reference CLI does. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zaptex.h"

static int usage(void) {
    fprintf(stderr, "zaptex-tex [-f bc1|bc3|bc4|bc5|bc7|astc] [-r rdo] [-m] [-S] w h in.rgba out.dds|.astc\n");
    return 2;
}

static unsigned char *load(const char *path, size_t *n) {
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); exit(1); }
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    unsigned char *p = malloc(len > 0 ? (size_t)len : 1);
    if (!p || (len > 0 && fread(p, 1, (size_t)len, f) != (size_t)len)) { perror(path); exit(1); }
    fclose(f);
    *n = (size_t)len;
    return p;
}

int main(int argc, char **argv) {
    static const char *names[] = { "bc1", "bc3", "bc4", "bc5", "bc7", "bc6h", "astc" };
    const char *fmt = "bc1";
    float rdo = 0;
    int mips = 0, srgb = 0, i = 1;
    for (; i + 1 < argc && argv[i][0] == '-'; i += 2) {
        if (argv[i][1] == 'm' || argv[i][1] == 'S') { /* flags without a value */
            if (argv[i][1] == 'm') mips = 1; else srgb = 1;
            i--;
            continue;
        }
        if (argv[i][1] == 'f') fmt = argv[i + 1];
        else if (argv[i][1] == 'r') rdo = (float)atof(argv[i + 1]);
        else return usage();
    }
    if (i + 4 != argc) return usage();
    int w = atoi(argv[i]), h = atoi(argv[i + 1]), f = 0;
    while (f < 7 && (f == ZAPTEX_BC6H || strcmp(fmt, names[f]))) f++;
    if (f == 7 || w < 1 || h < 1) return usage();
    size_t n;
    unsigned char *img = load(argv[i + 2], &n);
    if (n < (size_t)w * h * 4) { fprintf(stderr, "%s: expected %dx%d RGBA8\n", argv[i + 2], w, h); return 1; }
    if (f == ZAPTEX_ASTC && mips) fprintf(stderr, "note: .astc holds one level; -m ignored\n");
    size_t bytes = zaptex_tex_file_size(w, h, f, mips, srgb);
    unsigned char *out = malloc(bytes ? bytes : 1);
    if (!bytes || !out) { fprintf(stderr, "bad size %dx%d\n", w, h); return 1; }
    int rc = zaptex_tex_file(img, w, h, (size_t)w * 4, f, rdo, mips, srgb, out, bytes);
    if (rc != ZAPTEX_OK) { fprintf(stderr, "zaptex_tex_file failed (status %d): %s\n", rc, zaptex_last_error()); return 1; }
    FILE *fo = fopen(argv[i + 3], "wb");
    if (!fo || fwrite(out, 1, bytes, fo) != bytes || fclose(fo)) { perror(argv[i + 3]); return 1; }
    if (f == ZAPTEX_ASTC) {
        printf("%dx%d astc 4x4 -> %zu bytes\n", w, h, bytes);
    } else {
        int levels = 1, lw = w, lh = h;
        while (mips && (lw > 1 || lh > 1)) { lw = lw > 1 ? lw / 2 : 1; lh = lh > 1 ? lh / 2 : 1; levels++; }
        int s = srgb && f != ZAPTEX_BC4 && f != ZAPTEX_BC5;
        printf("%dx%d %s%s, %d level%s -> %zu bytes\n", w, h, names[f], s ? " sRGB" : "", levels, levels > 1 ? "s" : "", bytes);
    }
    free(img);
    free(out);
    return 0;
}
