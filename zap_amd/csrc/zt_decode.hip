// zt_decode.hip — BCn decoders (zap_tex.h:1508-1547, BC7 all 8 modes :378-417) and the block-field plane
// split / merge (zap_tex.h:1549-1571) for gfx950, one lane per block.
#include "zt_common.h"
#include "zt_kernels.h"
#include "zt_bc7_tables.h"
#include "zt_bc7_dec.h"

namespace {

ZT_DEV void bc1_decode(const uint8_t *b, int force4, uint32_t out[16]) {
    uint2 raw = *(const uint2 *)b;
    uint32_t c0 = raw.x & 0xFFFF, c1 = raw.x >> 16, idx = raw.y;
    int pal[4][3], e[2][3];
    uint32_t cc[2] = { c0, c1 };
    for (int k = 0; k < 2; k++) {
        int R = (cc[k] >> 11) & 31, G = (cc[k] >> 5) & 63, B = cc[k] & 31;
        e[k][0] = (R << 3) | (R >> 2); e[k][1] = (G << 2) | (G >> 4); e[k][2] = (B << 3) | (B >> 2);
    }
    bool four = force4 || c0 > c1;
    for (int k = 0; k < 3; k++) {
        int a = e[0][k], bb = e[1][k];
        pal[0][k] = a; pal[1][k] = bb;
        pal[2][k] = four ? (2 * a + bb + 1) / 3 : (a + bb + 1) / 2;
        pal[3][k] = four ? (a + 2 * bb + 1) / 3 : 0;
    }
    for (int i = 0; i < 16; i++) {
        int k = (idx >> (2 * i)) & 3;
        uint32_t alpha = (!four && k == 3) ? 0u : 255u;
        out[i] = (uint32_t)pal[k][0] | (uint32_t)pal[k][1] << 8 | (uint32_t)pal[k][2] << 16 | alpha << 24;
    }
}

ZT_DEV void bc4_decode(const uint8_t *b, int v[16]) {
    uint2 raw = *(const uint2 *)b;
    uint64_t all = (uint64_t)raw.x | (uint64_t)raw.y << 32;
    int a0 = (int)(all & 255), a1 = (int)((all >> 8) & 255), pal[8];
    uint64_t idx = all >> 16;
    pal[0] = a0; pal[1] = a1;
    if (a0 > a1) for (int i = 2; i < 8; i++) pal[i] = ((8 - i) * a0 + (i - 1) * a1 + 3) / 7;
    else { for (int i = 2; i < 6; i++) pal[i] = ((6 - i) * a0 + (i - 1) * a1 + 2) / 5; pal[6] = 0; pal[7] = 255; }
    for (int i = 0; i < 16; i++) v[i] = pal[(idx >> (3 * i)) & 7];
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256) zt_bc_decode(const uint8_t *bc, int w, int h, int fmt, uint8_t *rgba, uint64_t stride) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bw * bh) return;
    const int bx = g % bw, by = g / bw;
    const uint64_t bs = fmt == 0 || fmt == 2 ? 8 : 16;
    const uint8_t *b = bc + (uint64_t)g * bs;
    uint32_t px[16];
    if (fmt == 0) bc1_decode(b, 0, px);
    else if (fmt == 4) bc7_decode(b, px);
    else if (fmt == 1) {
        int a[16];
        bc1_decode(b + 8, 1, px);
        bc4_decode(b, a);
        for (int i = 0; i < 16; i++) px[i] = (px[i] & 0xFFFFFFu) | (uint32_t)a[i] << 24;
    } else {
        int r[16], gg[16];
        bc4_decode(b, r);
        if (fmt == 3) bc4_decode(b + 8, gg); else for (int i = 0; i < 16; i++) gg[i] = 0;
        for (int i = 0; i < 16; i++) px[i] = (uint32_t)r[i] | (uint32_t)gg[i] << 8 | 0xFF000000u;
    }
    for (int i = 0; i < 16; i++) {
        int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
        if (x < w && y < h) *(uint32_t *)(rgba + (uint64_t)y * stride + (uint64_t)x * 4) = px[i];
    }
}

// plane k holds field k of every block (zap_tex.h:1551-1568); the tail bytes are copied by the caller
extern "C" __global__ void __launch_bounds__(256) zt_bc_shuffle(const uint8_t *in, uint8_t *out, uint64_t nb, int fmt, int split) {
    const int F[5][5] = { { 2, 4, 4 }, { 4, 2, 6, 4, 4 }, { 2, 2, 6 }, { 4, 2, 6, 2, 6 }, { 3, 8, 1, 7 } };
    const uint64_t bs = fmt == 0 || fmt == 2 ? 8 : 16;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += step) {
        uint64_t plane = 0, off = 0;
        for (int k = 0; k < F[fmt][0]; k++) {
            const uint64_t sz = (uint64_t)F[fmt][k + 1];
            const uint8_t *s = split ? in + b * bs + off : in + plane + b * sz;
            uint8_t *d = split ? out + plane + b * sz : out + b * bs + off;
            for (uint64_t i = 0; i < sz; i++) d[i] = s[i];
            plane += nb * sz; off += sz;
        }
    }
}
