// zt_decode.hip — BCn decoders (zap_tex.h:1508-1547, BC7 all 8 modes :378-417) and the block-field plane
// split / merge (zap_tex.h:1549-1571) for gfx950, one lane per block.
#include "zt_common.h"
#include "zt_kernels.h"
#include "zt_bc7_tables.h"

namespace {

// per mode: subsets, partition bits, rotation bits, index-selector bits, colour bits, alpha bits,
// per-endpoint p-bit, shared p-bit, index bits, secondary index bits
__constant__ uint8_t kMode[8][10] = {
    { 3, 4, 0, 0, 4, 0, 1, 0, 3, 0 }, { 2, 6, 0, 0, 6, 0, 0, 1, 3, 0 }, { 3, 6, 0, 0, 5, 0, 0, 0, 2, 0 },
    { 2, 6, 0, 0, 7, 0, 1, 0, 2, 0 }, { 1, 0, 2, 1, 5, 6, 0, 0, 2, 3 }, { 1, 0, 2, 0, 7, 8, 0, 0, 2, 2 },
    { 1, 0, 0, 0, 7, 7, 1, 0, 4, 0 }, { 2, 6, 0, 0, 5, 5, 1, 0, 2, 0 } };

ZT_DEV int wgt(int bits, int j) { int D = (1 << bits) - 1; return (j * 64 + D / 2) / D; }
ZT_DEV int expand(int v, int n) { v <<= 8 - n; return v | v >> n; }

ZT_DEV void bc7_decode(const uint8_t *b, uint32_t out[16]) {
    BitReader R = reader_at(b);
    int mode = 0;
    while (mode < 8 && !((R.lo >> mode) & 1)) mode++;
    if (mode == 8) { for (int i = 0; i < 16; i++) out[i] = 0; return; }
    const int ns = kMode[mode][0], pbits = kMode[mode][1], rb = kMode[mode][2], isbb = kMode[mode][3];
    const int cbits = kMode[mode][4], abits = kMode[mode][5], epb = kMode[mode][6], spb = kMode[mode][7];
    const int ib = kMode[mode][8], ib2 = kMode[mode][9];
    R.pos = mode + 1;
    int part = R.take(pbits), rot = R.take(rb), isb = R.take(isbb), ne = 2 * ns;
    int ep[6][4];
    for (int c = 0; c < 3; c++) for (int e = 0; e < ne; e++) ep[e][c] = R.take(cbits);
    for (int e = 0; e < ne; e++) ep[e][3] = abits ? R.take(abits) : 255;
    int cb = cbits, ab = abits;
    if (epb || spb) {
        for (int e = 0; e < ne; e++) {
            if (epb || !(e & 1)) {
                int p = R.take(1);
                for (int k = e; k <= (epb ? e : e + 1); k++) {
                    for (int c = 0; c < 3; c++) ep[k][c] = ep[k][c] << 1 | p;
                    if (ab) ep[k][3] = ep[k][3] << 1 | p;
                }
            }
        }
        cb++; if (ab) ab++;
    }
    for (int e = 0; e < ne; e++) {
        for (int c = 0; c < 3; c++) ep[e][c] = expand(ep[e][c], cb);
        if (ab) ep[e][3] = expand(ep[e][3], ab);
    }
    auto subset = [&](int i) { return ns == 1 ? 0 : ns == 2 ? (int)((kP2all[part] >> i) & 1) : (int)((kP3[part] >> (2 * i)) & 3); };
    auto anchor = [&](int s) { return s == 0 ? 0 : ns == 2 ? (int)kA2all[part] : s == 1 ? (int)kA3a[part] : (int)kA3b[part]; };
    int i1[16], i2[16];
    for (int i = 0; i < 16; i++) i1[i] = R.take(ib - (i == anchor(subset(i))));
    for (int i = 0; i < 16; i++) i2[i] = ib2 ? R.take(ib2 - (i == 0)) : 0;
    for (int i = 0; i < 16; i++) {
        int sb = subset(i);
        int cw = ib2 && isb ? wgt(ib2, i2[i]) : wgt(ib, i1[i]);
        int aw = !ib2 ? cw : isb ? wgt(ib, i1[i]) : wgt(ib2, i2[i]);
        int px[4];
        for (int c = 0; c < 3; c++) px[c] = ((64 - cw) * ep[2 * sb][c] + cw * ep[2 * sb + 1][c] + 32) >> 6;
        px[3] = ((64 - aw) * ep[2 * sb][3] + aw * ep[2 * sb + 1][3] + 32) >> 6;
        if (rot) { int t = px[3]; px[3] = px[rot - 1]; px[rot - 1] = t; }
        out[i] = (uint32_t)px[0] | (uint32_t)px[1] << 8 | (uint32_t)px[2] << 16 | (uint32_t)px[3] << 24;
    }
}

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
