// zt_decode.hip — BCn decoders (zap_tex.h:1508-1547, BC7 all 8 modes :378-417) and the block-field plane
// split / merge (zap_tex.h:1549-1571) for gfx950, one lane per block.
#include "zt_common.h"
#include "zt_kernels.h"
#include "zt_bc7_tables.h"

namespace {

// BC7 modes as the specification lists them, with every field's bit position worked out at compile time: the
// decoder reads each field where it lies instead of walking the block.  Fields in block order: mode (mode + 1 bits,
// unary), partition, rotation, index selector, colour endpoints (R of every endpoint, then G, then B), alpha
// endpoints, p-bits (one per endpoint, or one per subset shared by its two endpoints), primary indices (one bit
// fewer at each subset's anchor texel), secondary indices (one bit fewer at texel 0).
struct Bc7Layout {
    uint8_t subsets, pb, rb, isb, cb, ab, pbits_per_subset, ib, ib2;   // pbits_per_subset: 0, 1 (shared) or 2
    uint8_t at_part, at_rot, at_isb, at_colour, at_alpha, at_pbit, at_idx, at_idx2;
};
constexpr Bc7Layout bc7_layout(int mode, int subsets, int pb, int rb, int isb, int cb, int ab, int pps, int ib, int ib2) {
    Bc7Layout L = { (uint8_t)subsets, (uint8_t)pb, (uint8_t)rb, (uint8_t)isb, (uint8_t)cb, (uint8_t)ab, (uint8_t)pps, (uint8_t)ib,
                    (uint8_t)ib2, 0, 0, 0, 0, 0, 0, 0, 0 };
    int at = mode + 1;
    L.at_part = (uint8_t)at; at += pb;
    L.at_rot = (uint8_t)at; at += rb;
    L.at_isb = (uint8_t)at; at += isb;
    L.at_colour = (uint8_t)at; at += 3 * 2 * subsets * cb;
    L.at_alpha = (uint8_t)at; at += 2 * subsets * ab;
    L.at_pbit = (uint8_t)at; at += subsets * pps;
    L.at_idx = (uint8_t)at; at += 16 * ib - subsets;
    L.at_idx2 = (uint8_t)at;
    return L;
}
__constant__ Bc7Layout kBc7[8] = {
    bc7_layout(0, 3, 4, 0, 0, 4, 0, 2, 3, 0), bc7_layout(1, 2, 6, 0, 0, 6, 0, 1, 3, 0), bc7_layout(2, 3, 6, 0, 0, 5, 0, 0, 2, 0),
    bc7_layout(3, 2, 6, 0, 0, 7, 0, 2, 2, 0), bc7_layout(4, 1, 0, 2, 1, 5, 6, 0, 2, 3), bc7_layout(5, 1, 0, 2, 0, 7, 8, 0, 2, 2),
    bc7_layout(6, 1, 0, 0, 0, 7, 7, 2, 4, 0), bc7_layout(7, 2, 6, 0, 0, 5, 5, 2, 2, 0) };
static_assert(bc7_layout(4, 1, 0, 2, 1, 5, 6, 0, 2, 3).at_idx2 + 16 * 3 - 1 == 128, "mode 4 fills 128 bits");
static_assert(bc7_layout(0, 3, 4, 0, 0, 4, 0, 2, 3, 0).at_idx2 == 128 && bc7_layout(7, 2, 6, 0, 0, 5, 5, 2, 2, 0).at_idx2 == 128,
              "modes 0 and 7 fill 128 bits");

ZT_DEV int wgt(int bits, int j) { int D = (1 << bits) - 1; return (j * 64 + D / 2) / D; }
ZT_DEV int expand(int v, int n) { v <<= 8 - n; return v | v >> n; }

ZT_DEV void bc7_decode(const uint8_t *b, uint32_t out[16]) {
    const uint4 raw = *(const uint4 *)b;
    const uint64_t lo = (uint64_t)raw.x | (uint64_t)raw.y << 32, hi = (uint64_t)raw.z | (uint64_t)raw.w << 32;
    if (!(lo & 0xFFu)) { for (int i = 0; i < 16; i++) out[i] = 0; return; }   // reserved mode
    const int mode = __builtin_ctzll(lo);
    const Bc7Layout L = kBc7[mode];
    auto field = [&](int at, int n) { return n ? (int)get_bits(lo, hi, at, n) : 0; };
    const int part = field(L.at_part, L.pb), rot = field(L.at_rot, L.rb), isb = field(L.at_isb, L.isb);
    const int ne = 2 * L.subsets, pb = L.pbits_per_subset ? 1 : 0;
    // endpoint e, channel c as 8 bits: the field, its p-bit below it, replicated upwards
    int ep[6][4];
    for (int e = 0; e < ne; e++) {
        const int p = L.pbits_per_subset == 2 ? field(L.at_pbit + e, 1) : L.pbits_per_subset == 1 ? field(L.at_pbit + (e >> 1), 1) : 0;
        for (int c = 0; c < 3; c++) ep[e][c] = expand(field(L.at_colour + (c * ne + e) * L.cb, L.cb) << pb | p, L.cb + pb);
        ep[e][3] = L.ab ? expand(field(L.at_alpha + e * L.ab, L.ab) << pb | p, L.ab + pb) : 255;
    }
    // subset of each texel and the anchors, whose indices are one bit short
    uint32_t sub = 0, anchors = 1;
    if (L.subsets == 2) {
        sub = 0;
        for (int i = 0; i < 16; i++) sub |= (uint32_t)((kP2all[part] >> i) & 1) << (2 * i);
        anchors |= 1u << kA2all[part];
    } else if (L.subsets == 3) {
        sub = kP3[part];
        anchors |= 1u << kA3a[part] | 1u << kA3b[part];
    }
    for (int i = 0; i < 16; i++) {
        const uint32_t below = anchors & ((1u << i) - 1u);
        const int short1 = (int)(anchors >> i & 1u), i1 = field(L.at_idx + i * L.ib - __popc(below), L.ib - short1);
        const int i2 = L.ib2 ? field(L.at_idx2 + i * L.ib2 - (i > 0), L.ib2 - (i == 0)) : 0;
        // colour and alpha weights: both from the primary indices, or (modes 4 and 5) one from each set, the
        // index selector choosing which set the colour takes
        int cw = wgt(L.ib, i1), aw = cw;
        if (L.ib2) { if (isb) cw = wgt(L.ib2, i2); else aw = wgt(L.ib2, i2); }
        const int s = (int)(sub >> (2 * i) & 3u);
        int px[4];
        for (int c = 0; c < 4; c++) {
            const int w = c == 3 ? aw : cw;
            px[c] = ((64 - w) * ep[2 * s][c] + w * ep[2 * s + 1][c] + 32) >> 6;
        }
        if (rot) { const int t = px[3]; px[3] = px[rot - 1]; px[rot - 1] = t; }
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
