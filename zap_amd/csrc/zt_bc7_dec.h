// zt_bc7_dec.h — the BC7 decoder (all 8 modes, zap_tex.h:378-417), shared by the decode kernels and the
// rate-distortion pass (which measures candidate blocks by decoding them, zap_tex.h:710-716).
#pragma once
#include "zt_common.h"
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

ZT_DEV int bc7_weight(int bits, int j) { int D = (1 << bits) - 1; return (j * 64 + D / 2) / D; }
ZT_DEV int bc7_expand(int v, int n) { v <<= 8 - n; return v | v >> n; }

ZT_DEV void bc7_decode_bits(uint64_t lo, uint64_t hi, uint32_t out[16]) {
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
        for (int c = 0; c < 3; c++) ep[e][c] = bc7_expand(field(L.at_colour + (c * ne + e) * L.cb, L.cb) << pb | p, L.cb + pb);
        ep[e][3] = L.ab ? bc7_expand(field(L.at_alpha + e * L.ab, L.ab) << pb | p, L.ab + pb) : 255;
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
        int cw = bc7_weight(L.ib, i1), aw = cw;
        if (L.ib2) { if (isb) cw = bc7_weight(L.ib2, i2); else aw = bc7_weight(L.ib2, i2); }
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

ZT_DEV void bc7_decode(const uint8_t *b, uint32_t out[16]) {
    const uint4 raw = *(const uint4 *)b;
    bc7_decode_bits((uint64_t)raw.x | (uint64_t)raw.y << 32, (uint64_t)raw.z | (uint64_t)raw.w << 32, out);
}

}  // namespace
