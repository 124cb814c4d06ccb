// zt_astc_setup.h — ASTC 4x4 LDR tables shared by the host table builder and the gfx950 kernels
// (zap_tex.h:977-1305).  Everything the encoder and decoder look up is built once per device from the ASTC
// specification's own data tables (integer-sequence ranges, colour-unquantisation patterns, trit/quint codings,
// block-mode layout) and kept in device memory: the kernels decode by table lookup, never by per-value bit logic.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ZT_HD __host__ __device__ inline
#else
#define ZT_HD inline
#endif

// ---- integer-sequence-encoding ranges, specification order (range index 0..20)
struct zt_ise_range { int levels, trit, quint, bits; };
// (levels, trit?, quint?, plain bits) per range index; every range is 2^bits, 3 * 2^bits or 5 * 2^bits levels
#define ZT_ISE_TABLE                                                                                                         \
    { { 2, 0, 0, 1 }, { 3, 1, 0, 0 }, { 4, 0, 0, 2 }, { 5, 0, 1, 0 }, { 6, 1, 0, 1 }, { 8, 0, 0, 3 }, { 10, 0, 1, 1 },     \
      { 12, 1, 0, 2 }, { 16, 0, 0, 4 }, { 20, 0, 1, 2 }, { 24, 1, 0, 3 }, { 32, 0, 0, 5 }, { 40, 0, 1, 3 }, { 48, 1, 0, 4 }, \
      { 64, 0, 0, 6 }, { 80, 0, 1, 4 }, { 96, 1, 0, 5 }, { 128, 0, 0, 7 }, { 160, 0, 1, 5 }, { 192, 1, 0, 6 }, { 256, 0, 0, 8 } }
ZT_HD zt_ise_range zt_ise(int i) {
    const zt_ise_range t[21] = ZT_ISE_TABLE;
    return t[i < 0 ? 0 : i > 20 ? 20 : i];
}
ZT_HD int zt_ceil_div(int a, int b) { return (a + b - 1) / b; }
// bits of an n-value sequence: n plain fields, plus ceil(8n/5) trit or ceil(7n/3) quint transfer bits
ZT_HD int zt_ise_bits(zt_ise_range r, int n) {
    return n * r.bits + (r.trit ? zt_ceil_div(8 * n, 5) : 0) + (r.quint ? zt_ceil_div(7 * n, 3) : 0);
}
// colour endpoints take the finest range (of at least 6 levels, range index >= 4) whose sequence fits the bits left
// (zap_tex.h:996-999); -1 when none fits
ZT_HD int zt_colour_range(int nvals, int avail) {
    int best = -1;
    for (int i = 4; i <= 20; i++)
        if (zt_ise_bits(zt_ise(i), nvals) <= avail) best = i;
    return best;
}
// an n-bit value repeated MSB-first to fill `to` bits (the ASTC bit-replication rule)
ZT_HD int zt_widen(int v, int n, int to) {
    int x = 0, have = 0;
    while (have < to) { x = x << n | v; have += n; }
    return x >> (have - to);
}
// a weight of n plain bits (n = 1..5) on the 0..64 scale: widened to 6 bits, values above 32 moved up by one
ZT_HD int zt_weight_value(int w, int n) {
    const int x = zt_widen(w, n, 6);
    return x + (x > 32);
}

// ---- the specification's tables in spec notation, turned into lookup tables by zt_astc_setup_build
namespace zt_spec {
// colour unquantisation of trit/quint ranges: T = D * C + B, then T ^ A, result (A & 0x80) | (T >> 2), where D is the
// trit/quint digit, A replicates the value's lowest plain bit to 9 bits and B is the 9-bit pattern below, MSB first,
// over the plain bits a (bit 0, in A) and b..f (bits 1..5)
struct UnquantRow { int levels, C; const char *B; };
const UnquantRow kUnquant[11] = { { 6, 204, "000000000" }, { 10, 113, "000000000" }, { 12, 93, "b000b0bb0" },
                                  { 20, 54, "b0000bb00" }, { 24, 44, "cb000cbcb" }, { 40, 26, "cb0000cbc" },
                                  { 48, 22, "dcb000dcb" }, { 80, 13, "dcb0000dc" }, { 96, 11, "edcb000ed" },
                                  { 160, 6, "edcb0000e" }, { 192, 5, "fedcb000f" } };
inline int field(int x, int hi, int lo) { return x >> lo & ((1 << (hi - lo + 1)) - 1); }
inline int colour_value(zt_ise_range r, int v) {
    if (!r.trit && !r.quint) return zt_widen(v, r.bits, 8);
    const int plain = v & ((1 << r.bits) - 1), D = v >> r.bits;
    const UnquantRow *row = kUnquant;
    while (row->levels != r.levels) row++;
    int B = 0;
    for (int k = 0; k < 9; k++) {
        const char ch = row->B[k];
        B = B << 1 | (ch == '0' ? 0 : plain >> (ch - 'a') & 1);
    }
    const int A = (plain & 1) ? 0x1FF : 0, T = (D * row->C + B) ^ A;
    return (A & 0x80) | (T >> 2);
}
// the five trits of an 8-bit trit block T[7:0] (decoding procedure of the specification's trit table)
inline void trit_digits(int T, int t[5]) {
    int C;
    if (field(T, 4, 2) == 7) {
        C = field(T, 7, 5) << 2 | field(T, 1, 0);
        t[4] = 2; t[3] = 2;
    } else {
        C = field(T, 4, 0);
        if (field(T, 6, 5) == 3) { t[4] = 2; t[3] = field(T, 7, 7); }
        else { t[4] = field(T, 7, 7); t[3] = field(T, 6, 5); }
    }
    if (field(C, 1, 0) == 3) {
        t[2] = 2; t[1] = field(C, 4, 4);
        t[0] = field(C, 3, 3) << 1 | (field(C, 2, 2) & (field(C, 3, 3) ^ 1));
    } else if (field(C, 3, 2) == 3) {
        t[2] = 2; t[1] = 2; t[0] = field(C, 1, 0);
    } else {
        t[2] = field(C, 4, 4); t[1] = field(C, 3, 2);
        t[0] = field(C, 1, 1) << 1 | (field(C, 0, 0) & (field(C, 1, 1) ^ 1));
    }
}
// the three quints of a 7-bit quint block Q[6:0] (decoding procedure of the specification's quint table)
inline void quint_digits(int Q, int q[3]) {
    if (field(Q, 2, 1) == 3 && field(Q, 6, 5) == 0) {
        q[2] = field(Q, 0, 0) << 2 | (field(Q, 4, 4) & (field(Q, 0, 0) ^ 1)) << 1 | (field(Q, 3, 3) & (field(Q, 0, 0) ^ 1));
        q[1] = 4; q[0] = 4;
        return;
    }
    int C;
    if (field(Q, 2, 1) == 3) { q[2] = 4; C = field(Q, 4, 3) << 3 | (field(Q, 6, 5) ^ 3) << 1 | field(Q, 0, 0); }
    else { q[2] = field(Q, 6, 5); C = field(Q, 4, 0); }
    if (field(C, 2, 0) == 5) { q[1] = 4; q[0] = field(C, 4, 3); }
    else { q[1] = field(C, 4, 3); q[0] = field(C, 2, 0); }
}
// weight bits per texel of an 11-bit block mode, for the modes this codec reads: a 4x4 weight grid, one plane,
// plain-bit weight ranges (2, 4, 8, 16 or 32 levels); 0 for every other mode.  Spec layout of modes with
// bits[1:0] != 0: R = {bit 1, bit 0, bit 4}, H = bit 9, dual plane = bit 10, and for bits[3:2] == 0 a grid of
// (bits[8:7] + 4) x (bits[6:5] + 2); (H, R) pick the weight range.
inline int block_mode_weight_bits(int mode) {
    if (field(mode, 1, 0) == 0 || field(mode, 3, 2) != 0) return 0;
    if (field(mode, 8, 7) + 4 != 4 || field(mode, 6, 5) + 2 != 4 || field(mode, 10, 10)) return 0;
    const int R = field(mode, 1, 1) << 2 | field(mode, 0, 0) << 1 | field(mode, 4, 4), H = field(mode, 9, 9);
    const int low[8] = { 0, 0, 2, 3, 4, 5, 6, 8 }, high[8] = { 0, 0, 10, 12, 16, 20, 24, 32 };
    const int levels = H ? high[R] : low[R];
    for (int b = 1; b <= 5; b++) if (levels == 1 << b) return b;
    return 0;
}
}  // namespace zt_spec

ZT_HD uint32_t zt_part_hash(uint32_t p) {  // spec C.2.21 (zap_tex.h:1098-1101)
    p ^= p >> 15; p -= p << 17; p += p << 7; p += p << 4; p ^= p >> 5; p += p << 16; p ^= p >> 7; p ^= p >> 3; p ^= p << 6; p ^= p >> 17;
    return p;
}
ZT_HD int zt_astc_partition(int seed, int x, int y, int count) {  // small-block partition (zap_tex.h:1102-1120)
    x <<= 1; y <<= 1;
    seed += (count - 1) * 1024;
    uint32_t rn = zt_part_hash((uint32_t)seed);
    int s[12] = { (int)(rn & 15), (int)(rn >> 4 & 15), (int)(rn >> 8 & 15), (int)(rn >> 12 & 15), (int)(rn >> 16 & 15), (int)(rn >> 20 & 15),
                  (int)(rn >> 24 & 15), (int)(rn >> 28 & 15), (int)(rn >> 18 & 15), (int)(rn >> 22 & 15), (int)(rn >> 26 & 15), (int)((rn >> 30 | rn << 2) & 15) };
    for (int i = 0; i < 12; i++) s[i] *= s[i];
    int sh1, sh2;
    if (seed & 1) { sh1 = seed & 2 ? 4 : 5; sh2 = count == 3 ? 6 : 5; }
    else { sh1 = count == 3 ? 6 : 5; sh2 = seed & 2 ? 4 : 5; }
    int sh3 = seed & 0x10 ? sh1 : sh2;
    for (int i = 0; i < 8; i++) s[i] >>= i & 1 ? sh2 : sh1;
    for (int i = 8; i < 12; i++) s[i] >>= sh3;
    int a = (s[0] * x + s[1] * y + (int)(rn >> 14)) & 0x3F, b = (s[2] * x + s[3] * y + (int)(rn >> 10)) & 0x3F;
    int c = (s[4] * x + s[5] * y + (int)(rn >> 6)) & 0x3F, d = (s[6] * x + s[7] * y + (int)(rn >> 2)) & 0x3F;
    if (count < 4) d = 0;
    if (count < 3) c = 0;
    return a >= b && a >= c && a >= d ? 0 : b >= c && b >= d ? 1 : c >= d ? 2 : 3;
}

// one trial configuration of the encoder (zap_tex.h:1170-1180, :1298-1300)
struct zt_astc_cfg { int range, levels, trit, quint, bits, wb, cem, nv, parts; };
struct zt_astc_setup {
    zt_astc_cfg cfg[8];
    uint8_t q[8][256];        // 8-bit value -> nearest ISE value of the cfg's colour range
    uint8_t unq[8][256];      // ISE value -> 0..255, per cfg
    uint8_t colour[21][256];  // ISE value -> 0..255, every range (the decoder's)
    uint8_t weight[6][32];    // plain-bit weight w of n bits (n = 1..5) -> 0..64
    uint16_t tdec[256];       // trit block -> its 5 trits, 2 bits each (trit 0 lowest)
    uint16_t qdec[128];       // quint block -> its 3 quints, 3 bits each
    uint8_t tenc[256];        // 5 trits (base-3 digits, t0 lowest) -> the smallest trit block decoding to them
    uint8_t qenc[128];        // 3 quints -> the smallest quint block
    uint8_t mode_wb[2048];    // block mode -> weight bits of the supported 4x4 modes, 0 otherwise
    uint16_t part2[1024];     // 2-partition pattern of every seed: bit i = texel i in partition 1
    uint16_t pmask[1024];     // usable 2-partition patterns (neither all 0 nor all 1), in seed order
    uint16_t pseed[1024];
    int np;
};

inline void zt_astc_setup_build(zt_astc_setup *s) {
    for (int T = 0; T < 256; T++) {
        int t[5];
        zt_spec::trit_digits(T, t);
        s->tdec[T] = (uint16_t)(t[0] | t[1] << 2 | t[2] << 4 | t[3] << 6 | t[4] << 8);
    }
    for (int Q = 0; Q < 128; Q++) {
        int q[3];
        zt_spec::quint_digits(Q, q);
        s->qdec[Q] = (uint16_t)(q[0] | q[1] << 3 | q[2] << 6);
    }
    // encode tables: the inverse of the decode tables, the smallest block where several decode alike
    bool tset[243] = {}, qset[125] = {};
    for (int T = 0; T < 256; T++) {
        int id = 0;
        for (int k = 4; k >= 0; k--) id = id * 3 + (s->tdec[T] >> (2 * k) & 3);
        if (!tset[id]) { tset[id] = true; s->tenc[id] = (uint8_t)T; }
    }
    for (int Q = 0; Q < 128; Q++) {
        int id = 0;
        for (int k = 2; k >= 0; k--) id = id * 5 + (s->qdec[Q] >> (3 * k) & 7);
        if (!qset[id]) { qset[id] = true; s->qenc[id] = (uint8_t)Q; }
    }
    for (int r = 0; r < 21; r++)
        for (int v = 0; v < 256; v++) s->colour[r][v] = (uint8_t)(v < zt_ise(r).levels ? zt_spec::colour_value(zt_ise(r), v) : 0);
    for (int n = 0; n < 6; n++)
        for (int w = 0; w < 32; w++) s->weight[n][w] = (uint8_t)(n && w < (1 << n) ? zt_weight_value(w, n) : 0);
    for (int m = 0; m < 2048; m++) s->mode_wb[m] = (uint8_t)zt_spec::block_mode_weight_bits(m);
    // the encoder's trial configurations: (colour endpoint mode, weight bits, partitions)
    const int spec[8][3] = { { 8, 4, 1 }, { 8, 3, 1 }, { 8, 5, 1 }, { 12, 4, 1 }, { 12, 3, 1 }, { 12, 2, 1 }, { 8, 3, 2 }, { 8, 2, 2 } };
    for (int k = 0; k < 8; k++) {
        zt_astc_cfg &c = s->cfg[k];
        c.cem = spec[k][0]; c.wb = spec[k][1]; c.parts = spec[k][2]; c.nv = c.cem == 8 ? 6 : 8;
        c.range = zt_colour_range(c.nv * c.parts, 128 - (c.parts == 1 ? 17 : 29) - 16 * c.wb);
        const zt_ise_range r = zt_ise(c.range);
        c.levels = r.levels; c.trit = r.trit; c.quint = r.quint; c.bits = r.bits;
        for (int v = 0; v < 256; v++) s->unq[k][v] = s->colour[c.range][v];
        for (int x = 0; x < 256; x++) {   // nearest level, the lowest ISE value on ties
            int best = 0, bd = 1 << 30;
            for (int v = 0; v < r.levels; v++) {
                const int d = s->colour[c.range][v] > x ? s->colour[c.range][v] - x : x - s->colour[c.range][v];
                if (d < bd) { bd = d; best = v; }
            }
            s->q[k][x] = (uint8_t)best;
        }
    }
    s->np = 0;
    for (int sd = 0; sd < 1024; sd++) {
        unsigned m = 0;
        for (int i = 0; i < 16; i++) m |= (unsigned)zt_astc_partition(sd, i & 3, i >> 2, 2) << i;
        s->part2[sd] = (uint16_t)m;
        if (m && m != 0xFFFFu) { s->pmask[s->np] = (uint16_t)m; s->pseed[s->np++] = (uint16_t)sd; }
    }
}
