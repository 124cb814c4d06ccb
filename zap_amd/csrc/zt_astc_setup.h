// zt_astc_setup.h — ASTC 4x4 LDR tables and spec functions shared by the host table builder and the
// gfx950 kernels (zap_tex.h:977-1305).  The encoder's per-call setup in the reference (ISE packing tables,
// 8 trial configurations with their 8-bit -> ISE quantisation maps, the 822 usable 2-partition patterns) is
// built once per device here and kept in device memory.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define ZT_HD __host__ __device__ inline
#else
#define ZT_HD inline
#endif

// integer-sequence-encoding ranges, ASTC specification order: levels, trits, quints, bits
struct zt_ise_range { int levels, trit, quint, bits; };
ZT_HD zt_ise_range zt_ise(int i) {
    switch (i) {
    case 0: return { 2, 0, 0, 1 };    case 1: return { 3, 1, 0, 0 };    case 2: return { 4, 0, 0, 2 };
    case 3: return { 5, 0, 1, 0 };    case 4: return { 6, 1, 0, 1 };    case 5: return { 8, 0, 0, 3 };
    case 6: return { 10, 0, 1, 1 };   case 7: return { 12, 1, 0, 2 };   case 8: return { 16, 0, 0, 4 };
    case 9: return { 20, 0, 1, 2 };   case 10: return { 24, 1, 0, 3 };  case 11: return { 32, 0, 0, 5 };
    case 12: return { 40, 0, 1, 3 };  case 13: return { 48, 1, 0, 4 };  case 14: return { 64, 0, 0, 6 };
    case 15: return { 80, 0, 1, 4 };  case 16: return { 96, 1, 0, 5 };  case 17: return { 128, 0, 0, 7 };
    case 18: return { 160, 0, 1, 5 }; case 19: return { 192, 1, 0, 6 }; default: return { 256, 0, 0, 8 };
    }
}
ZT_HD int zt_ise_bits(zt_ise_range r, int n) { return n * r.bits + (r.trit ? (8 * n + 4) / 5 : r.quint ? (7 * n + 2) / 3 : 0); }
// largest colour range (>= 6 levels) whose encoding fits (zap_tex.h:996-999); -1 if none
ZT_HD int zt_colour_range(int nvals, int avail) {
    for (int i = 20; i >= 4; i--) if (zt_ise_bits(zt_ise(i), nvals) <= avail) return i;
    return -1;
}
ZT_HD int zt_replicate(int v, int n, int to) {  // zap_tex.h:1001-1005
    int x = 0;
    for (int sh = to - n; sh > -n; sh -= n) x |= sh >= 0 ? v << sh : v >> -sh;
    return x & ((1 << to) - 1);
}
ZT_HD int zt_colour_unquant(zt_ise_range r, int v) {  // zap_tex.h:1008-1027
    if (!r.trit && !r.quint) return zt_replicate(v, r.bits, 8);
    int n = r.bits, m = v & ((1 << n) - 1), D = v >> n;
    int a = m & 1, b = m >> 1 & 1, c = m >> 2 & 1, d = m >> 3 & 1, e = m >> 4 & 1, f = m >> 5 & 1, A = a ? 0x1FF : 0, B = 0, C = 0;
    switch (r.levels) {
    case 6: C = 204; break;
    case 10: C = 113; break;
    case 12: B = b << 8 | b << 4 | b << 2 | b << 1; C = 93; break;
    case 20: B = b << 8 | b << 3 | b << 2; C = 54; break;
    case 24: B = c << 8 | b << 7 | c << 3 | b << 2 | c << 1 | b; C = 44; break;
    case 40: B = c << 8 | b << 7 | c << 2 | b << 1 | c; C = 26; break;
    case 48: B = d << 8 | c << 7 | b << 6 | d << 2 | c << 1 | b; C = 22; break;
    case 80: B = d << 8 | c << 7 | b << 6 | d << 1 | c; C = 13; break;
    case 96: B = e << 8 | d << 7 | c << 6 | b << 5 | e << 1 | d; C = 11; break;
    case 160: B = e << 8 | d << 7 | c << 6 | b << 5 | e; C = 6; break;
    case 192: B = f << 8 | e << 7 | d << 6 | c << 5 | b << 4 | f; C = 5; break;
    }
    int T = (D * C + B) ^ A;
    return (A & 0x80) | (T >> 2);
}
ZT_HD int zt_weight_unquant(int w, int n) { int x = zt_replicate(w, n, 6); return x > 32 ? x + 1 : x; }

ZT_HD void zt_trits(int T, int t[5]) {  // 8 bits -> 5 trits (spec C.2.12; zap_tex.h:1032-1042)
    int C;
    if ((T >> 2 & 7) == 7) { C = (T >> 5 & 7) << 2 | (T & 3); t[4] = 2; t[3] = 2; }
    else {
        C = T & 31;
        if ((T >> 5 & 3) == 3) { t[4] = 2; t[3] = T >> 7 & 1; } else { t[4] = T >> 7 & 1; t[3] = T >> 5 & 3; }
    }
    if ((C & 3) == 3) { t[2] = 2; t[1] = C >> 4 & 1; t[0] = (C >> 3 & 1) << 1 | ((C >> 2 & 1) & ~(C >> 3) & 1); }
    else if ((C >> 2 & 3) == 3) { t[2] = 2; t[1] = 2; t[0] = C & 3; }
    else { t[2] = C >> 4 & 1; t[1] = C >> 2 & 3; t[0] = (C >> 1 & 1) << 1 | ((C & 1) & ~(C >> 1) & 1); }
}
ZT_HD void zt_quints(int Q, int q[3]) {  // 7 bits -> 3 quints (zap_tex.h:1044-1053)
    if ((Q >> 1 & 3) == 3 && (Q >> 5 & 3) == 0) {
        q[2] = (Q & 1) << 2 | ((Q >> 4 & 1) & ~Q & 1) << 1 | ((Q >> 3 & 1) & ~Q & 1); q[1] = 4; q[0] = 4;
        return;
    }
    int C;
    if ((Q >> 1 & 3) == 3) { q[2] = 4; C = (Q >> 3 & 3) << 3 | (~(Q >> 5) & 3) << 1 | (Q & 1); }
    else { q[2] = Q >> 5 & 3; C = Q & 31; }
    if ((C & 7) == 5) { q[1] = 4; q[0] = C >> 3 & 3; } else { q[1] = C >> 3 & 3; q[0] = C & 7; }
}
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
    uint8_t q[8][256];     // 8-bit value -> nearest ISE value of the cfg's colour range
    uint8_t unq[8][256];   // ISE value -> 0..255
    uint8_t tenc[256];     // 5 trits (base-3 digits, t0 lowest) -> packed 8-bit code (smallest code wins)
    uint8_t qenc[128];     // 3 quints -> 7-bit code
    uint16_t pmask[1024];  // usable 2-partition patterns: bit i = texel i in partition 1
    uint16_t pseed[1024];
    int np;
};

inline void zt_astc_setup_build(zt_astc_setup *s) {
    for (int T = 255; T >= 0; T--) { int t[5]; zt_trits(T, t); s->tenc[t[0] + 3 * t[1] + 9 * t[2] + 27 * t[3] + 81 * t[4]] = (uint8_t)T; }
    for (int Q = 127; Q >= 0; Q--) { int q[3]; zt_quints(Q, q); s->qenc[q[0] + 5 * q[1] + 25 * q[2]] = (uint8_t)Q; }
    const int spec[8][3] = { { 8, 4, 1 }, { 8, 3, 1 }, { 8, 5, 1 }, { 12, 4, 1 }, { 12, 3, 1 }, { 12, 2, 1 }, { 8, 3, 2 }, { 8, 2, 2 } };
    for (int k = 0; k < 8; k++) {
        zt_astc_cfg &c = s->cfg[k];
        c.cem = spec[k][0]; c.wb = spec[k][1]; c.parts = spec[k][2]; c.nv = c.cem == 8 ? 6 : 8;
        c.range = zt_colour_range(c.nv * c.parts, 128 - (c.parts == 1 ? 17 : 29) - 16 * c.wb);
        zt_ise_range r = zt_ise(c.range);
        c.levels = r.levels; c.trit = r.trit; c.quint = r.quint; c.bits = r.bits;
        for (int v = 0; v < 256; v++) s->unq[k][v] = (uint8_t)(v < r.levels ? zt_colour_unquant(r, v) : 0);
        for (int x = 0; x < 256; x++) {
            int best = 0, bd = 1 << 30;
            for (int v = 0; v < r.levels; v++) { int d = zt_colour_unquant(r, v) - x; d = d < 0 ? -d : d; if (d < bd) { bd = d; best = v; } }
            s->q[k][x] = (uint8_t)best;
        }
    }
    s->np = 0;
    for (int sd = 0; sd < 1024; sd++) {
        unsigned m = 0;
        for (int i = 0; i < 16; i++) m |= (unsigned)zt_astc_partition(sd, i & 3, i >> 2, 2) << i;
        if (m && m != 0xFFFFu) { s->pmask[s->np] = (uint16_t)m; s->pseed[s->np++] = (uint16_t)sd; }
    }
}
