// zt_astc.hip — ASTC 4x4 LDR encoder and decoder for gfx950 (zap_tex.h:977-1356).
//
// Per block: three 1-partition trial configurations (CEM 8 for opaque blocks, CEM 12 otherwise), each a PCA
// start, quantisation through the configuration's 8-bit -> ISE map, a weight argmin and <= 3 least-squares
// refinements; opaque blocks with SSE > 192 then try the 4 stored 2-partition patterns nearest to a 2-means
// split.  Configurations are compile-time (their colour ranges are fixed by the spec rule, checked at set-up),
// so every per-texel loop unrolls into registers.  The encoder runs in stages (zt_kernels.h): the trials are
// items (configuration, listed block) or (configuration, block, pattern), configuration-major, so each lane holds
// one trial's registers and a wave runs one configuration's code; picks replay the reference's comparison order.
#include "zt_astc_setup.h"
#include "zt_common.h"
#include "zt_kernels.h"
#include "zt_pca.h"

namespace {

__device__ zt_astc_setup g_astc;

// trial configuration K: CEM, weight bits, partitions, and its colour range (trits, quints, bits)
template <int K> struct AC;
template <> struct AC<0> { static constexpr int CEM = 8, WB = 4, PARTS = 1, TRIT = 1, QUINT = 0, BITS = 6; };
template <> struct AC<1> { static constexpr int CEM = 8, WB = 3, PARTS = 1, TRIT = 0, QUINT = 0, BITS = 8; };
template <> struct AC<2> { static constexpr int CEM = 8, WB = 5, PARTS = 1, TRIT = 0, QUINT = 0, BITS = 5; };
template <> struct AC<3> { static constexpr int CEM = 12, WB = 4, PARTS = 1, TRIT = 1, QUINT = 0, BITS = 4; };
template <> struct AC<4> { static constexpr int CEM = 12, WB = 3, PARTS = 1, TRIT = 1, QUINT = 0, BITS = 6; };
template <> struct AC<5> { static constexpr int CEM = 12, WB = 2, PARTS = 1, TRIT = 0, QUINT = 0, BITS = 8; };
template <> struct AC<6> { static constexpr int CEM = 8, WB = 3, PARTS = 2, TRIT = 0, QUINT = 0, BITS = 4; };
template <> struct AC<7> { static constexpr int CEM = 8, WB = 2, PARTS = 2, TRIT = 0, QUINT = 1, BITS = 3; };

ZT_DEV int lerp(int a, int b, int w) { return (((a * 257) * (64 - w) + (b * 257) * w + 32) >> 6) >> 8; }

// decoded endpoints of one partition (zap_tex.h:1125-1129); packed RGBA8 in e0/e1
template <int CEM> ZT_DEV void endpoints(const int *v, int e0[4], int e1[4]) {
    int a0 = CEM == 12 ? v[6] : 255, a1 = CEM == 12 ? v[7] : 255;
    if (v[1] + v[3] + v[5] >= v[0] + v[2] + v[4]) {
        e0[0] = v[0]; e0[1] = v[2]; e0[2] = v[4]; e0[3] = a0; e1[0] = v[1]; e1[1] = v[3]; e1[2] = v[5]; e1[3] = a1;
    } else {
        e0[0] = (v[1] + v[5]) >> 1; e0[1] = (v[3] + v[5]) >> 1; e0[2] = v[5]; e0[3] = a1;
        e1[0] = (v[0] + v[4]) >> 1; e1[1] = (v[2] + v[4]) >> 1; e1[2] = v[4]; e1[3] = a0;
    }
}

// the configurations' 8-bit -> ISE maps and ISE -> 8-bit values, staged into LDS by each workgroup (every
// endpoint quantisation looks them up per lane)
struct AstcLds { uint8_t q[8][256]; uint8_t unq[8][256]; };

// a trial's colour endpoint values, one byte each (value v of partition k at byte k * NV + v): the ISE values in
// q, their unquantised 8-bit values in u
struct Ends { uint32_t q[3], u[3]; };
ZT_DEV int byte_at(const uint32_t *a, int i) { return (int)((a[i >> 2] >> (8 * (i & 3))) & 255u); }
ZT_DEV void byte_put(uint32_t *a, int i, int v) { a[i >> 2] = (a[i >> 2] & ~(255u << (8 * (i & 3)))) | (uint32_t)v << (8 * (i & 3)); }

// float endpoints -> partition k's ISE values and their unquantised values (zap_tex.h:1183-1194)
template <int K> ZT_DEV void quantise(const float lo[4], const float hi[4], Ends &E, int k, const AstcLds &T) {
    constexpr int NV = AC<K>::CEM == 8 ? 6 : 8;
    float s0 = lo[0] + lo[1] + lo[2], s1 = hi[0] + hi[1] + hi[2];
    const bool keep = s1 >= s0;
    int ise[NV], uv[NV];
#pragma unroll
    for (int c = 0; c < NV / 2; c++) {
        float a = keep ? lo[c] : hi[c], b = keep ? hi[c] : lo[c];
        int x0 = clampi(cvt_i32(a + 0.5f), 0, 255), x1 = clampi(cvt_i32(b + 0.5f), 0, 255);
        ise[2 * c] = T.q[K][x0]; ise[2 * c + 1] = T.q[K][x1];
        uv[2 * c] = T.unq[K][ise[2 * c]]; uv[2 * c + 1] = T.unq[K][ise[2 * c + 1]];
    }
    const int sw = uv[1] + uv[3] + uv[5] < uv[0] + uv[2] + uv[4];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        byte_put(E.q, k * NV + v, sw ? ise[v ^ 1] : ise[v]);
        byte_put(E.u, k * NV + v, sw ? uv[v ^ 1] : uv[v]);
    }
}

// weights for fixed endpoints (zap_tex.h:1197-1210); pm: bit i = texel i in partition 1; w: texel i's weight in
// byte i % 4 of w[i / 4].  key_j = 2^(IB+1) dot(p, pal_j) - 2^IB |pal_j|^2 + (2^IB - 1 - j): the max is the min
// SSE, lowest j on ties; IB index bits hold up to 32 entries (5-bit weights), keys stay below 2^25.  The loop runs
// over palette entries outside and texels inside, so only the 16 running maxima stay live.  The palette's unused
// alpha byte is 0, so texel alpha drops out of dot(p, pal_j); spp = sum |p|^2 over the used channels.
template <int K> ZT_DEV int fit_weights(const uint32_t px[16], int spp, const Ends &E, uint32_t pm, uint32_t w[4]) {
    constexpr int CEM = AC<K>::CEM, WB = AC<K>::WB, P = AC<K>::PARTS, L = 1 << WB, NV = CEM == 8 ? 6 : 8;
    constexpr int IB = WB > 4 ? WB : 4, IM = (1 << IB) - 1;
    constexpr uint32_t CM = CEM == 12 ? 0xFFFFFFFFu : 0x00FFFFFFu;
    int e0[P][4], e1[P][4];
#pragma unroll
    for (int k = 0; k < P; k++) {
        int v[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) v[i] = byte_at(E.u, k * NV + i);
        endpoints<CEM>(v, e0[k], e1[k]);
    }
    int best[16];
#pragma unroll
    for (int j = 0; j < L; j++) {
        const int wt = zt_weight_value(j, WB);
        uint32_t pj[P];
        int kj[P];
#pragma unroll
        for (int k = 0; k < P; k++) {
            uint32_t v = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) v |= (uint32_t)lerp(e0[k][c], e1[k][c], wt) << (8 * c);
            pj[k] = v & CM;
            kj[k] = no_fold(IM - j - (1 << IB) * (int)dot4(pj[k], pj[k]));
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const bool second = P == 2 && ((pm >> i) & 1);
            const int key = (int)(dot4(px[i], second ? pj[P - 1] : pj[0]) << (IB + 1)) + (second ? kj[P - 1] : kj[0]);
            best[i] = j ? max(best[i], key) : key;
        }
    }
    int sk = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        sk += best[i] >> IB;
        w[i >> 2] |= (uint32_t)(IM - (best[i] & IM)) << (8 * (i & 3));
    }
    return spp - sk;
}
ZT_DEV int wbyte(const uint32_t w[4], int i) { return (int)((w[i >> 2] >> (8 * (i & 3))) & 255u); }
// four WB-bit weight indices (one per byte) -> their values on the 0..64 scale, per byte: bit replication to 6 bits,
// then +1 above 32 (every byte stays below 128, so no step carries into the next byte)
template <int WB> ZT_DEV uint32_t wvals4(uint32_t x) {
    uint32_t v;
    if (WB == 2) v = x * 21u;
    else if (WB == 3) v = x * 9u;
    else if (WB == 4) v = x << 2 | ((x >> 2) & 0x03030303u);
    else v = x << 1 | ((x >> 4) & 0x01010101u);
    return v + (((v + 0x5F5F5F5Fu) >> 7) & 0x01010101u);
}
// channel planes of pixels 4r..4r+3: byte j of pl[c] = channel c of pixel 4r + j
ZT_DEV void planes(const uint32_t px[16], int r, uint32_t pl[4]) {
    const uint32_t a = px[4 * r], b = px[4 * r + 1], c = px[4 * r + 2], d = px[4 * r + 3];
    const uint32_t ab01 = perm(b, a, 0x05010400u), ab23 = perm(b, a, 0x07030602u);
    const uint32_t cd01 = perm(d, c, 0x05010400u), cd23 = perm(d, c, 0x07030602u);
    pl[0] = perm(cd01, ab01, 0x05040100u); pl[1] = perm(cd01, ab01, 0x07060302u);
    pl[2] = perm(cd23, ab23, 0x05040100u); pl[3] = perm(cd23, ab23, 0x07060302u);
}

// one trial configuration (zap_tex.h:1212-1244).  The least-squares refit's sums are exact in fp32: with f = wv/64
// and g = 1 - f for a weight value wv in 0..64, each g*g, g*f, f*f is a multiple of 1/4096 and each g*p, f*p a
// multiple of 1/64, and every partial sum stays below 2^24 of those units, so summing in any order gives the
// reference's floats.  They are summed here as integers, four texels per dot4 over byte planes of the pixels.
template <int K> ZT_DEV int trial(uint32_t px[16], uint32_t pm, Ends &E, uint32_t w[4], const AstcLds &T) {
    constexpr int CEM = AC<K>::CEM, WB = AC<K>::WB, P = AC<K>::PARTS, NC = CEM == 12 ? 4 : 3;
    constexpr uint32_t CM = NC == 4 ? 0xFFFFFFFFu : 0x00FFFFFFu;
    int spp = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) spp += (int)dot4(px[i] & CM, px[i] & CM);
#pragma unroll
    for (int q = 0; q < 3; q++) { E.q[q] = 0; E.u[q] = 0; }
#pragma unroll
    for (int k = 0; k < P; k++) {
        float lo[4], hi[4];
        if (P == 1) pca<NC, true, true>(px, 0xFFFFu, 16, lo, hi);
        else {
            uint32_t m = k ? pm : (~pm & 0xFFFFu);
            int cnt = __popc(m);
            if (!cnt) { m = 1; cnt = 1; }
            pca<NC, false, true>(px, m, cnt, lo, hi);
        }
        quantise<K>(lo, hi, E, k, T);
    }
    int err = fit_weights<K>(px, spp, E, pm, w);
#pragma unroll 1
    for (int it = 0; it < 3 && err > 0; it++) {
        Ends t = E;
        uint32_t tw[4];
#pragma unroll
        for (int k = 0; k < P; k++) {
            const uint32_t m = P == 1 ? 0xFFFFu : k ? pm : (~pm & 0xFFFFu);
            uint32_t s00 = 0, s01 = 0, s11 = 0, s0[NC], s1[NC];   // sums of g*g, g*f, f*f, g*p, f*p, scaled to integers
#pragma unroll
            for (int c = 0; c < NC; c++) { s0[c] = 0; s1[c] = 0; }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // texels 4r..4r+3: 64 f and 64 g per byte, zero outside the partition
                const uint32_t bm = P == 1 ? 0xFFFFFFFFu : ((((m >> (4 * r)) & 15u) * 0x00204081u) & 0x01010101u) * 0xFFu;
                const uint32_t wv = wvals4<WB>(w[r]), v1 = wv & bm, v0 = (0x40404040u - wv) & bm;
                s00 = dot4acc(v0, v0, s00); s01 = dot4acc(v0, v1, s01); s11 = dot4acc(v1, v1, s11);
                uint32_t pl[4];
                planes(px, r, pl);
#pragma unroll
                for (int c = 0; c < NC; c++) { s0[c] = dot4acc(v0, pl[c], s0[c]); s1[c] = dot4acc(v1, pl[c], s1[c]); }
            }
            const float aa = (float)s00 / 4096.0f, ab = (float)s01 / 4096.0f, bb = (float)s11 / 4096.0f;
            const float det = aa * bb - ab * ab;
            if (det < 1e-6f && det > -1e-6f) continue;
            float nlo[4], nhi[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float xa = (float)s0[c < NC ? c : 0] / 64.0f, xb = (float)s1[c < NC ? c : 0] / 64.0f;
                if (c < NC) { nlo[c] = (xa * bb - xb * ab) / det; nhi[c] = (xb * aa - xa * ab) / det; }
                else { nlo[c] = 255.0f; nhi[c] = 255.0f; }
            }
            quantise<K>(nlo, nhi, t, k, T);
        }
        int e = fit_weights<K>(px, spp, t, pm, tw);
        if (e >= err) break;
        err = e;
        E = t;
#pragma unroll
        for (int q = 0; q < 4; q++) w[q] = tw[q];
    }
    return err;
}

// integer-sequence encode of N values (zap_tex.h:1080-1095)
template <int K, int N> ZT_DEV void ise_put(Bits128 &b, int pos, const uint32_t *v) {
    constexpr int TR = AC<K>::TRIT, QU = AC<K>::QUINT, BITS = AC<K>::BITS;
    constexpr int GS = TR ? 5 : QU ? 3 : 1, BASE = TR ? 3 : 5;
    const int xb[5] = { TR ? 2 : 3, 2, TR ? 1 : 2, 2, 1 };
#pragma unroll
    for (int g = 0; g < N; g += GS) {
        const int cnt = N - g < GS ? N - g : GS;
        int X = 0, xp = 0;
        if (GS > 1) {
            int id = 0;
#pragma unroll
            for (int k = GS - 1; k >= 0; k--) id = id * BASE + (k < cnt ? byte_at(v, g + k) >> BITS : 0);
            X = TR ? g_astc.tenc[id] : g_astc.qenc[id];
        }
#pragma unroll
        for (int k = 0; k < cnt; k++) {
            if (BITS) { b.put(pos, (uint64_t)(byte_at(v, g + k) & ((1 << BITS) - 1)), BITS); pos += BITS; }
            if (GS > 1) { b.put(pos, (uint64_t)((X >> xp) & ((1 << xb[k]) - 1)), xb[k]); pos += xb[k]; xp += xb[k]; }
        }
    }
}

template <int K> ZT_DEV void pack(int seed, const Ends &E, const uint32_t w[4], uint64_t &olo, uint64_t &ohi) {  // zap_tex.h:1246-1260
    constexpr int CEM = AC<K>::CEM, WB = AC<K>::WB, P = AC<K>::PARTS, NV = CEM == 8 ? 6 : 8;
    constexpr int R = WB == 2 ? 4 : WB == 3 ? 7 : WB == 4 ? 4 : 7, H = WB >= 4 ? 1 : 0;
    constexpr uint32_t mode = (R >> 1 & 1) | (R >> 2 & 1) << 1 | (R & 1) << 4 | 2 << 5 | H << 9;
    Bits128 b{ mode | (uint64_t)(P - 1) << 11, 0 };
    int pos;
    if (P == 2) { b.put(13, (uint64_t)seed, 10); b.put(23, (uint64_t)CEM << 2, 6); pos = 29; }
    else { b.put(13, (uint64_t)CEM, 4); pos = 17; }
    ise_put<K, NV * P>(b, pos, E.q);
    Bits128 wb{ 0, 0 };
#pragma unroll
    for (int i = 0; i < 16; i++) wb.put(i * WB, (uint64_t)wbyte(w, i), WB);
    // weight bit i lands at block bit 127 - i
    olo = b.lo | __brevll(wb.hi);
    ohi = b.hi | __brevll(wb.lo);
}

ZT_DEV uint32_t two_means(const uint32_t px[16]) {  // zap_tex.h:1263-1283
    int lo = 0, hi = 0, slo = chan(px[0], 0) + chan(px[0], 1) + chan(px[0], 2), shi = slo;
#pragma unroll
    for (int i = 1; i < 16; i++) {
        int s = chan(px[i], 0) + chan(px[i], 1) + chan(px[i], 2);
        if (s < slo) { lo = i; slo = s; }
        if (s > shi) { hi = i; shi = s; }
    }
    float c[2][3];
    uint32_t plo = 0, phi = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) { if (i == lo) plo = px[i]; if (i == hi) phi = px[i]; }
#pragma unroll
    for (int k = 0; k < 3; k++) { c[0][k] = chanf(plo, k); c[1][k] = chanf(phi, k); }
    uint32_t m = 0;
    for (int it = 0; it < 4; it++) {
        float sum[2][3] = { { 0.0f, 0.0f, 0.0f }, { 0.0f, 0.0f, 0.0f } };
        int n[2] = { 0, 0 };
        m = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float d0 = 0.0f, d1 = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) { float a = chanf(px[i], k) - c[0][k], b = chanf(px[i], k) - c[1][k]; d0 += a * a; d1 += b * b; }
            const bool s = d1 < d0;
            m |= (uint32_t)s << i;
            if (s) { n[1]++; } else { n[0]++; }
#pragma unroll
            for (int k = 0; k < 3; k++) { if (s) sum[1][k] += chanf(px[i], k); else sum[0][k] += chanf(px[i], k); }
        }
#pragma unroll
        for (int s = 0; s < 2; s++) if (n[s]) {
#pragma unroll
            for (int k = 0; k < 3; k++) c[s][k] = sum[s][k] / (float)n[s];
        }
    }
    return m;
}

// ---- decoder (zap_tex.h:1135-1168): void-extent blocks and the 4x4 single-plane modes with plain-bit weights and
// colour endpoint modes 8 / 12 (one or two partitions); anything else decodes to magenta.  Every lookup is a set-up
// table: the block mode's weight bits, the trit/quint digits of a transfer block, each range's unquantised colour
// values, the weight values, the 2-partition pattern of the seed.

// n values of ISE range r starting at block bit `pos`: each group of 5 (trits) or 3 (quints) values interleaves
// its plain fields with the transfer bits of one trit/quint block, 2,2,1,2,1 or 3,2,2 bits after the fields
ZT_DEV void ise_read(uint64_t lo, uint64_t hi, int pos, const zt_ise_range &r, int n, int v[16]) {
    const int gs = r.trit ? 5 : r.quint ? 3 : 1;
    const uint32_t xbits = r.trit ? 0x12122u : 0x223u;   // transfer bits after field j: nibble j
    for (int g = 0; g < n; g += gs) {
        const int cnt = min(gs, n - g);
        int X = 0, xp = 0;
        for (int j = 0; j < cnt; j++) {
            v[g + j] = (int)get_bits(lo, hi, pos, r.bits);
            pos += r.bits;
            if (gs > 1) {
                const int xb = (int)(xbits >> (4 * j) & 15u);
                X |= (int)get_bits(lo, hi, pos, xb) << xp;
                pos += xb; xp += xb;
            }
        }
        if (gs == 1) continue;
        const uint32_t dig = r.trit ? g_astc.tdec[X] : g_astc.qdec[X];
        for (int j = 0; j < cnt; j++) v[g + j] |= (int)(r.trit ? dig >> (2 * j) & 3u : dig >> (3 * j) & 7u) << r.bits;
    }
}

ZT_DEV void decode_block(const uint8_t *blk, uint32_t out[16]) {
    const uint4 raw = *(const uint4 *)blk;
    const uint64_t lo = (uint64_t)raw.x | (uint64_t)raw.y << 32, hi = (uint64_t)raw.z | (uint64_t)raw.w << 32;
    for (int i = 0; i < 16; i++) out[i] = 0xFFFF00FFu;
    const int mode = (int)(lo & 0x7FFu);
    if ((mode & 0x1FF) == 0x1FC) {   // void extent: the UNORM16 colour at bits 64..127, high byte per channel
        uint32_t v = 0;
        for (int c = 0; c < 4; c++) v |= (uint32_t)(hi >> (16 * c + 8) & 0xFFu) << (8 * c);
        for (int i = 0; i < 16; i++) out[i] = v;
        return;
    }
    const int wb = g_astc.mode_wb[mode], parts = (int)(lo >> 11 & 3u) + 1;
    if (!wb || parts > 2) return;
    int seed = 0, cem, pos;
    if (parts == 2) {
        const int cf = (int)(lo >> 23 & 0x3Fu);
        if (cf & 3) return;   // mixed endpoint modes are outside the subset
        seed = (int)(lo >> 13 & 0x3FFu); cem = cf >> 2; pos = 29;
    } else {
        cem = (int)(lo >> 13 & 15u); pos = 17;
    }
    const int nv = cem == 8 ? 6 : cem == 12 ? 8 : 0;
    if (!nv) return;
    const int ri = zt_colour_range(nv * parts, 128 - pos - 16 * wb);
    if (ri < 0) return;
    int v[16];
    ise_read(lo, hi, pos, zt_ise(ri), nv * parts, v);
    for (int i = 0; i < nv * parts; i++) v[i] = g_astc.colour[ri][v[i]];
    int e0[2][4], e1[2][4];
    for (int k = 0; k < parts; k++) {
        if (cem == 8) endpoints<8>(v + k * nv, e0[k], e1[k]);
        else endpoints<12>(v + k * nv, e0[k], e1[k]);
    }
    // weights fill the block from bit 127 down: reversed, they read forward from bit 0
    const uint64_t wlo = __brevll(hi), whi = __brevll(lo);
    const uint32_t pm = parts == 2 ? g_astc.part2[seed] : 0u;
    for (int i = 0; i < 16; i++) {
        const int k = (int)(pm >> i & 1u), w = g_astc.weight[wb][get_bits(wlo, whi, i * wb, wb)];
        uint32_t o = 0;
        for (int c = 0; c < 4; c++) o |= (uint32_t)lerp(e0[k][c], e1[k][c], w) << (8 * c);
        out[i] = o;
    }
}

}  // namespace

hipError_t zt_astc_upload_tables(const void *setup, size_t bytes) {
    const zt_astc_setup *s = (const zt_astc_setup *)setup;
    // the kernels' compile-time colour ranges must be the ones the spec rule picks
    const int tr[8] = { AC<0>::TRIT, AC<1>::TRIT, AC<2>::TRIT, AC<3>::TRIT, AC<4>::TRIT, AC<5>::TRIT, AC<6>::TRIT, AC<7>::TRIT };
    const int qu[8] = { AC<0>::QUINT, AC<1>::QUINT, AC<2>::QUINT, AC<3>::QUINT, AC<4>::QUINT, AC<5>::QUINT, AC<6>::QUINT, AC<7>::QUINT };
    const int bi[8] = { AC<0>::BITS, AC<1>::BITS, AC<2>::BITS, AC<3>::BITS, AC<4>::BITS, AC<5>::BITS, AC<6>::BITS, AC<7>::BITS };
    for (int k = 0; k < 8; k++)
        if (s->cfg[k].trit != tr[k] || s->cfg[k].quint != qu[k] || s->cfg[k].bits != bi[k]) return hipErrorInvalidValue;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_astc), setup, bytes);
}

// ---- the staged encoder (declared with its stages in zt_kernels.h)
namespace {
ZT_DEV void stage_tables(AstcLds &T) {
    for (int i = threadIdx.x; i < 8 * 256 / 4; i += blockDim.x) {
        ((uint32_t *)T.q)[i] = ((const uint32_t *)g_astc.q)[i];
        ((uint32_t *)T.unq)[i] = ((const uint32_t *)g_astc.unq)[i];
    }
    __syncthreads();
}
ZT_DEV void load_px(const zt_astc_ws &W, uint32_t g, uint32_t px[16]) {
#pragma unroll
    for (int r = 0; r < 4; r++) { const uint4 v = W.px[(uint64_t)g * 4 + r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
}
// the sharded lists' counters: list k's shards start at counters + k * ZT_SHARDS * ZT_SHARD_STRIDE
ZT_DEV unsigned *list_counts(const zt_astc_ws &W, int k) { return W.counters + k * ZT_SHARDS * ZT_SHARD_STRIDE; }
// append g to sharded list k where `pred` holds (item0: the wave's first item)
ZT_DEV void list_append(const zt_astc_ws &W, int k, uint32_t *list, bool pred, uint32_t g, uint32_t item0) {
    unsigned *c = shard_counter(list_counts(W, k), item0);
    const int slot = wave_append(pred, c);
    if (slot >= 0) list[(uint32_t)((c - list_counts(W, k)) / ZT_SHARD_STRIDE) * W.cap + (uint32_t)slot] = g;
}
// block of opaque-list entry e
ZT_DEV uint32_t opaque_block(const zt_astc_ws &W, uint32_t e) { return W.lo[shard_slot(list_counts(W, 0), W.cap, e)]; }
// configuration K's trial of list entry e (opaque list for K < 3, translucent otherwise; n0 opaque entries)
template <int K> ZT_DEV void single_item(const zt_astc_ws &W, const AstcLds &T, uint32_t e, uint32_t n0) {
    const uint32_t g = (K < 3 ? W.lo : W.lt)[shard_slot(list_counts(W, K < 3 ? 0 : 1), W.cap, e)];
    uint32_t px[16];
    load_px(W, g, px);
    Ends E;
    uint32_t w[4];
    uint64_t lo, hi;
    const int err = trial<K>(px, 0, E, w, T);
    pack<K>(0, E, w, lo, hi);
    const uint64_t o = (uint64_t)(K % 3) * W.n + (K < 3 ? e : n0 + e);
    W.err1[o] = err;
    W.blk1[o] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}
// configuration K's trial of pattern candidate i (= listed block i / 4, its pattern i % 4)
template <int K> ZT_DEV void pair_item(const zt_astc_ws &W, const AstcLds &T, uint32_t i) {
    const uint32_t c = W.cand[i], o = (i >> 2) * 8 + (i & 3) * 2 + (K - 6);
    if (c == 0xFFFFFFFFu) { W.err2[o] = 0x7FFFFFFF; return; }
    uint32_t px[16];
    load_px(W, opaque_block(W, W.lp[shard_slot(list_counts(W, 2), W.cap, i >> 2)]), px);
    Ends E;
    uint32_t w[4];
    uint64_t lo, hi;
    const int err = trial<K>(px, c & 0xFFFFu, E, w, T);
    pack<K>((int)(c >> 16), E, w, lo, hi);
    W.err2[o] = err;
    W.blk2[o] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}
// the 1-partition winner of entry u: first smallest error in configuration order; returns its error, `at` its result
ZT_DEV int first_best(const zt_astc_ws &W, uint32_t u, uint64_t &at) {
    int best = W.err1[u];
    at = u;
#pragma unroll
    for (int k = 1; k < 3; k++) {
        const uint64_t o = (uint64_t)k * W.n + u;
        const int e = W.err1[o];
        if (e < best) { best = e; at = o; }
    }
    return best;
}
ZT_DEV void store_block(const zt_jobs &jobs, uint32_t g, uint4 v) {
    const zt_job J = job_at(jobs, zt_find_job(jobs, g));
    store16(J.dst + (uint64_t)(g - J.first) * 16, (uint64_t)v.x | (uint64_t)v.y << 32, (uint64_t)v.z | (uint64_t)v.w << 32);
}
}  // namespace

extern "C" __global__ void __launch_bounds__(256) zt_astc_classify(zt_jobs jobs, uint32_t total, zt_astc_ws W) {
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t g0 = blockIdx.x * blockDim.x; g0 < total; g0 += step) {   // whole waves stay in the loop
        const uint32_t g = g0 + threadIdx.x;
        bool opaque = false;
        const bool valid = g < total;
        if (valid) {
            const zt_job J = job_at(jobs, zt_find_job(jobs, g));
            const uint32_t b = g - J.first;
            uint32_t px[16];
            gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
            uint32_t amin = 255;
#pragma unroll
            for (int i = 0; i < 16; i++) amin = min(amin, px[i] >> 24);
            opaque = amin == 255;
#pragma unroll
            for (int r = 0; r < 4; r++) W.px[(uint64_t)g * 4 + r] = make_uint4(px[4 * r], px[4 * r + 1], px[4 * r + 2], px[4 * r + 3]);
        }
        const uint32_t item0 = g0 + (threadIdx.x & ~63u);
        list_append(W, 0, W.lo, valid && opaque, g, item0);
        list_append(W, 1, W.lt, valid && !opaque, g, item0);
    }
}

// every 1-partition trial in one launch: items [0, 3 n0) are configurations 0..2 over the opaque list, items
// [3 n0, 3 n0 + 3 n1) configurations 3..5 over the translucent list, configuration-major, so a wave runs one
// configuration's code (at most a few waves straddle two).  Each item holds one trial's registers only; at least 4
// waves per SIMD (a handful of values spill to reach them).
extern "C" __global__ void __launch_bounds__(256, 4) zt_astc_single(zt_astc_ws W) {
    __shared__ AstcLds T;
    stage_tables(T);
    const uint32_t n0 = shard_total(list_counts(W, 0)), n1 = shard_total(list_counts(W, 1));
    const uint32_t n = 3 * n0 + 3 * n1, step = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const bool op = i < 3 * n0;
        const uint32_t j = op ? i : i - 3 * n0, m = op ? n0 : n1, c = j / m, e = j - c * m;
        switch (op ? c : 3 + c) {
        case 0: single_item<0>(W, T, e, n0); break;
        case 1: single_item<1>(W, T, e, n0); break;
        case 2: single_item<2>(W, T, e, n0); break;
        case 3: single_item<3>(W, T, e, n0); break;
        case 4: single_item<4>(W, T, e, n0); break;
        default: single_item<5>(W, T, e, n0); break;
        }
    }
}

// over entries u: the opaque list's entries, then the translucent list's
extern "C" __global__ void __launch_bounds__(256) zt_astc_pick1(zt_jobs jobs, uint32_t total, zt_astc_ws W) {
    const uint32_t step = gridDim.x * blockDim.x, n0 = shard_total(list_counts(W, 0));
    for (uint32_t u0 = blockIdx.x * blockDim.x; u0 < total; u0 += step) {
        const uint32_t u = u0 + threadIdx.x;
        bool search = false;
        if (u < total) {
            const bool opaque = u < n0;
            const uint32_t g = opaque ? W.lo[shard_slot(list_counts(W, 0), W.cap, u)] : W.lt[shard_slot(list_counts(W, 1), W.cap, u - n0)];
            uint64_t at;
            const int best = first_best(W, u, at);
            search = opaque && best > 16 * 3 * 4;   // opaque and not good enough yet (zap_tex.h:1327)
            if (!search) store_block(jobs, g, W.blk1[at]);
        }
        list_append(W, 2, W.lp, search, u, u0 + (threadIdx.x & ~63u));
    }
}

// 16 lanes per listed block (4 blocks per wave): each lane computes the block's 2-means split (zap_tex.h:1263-1283)
// and scores patterns lane, lane + 16, ... as (Hamming distance under either labelling) << 10 | pattern, keeping
// its 4 smallest; the 16 lanes then merge theirs.  The 4 smallest keys are the reference's scan result
// (zap_tex.h:1328-1336: ascending distance, earlier patterns first among equal distances).
extern "C" __global__ void __launch_bounds__(256) zt_astc_search(zt_astc_ws W) {
    const unsigned *counts = list_counts(W, 2);
    const uint32_t n = shard_total(counts), groups = gridDim.x * (blockDim.x >> 4);
    const int sub = threadIdx.x & 15, np = g_astc.np;
    const uint32_t e0 = blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);
    const uint32_t trips = (n + groups - 1) / groups;   // the same for every lane: the shuffles below need whole waves
    for (uint32_t t = 0; t < trips; t++) {
        const uint32_t e = e0 + t * groups;
        const bool valid = e < n;
        if (!__any(valid)) continue;   // whole idle waves skip the work (a wave's groups hold consecutive entries)
        uint32_t px[16];
        load_px(W, valid ? opaque_block(W, W.lp[shard_slot(counts, W.cap, e)]) : 0u, px);   // idle groups read block 0, store nothing
        const uint32_t km = two_means(px);
        uint32_t b0 = 0xFFFFFFFFu, b1 = 0xFFFFFFFFu, b2 = 0xFFFFFFFFu, b3 = 0xFFFFFFFFu;   // ascending
#pragma unroll 8
        for (int j = 0; j < 1024 / 16; j++) {
            const int pi = sub + 16 * j;
            if (pi >= np) break;
            const int dd = __popc((uint32_t)g_astc.pmask[pi] ^ km), dist = min(dd, 16 - dd);
            const uint32_t x = (uint32_t)dist << 10 | (uint32_t)pi;
            b3 = max(b2, min(b3, x));
            b2 = max(b1, min(b2, x));
            b1 = max(b0, min(b1, x));
            b0 = min(b0, x);
        }
        uint32_t out[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            uint32_t m = b0;
#pragma unroll
            for (int off = 8; off >= 1; off >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, off));
            out[r] = m;
            const bool pop = b0 == m;
            b0 = pop ? b1 : b0; b1 = pop ? b2 : b1; b2 = pop ? b3 : b2; b3 = pop ? 0xFFFFFFFFu : b3;
        }
        if (valid && sub < 4) {
            const uint32_t k = sub == 0 ? out[0] : sub == 1 ? out[1] : sub == 2 ? out[2] : out[3];
            const int pi = (int)(k & 1023u);
            W.cand[(uint64_t)e * 4 + sub] = k == 0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)g_astc.pmask[pi] | (uint32_t)g_astc.pseed[pi] << 16;
        }
    }
}

// every 2-partition trial in one launch: items [0, 4 n) are configuration 6 over the (listed block, pattern)
// candidates, items [4 n, 8 n) configuration 7
extern "C" __global__ void __launch_bounds__(256, 4) zt_astc_pair(zt_astc_ws W) {
    __shared__ AstcLds T;
    stage_tables(T);
    const uint32_t n4 = shard_total(list_counts(W, 2)) * 4u, step = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n4; i += step) {
        if (i < n4) pair_item<6>(W, T, i);
        else pair_item<7>(W, T, i - n4);
    }
}

extern "C" __global__ void __launch_bounds__(256) zt_astc_pick2(zt_jobs jobs, zt_astc_ws W) {
    const uint32_t n = shard_total(list_counts(W, 2)), step = gridDim.x * blockDim.x;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += step) {
        const uint32_t u = W.lp[shard_slot(list_counts(W, 2), W.cap, e)], g = opaque_block(W, u);
        uint64_t at;
        int best = first_best(W, u, at);
        uint4 v = W.blk1[at];
        for (int j = 0; j < 4; j++) {   // the reference's order: pattern j with configuration 6, then 7
            if (W.cand[(uint64_t)e * 4 + j] == 0xFFFFFFFFu) break;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int err = W.err2[(uint64_t)e * 8 + j * 2 + k];
                if (err < best) { best = err; v = W.blk2[(uint64_t)e * 8 + j * 2 + k]; }
            }
        }
        store_block(jobs, g, v);
    }
}

extern "C" __global__ void __launch_bounds__(256) zt_astc_decode(const uint8_t *bc, int w, int h, uint8_t *rgba, uint64_t stride) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bw * bh) return;
    const int bx = g % bw, by = g / bw;
    uint32_t px[16];
    decode_block(bc + (uint64_t)g * 16, px);
    for (int i = 0; i < 16; i++) {
        int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
        if (x < w && y < h) *(uint32_t *)(rgba + (uint64_t)y * stride + (uint64_t)x * 4) = px[i];
    }
}
