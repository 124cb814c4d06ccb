// zt_bc7_core.h — BC7 device building blocks shared by the reference search (zt_bc7.hip) and the
// exhaustive / learned-proposal search (zt_bc7_full.hip): spec tables, p-bit and plain quantisers, exact
// palette argmins, subset encoders (PCA start, fit, least-squares refinement) and the mode 6/5/1 encoders
// of zap_tex.h:419-748.
#pragma once
#include "zt_common.h"
#include "zt_pca.h"

namespace {

// BC7 format-specification tables (2-subset partitions as bit masks, second-subset anchors, weights)
__constant__ uint16_t kP2[64] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8, 0xFF00, 0xFFF0, 0xF000,
    0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110, 0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C,
    0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696, 0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660,
    0x0272, 0x04E4, 0x4E40, 0x2720, 0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22 };
__constant__ uint8_t kA2[64] = {
    15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2, 8, 2, 2, 8, 8, 15, 2, 8, 2, 2, 8, 8, 2, 2,
    15, 15, 6, 8, 2, 8, 15, 15, 2, 8, 2, 2, 2, 15, 15, 6, 6, 2, 6, 8, 15, 15, 2, 2, 15, 15, 15, 15, 15, 2, 2, 15 };

// BC7 weights: W2 {0,21,43,64}, W3 {0,9,..,64}, W4 {0,4,..,64} are round(64 j / (2^IB - 1))
template <int IB> ZT_DEV int wgt(int j) {
    constexpr int D = (1 << IB) - 1;
    return (j * 64 + D / 2) / D;
}
ZT_DEV int expand_bits(int v, int n) { v <<= 8 - n; return v | v >> n; }

// channel-masked pixel: NC channels starting at C0, packed from byte 0
template <int NC, int C0> ZT_DEV uint32_t sub_px(uint32_t p) {
    uint32_t v = p >> (8 * C0);
    return v & (uint32_t)((1ull << (8 * NC)) - 1ull);
}

// ---- p-bit quantiser for one channel (zap_tex.h:426-438); N = bits including the p-bit, t = v / 255 * (2^N - 1)
// (shared by both p-bits).  The reference tries g-1, g, g+1, each clamped to [0, mx], and keeps the first of the
// smallest errors.  Here g is clamped to [1, mx-1] first, so the three candidates need no clamp: where that moves
// the set (g < 1 or g > mx-1), the candidate it adds lies beyond the ones the reference tries and its error is
// strictly larger, so the choice is the same.  A candidate's key is err << 2 | dt; the smallest key is the
// reference's choice.  Checked exhaustively against the reference's arithmetic for every float |v| <= 2048 and
// N = 5..8 (tests/quant_exhaustive.c); endpoint values here stay within [-1020, 1275] (PCA: the mean plus a
// projection of at most 4 * 255 along an axis whose largest component is 1; least squares: clamped to [0, 255]).
template <int N> ZT_DEV int quant_pbit1(float v, float t, int p, int &q) {
    const int mx = (1 << (N - 1)) - 1, g = clampi((int)((t - (float)p) * 0.5f + 0.5f), 1, mx - 1), vm = g << 1 | p;
    uint32_t k[3];
    if (N == 8) {   // x = tq << 1 | p: the neighbours are 2 apart, exactly, in float too
        const float xm = (float)vm, xs[3] = { xm - 2.0f, xm, xm + 2.0f };
#pragma unroll
        for (int dt = 0; dt < 3; dt++) { const float d = xs[dt] - v; k[dt] = (uint32_t)(int)(d * d) << 2 | (uint32_t)dt; }
    } else {
#pragma unroll
        for (int dt = 0; dt < 3; dt++) {
            const float d = (float)expand_bits(vm + 2 * (dt - 1), N) - v;
            k[dt] = (uint32_t)(int)(d * d) << 2 | (uint32_t)dt;
        }
    }
    const uint32_t best = min(min(k[0], k[1]), k[2]);
    q = g + (int)(best & 3u) - 1;
    return (int)(best >> 2);
}
template <int N> ZT_DEV float quant_scale(float v) { return div255(v) * (float)((1 << N) - 1); }

// ---- argmin over a palette of NP packed entries (ties -> lowest index).  key_j = 32 p.pal_j - 16 |pal_j|^2 +
// (15 - j) = 16 (|p|^2 - |p - pal_j|^2) + (15 - j), one v_lshl_add per entry (|key| < 2^24): the largest key is
// the reference's choice, (key & 15) ^ 15 = j and key >> 4 = |p|^2 - SSE.  Over a subset, sum (key >> 4) =
// (sum key - sum (key & 15)) / 16 exactly, so a fit's SSE is spp minus that, with spp = sum |p|^2 of the subset.
template <int NP> ZT_DEV int best_key(uint32_t p, const uint32_t pal[NP], const int K[NP]) {
    int best = (int)(dot4(p, pal[0]) << 5) + K[0];
#pragma unroll
    for (int j = 1; j < NP; j++) best = max(best, (int)(dot4(p, pal[j]) << 5) + K[j]);
    return best;
}
// the same key for a one-channel palette (entries are plain values, v a plain byte): 32 v p_j + K_j as one 24-bit
// multiply-add per entry instead of a v_dot4_u32_u8 and a shift-add
template <int NP> ZT_DEV int best_key1(int v, const int P32[NP], const int K[NP]) {
    int best = __mul24(v, P32[0]) + K[0];
#pragma unroll
    for (int j = 1; j < NP; j++) best = max(best, __mul24(v, P32[j]) + K[j]);
    return best;
}
template <int NP> ZT_DEV void palette_keys(const uint32_t pal[NP], int K[NP]) {
#pragma unroll
    for (int j = 0; j < NP; j++) K[j] = no_fold(15 - j - 16 * (int)dot4(pal[j], pal[j]));
}
// 0xF in the nibble of every pixel of `mask`
ZT_DEV uint64_t subset_nibbles(uint32_t mask) {
    uint64_t m = mask;
    m = (m | m << 24) & 0x000000FF000000FFull;
    m = (m | m << 12) & 0x000F000F000F000Full;
    m = (m | m << 6) & 0x0303030303030303ull;
    m = (m | m << 3) & 0x1111111111111111ull;
    return m * 15u;
}
// sum of the 16 4-bit fields of b / of the 16 2-bit fields of w
ZT_DEV int nibble_sum(uint64_t b) {
    const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32), m = 0x0F0F0F0Fu, one = 0x01010101u;
    return (int)dot4acc((hi >> 4) & m, one, dot4acc(hi & m, one, dot4acc((lo >> 4) & m, one, dot4(lo & m, one))));
}
ZT_DEV int pair_sum(uint32_t w) {
    const uint32_t x = (w & 0x33333333u) + ((w >> 2) & 0x33333333u);   // nibbles of at most 6
    return (int)dot4acc((x >> 4) & 0x0F0F0F0Fu, 0x01010101u, dot4(x & 0x0F0F0F0Fu, 0x01010101u));
}
// sum of |p|^2 over the subset (channels [C0, C0 + NC))
template <int NC, int C0> ZT_DEV int subset_sq(const uint32_t px[16], uint32_t mask) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (!((mask >> k) & 1)) continue;
        const uint32_t p = sub_px<NC, C0>(px[k]);
        s += (int)dot4(p, p);
    }
    return s;
}

// ---- palette of an endpoint pair: entry j = ((64 - w_j) e0 + w_j e1 + 32) >> 6 per channel (zap_tex.h:477-478,
// :539-541, :599), NC channels packed from byte 0.  Two channels share a 32-bit word in 16-bit lanes; with both
// weights scaled by 4 a lane's sum, 4 * ((64 - w) e0 + w e1 + 32) <= 65408, cannot carry into the next lane and
// holds the entry in its high byte, so one v_perm gathers an entry's four bytes.
template <int NC, int IB> ZT_DEV void build_palette(const int e0[], const int e1[], uint32_t pal[1 << IB]) {
    constexpr int NP = 1 << IB;
    if (NC == 1) {
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int w = wgt<IB>(j);
            pal[j] = (__umul24(64 - w, e0[0]) + __umul24(w, e1[0]) + 32) >> 6;
        }
        return;
    }
    const uint32_t a_rb = (uint32_t)e0[0] | (uint32_t)e0[NC > 2 ? 2 : 0] << 16 * (NC > 2);
    const uint32_t a_ga = (uint32_t)e0[1] | (NC > 3 ? (uint32_t)e0[NC - 1] << 16 : 0u);
    const uint32_t b_rb = (uint32_t)e1[0] | (uint32_t)e1[NC > 2 ? 2 : 0] << 16 * (NC > 2);
    const uint32_t b_ga = (uint32_t)e1[1] | (NC > 3 ? (uint32_t)e1[NC - 1] << 16 : 0u);
#pragma unroll
    for (int j = 0; j < NP; j++) {
        const uint32_t w = (uint32_t)wgt<IB>(j);
        const uint32_t rb = __umul24(b_rb, 4u * w) + __umul24(a_rb, 4u * (64u - w)) + 0x00800080u;
        const uint32_t ga = __umul24(b_ga, 4u * w) + __umul24(a_ga, 4u * (64u - w)) + 0x00800080u;
        pal[j] = perm(ga, rb, 0x07030501u);   // with NC = 3 the alpha lane holds 0x80 >> 8 = 0
    }
}

// ---- exact least-squares sums (zap_tex.h:497-501, :627-632).  A BC7 weight is w = k/64 with k an integer in
// 0..64 and pixels are 8-bit integers, so every product the reference accumulates in float is exact and every
// partial sum is a multiple of 2^-12 (aa, ab, bb; at most 16 * 4096 units) or of 2^-6 (ax, bx; at most
// 16 * 64 * 255 units) below 2^24 units: the float sums are exact whatever their order, and equal
//   aa = sum (64-k)^2 / 4096,  ab = sum (64-k) k / 4096,  bb = sum k^2 / 4096,
//   ax_c = sum (64-k) p_c / 64 = (64 P_c - S_c) / 64,  bx_c = S_c / 64,  with P_c = sum p_c and S_c = sum k p_c.
// The integer sums take four pixels per v_dot4_u32_u8: channel words (byte t = channel c of pixel 4g+t) against
// weight words (byte t = k of pixel 4g+t; 0 for pixels outside the subset).
struct ChanWords { uint32_t w[4][4]; };   // w[g][c]
ZT_DEV void chan_words(const uint32_t px[16], ChanWords &T) {
#pragma unroll
    for (int g = 0; g < 4; g++) {   // 4x4 byte transpose of pixels 4g..4g+3
        const uint32_t a = px[4 * g], b = px[4 * g + 1], c = px[4 * g + 2], d = px[4 * g + 3];
        const uint32_t rg01 = perm(b, a, 0x05010400u), ba01 = perm(b, a, 0x07030602u);
        const uint32_t rg23 = perm(d, c, 0x05010400u), ba23 = perm(d, c, 0x07030602u);
        T.w[g][0] = perm(rg23, rg01, 0x05040100u); T.w[g][1] = perm(rg23, rg01, 0x07060302u);
        T.w[g][2] = perm(ba23, ba01, 0x05040100u); T.w[g][3] = perm(ba23, ba01, 0x07060302u);
    }
}
// P_c over the subset `mask` (all 16 pixels when FULL)
template <int NC, int C0, bool FULL> ZT_DEV void chan_sums(const ChanWords &T, uint32_t mask, int P[NC]) {
#pragma unroll
    for (int c = 0; c < NC; c++) P[c] = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        // one byte per pixel: bit t of the group's mask nibble moves to bit 8t (the shifted copies never overlap)
        const uint32_t sel = FULL ? 0x01010101u : (__umul24((mask >> (4 * g)) & 15u, 0x204081u) & 0x01010101u);
#pragma unroll
        for (int c = 0; c < NC; c++) P[c] = (int)dot4acc(T.w[g][C0 + c], sel, (uint32_t)P[c]);
    }
}
// index bytes of pixels 4g..4g+3 from 4-bit fields (pixel k at bit 4k) or from 2-bit fields (pixel k at bit 2k)
ZT_DEV uint32_t idx_bytes4(uint64_t idx, int g) {
    const uint32_t u = perm((uint32_t)(idx >> 32), (uint32_t)idx, 0x0C000C00u | (2u * g + 1u) << 16 | 2u * g);
    return (u & 0x000F000Fu) | ((u << 4) & 0x0F000F00u);
}
ZT_DEV uint32_t idx_bytes2(uint32_t idx, int g) {
    const uint32_t b = (idx >> (8 * g)) & 255u, u = b | b << 12;
    return (u | u << 6) & 0x03030303u;
}
// k = round(64 j / (2^IB - 1)) of each index byte j: 4j + (j+2)/4 (IB 4), 9j + j/4 (IB 3), 21j + j/2 (IB 2)
template <int IB> ZT_DEV uint32_t weight_bytes(uint32_t J) {
    if (IB == 4) return (J << 2) + (((J + 0x02020202u) >> 2) & 0x3F3F3F3Fu);
    if (IB == 3) return (J << 3) + J + ((J >> 2) & 0x01010101u);
    return (J << 4) + (J << 2) + J + ((J >> 1) & 0x01010101u);
}
template <int NC, int C0>
ZT_DEV void ls_sums(const ChanWords &T, const uint32_t KW[4], const int P[NC], int n, float &aa, float &ab, float &bb,
                    float sa[NC], float sb[NC]) {
    uint32_t k1 = 0, k2 = 0, s[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) s[c] = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        k1 = dot4acc(KW[g], 0x01010101u, k1);
        k2 = dot4acc(KW[g], KW[g], k2);
#pragma unroll
        for (int c = 0; c < NC; c++) s[c] = dot4acc(T.w[g][C0 + c], KW[g], s[c]);
    }
    aa = (float)(4096 * n - 128 * (int)k1 + (int)k2) * (1.0f / 4096.0f);
    ab = (float)(64 * (int)k1 - (int)k2) * (1.0f / 4096.0f);
    bb = (float)k2 * (1.0f / 4096.0f);
#pragma unroll
    for (int c = 0; c < NC; c++) { sa[c] = (float)(64 * P[c] - (int)s[c]) * (1.0f / 64.0f); sb[c] = (float)s[c] * (1.0f / 64.0f); }
}

// ---- one subset: quantised endpoints + indices (zap_tex.h:466-486).  idx: 4 bits per pixel at 4*i.
// The quantised endpoints with the best p-bit pair (zap_tex.h:468-475)
template <int NC, int N, bool SHARED> ZT_DEV void quant_pair(const float lo[4], const float hi[4], int q[2][4], int pb[2]) {
    int ql[2][NC], qh[2][NC], el[2], eh[2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
        el[p] = 0; eh[p] = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            el[p] += quant_pbit1<N>(lo[c], quant_scale<N>(lo[c]), p, ql[p][c]);
            eh[p] += quant_pbit1<N>(hi[c], quant_scale<N>(hi[c]), p, qh[p][c]);
        }
    }
    int best = 1 << 30, b0 = 0, b1 = 0;
#pragma unroll
    for (int p0 = 0; p0 < 2; p0++)
#pragma unroll
        for (int p1 = 0; p1 < 2; p1++) {
            if (SHARED && p0 != p1) continue;
            int e = el[p0] + eh[p1];
            if (e < best) { best = e; b0 = p0; b1 = p1; }
        }
    pb[0] = b0; pb[1] = b1;
#pragma unroll
    for (int c = 0; c < NC; c++) { q[0][c] = b0 ? ql[1][c] : ql[0][c]; q[1][c] = b1 ? qh[1][c] : qh[0][c]; }
}
// SSE and indices of quantised endpoints over the subset: the palette and its argmin (zap_tex.h:476-485)
template <int NC, int N, int IB>
ZT_DEV int fit_quant(const uint32_t px[16], uint32_t mask, int spp, uint64_t snib, const int q[2][4], const int pb[2], uint64_t &idx) {
    int e0[NC], e1[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) { e0[c] = expand_bits(q[0][c] << 1 | pb[0], N); e1[c] = expand_bits(q[1][c] << 1 | pb[1], N); }
    constexpr int NP = 1 << IB;
    uint32_t pal[NP];
    int K[NP];
    build_palette<NC, IB>(e0, e1, pal);
    palette_keys<NP>(pal, K);
    int sk = 0;
    uint64_t bits = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (!((mask >> k) & 1)) continue;
        const int key = best_key<NP>(sub_px<NC, 0>(px[k]), pal, K);
        sk += key;
        bits |= (uint64_t)(key & 15) << (4 * k);
    }
    idx = bits ^ snib;
    return spp - ((sk - nibble_sum(bits)) >> 4);
}
template <int NC, int N, int IB, bool SHARED>
ZT_DEV int fit_subset(const uint32_t px[16], uint32_t mask, int spp, uint64_t snib, const float lo[4], const float hi[4], int q[2][4],
                      int pb[2], uint64_t &idx) {
    quant_pair<NC, N, SHARED>(lo, hi, q, pb);
    return fit_quant<NC, N, IB>(px, mask, spp, snib, q, pb, idx);
}

// ---- one least-squares step (zap_tex.h:495-507): endpoints for the indices idx, clamped to [0, 255]; false when
// the system is singular (|det| < 1e-6), which ends the reference's refinement.  RT: the channel words again from
// the pixels, so they are not live across the fits.
template <int NC, int IB, bool RT, typename PX>
ZT_DEV bool ls_step(PX *px, const ChanWords &T, const int P[NC], int cnt, uint64_t idx, float nlo[4], float nhi[4]) {
    // pixels outside the subset have index 0, hence weight 0
    uint32_t KW[4];
#pragma unroll
    for (int g = 0; g < 4; g++) KW[g] = weight_bytes<IB>(idx_bytes4(idx, g));
    float aa, ab, bb, sa[NC], sb[NC];
    if constexpr (RT) {
        uint32_t pr[16];
#pragma unroll
        for (int k = 0; k < 16; k++) pr[k] = px[k];
        refresh16(pr);
        ChanWords T2;
        chan_words(pr, T2);
        ls_sums<NC, 0>(T2, KW, P, cnt, aa, ab, bb, sa, sb);
    } else ls_sums<NC, 0>(T, KW, P, cnt, aa, ab, bb, sa, sb);
    float det = aa * bb - ab * ab;
    if (det < 1e-6f && det > -1e-6f) return false;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c < NC) {
            float l = (sa[c] * bb - sb[c] * ab) / det, h = (sb[c] * aa - sa[c] * ab) / det;
            nlo[c] = l < 0 ? 0 : l > 255 ? 255 : l; nhi[c] = h < 0 ? 0 : h > 255 ? 255 : h;
        } else { nlo[c] = 255.0f; nhi[c] = 255.0f; }
    }
    return true;
}

// ---- PCA start, fit, at most 2 least-squares refinements (zap_tex.h:489-515)
// R: refresh the (then non-const) pixel registers between the PCA passes (pca's REFRESH)
template <int NC, int N, int IB, bool SHARED, bool FULL, bool R = false, bool RT = false, typename PX>
ZT_DEV int encode_subset(PX *px, const ChanWords &T, uint32_t mask, int cnt, int q[2][4], int pb[2], uint64_t &idx) {
    int P[NC];
    chan_sums<NC, 0, FULL>(T, mask, P);
    float lo[4], hi[4];
    pca<NC, FULL, R>(px, mask, cnt, lo, hi, P);
    const int spp = subset_sq<NC, 0>(px, mask);
    const uint64_t snib = FULL ? ~0ull : subset_nibbles(mask);
    int err = fit_subset<NC, N, IB, SHARED>(px, mask, spp, snib, lo, hi, q, pb, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float nlo[4], nhi[4];
        if (!ls_step<NC, IB, RT>(px, T, P, cnt, idx, nlo, nhi)) break;
        int nq[2][4], np[2];
        uint64_t ni;
        int ne = fit_subset<NC, N, IB, SHARED>(px, mask, spp, snib, nlo, nhi, nq, np, ni);
        if (ne >= err) break;
        err = ne; pb[0] = np[0]; pb[1] = np[1]; idx = ni;
#pragma unroll
        for (int c = 0; c < NC; c++) { q[0][c] = nq[0][c]; q[1][c] = nq[1][c]; }
    }
    return err;
}

// ---- mode 6 (zap_tex.h:582-593, pack :520-527)
// the block: endpoints, p-bits and indices as the reference stores them (anchor index MSB clear)
ZT_DEV void pack6(int q[2][4], int pb[2], uint64_t idx, uint64_t &lo, uint64_t &hi) {
    if (idx & 8) {  // anchor MSB set: swap endpoints and p-bits, invert indices
#pragma unroll
        for (int c = 0; c < 4; c++) { int t = q[0][c]; q[0][c] = q[1][c]; q[1][c] = t; }
        int t = pb[0]; pb[0] = pb[1]; pb[1] = t;
        idx ^= 0xFFFFFFFFFFFFFFFFull;
    }
    uint64_t l = 0x40;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) l |= (uint64_t)q[e][c] << (7 + 7 * (2 * c + e));
    l |= (uint64_t)pb[0] << 63;
    uint64_t field = (idx & 7) | ((idx >> 1) & ~7ull);  // pixel 0 keeps 3 bits
    lo = l;
    hi = (uint64_t)pb[1] | field << 1;
}
ZT_DEV int mode6(const uint32_t px[16], uint64_t &lo, uint64_t &hi) {
    int q[2][4], pb[2];
    uint64_t idx;
    ChanWords T;
    chan_words(px, T);
    int err = encode_subset<4, 8, 4, false, true, false, true>(px, T, 0xFFFFu, 16, q, pb, idx);
    pack6(q, pb, idx, lo, hi);
    return err;
}
// Mode 6 in passes.  A refinement step whose least-squares endpoints quantise to the endpoints and p-bits the
// block already has refits nothing: the same palette gives the same SSE, which the reference's loop rejects
// (ne >= err) and stops.  So each pass ends every block it can, and hands on only the blocks whose next fit has new
// endpoints.  q and p-bits travel as one word: q[e][c] at 7 (2c + e), p-bits at 56 and 57.
ZT_DEV uint64_t qp_word(const int q[2][4], const int pb[2]) {
    uint64_t w = (uint64_t)pb[0] << 56 | (uint64_t)pb[1] << 57;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) w |= (uint64_t)q[e][c] << (7 * (2 * c + e));
    return w;
}
ZT_DEV void qp_unword(uint64_t w, int q[2][4], int pb[2]) {
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) q[e][c] = (int)((w >> (7 * (2 * c + e))) & 127u);
    pb[0] = (int)((w >> 56) & 1u); pb[1] = (int)((w >> 57) & 1u);
}
// the next refinement step after a fit with (q, pb, idx) and SSE err: true with its quantised endpoints in nqp when
// they differ from (q, pb) (the step's fit may improve), false when the loop ends here
ZT_DEV bool mode6_next(const uint32_t px[16], const ChanWords &T, const int P[4], int err, const int q[2][4], const int pb[2],
                       uint64_t idx, uint64_t &nqp) {
    if (err <= 0) return false;
    float nlo[4], nhi[4];
    if (!ls_step<4, 4, true>(px, T, P, 16, idx, nlo, nhi)) return false;
    int nq[2][4], np[2];
    quant_pair<4, 8, false>(nlo, nhi, nq, np);
    nqp = qp_word(nq, np);
    return nqp != qp_word(q, pb);
}
// first pass: the PCA start, its fit and the first step's endpoints; returns the SSE, the block in (lo, hi), and
// whether the block goes on (nqp = the endpoints of its first refit)
ZT_DEV int mode6_first(const uint32_t px[16], uint64_t &lo, uint64_t &hi, bool &more, uint64_t &nqp) {
    ChanWords T;
    chan_words(px, T);
    int P[4];
    chan_sums<4, 0, true>(T, 0xFFFFu, P);
    float plo[4], phi[4];
    pca<4, true, false>(px, 0xFFFFu, 16, plo, phi, P);
    const int spp = subset_sq<4, 0>(px, 0xFFFFu);
    int q[2][4], pb[2];
    uint64_t idx;
    const int err = fit_subset<4, 8, 4, false>(px, 0xFFFFu, spp, ~0ull, plo, phi, q, pb, idx);
    more = mode6_next(px, T, P, err, q, pb, idx, nqp);
    pack6(q, pb, idx, lo, hi);
    return err;
}
// a later pass: the fit of the endpoints nqp against the block's SSE err.  An improvement is written to (lo, hi)
// (improved = true) and, unless LAST, the next step's endpoints decide whether the block goes on (more = true, nqp =
// those endpoints).
template <bool LAST>
ZT_DEV int mode6_refit(const uint32_t px[16], int err, uint64_t &nqp, uint64_t &lo, uint64_t &hi, bool &improved, bool &more) {
    int q[2][4], pb[2];
    qp_unword(nqp, q, pb);
    uint64_t idx;
    const int ne = fit_quant<4, 8, 4>(px, 0xFFFFu, subset_sq<4, 0>(px, 0xFFFFu), ~0ull, q, pb, idx);
    improved = ne < err;
    more = false;
    if (!improved) return err;
    if (!LAST) {
        ChanWords T;
        chan_words(px, T);
        int P[4];
        chan_sums<4, 0, true>(T, 0xFFFFu, P);
        more = mode6_next(px, T, P, ne, q, pb, idx, nqp);
    }
    pack6(q, pb, idx, lo, hi);
    return ne;
}

// ---- mode 5 pieces (zap_tex.h:597-657); NC = 3 (colour, 7 bits) or 1 (alpha, 8 bits)
// plain n-bit quantiser (zap_tex.h:613-624): candidates and keys as in quant_pbit1 (checked exhaustively for
// every float |v| <= 2048 and n = 4..8)
template <int NC> ZT_DEV void quant_plain(const float *v, int n, int q[], int e[]) {
    const int mx = (1 << n) - 1;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int g = clampi((int)(div255(v[c] * (float)mx) + 0.5f), 1, mx - 1);
        uint32_t k[3];
        if (n == 8) {
            const float xm = (float)g, xs[3] = { xm - 1.0f, xm, xm + 1.0f };
#pragma unroll
            for (int dt = 0; dt < 3; dt++) k[dt] = (uint32_t)(int)fabsf(xs[dt] - v[c]) << 2 | (uint32_t)dt;
        } else {
#pragma unroll
            for (int dt = 0; dt < 3; dt++) k[dt] = (uint32_t)(int)fabsf((float)expand_bits(g + dt - 1, n) - v[c]) << 2 | (uint32_t)dt;
        }
        const uint32_t best = min(min(k[0], k[1]), k[2]);
        q[c] = g + (int)(best & 3u) - 1; e[c] = expand_bits(q[c], n);
    }
}
template <int NC, int C0> ZT_DEV int fit_2bit(const uint32_t r[16], int spp, const int e0[], const int e1[], uint32_t &idx) {
    uint32_t pal[4];
    int K[4];
    build_palette<NC, 2>(e0, e1, pal);
    palette_keys<4>(pal, K);
    int sk = 0, P32[4];
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) P32[j] = (int)pal[j] << 5;   // used when NC == 1
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int key = NC == 1 ? best_key1<4>((int)sub_px<NC, C0>(r[i]), P32, K) : best_key<4>(sub_px<NC, C0>(r[i]), pal, K);   // low nibble 15 - j, j < 4
        sk += key;
        bits |= (uint32_t)(key & 3) << (2 * i);
    }
    idx = ~bits;                                                      // (key & 3) ^ 3 = j
    return spp - ((sk - 12 * 16 - pair_sum(bits)) >> 4);              // key & 15 = 12 + (key & 3)
}
template <int NC, int C0> ZT_DEV bool lsq_2bit(const ChanWords &T, const int P[4], uint32_t idx, float lo[], float hi[]) {
    uint32_t KW[4];
#pragma unroll
    for (int g = 0; g < 4; g++) KW[g] = weight_bytes<2>(idx_bytes2(idx, g));
    float aa, ab, bb, sa[NC], sb[NC];
    ls_sums<NC, C0>(T, KW, P + C0, 16, aa, ab, bb, sa, sb);
    float det = aa * bb - ab * ab;
    if (det < 1e-6f && det > -1e-6f) return false;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        float l = (sa[c] * bb - sb[c] * ab) / det, h = (sb[c] * aa - sa[c] * ab) / det;
        lo[c] = l < 0 ? 0 : l > 255 ? 255 : l; hi[c] = h < 0 ? 0 : h > 255 ? 255 : h;
    }
    return true;
}
template <int NC, int C0>
ZT_DEV int part_mode5(const uint32_t r[16], const ChanWords &T, const int P[4], int n, const float lo[], const float hi[], int q[2][3],
                      uint32_t &idx) {
    int e0[NC], e1[NC];
    quant_plain<NC>(lo, n, q[0], e0); quant_plain<NC>(hi, n, q[1], e1);
    const int spp = subset_sq<NC, C0>(r, 0xFFFFu);
    int err = fit_2bit<NC, C0>(r, spp, e0, e1, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float nlo[NC], nhi[NC];
        if (!lsq_2bit<NC, C0>(T, P, idx, nlo, nhi)) break;
        int tq[2][3], te0[NC], te1[NC];
        quant_plain<NC>(nlo, n, tq[0], te0); quant_plain<NC>(nhi, n, tq[1], te1);
        uint32_t ti;
        int e = fit_2bit<NC, C0>(r, spp, te0, te1, ti);
        if (e >= err) break;
        err = e; idx = ti;
#pragma unroll
        for (int c = 0; c < NC; c++) { q[0][c] = tq[0][c]; q[1][c] = tq[1][c]; }
    }
    return err;
}
// swap alpha with channel rot-1 (rot 0 = identity)
ZT_DEV uint32_t rotate_px(uint32_t p, int rot) {
    if (!rot) return p;
    int s = 8 * (rot - 1);
    uint32_t a = p >> 24, c = (p >> s) & 255u;
    p &= ~(0xFFu << s) & 0x00FFFFFFu;
    return p | a << s | c << 24;
}
ZT_DEV int mode5_rotation(const uint32_t px[16], int rot, int cq[2][3], int aq[2][3], uint32_t &ci, uint32_t &ai) {
    uint32_t r[16];
#pragma unroll
    for (int i = 0; i < 16; i++) r[i] = rotate_px(px[i], rot);
    ChanWords T;
    chan_words(r, T);
    int P[4];
    chan_sums<4, 0, true>(T, 0xFFFFu, P);
    float lo[4], hi[4], alo[1] = { 255.0f }, ahi[1] = { 0.0f };
    pca<3, true>(r, 0xFFFFu, 16, lo, hi, P);
    int err = part_mode5<3, 0>(r, T, P, 7, lo, hi, cq, ci);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float a = (float)(r[i] >> 24);
        if (a < alo[0]) alo[0] = a;
        if (a > ahi[0]) ahi[0] = a;
    }
    err += part_mode5<1, 3>(r, T, P, 8, alo, ahi, aq, ai);
    return err;
}
ZT_DEV void pack_mode5(int rot, int cq[2][3], int aq[2][3], uint32_t ci, uint32_t ai, uint64_t &lo, uint64_t &hi) {
    if (ci & 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) { int t = cq[0][c]; cq[0][c] = cq[1][c]; cq[1][c] = t; }
        ci ^= 0xFFFFFFFFu;
    }
    if (ai & 2) { int t = aq[0][0]; aq[0][0] = aq[1][0]; aq[1][0] = t; ai ^= 0xFFFFFFFFu; }
    Bits128 b{ 0x20ull | (uint64_t)rot << 6, 0 };
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) b.put(8 + 7 * (2 * c + e), (uint64_t)cq[e][c], 7);
    b.put(50, (uint64_t)aq[0][0], 8); b.put(58, (uint64_t)aq[1][0], 8);
    uint64_t cf = (ci & 1) | ((ci >> 1) & ~1u), af = (ai & 1) | ((ai >> 1) & ~1u);  // pixel 0 keeps 1 bit
    b.put(66, cf, 31); b.put(97, af, 31);
    lo = b.lo; hi = b.hi;
}

// ---- mode-1 partition estimate (zap_tex.h:722-748)
ZT_DEV float mode1_estimate(const uint32_t px[16], const float tot[9], int part) {
    float s1[9];
    int n1 = 0;
    const uint32_t m = kP2[part];
#pragma unroll
    for (int j = 0; j < 9; j++) s1[j] = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (!((m >> i) & 1)) continue;
        float r = chanf(px[i], 0), g = chanf(px[i], 1), b = chanf(px[i], 2);
        float v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
        n1++;
#pragma unroll
        for (int j = 0; j < 9; j++) s1[j] += v[j];
    }
    float err = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        float v[9];
#pragma unroll
        for (int j = 0; j < 9; j++) v[j] = k ? s1[j] : tot[j] - s1[j];
        float n = (float)(k ? n1 : 16 - n1);
        if (n < 2) continue;
        float S[3][3];
        S[0][0] = v[3] - v[0] * v[0] / n; S[0][1] = v[6] - v[0] * v[1] / n; S[0][2] = v[7] - v[0] * v[2] / n;
        S[1][1] = v[4] - v[1] * v[1] / n; S[1][2] = v[8] - v[1] * v[2] / n; S[2][2] = v[5] - v[2] * v[2] / n;
        S[1][0] = S[0][1]; S[2][0] = S[0][2]; S[2][1] = S[1][2];
        float tr = S[0][0] + S[1][1] + S[2][2], ax[3];
        int r0 = S[0][0] >= S[1][1] && S[0][0] >= S[2][2] ? 0 : S[1][1] >= S[2][2] ? 1 : 2;
#pragma unroll
        for (int c = 0; c < 3; c++) ax[c] = r0 == 0 ? S[0][c] : r0 == 1 ? S[1][c] : S[2][c];
        for (int it = 0; it < 3; it++) {
            float t[3], mx = 1e-12f;
#pragma unroll
            for (int c = 0; c < 3; c++) { t[c] = S[c][0] * ax[0] + S[c][1] * ax[1] + S[c][2] * ax[2]; mx = t[c] > mx ? t[c] : -t[c] > mx ? -t[c] : mx; }
#pragma unroll
            for (int c = 0; c < 3; c++) ax[c] = t[c] / mx;
        }
        float vv = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], vsv = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) vsv += ax[c] * (S[c][0] * ax[0] + S[c][1] * ax[1] + S[c][2] * ax[2]);
        float lam = vv > 1e-12f ? vsv / vv : 0;
        err += tr - lam * (48.0f / 49.0f);
    }
    return err;
}

// ---- mode 1 fit of one partition (zap_tex.h:687-708); returns SSE, optionally packs
ZT_DEV int mode1(const uint32_t px[16], int part, bool pack, uint64_t &lo, uint64_t &hi) {
    const uint32_t m1 = kP2[part], m0 = ~m1 & 0xFFFFu;
    int q[2][2][4], pb[2][2];
    uint64_t idx[2];
    ChanWords T;
    chan_words(px, T);
    int err = encode_subset<3, 7, 3, true, false>(px, T, m0, __popc(m0), q[0], pb[0], idx[0]);
    err += encode_subset<3, 7, 3, true, false>(px, T, m1, __popc(m1), q[1], pb[1], idx[1]);
    if (!pack) return err;
    const int anc = kA2[part];
    uint64_t full = idx[0] | idx[1];  // disjoint pixel sets, 4 bits per pixel
    // anchor fix per subset: swap its endpoints and invert its indices (7 - i = i ^ 7 on the subset's pixels)
    uint64_t sel0 = 0, sel1 = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) { if ((m0 >> i) & 1) sel0 |= 7ull << (4 * i); else sel1 |= 7ull << (4 * i); }
    if (full & 4) {
#pragma unroll
        for (int c = 0; c < 3; c++) { int t = q[0][0][c]; q[0][0][c] = q[0][1][c]; q[0][1][c] = t; }
        full ^= sel0;
    }
    if ((full >> (4 * anc)) & 4) {
#pragma unroll
        for (int c = 0; c < 3; c++) { int t = q[1][0][c]; q[1][0][c] = q[1][1][c]; q[1][1][c] = t; }
        full ^= sel1;
    }
    Bits128 b{ 2ull | (uint64_t)part << 2, 0 };
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int e = 0; e < 2; e++) b.put(8 + 6 * (4 * c + 2 * s + e), (uint64_t)q[s][e][c], 6);
    b.put(80, (uint64_t)pb[0][0], 1); b.put(81, (uint64_t)pb[1][0], 1);
    int pos = 82;
    for (int i = 0; i < 16; i++) {
        int nb = 3 - (i == 0 || i == anc);
        b.put(pos, (full >> (4 * i)) & 7, nb);
        pos += nb;
    }
    lo = b.lo; hi = b.hi;
    return err;
}

}  // namespace
