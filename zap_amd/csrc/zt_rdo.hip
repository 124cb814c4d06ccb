// zt_rdo.hip — rate-distortion encoding (rdo > 0) for gfx950: zap_tex.h:223-306 (rate model, BC1 and BC4 blocks),
// :520-579 (mode-6 refits), :786-815 (BC7), driven like :1474-1506.
//
// The reference walks the blocks in raster order and, per block, weighs its plain encoding against 48 candidates
// built from the previous 16 blocks' final bytes (3 kinds each), keeping the smallest J = SSE + lambda * rate.  The
// rate of a field depends on those 16 blocks and on the match distances the previous block chose, so every block
// waits on the one before it.  Here one wave owns one stream (an image's BC1, BC4 or BC7 blocks; BC3 and BC5
// images hold two streams, one per 8-byte half, which never read each other) and walks it in the reference's
// order: per block, lanes 0..47 build the 48 candidates in parallel (lane = 3 (j - 1) + kind, predecessor j),
// lane 48 re-scores the plain block the encoder already wrote in place, every lane prices its candidate against
// the last 16 final blocks held in LDS, and one wave-wide minimum of (J, order) picks the block.  The reference
// skips a candidate whose SSE alone reaches the best J so far; rates are never negative, so such a candidate can
// never win, and the first smallest J in candidate order is exactly the block the reference keeps.
#include "zt_bc7_core.h"
#include "zt_bc7_dec.h"
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

// this stream's 8 or 16 bytes of block b in a block array laid out like the image's output
ZT_DEV uint4 read_block(const zt_rdo_stream &st, const uint8_t *base, uint32_t b) {
    const uint8_t *p = base + (uint64_t)b * st.bs + st.off;
    if (st.kind == ZT_RDO_BC7) return *(const uint4 *)p;
    const uint2 v = *(const uint2 *)p;
    return make_uint4(v.x, v.y, 0u, 0u);
}
ZT_DEV void write_block(const zt_rdo_stream &st, uint8_t *base, uint32_t b, uint4 v) {
    uint8_t *p = base + (uint64_t)b * st.bs + st.off;
    if (st.kind == ZT_RDO_BC7) *(uint4 *)p = v;
    else *(uint2 *)p = make_uint2(v.x, v.y);
}

// ---- one block's pixels and its plain bytes: lanes 0..15 load a pixel each (edge-clamped like
// zap_tex.h:1482-1488), every lane the plain block from the stream's copy of the plain encoding
ZT_DEV void load_block(const zt_rdo_stream &st, uint32_t b, int lane, uint32_t &pix, uint4 &plain) {
    const int bx = (int)(b % (uint32_t)st.bw), by = (int)(b / (uint32_t)st.bw);
    if (lane < 16) {
        const int x = min(bx * 4 + (lane & 3), st.w - 1), y = min(by * 4 + (lane >> 2), st.h - 1);
        pix = *(const uint32_t *)(st.src + (uint64_t)y * st.stride + (uint64_t)x * 4);
    }
    plain = read_block(st, st.plain, b);
}

// the two fields a stream prices (zap_tex.h:245-250, :281-285, :788-790) as integers, and their sizes in bytes
template <int K> ZT_DEV uint64_t field0(uint4 b) {
    if (K == ZT_RDO_BC1) return b.x;                                       // endpoints, bytes 0-3
    if (K == ZT_RDO_BC4) return b.x & 0xFFFFu;                             // a0 a1, bytes 0-1
    return (uint64_t)b.x | (uint64_t)b.y << 32;                            // BC7 bytes 0-7
}
template <int K> ZT_DEV uint64_t field1(uint4 b) {
    if (K == ZT_RDO_BC1) return b.y;                                       // indices, bytes 4-7
    if (K == ZT_RDO_BC4) return (uint64_t)(b.x >> 16) | (uint64_t)b.y << 16; // indices, bytes 2-7
    return ((uint64_t)b.z | (uint64_t)b.w << 32) >> 8;                     // BC7 bytes 9-15 (byte 8 is always paid)
}
template <int K> struct FieldBytes;
template <> struct FieldBytes<ZT_RDO_BC1> { static constexpr int N0 = 4, N1 = 4; };
template <> struct FieldBytes<ZT_RDO_BC4> { static constexpr int N0 = 2, N1 = 6; };
template <> struct FieldBytes<ZT_RDO_BC7> { static constexpr int N0 = 8, N1 = 7; };

// zap_tex.h:234-242: 0 if the field continues the match at the current distance, the new-match cost (min(n, 3))
// at the nearest equal field of the previous 16 blocks, else n literal bytes; nd receives the distance kept.  m has
// bit d - 1 set when block cur - d exists and its field equals this one.
template <int N> ZT_DEV float field_rate(uint32_t m, int dist, int &nd) {
    nd = dist;
    if ((m >> (dist - 1)) & 1u) return 0.0f;          // dist is always 1..16
    if (m) { nd = __ffs(m); return (float)(N < 3 ? N : 3); }
    return (float)N;
}

// ---- BC1 pieces (zap_tex.h:41-94): palettes of packed RGB bytes, fixed-endpoint fit, fixed-index score
ZT_DEV uint32_t rgb565(uint32_t c) {
    const uint32_t R = (c >> 11) & 31u, G = (c >> 5) & 63u, B = c & 31u;
    return (R << 3 | R >> 2) | (G << 2 | G >> 4) << 8 | (B << 3 | B >> 2) << 16;
}
ZT_DEV void bc1_palette(uint32_t c0, uint32_t c1, bool four, uint32_t pal[4]) {
    const uint32_t a = rgb565(c0), b = rgb565(c1);
    pal[0] = a; pal[1] = b; pal[2] = 0; pal[3] = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int x = chan(a, c), y = chan(b, c);
        pal[2] |= (uint32_t)(four ? (2 * x + y + 1) / 3 : (x + y + 1) / 2) << (8 * c);
        pal[3] |= (uint32_t)(four ? (x + 2 * y + 1) / 3 : 0) << (8 * c);
    }
}
ZT_DEV int sse_rgb(uint32_t p, uint32_t q) {
    int s = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) { const int d = chan(p, c) - chan(q, c); s += d * d; }
    return s;
}
// best index per pixel over the palette's 4 (or 3) colours, first smallest SSE (zap_tex.h:70-80)
ZT_DEV int bc1_fit(const uint32_t px[16], uint32_t c0, uint32_t c1, bool force4, uint32_t &idx) {
    const bool four = force4 || c0 > c1;
    uint32_t pal[4];
    bc1_palette(c0, c1, four, pal);
    int err = 0;
    idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int be = sse_rgb(px[i], pal[0]), bi = 0;
#pragma unroll
        for (int k = 1; k < 4; k++) {
            const int e = sse_rgb(px[i], pal[k]);
            if ((k < 3 || four) && e < be) { be = e; bi = k; }
        }
        idx |= (uint32_t)bi << (2 * i);
        err += be;
    }
    return err;
}
ZT_DEV int bc1_score(const uint32_t px[16], uint32_t c0, uint32_t c1, bool force4, uint32_t idx) {   // zap_tex.h:82-87
    uint32_t pal[4];
    bc1_palette(c0, c1, force4 || c0 > c1, pal);
    int err = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t k = (idx >> (2 * i)) & 3u;
        err += sse_rgb(px[i], k == 0 ? pal[0] : k == 1 ? pal[1] : k == 2 ? pal[2] : pal[3]);
    }
    return err;
}

// ---- BC4 pieces (zap_tex.h:172-196) on one channel
ZT_DEV int bc4_entry(int a0, int a1, int k) {
    if (k < 2) return k ? a1 : a0;
    if (a0 > a1) return ((8 - k) * a0 + (k - 1) * a1 + 3) / 7;
    return k < 6 ? ((6 - k) * a0 + (k - 1) * a1 + 2) / 5 : (k == 6 ? 0 : 255);
}
ZT_DEV int bc4_fit(const int v[16], int a0, int a1, uint64_t &idx) {
    int pal[8], err = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) pal[k] = bc4_entry(a0, a1, k);
    idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int be = (v[i] - pal[0]) * (v[i] - pal[0]), bi = 0;
#pragma unroll
        for (int k = 1; k < 8; k++) { const int e = (v[i] - pal[k]) * (v[i] - pal[k]); if (e < be) { be = e; bi = k; } }
        idx |= (uint64_t)bi << (3 * i);
        err += be;
    }
    return err;
}
ZT_DEV int bc4_score(const int v[16], int a0, int a1, uint64_t idx) {
    int pal[8], err = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) pal[k] = bc4_entry(a0, a1, k);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int k = (int)((idx >> (3 * i)) & 7u);
        int p = pal[0];
#pragma unroll
        for (int t = 1; t < 8; t++) p = k == t ? pal[t] : p;
        err += (v[i] - p) * (v[i] - p);
    }
    return err;
}

// ---- BC7 pieces: decoded SSE (zap_tex.h:710-716) and the two mode-6 refits (zap_tex.h:529-579)
ZT_DEV int sse_rgba(const uint32_t a[16], const uint32_t b[16]) {
    int s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) { const int d = chan(a[i], c) - chan(b[i], c); s += d * d; }
    return s;
}
ZT_DEV int decoded_sse(const uint32_t px[16], uint4 blk) {
    uint32_t d[16];
    bc7_decode_bits((uint64_t)blk.x | (uint64_t)blk.y << 32, (uint64_t)blk.z | (uint64_t)blk.w << 32, d);
    return sse_rgba(px, d);
}
ZT_DEV bool is_mode6(uint4 b) { return (b.x & 0x7Fu) == 0x40u; }
// mode-6 fields: q[e][c] 7 bits at 7 + 7 (2c + e), p-bits at 63 and 64, indices from 65 (pixel 0: 3 bits), 4 bits each
ZT_DEV void mode6_fields(uint4 b, int q[2][4], int pb[2], uint64_t &idx) {
    const uint64_t lo = (uint64_t)b.x | (uint64_t)b.y << 32, hi = (uint64_t)b.z | (uint64_t)b.w << 32;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) q[e][c] = (int)((lo >> (7 + 7 * (2 * c + e))) & 127u);
    pb[0] = (int)(lo >> 63); pb[1] = (int)(hi & 1u);
    const uint64_t f = hi >> 1;               // pixel 0's 3 bits, then 4 per pixel
    idx = (f & 7u) | ((f >> 3) << 4);
}
ZT_DEV uint4 mode6_pack(const int q[2][4], const int pb[2], uint64_t idx) {
    uint64_t lo = 0x40;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) lo |= (uint64_t)q[e][c] << (7 + 7 * (2 * c + e));
    lo |= (uint64_t)pb[0] << 63;
    const uint64_t hi = (uint64_t)pb[1] | ((idx & 7u) | ((idx >> 1) & ~7ull)) << 1;
    return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}
ZT_DEV void mode6_palette(const int q[2][4], const int pb[2], uint32_t pal[16], int K[16]) {
    int e0[4], e1[4];
#pragma unroll
    for (int c = 0; c < 4; c++) { e0[c] = q[0][c] << 1 | pb[0]; e1[c] = q[1][c] << 1 | pb[1]; }
    build_palette<4, 4>(e0, e1, pal);
    palette_keys<16>(pal, K);
}
// the endpoints kept, indices refitted (pixel 0 limited to 0..7, so the anchor stays valid without a swap)
ZT_DEV int mode6_refit_indices(const uint32_t px[16], const int q[2][4], const int pb[2], uint64_t &idx) {
    uint32_t pal[16];
    int K[16], err = 0;
    mode6_palette(q, pb, pal, K);
    idx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int key = i ? best_key<16>(px[i], pal, K) : best_key<8>(px[i], pal, K);
        err += (int)dot4(px[i], px[i]) - (key >> 4);
        idx |= (uint64_t)(15 - (key & 15)) << (4 * i);
    }
    return err;
}
// the indices kept, endpoints re-solved by least squares and quantised, each endpoint with its own p-bit
ZT_DEV int mode6_refit_endpoints(const uint32_t px[16], uint64_t idx, int q[2][4], int pb[2]) {
    ChanWords T;
    chan_words(px, T);
    int P[4];
    chan_sums<4, 0, true>(T, 0xFFFFu, P);
    uint32_t KW[4];
#pragma unroll
    for (int g = 0; g < 4; g++) KW[g] = weight_bytes<4>(idx_bytes4(idx, g));
    float aa, ab, bb, sa[4], sb[4];
    ls_sums<4, 0>(T, KW, P, 16, aa, ab, bb, sa, sb);
    const float det = aa * bb - ab * ab;
    float lo[4], hi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float l, h;
        if (det < 1e-6f && det > -1e-6f) l = h = (sa[c] + sb[c]) / 16.0f;   // one index everywhere: flat
        else { l = (sa[c] * bb - sb[c] * ab) / det; h = (sb[c] * aa - sa[c] * ab) / det; }
        lo[c] = l < 0 ? 0 : l > 255 ? 255 : l; hi[c] = h < 0 ? 0 : h > 255 ? 255 : h;
    }
    int el[2] = { 0, 0 }, eh[2] = { 0, 0 }, ql[2][4], qh[2][4];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            el[p] += quant_pbit1<8>(lo[c], quant_scale<8>(lo[c]), p, ql[p][c]);
            eh[p] += quant_pbit1<8>(hi[c], quant_scale<8>(hi[c]), p, qh[p][c]);
        }
    pb[0] = el[1] < el[0]; pb[1] = eh[1] < eh[0];
#pragma unroll
    for (int c = 0; c < 4; c++) { q[0][c] = pb[0] ? ql[1][c] : ql[0][c]; q[1][c] = pb[1] ? qh[1][c] : qh[0][c]; }
    int err = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int w = wgt<4>((int)((idx >> (4 * i)) & 15u));
        uint32_t e = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
            e |= (uint32_t)(((64 - w) * (q[0][c] << 1 | pb[0]) + w * (q[1][c] << 1 | pb[1]) + 32) >> 6) << (8 * c);
        err += (int)dot4(px[i], px[i]) - 2 * (int)dot4(px[i], e) + (int)dot4(e, e);
    }
    return err;
}

// wave-wide minimum of a 64-bit key (every lane receives it)
ZT_DEV uint64_t wave_min(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off), hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off);
        const uint64_t o = (uint64_t)hi << 32 | lo;
        v = o < v ? o : v;
    }
    return v;
}

// ---- one block's decision (zap_tex.h:245-276, :281-306, :792-815).  Every lane holds the block's pixels; `plain` is
// its plain encoding, ring[b & 15] the final bytes of block b for the 16 blocks before cur (those that exist), and
// dist0 / dist1 the match distances the previous block left.  Returns the bytes kept (on every lane) and updates the
// distances.
// mode: SPEC stores every lane's candidate in the stream's cache; REPAIR takes a lane's candidate from the cache when its
// predecessor's bytes equal those of the speculative walk (spec_ring), since a candidate depends only on its
// predecessor, the block's plain bytes and its pixels.
// PLAIN neither stores nor reads the cache (warm-up blocks, streams of one segment).
enum { DECIDE_SPEC = 0, DECIDE_REPAIR = 1, DECIDE_PLAIN = 2 };
template <int K, int MODE>
ZT_DEV uint4 decide(const zt_rdo_stream &st, float lambda, uint32_t cur, const uint32_t px[16], uint4 plain, const uint4 *ring,
                    const uint4 *spec_ring, int &dist0, int &dist1, int cache_e = 0, uint4 cache_cand = uint4()) {
    const int lane = (int)threadIdx.x & 63;
    const int j = lane / 3 + 1, kind = lane - 3 * (j - 1);   // lanes 0..47: predecessor j, candidate kind
    const bool force4 = st.force4 != 0, is_plain = lane == 48;
    bool valid = is_plain || (lane < 48 && (uint32_t)j <= cur);
    const uint4 pred = lane < 48 ? ring[(cur - (uint32_t)j) & 15u] : plain;
    uint4 cand = plain;
    int e = 0;
    bool cached = false;
    if (MODE == DECIDE_REPAIR && lane < ZT_RDO_LANES) {
        const uint4 sp = lane < 48 ? spec_ring[(cur - (uint32_t)j) & 15u] : plain;
        if (pred.x == sp.x && pred.y == sp.y && pred.z == sp.z && pred.w == sp.w) {
            e = cache_e;
            cand = cache_cand;
            valid = valid && e >= 0;
            cached = true;
        }
    }
    if (cached) {
    } else if (K == ZT_RDO_BC1) {
        const uint32_t c0 = plain.x & 0xFFFFu, c1 = plain.x >> 16, p0 = pred.x & 0xFFFFu, p1 = pred.x >> 16;
        if (is_plain || kind == 2) { e = bc1_score(px, p0, p1, force4, pred.y); cand = pred; }          // (plain) whole block
        else if (kind == 0) { uint32_t fi; e = bc1_fit(px, p0, p1, force4, fi); cand = make_uint4(pred.x, fi, 0u, 0u); }
        else { e = bc1_score(px, c0, c1, force4, pred.y); cand = make_uint4(plain.x, pred.y, 0u, 0u); }   // its indices
        if (!force4 && !is_plain && (cand.x & 0xFFFFu) <= (cand.x >> 16)) {   // 3-colour: index 3 is transparent
            const uint32_t y = cand.y;
            valid = valid && !((y & (y >> 1)) & 0x55555555u);
        }
    } else if (K == ZT_RDO_BC4) {
        int v[16];
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = chan(px[i], st.chan);
        const int a0 = (int)(plain.x & 255u), a1 = (int)((plain.x >> 8) & 255u), p0 = (int)(pred.x & 255u), p1 = (int)((pred.x >> 8) & 255u);
        const uint64_t pi = field1<ZT_RDO_BC4>(pred);
        if (is_plain || kind == 2) { e = bc4_score(v, p0, p1, pi); cand = pred; }
        else if (kind == 0) {
            uint64_t fi;
            e = bc4_fit(v, p0, p1, fi);
            cand = make_uint4((uint32_t)(p0 | p1 << 8) | (uint32_t)fi << 16, (uint32_t)(fi >> 16), 0u, 0u);
        } else { e = bc4_score(v, a0, a1, pi); cand = make_uint4((plain.x & 0xFFFFu) | (pred.x & 0xFFFF0000u), pred.y, 0u, 0u); }
    } else {
        if (is_plain || kind == 0) { e = decoded_sse(px, pred); cand = pred; }
        else if (!is_mode6(pred)) valid = false;
        else {
            int q[2][4], pb[2];
            uint64_t idx;
            mode6_fields(pred, q, pb, idx);
            if (kind == 1) { uint64_t ni; e = mode6_refit_indices(px, q, pb, ni); cand = mode6_pack(q, pb, ni); }
            else { int nq[2][4], np[2]; e = mode6_refit_endpoints(px, idx, nq, np); cand = mode6_pack(nq, np, idx); }
        }
    }
    if (MODE == DECIDE_SPEC && lane < ZT_RDO_LANES) {
        const uint64_t c = (uint64_t)cur * ZT_RDO_LANES + (uint64_t)lane;
        st.cand_err[c] = valid ? e : -1;
        st.cand[c] = cand;
    }
    // which of the previous 16 blocks hold the same fields: 16 independent compares, bit d - 1 for block cur - d
    const uint64_t f0 = field0<K>(cand), f1 = field1<K>(cand);
    uint32_t m0 = 0, m1 = 0;
#pragma unroll
    for (int d = 1; d <= 16; d++) {
        const uint4 b = ring[(cur - (uint32_t)d) & 15u];
        const bool exists = (uint32_t)d <= cur;
        m0 |= (uint32_t)(exists && field0<K>(b) == f0) << (d - 1);
        m1 |= (uint32_t)(exists && field1<K>(b) == f1) << (d - 1);
    }
    // J = SSE + lambda * rate, in the reference's float order
    int nd0, nd1;
    float rate = field_rate<FieldBytes<K>::N0>(m0, dist0, nd0) + field_rate<FieldBytes<K>::N1>(m1, dist1, nd1);
    if (K == ZT_RDO_BC7) rate = rate + 1.0f;
    const float J = (float)e + lambda * rate;
    const uint32_t order = is_plain ? 0u : (uint32_t)lane + 1u;
    // the reference's strict compares: a NaN candidate is never taken, a NaN plain J keeps the plain block
    uint64_t key = valid && J == J ? ((uint64_t)__float_as_uint(J) << 32 | order) : ~0ull;
    if (is_plain && J != J) key = 0;
    const uint32_t wo = (uint32_t)wave_min(key);
    const int win = wo == 0 ? 48 : (int)wo - 1;
    dist0 = __shfl(nd0, win);
    dist1 = __shfl(nd1, win);
    return make_uint4((uint32_t)__shfl((int)cand.x, win), (uint32_t)__shfl((int)cand.y, win), (uint32_t)__shfl((int)cand.z, win),
                      (uint32_t)__shfl((int)cand.w, win));
}

// Speculative walk of one segment [k S, k S + S) of a stream: start `warm` blocks earlier from a guessed state (the
// plain bytes of the 16 blocks before that, distances 1 and 1), decide the warm-up blocks into the ring only, record the
// state the segment starts from (its 16 predecessors and the distances), then decide the segment's blocks into the
// output with the distances each leaves.  Segment 0 starts at block 0, where the guess is the reference's own start,
// so it is exact; zt_rdo_fix checks the others.
template <int K> ZT_DEV void speculate(const zt_rdo_stream &st, float lambda, uint32_t k) {
    __shared__ uint4 ring[16];
    __shared__ uint32_t spx[16];
    const int lane = (int)threadIdx.x;
    const uint32_t n = (uint32_t)st.bw * (uint32_t)st.bh, S = (uint32_t)st.seg;
    const uint32_t s_begin = k * S, s_end = min(n, s_begin + S), s0 = s_begin > (uint32_t)st.warm ? s_begin - (uint32_t)st.warm : 0u;
    if (lane < 16 && s0 >= 16u - (uint32_t)lane) {
        const uint32_t b = s0 - 16u + (uint32_t)lane;
        ring[b & 15u] = read_block(st, st.plain, b);
    }
    int dist0 = 1, dist1 = 1;
    uint32_t pf = 0;
    uint4 pp;
    load_block(st, s0, lane, pf, pp);
    __syncthreads();
    for (uint32_t cur = s0; cur < s_end; cur++) {
        if (cur == s_begin && k) {   // the state this segment assumes
            zt_rdo_state *S0 = st.states + (k - 1);
            if (lane < 16) S0->ring[lane] = ring[(cur - 16u + (uint32_t)lane) & 15u];
            if (lane == 0) S0->dist = (uint32_t)dist0 | (uint32_t)dist1 << 8;
        }
        if (lane < 16) spx[lane] = pf;
        const uint4 plain = pp;
        __syncthreads();
        if (cur + 1 < s_end) load_block(st, cur + 1, lane, pf, pp);   // the next block's loads run under this one's work
        uint32_t px[16];
#pragma unroll
        for (int i = 0; i < 16; i++) px[i] = spx[i];
        const uint4 w = cur >= s_begin && st.nseg > 1 ? decide<K, DECIDE_SPEC>(st, lambda, cur, px, plain, ring, ring, dist0, dist1)
                                                      : decide<K, DECIDE_PLAIN>(st, lambda, cur, px, plain, ring, ring, dist0, dist1);
        if (lane == 0) {
            if (cur >= s_begin) { write_block(st, st.out, cur, w); st.dists[cur] = (uint16_t)(dist0 | dist1 << 8); }
            ring[cur & 15u] = w;
        }
        __syncthreads();
    }
    if (lane == 0) st.seg_end[k] = (uint32_t)dist0 | (uint32_t)dist1 << 8;
}

// Ordered check of a stream's segments 1.. against the true state: the final bytes of the 16 blocks before the
// segment (in the output, exact by induction) and the distances the previous segment leaves.  A segment whose
// assumed state equals it is exact as speculated.  Otherwise it is decided again from the true state until the
// decisions rejoin the speculative ones: 16 blocks in a row equal to the speculated bytes with the same distances
// after them mean the rest of the segment saw exactly the speculative run's inputs.
template <int K> ZT_DEV void fixup(const zt_rdo_stream &st, float lambda) {
    __shared__ uint4 ring[16], spec_ring[16];   // true and speculative final bytes of the 16 blocks before cur
    __shared__ uint32_t spx[16];
    const int lane = (int)threadIdx.x;
    const uint32_t n = (uint32_t)st.bw * (uint32_t)st.bh, S = (uint32_t)st.seg, nseg = (n + S - 1) / S;
    uint32_t tdist = st.seg_end[0];
    bool tail_in_ring = false;   // the previous segment was re-walked to its end: its last 16 blocks are the ring
    for (uint32_t k = 1; k < nseg; k++) {
        const uint32_t s_begin = k * S, s_end = min(n, s_begin + S);
        const zt_rdo_state *S0 = st.states + (k - 1);
        // the true 16 predecessors: the ring when this wave just wrote them, else the output (written by the
        // speculative launch, or equal to it)
        const uint4 t = lane >= 16 ? make_uint4(0u, 0u, 0u, 0u)
                      : tail_in_ring ? ring[(s_begin - 16u + (uint32_t)lane) & 15u] : read_block(st, st.out, s_begin - 16u + (uint32_t)lane);
        __syncthreads();
        const uint4 a = lane < 16 ? S0->ring[lane] : make_uint4(0u, 0u, 0u, 0u);
        const bool same = t.x == a.x && t.y == a.y && t.z == a.z && t.w == a.w;
        if (__all(same) && S0->dist == tdist) { tdist = st.seg_end[k]; tail_in_ring = false; continue; }
        if (lane < 16) { ring[(s_begin - 16u + (uint32_t)lane) & 15u] = t; spec_ring[(s_begin - 16u + (uint32_t)lane) & 15u] = a; }
        int dist0 = (int)(tdist & 255u), dist1 = (int)(tdist >> 8);
        uint32_t run = 0, pf = 0;
        bool joined = false;
        // the walk's global reads (pixels, plain and speculative bytes, distances, this lane's cached candidate) are
        // issued one block ahead, so only the decision itself is on the sequential path
        uint4 pp, ps, pc = make_uint4(0u, 0u, 0u, 0u);
        uint32_t pd;
        int pe = -1;
        auto fetch = [&](uint32_t b) {
            load_block(st, b, lane, pf, pp);
            ps = read_block(st, st.out, b);
            pd = st.dists[b];
            if (lane < ZT_RDO_LANES) { pe = st.cand_err[(uint64_t)b * ZT_RDO_LANES + lane]; pc = st.cand[(uint64_t)b * ZT_RDO_LANES + lane]; }
        };
        fetch(s_begin);
        __syncthreads();
        for (uint32_t cur = s_begin; cur < s_end; cur++) {
            if (lane < 16) spx[lane] = pf;
            const uint4 plain = pp, spec = ps, cc = pc;
            const uint32_t spec_dist = pd;
            const int ce = pe;
            __syncthreads();
            if (cur + 1 < s_end) fetch(cur + 1);
            uint32_t px[16];
#pragma unroll
            for (int i = 0; i < 16; i++) px[i] = spx[i];
            const uint4 w = decide<K, DECIDE_REPAIR>(st, lambda, cur, px, plain, ring, spec_ring, dist0, dist1, ce, cc);
            const bool eq = w.x == spec.x && w.y == spec.y && w.z == spec.z && w.w == spec.w;
            run = eq ? run + 1 : 0;
            if (lane == 0) {
                if (!eq) write_block(st, st.out, cur, w);
                ring[cur & 15u] = w;
                spec_ring[cur & 15u] = spec;
            }
            __syncthreads();
            if (run >= 16 && ((uint32_t)dist0 | (uint32_t)dist1 << 8) == spec_dist) { joined = true; break; }
        }
        tdist = joined ? st.seg_end[k] : ((uint32_t)dist0 | (uint32_t)dist1 << 8);
        tail_in_ring = !joined;
        __syncthreads();
    }
}

}  // namespace

// one wave per (stream, segment): blockIdx.x counts the segments of the launch's streams in order
extern "C" __global__ void __launch_bounds__(64) zt_rdo_spec(zt_rdo_streams S, float lambda) {
    uint32_t g = blockIdx.x;
    int i = 0;
    while (g >= S.s[i].nseg) { g -= S.s[i].nseg; i++; }
    const zt_rdo_stream st = S.s[i];
    if (st.kind == ZT_RDO_BC1) speculate<ZT_RDO_BC1>(st, lambda, g);
    else if (st.kind == ZT_RDO_BC4) speculate<ZT_RDO_BC4>(st, lambda, g);
    else speculate<ZT_RDO_BC7>(st, lambda, g);
}
// one wave per stream with more than one segment
extern "C" __global__ void __launch_bounds__(64) zt_rdo_fix(zt_rdo_streams S, float lambda) {
    const zt_rdo_stream st = S.s[blockIdx.x];
    if (st.nseg < 2) return;
    if (st.kind == ZT_RDO_BC1) fixup<ZT_RDO_BC1>(st, lambda);
    else if (st.kind == ZT_RDO_BC4) fixup<ZT_RDO_BC4>(st, lambda);
    else fixup<ZT_RDO_BC7>(st, lambda);
}
