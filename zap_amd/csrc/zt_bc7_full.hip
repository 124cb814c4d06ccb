// zt_bc7_full.hip — exhaustive all-mode BC7 search for gfx950, and the per-candidate encoder the learned
// proposal uses.  Project-defined (the reference encoder only tries modes 6, 5 and 1): 285 candidates per
// block, numbered mode*64 + slot (partition for modes 0-3 and 7, rot*2 + index selector for mode 4,
// rotation for mode 5, 0 for mode 6).  Every subset is encoded the reference's way (PCA start, fit, <= 2
// least-squares refinements); modes 6, 5 and 1 are the reference encoders, so the exhaustive result is never
// worse than the reference search.  SSEs count all four channels of the decoded block, exactly.
#include "zt_bc7_core.h"
#include "zt_bc7_tables.h"
#include "zt_kernels.h"

namespace {

// mode table (BC7 specification): subsets, partition bits, colour bits, alpha bits, endpoint p-bit, shared
// p-bit, index bits
template <int M> struct B7;
template <> struct B7<0> { static constexpr int NS = 3, PB = 4, CB = 4, AB = 0, EPB = 1, SPB = 0, IB = 3; };
template <> struct B7<1> { static constexpr int NS = 2, PB = 6, CB = 6, AB = 0, EPB = 0, SPB = 1, IB = 3; };
template <> struct B7<2> { static constexpr int NS = 3, PB = 6, CB = 5, AB = 0, EPB = 0, SPB = 0, IB = 2; };
template <> struct B7<3> { static constexpr int NS = 2, PB = 6, CB = 7, AB = 0, EPB = 1, SPB = 0, IB = 2; };
template <> struct B7<6> { static constexpr int NS = 1, PB = 0, CB = 7, AB = 7, EPB = 1, SPB = 0, IB = 4; };
template <> struct B7<7> { static constexpr int NS = 2, PB = 6, CB = 5, AB = 5, EPB = 1, SPB = 0, IB = 2; };

// subset without p-bits: plain n-bit endpoints, palette argmin (idx: 4 bits per pixel)
template <int NC, int N, int IB>
ZT_DEV int fit_subset_plain(const uint32_t px[16], uint32_t mask, int spp, uint64_t snib, const float lo[4], const float hi[4], int q[2][4],
                            uint64_t &idx) {
    int e0[NC], e1[NC];
    quant_plain<NC>(lo, N, q[0], e0); quant_plain<NC>(hi, N, q[1], e1);
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
template <int NC, int N, int IB, bool R = false, typename PX>
ZT_DEV int encode_subset_plain(PX *px, const ChanWords &T, uint32_t mask, int cnt, int q[2][4], uint64_t &idx) {
    int P[NC];
    chan_sums<NC, 0, false>(T, mask, P);
    float lo[4], hi[4];
    pca<NC, false, R>(px, mask, cnt, lo, hi);
    const int spp = subset_sq<NC, 0>(px, mask);
    const uint64_t snib = subset_nibbles(mask);
    int err = fit_subset_plain<NC, N, IB>(px, mask, spp, snib, lo, hi, q, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        uint32_t KW[4];   // pixels outside the subset have index 0, hence weight 0
#pragma unroll
        for (int g = 0; g < 4; g++) KW[g] = weight_bytes<IB>(idx_bytes4(idx, g));
        float aa, ab, bb, sa[NC], sb[NC];
        ls_sums<NC, 0>(T, KW, P, cnt, aa, ab, bb, sa, sb);
        float det = aa * bb - ab * ab;
        if (det < 1e-6f && det > -1e-6f) break;
        float nlo[4], nhi[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (c < NC) {
                float l = (sa[c] * bb - sb[c] * ab) / det, h = (sb[c] * aa - sa[c] * ab) / det;
                nlo[c] = l < 0 ? 0 : l > 255 ? 255 : l; nhi[c] = h < 0 ? 0 : h > 255 ? 255 : h;
            } else { nlo[c] = 255.0f; nhi[c] = 255.0f; }
        }
        int nq[2][4];
        uint64_t ni;
        int ne = fit_subset_plain<NC, N, IB>(px, mask, spp, snib, nlo, nhi, nq, ni);
        if (ne >= err) break;
        err = ne; idx = ni;
#pragma unroll
        for (int c = 0; c < NC; c++) { q[0][c] = nq[0][c]; q[1][c] = nq[1][c]; }
    }
    return err;
}

ZT_DEV uint32_t subset_mask(int ns, int part, int s) {
    if (ns == 1) return 0xFFFFu;
    if (ns == 2) { uint32_t m = kP2[part]; return s ? m : (~m & 0xFFFFu); }
    uint32_t p3 = kP3[part], m = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) m |= (uint32_t)(((p3 >> (2 * i)) & 3) == (uint32_t)s) << i;
    return m;
}
ZT_DEV int anchor_pixel(int ns, int part, int s) {
    return s == 0 ? 0 : ns == 2 ? (int)kA2[part] : s == 1 ? (int)kA3a[part] : (int)kA3b[part];
}
// nibble mask selecting the low `bits` bits of every pixel in `mask`
ZT_DEV uint64_t nibble_fill(uint32_t mask, int bits) {
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) if ((mask >> i) & 1) r |= (uint64_t)((1 << bits) - 1) << (4 * i);
    return r;
}

// modes 0, 1, 2, 3, 6, 7 for one partition: SSE, packed block.  R: refresh the pixel registers between the PCA
// passes (fewer VGPRs: the search kernels run at higher occupancy; px must then be modifiable)
template <int M, bool R = false, typename PX> ZT_DEV int cand_subsets(PX *px, int part, uint64_t &olo, uint64_t &ohi) {
    using T = B7<M>;
    constexpr int NC = T::AB ? 4 : 3, HASP = T::EPB || T::SPB, N = T::CB + HASP;
    int q[T::NS][2][4], pb[T::NS][2];
    uint64_t full = 0;
    uint32_t masks[T::NS];
    int err = 0;
    ChanWords CW;
    chan_words(px, CW);
#pragma unroll
    for (int s = 0; s < T::NS; s++) {
        masks[s] = subset_mask(T::NS, part, s);
        const int cnt = __popc(masks[s]);
        uint64_t idx;
        pb[s][0] = pb[s][1] = 0;
        if (HASP) err += encode_subset<NC, N, T::IB, (bool)T::SPB, (T::NS == 1), R>(px, CW, masks[s], cnt, q[s], pb[s], idx);
        else err += encode_subset_plain<NC, N, T::IB, R>(px, CW, masks[s], cnt, q[s], idx);
        full |= idx;
    }
    if (NC == 3) {
#pragma unroll
        for (int i = 0; i < 16; i++) { int d = 255 - (int)(px[i] >> 24); err += d * d; }  // decoded alpha is 255
    }
    int anc[T::NS];
#pragma unroll
    for (int s = 0; s < T::NS; s++) {
        anc[s] = anchor_pixel(T::NS, part, s);
        if ((full >> (4 * anc[s])) & (1u << (T::IB - 1))) {
#pragma unroll
            for (int c = 0; c < NC; c++) { int t = q[s][0][c]; q[s][0][c] = q[s][1][c]; q[s][1][c] = t; }
            if (T::EPB) { int t = pb[s][0]; pb[s][0] = pb[s][1]; pb[s][1] = t; }
            full ^= nibble_fill(masks[s], T::IB);
        }
    }
    Bits128 b{ 1ull << M, 0 };
    int pos = M + 1;
    b.put(pos, (uint64_t)part, T::PB); pos += T::PB;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int s = 0; s < T::NS; s++)
#pragma unroll
            for (int e = 0; e < 2; e++) { b.put(pos, (uint64_t)q[s][e][c], T::CB); pos += T::CB; }
    if (T::AB) {
#pragma unroll
        for (int s = 0; s < T::NS; s++)
#pragma unroll
            for (int e = 0; e < 2; e++) { b.put(pos, (uint64_t)q[s][e][3], T::AB); pos += T::AB; }
    }
    if (T::EPB) {
#pragma unroll
        for (int s = 0; s < T::NS; s++)
#pragma unroll
            for (int e = 0; e < 2; e++) { b.put(pos, (uint64_t)pb[s][e], 1); pos += 1; }
    }
    if (T::SPB) {
#pragma unroll
        for (int s = 0; s < T::NS; s++) { b.put(pos, (uint64_t)pb[s][0], 1); pos += 1; }
    }
    for (int i = 0; i < 16; i++) {
        bool is_anchor = i == 0;
#pragma unroll
        for (int s = 1; s < T::NS; s++) is_anchor |= i == anc[s];
        const int nb = T::IB - (is_anchor ? 1 : 0);
        b.put(pos, (full >> (4 * i)) & ((1u << nb) - 1), nb);
        pos += nb;
    }
    olo = b.lo; ohi = b.hi;
    return err;
}

// n-bit endpoints, IB-bit indices on channels [C0, C0 + NC) of the rotated block (generalises the mode-5 part)
template <int NC, int C0, int IB> ZT_DEV int fit_nbit(const uint32_t r[16], int spp, const int e0[], const int e1[], uint64_t &idx) {
    constexpr int NP = 1 << IB;
    uint32_t pal[NP];
    int K[NP];
    build_palette<NC, IB>(e0, e1, pal);
    palette_keys<NP>(pal, K);
    int sk = 0, P32[NP];
    uint64_t bits = 0;
#pragma unroll
    for (int j = 0; j < NP; j++) P32[j] = (int)pal[j] << 5;   // used when NC == 1
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int key = NC == 1 ? best_key1<NP>((int)sub_px<NC, C0>(r[i]), P32, K) : best_key<NP>(sub_px<NC, C0>(r[i]), pal, K);
        sk += key;
        bits |= (uint64_t)(key & 15) << (4 * i);
    }
    idx = ~bits;
    return spp - ((sk - nibble_sum(bits)) >> 4);
}
template <int NC, int C0, int IB> ZT_DEV bool lsq_nbit(const ChanWords &T, const int P[4], uint64_t idx, float lo[], float hi[]) {
    uint32_t KW[4];
#pragma unroll
    for (int g = 0; g < 4; g++) KW[g] = weight_bytes<IB>(idx_bytes4(idx, g));
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
template <int NC, int C0, int IB>
ZT_DEV int part_nbit(const uint32_t r[16], const ChanWords &T, const int P[4], int n, const float lo[], const float hi[], int q[2][3],
                     uint64_t &idx) {
    int e0[NC], e1[NC];
    quant_plain<NC>(lo, n, q[0], e0); quant_plain<NC>(hi, n, q[1], e1);
    const int spp = subset_sq<NC, C0>(r, 0xFFFFu);
    int err = fit_nbit<NC, C0, IB>(r, spp, e0, e1, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float nlo[NC], nhi[NC];
        if (!lsq_nbit<NC, C0, IB>(T, P, idx, nlo, nhi)) break;
        int tq[2][3], te0[NC], te1[NC];
        quant_plain<NC>(nlo, n, tq[0], te0); quant_plain<NC>(nhi, n, tq[1], te1);
        uint64_t ti;
        int e = fit_nbit<NC, C0, IB>(r, spp, te0, te1, ti);
        if (e >= err) break;
        err = e; idx = ti;
#pragma unroll
        for (int c = 0; c < NC; c++) { q[0][c] = tq[0][c]; q[1][c] = tq[1][c]; }
    }
    return err;
}

// modes 4 (ISB = index selector) and 5: rotation `rot`
template <int M, int ISB, bool R = false> ZT_DEV int cand_rotation(const uint32_t px[16], int rot, uint64_t &olo, uint64_t &ohi) {
    constexpr int CB = M == 4 ? 5 : 7, AB = M == 4 ? 6 : 8, IBC = (M == 4 && ISB) ? 3 : 2, IBA = (M == 4 && !ISB) ? 3 : 2;
    uint32_t r[16];
#pragma unroll
    for (int i = 0; i < 16; i++) r[i] = rotate_px(px[i], rot);
    float lo[4], hi[4], alo[1] = { 255.0f }, ahi[1] = { 0.0f };
    int cq[2][3] = { { 0, 0, 0 }, { 0, 0, 0 } }, aq[2][3] = { { 0, 0, 0 }, { 0, 0, 0 } };
    uint64_t ci, ai;
    ChanWords CW;
    chan_words(r, CW);
    int P[4];
    chan_sums<4, 0, true>(CW, 0xFFFFu, P);
    pca<3, true, R>(r, 0xFFFFu, 16, lo, hi, P);
    int err = part_nbit<3, 0, IBC>(r, CW, P, CB, lo, hi, cq, ci);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float a = (float)(r[i] >> 24);
        if (a < alo[0]) alo[0] = a;
        if (a > ahi[0]) ahi[0] = a;
    }
    err += part_nbit<1, 3, IBA>(r, CW, P, AB, alo, ahi, aq, ai);
    if (ci & (1u << (IBC - 1))) {
#pragma unroll
        for (int c = 0; c < 3; c++) { int t = cq[0][c]; cq[0][c] = cq[1][c]; cq[1][c] = t; }
        ci ^= nibble_fill(0xFFFFu, IBC);
    }
    if (ai & (1u << (IBA - 1))) { int t = aq[0][0]; aq[0][0] = aq[1][0]; aq[1][0] = t; ai ^= nibble_fill(0xFFFFu, IBA); }
    Bits128 b{ 1ull << M, 0 };
    int pos = M + 1;
    b.put(pos, (uint64_t)rot, 2); pos += 2;
    if (M == 4) { b.put(pos, (uint64_t)ISB, 1); pos += 1; }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < 2; e++) { b.put(pos, (uint64_t)cq[e][c], CB); pos += CB; }
#pragma unroll
    for (int e = 0; e < 2; e++) { b.put(pos, (uint64_t)aq[e][0], AB); pos += AB; }
    const uint64_t first = ISB ? ai : ci, second = ISB ? ci : ai;   // the 2-bit index set is stored first
    constexpr int B1 = 2, B2 = M == 4 ? 3 : 2;
#pragma unroll
    for (int i = 0; i < 16; i++) { const int nb = B1 - (i == 0); b.put(pos, (first >> (4 * i)) & ((1u << nb) - 1), nb); pos += nb; }
#pragma unroll
    for (int i = 0; i < 16; i++) { const int nb = B2 - (i == 0); b.put(pos, (second >> (4 * i)) & ((1u << nb) - 1), nb); pos += nb; }
    olo = b.lo; ohi = b.hi;
    return err;
}

}  // namespace

// one candidate of one block (the learned proposal's exact evaluation); SSE, or -1 for an unused slot
__device__ int zt_bc7_eval_candidate(const uint32_t px[16], int cand, uint64_t &lo, uint64_t &hi) {
    const int mode = cand >> 6, slot = cand & 63;
    switch (mode) {
    case 0: return slot < 16 ? cand_subsets<0>(px, slot, lo, hi) : -1;
    case 1: return cand_subsets<1>(px, slot, lo, hi);
    case 2: return cand_subsets<2>(px, slot, lo, hi);
    case 3: return cand_subsets<3>(px, slot, lo, hi);
    case 4: return slot >= 8 ? -1 : (slot & 1) ? cand_rotation<4, 1>(px, slot >> 1, lo, hi) : cand_rotation<4, 0>(px, slot >> 1, lo, hi);
    case 5: return slot < 4 ? cand_rotation<5, 0>(px, slot, lo, hi) : -1;
    case 6: return slot == 0 ? cand_subsets<6>(px, 0, lo, hi) : -1;
    default: return cand_subsets<7>(px, slot, lo, hi);
    }
}

namespace {
// the slots of mode M handled by one launch (mode 4 is split by its index selector ISB: slots 2 r + ISB)
template <int M, int ISB> constexpr int kSlots = M == 0 ? 16 : (M == 4 || M == 5) ? 4 : M == 6 ? 1 : 64;
template <int M, int ISB> ZT_DEV int slot_of(int s) { return M == 4 ? 2 * s + ISB : s; }
// the search kernels' evaluation: pixel registers refreshed between PCA passes (px is the caller's own copy)
template <int M, int ISB> ZT_DEV int eval_mode(uint32_t px[16], int slot, uint64_t &lo, uint64_t &hi) {
    if constexpr (M == 4) return cand_rotation<4, ISB, true>(px, slot >> 1, lo, hi);
    else if constexpr (M == 5) return cand_rotation<5, 0, true>(px, slot, lo, hi);
    else return cand_subsets<M, true>(px, slot, lo, hi);
}
ZT_DEV void load_block(const uint4 *blocks, uint32_t g, uint32_t px[16]) {
    const uint4 *s = blocks + (uint64_t)g * 4;
#pragma unroll
    for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
}
// SSE first, then the lowest candidate; SSE < 2^22 for a 4x4 RGBA8 block
ZT_DEV uint32_t cand_key(int err, int cand) { return err < 0 ? 0xFFFFFFFFu : (uint32_t)err << 9 | (uint32_t)cand; }
}  // namespace

// Exhaustive search, one mode (mode 4: one index selector) per launch.  Lane = (block, slot), the S slots of a
// block in adjacent lanes of one wave, so a wave never mixes modes and each launch holds only its mode's code;
// the S keys reduce inside the wave and one atomicMin per block folds them into key[g] (initialised to ~0).
// `list`/`nlist` (optional) restrict the search to listed blocks (the learned path's fallback); `errs`
// (optional, 512 per block, pre-filled with ~0) receive every candidate's SSE.
template <int M, int ISB>
ZT_DEV void exh_body(const uint4 *blocks, uint32_t total, const uint32_t *list, const unsigned *nlist, uint32_t *key, uint32_t *errs) {
    constexpr int S = kSlots<M, ISB>;
    const uint64_t count = (uint64_t)(list ? *nlist : total) * S;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < count; base += step) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < count;  // count is a multiple of S, so a block's S lanes are all valid or all not
        uint32_t k = 0xFFFFFFFFu, g = 0;
        if (valid) {
            const uint32_t e = (uint32_t)(i / S);
            const int slot = slot_of<M, ISB>((int)(i % S));
            g = list ? list[e] : e;
            uint32_t px[16];
            load_block(blocks, g, px);
            uint64_t lo, hi;
            const int err = eval_mode<M, ISB>(px, slot, lo, hi);
            k = cand_key(err, M * 64 + slot);
            if (errs) errs[(uint64_t)g * 512 + M * 64 + slot] = err < 0 ? 0xFFFFFFFFu : (uint32_t)err;
        }
#pragma unroll
        for (int off = 1; off < S; off <<= 1) k = min(k, (uint32_t)__shfl_xor(k, off));
        if (valid && i % S == 0) atomicMin(key + g, k);
    }
}
#define ZT_EXH(NAME, M, ISB)                                                                                       \
    extern "C" __global__ void __launch_bounds__(256) NAME(const uint4 *blocks, uint32_t total, const uint32_t *list,          \
                                                           const unsigned *nlist, uint32_t *key, uint32_t *errs) {            \
        exh_body<M, ISB>(blocks, total, list, nlist, key, errs);                                                               \
    }
ZT_EXH(zt_bc7_exh_m0, 0, 0)
ZT_EXH(zt_bc7_exh_m1, 1, 0)
ZT_EXH(zt_bc7_exh_m2, 2, 0)
ZT_EXH(zt_bc7_exh_m3, 3, 0)
ZT_EXH(zt_bc7_exh_m4i0, 4, 0)
ZT_EXH(zt_bc7_exh_m4i1, 4, 1)
ZT_EXH(zt_bc7_exh_m5, 5, 0)
ZT_EXH(zt_bc7_exh_m6, 6, 0)
ZT_EXH(zt_bc7_exh_m7, 7, 0)
#undef ZT_EXH

// The winner of every block (or, with `items`, of the count[0] blocks items[i].x, best sorted by winning mode so a
// wave runs one mode's code): its candidate re-encoded, the 16 bytes stored, and the candidate slot written to
// best_cand (optional).
extern "C" __global__ void __launch_bounds__(256) zt_bc7_finish(zt_jobs jobs, const uint4 *blocks, uint32_t total, const uint32_t *key,
                                                                  uint32_t *best_cand, const uint2 *items, const unsigned *count) {
    const uint32_t step = gridDim.x * blockDim.x, n = items ? count[0] : total;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint32_t g = items ? items[i].x : i;
        const uint32_t k = key[g];
        if (k == 0xFFFFFFFFu) continue;
        uint32_t px[16];
        load_block(blocks, g, px);
        uint64_t lo = 0, hi = 0;
        zt_bc7_eval_candidate(px, (int)(k & 511u), lo, hi);
        const zt_job J = job_at(jobs, zt_find_job(jobs, g));
        store16(J.dst + (uint64_t)(g - J.first) * 16, lo, hi);
        if (best_cand) best_cand[g] = k & 511u;
    }
}

// deltas[g][c] = min(65535, errs[g][c] - best SSE of block g), 65535 where the slot has no candidate
extern "C" __global__ void __launch_bounds__(256) zt_bc7_deltas(const uint32_t *errs, const uint32_t *key, uint32_t total, uint16_t *deltas) {
    const uint64_t n = (uint64_t)total * 512, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint32_t e = errs[i], best = key[i >> 9] >> 9;
        deltas[i] = e == 0xFFFFFFFFu ? (uint16_t)65535 : (uint16_t)min(65535u, e - best);
    }
}

// ===================================================================================== learned proposal
// The network proposes K candidates per block (zt_mlp.hip); they are bucketed by mode so that every wave of
// the exact evaluation runs one mode's code, and each evaluation folds its key into the block's key.  Blocks
// whose predicted SSE cost of trusting the K (the network's risk output) exceeds the threshold get the
// exhaustive launches over a list instead; zt_bc7_finish then encodes every block's winner.

// gather every block of the jobs into a dense array (16 RGBA8 pixels = 4 x uint4 per block)
extern "C" __global__ void __launch_bounds__(256) zt_bc7_gather(zt_jobs jobs, uint32_t total, uint4 *blocks) {
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const zt_job J = job_at(jobs, zt_find_job(jobs, g));
        const uint32_t b = g - J.first;
        uint32_t px[16];
        gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
        uint4 *d = blocks + (uint64_t)g * 4;
        d[0] = make_uint4(px[0], px[1], px[2], px[3]); d[1] = make_uint4(px[4], px[5], px[6], px[7]);
        d[2] = make_uint4(px[8], px[9], px[10], px[11]); d[3] = make_uint4(px[12], px[13], px[14], px[15]);
    }
}

// This lane's slot among the workgroup's items of class m (0 <= m < NM) in LDS tallies cnt[]: one LDS atomic per wave
// and class present (a ballot per class), not one per item on a handful of shared addresses.
template <int NM>
ZT_DEV unsigned lds_tally(unsigned *cnt, int m, bool valid) {
    const int lane = __lane_id();
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    unsigned local = 0;
#pragma unroll
    for (int mm = 0; mm < NM; mm++) {
        const uint64_t b = __ballot(valid && m == mm);
        if (!b) continue;
        const int leader = __ffsll((unsigned long long)b) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&cnt[mm], (unsigned)__popcll(b));
        base = __shfl(base, leader);
        if (valid && m == mm) local = base + (unsigned)__popcll(b & below);
    }
    return local;
}

// Bucketing in two passes over the same grid.  Pass 0 tallies each workgroup's items per class (mode 0..7; class 8:
// fallback blocks) in LDS and adds the tallies to the class's counter of shard blockIdx % 8 -- each (class, shard)
// counter on its own 256-byte line, since atomics on one line complete one after another.  zt_bc7_bucket_scan
// turns the counters into run starts (class-contiguous runs, shard sub-runs in order) and fills the summary words;
// pass 1 recomputes the same tallies and reserves its ranges from the same counters.
ZT_DEV unsigned *bucket_shard(unsigned *counters, int m) {
    return counters + ZT_BUCKET_SHARDS_AT + (m * ZT_SHARDS + (int)(blockIdx.x & (ZT_SHARDS - 1))) * ZT_SHARD_STRIDE;
}
// the proposals of every block: K candidates each, or none for a fallback block (risk > threshold), which goes to the
// fallback list instead; pass 1 also records each proposal's item index in slot[block * K + k] (0xFFFFFFFF for a
// fallback block)
extern "C" __global__ void __launch_bounds__(256) zt_bc7_bucket(const uint32_t *cand, const float *risk, uint32_t total, int K,
                                                                  float threshold, unsigned *counters, uint2 *items, uint32_t pass,
                                                                  uint32_t *fallback, uint32_t *slot) {
    __shared__ unsigned cnt[9], base[9];
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < total; b0 += step) {
        if (threadIdx.x < 9) cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t g = b0 + threadIdx.x;
        const bool valid = g < total;
        const bool fb = valid && risk[g] > threshold;
        uint32_t c[16];
        unsigned local[16];
        for (int k = 0; k < K; k++) {
            c[k] = valid && !fb ? cand[(uint64_t)g * K + k] : 0xFFFFFFFFu;
            local[k] = lds_tally<8>(cnt, (int)(c[k] >> 6), c[k] != 0xFFFFFFFFu);
        }
        const unsigned lfb = lds_tally<1>(cnt + 8, 0, fb);
        __syncthreads();
        if (threadIdx.x < 9 && cnt[threadIdx.x]) {
            unsigned *ctr = bucket_shard(counters, threadIdx.x);
            if (pass == 0) atomicAdd(ctr, cnt[threadIdx.x]);
            else base[threadIdx.x] = atomicAdd(ctr, cnt[threadIdx.x]);
        }
        __syncthreads();
        if (pass == 1) {
            for (int k = 0; k < K; k++) {
                if (c[k] != 0xFFFFFFFFu) items[base[c[k] >> 6] + local[k]] = make_uint2(g, c[k]);
                if (valid) slot[(uint64_t)g * K + k] = c[k] != 0xFFFFFFFFu ? base[c[k] >> 6] + local[k] : 0xFFFFFFFFu;
            }
            if (fb) fallback[base[8] + lfb] = g;
        }
        __syncthreads();
    }
}
// the wide proposals of the fallback list: item e < count[0] * K is candidate cand[e] of block list[e / K]; without
// cand (K = 1), the block's current winner key[list[e]] & 511
extern "C" __global__ void __launch_bounds__(256) zt_bc7_bucket_wide(const uint32_t *cand, const uint32_t *key, const uint32_t *list,
                                                                       const unsigned *count, int K, unsigned *counters, uint2 *items,
                                                                       uint32_t pass) {
    __shared__ unsigned cnt[8], base[8];
    const uint64_t n = (uint64_t)count[0] * (uint64_t)K, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x; b0 < n; b0 += step) {
        if (threadIdx.x < 8) cnt[threadIdx.x] = 0;
        __syncthreads();
        const uint64_t e = b0 + threadIdx.x;
        const bool valid = e < n;
        const uint32_t c = !valid ? 0u : cand ? cand[e] : key[list[e]] & 511u;
        const uint32_t local = lds_tally<8>(cnt, (int)(c >> 6), valid);
        __syncthreads();
        if (threadIdx.x < 8 && cnt[threadIdx.x]) {
            unsigned *ctr = bucket_shard(counters, threadIdx.x);
            if (pass == 0) atomicAdd(ctr, cnt[threadIdx.x]);
            else base[threadIdx.x] = atomicAdd(ctr, cnt[threadIdx.x]);
        }
        __syncthreads();
        if (pass == 1 && valid) items[base[c >> 6] + local] = make_uint2(list[e / (uint64_t)K], c);
        __syncthreads();
    }
}
// shard counts -> run starts: counters[m] = items of mode m, counters[24 + m] = its run's start, counters[8] =
// fallback blocks; each shard counter becomes its sub-run's start (the fallback list starts at 0)
extern "C" __global__ void zt_bc7_bucket_scan(unsigned *counters) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned run = 0;
        for (int m = 0; m < 9; m++) {
            if (m == 8) run = 0;
            const unsigned start = run;
            for (int s = 0; s < ZT_SHARDS; s++) {
                unsigned *ctr = counters + ZT_BUCKET_SHARDS_AT + (m * ZT_SHARDS + s) * ZT_SHARD_STRIDE;
                const unsigned c = *ctr;
                *ctr = run;
                run += c;
            }
            counters[m] = run - start;
            if (m < 8) counters[24 + m] = start;
        }
    }
}

// exact evaluation of the proposals of mode M (lane per proposal), folded into the blocks' keys; the encoded block
// is kept per item (enc, optional) for zt_bc7_finish_learned
template <int M>
ZT_DEV void prop_body(const uint4 *blocks, const uint2 *items, const unsigned *counters, uint32_t *key, uint4 *enc) {
    const uint32_t begin = counters[24 + M], n = counters[M], step = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const uint2 it = items[begin + i];
        uint32_t px[16];
        load_block(blocks, it.x, px);
        uint64_t lo, hi;
        const int slot = (int)(it.y & 63u);
        int err = -1;
        if constexpr (M == 4) {
            if (slot < 8) err = (slot & 1) ? eval_mode<4, 1>(px, slot, lo, hi) : eval_mode<4, 0>(px, slot, lo, hi);
        } else if (slot < kSlots<M, 0>) {
            err = eval_mode<M, 0>(px, slot, lo, hi);
        }
        if (enc && err >= 0) enc[begin + i] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
        atomicMin(key + it.x, cand_key(err, (int)it.y));
    }
}
#define ZT_PROP(M)                                                                                                   \
    extern "C" __global__ void __launch_bounds__(256) zt_bc7_prop_m##M(const uint4 *blocks, const uint2 *items,               \
                                                                        const unsigned *counters, uint32_t *key, uint4 *enc) {\
        prop_body<M>(blocks, items, counters, key, enc);                                                                       \
    }
ZT_PROP(0)
ZT_PROP(1)
ZT_PROP(2)
ZT_PROP(3)
ZT_PROP(4)
ZT_PROP(5)
ZT_PROP(6)
ZT_PROP(7)
#undef ZT_PROP

// every proposal-searched block's winner: the stored encoding of its winning proposal (fallback blocks, slot
// 0xFFFFFFFF, are encoded by zt_bc7_finish over the fallback list)
extern "C" __global__ void __launch_bounds__(256) zt_bc7_finish_learned(zt_jobs jobs, uint32_t total, int K, const uint32_t *cand,
                                                                          const uint32_t *slot, const uint4 *enc, const uint32_t *key) {
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const uint32_t win = key[g] & 511u;
        uint32_t it = 0xFFFFFFFFu;
        for (int k = 0; k < K; k++) {
            const uint32_t sl = slot[(uint64_t)g * K + k];
            if (sl != 0xFFFFFFFFu && cand[(uint64_t)g * K + k] == win) it = sl;
        }
        if (it == 0xFFFFFFFFu) continue;
        const zt_job J = job_at(jobs, zt_find_job(jobs, g));
        *(uint4 *)(J.dst + (uint64_t)(g - J.first) * 16) = enc[it];
    }
}
