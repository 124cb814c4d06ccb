// zt_bc7.hip — BC7 encoder for gfx950: the reference search (zap_tex.h:755-784) as three compacted stages.
//
//   stage 6  (all blocks)       one lane per block: mode 6 (PCA, p-bit fit, 2x least squares). Blocks whose
//                               SSE exceeds 64 are appended (wave ballot, one atomic per wave) to list 5.
//   stage 5  (list 5)           one lane per (block, rotation): mode 5; a 4-lane shuffle picks the rotation
//                               (strict <, lowest rotation wins). Opaque blocks are appended to list 1.
//   stage 1  (list 1)           one wave per block, one lane per partition: mode-1 partition estimate, wave
//                               rank of (estimate, partition), exact 2-subset fits on the top-K lanes, and a
//                               wave argmin over (fit SSE, rank).
// The modes-5/1 scores are the fit SSEs: for those encoders the decoded block's SSE equals the fit SSE
// (endpoint swap with index inversion is symmetric for W2/W3, rotation permutes channels back), so the
// reference's decode-and-measure (zap_tex.h:761, :780) compares the same integers.
#include "zt_bc7_core.h"
#include "zt_kernels.h"


// ===================================================================================== kernels

// The kernels' bodies take the workgroup index and grid size as arguments (bid, nblk), so the persistent batch
// kernel can run any workgroup of them as a task.

// Stage 6: every block. Writes the mode-6 block; blocks with SSE > 64 go to list 5 with their pixels.
namespace {
ZT_DEV void stage6_body(const zt_jobs &jobs, uint32_t total, const zt_bc7_lists &L, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t step = nblk * blockDim.x;
    for (uint32_t base = bid * blockDim.x; base < total; base += step) {
        uint32_t g = base + tid;
        bool valid = g < total, pred = false;
        int err = 0, opaque = 0, j = 0;
        uint32_t px[16];
        if (valid) {
            j = zt_find_job(jobs, g);
            const zt_job J = job_at(jobs, j);
            uint32_t b = g - J.first;
            int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw);
            gather_block(J.src, J.w, J.h, J.stride, bx, by, px);
            uint64_t lo, hi;
            err = mode6(px, lo, hi);
            store16(J.dst + (uint64_t)b * 16, lo, hi);
            uint32_t amin = 255;
#pragma unroll
            for (int i = 0; i < 16; i++) amin = min(amin, px[i] >> 24);
            opaque = amin == 255;
            pred = err > 64;
        }
        unsigned *c = shard_counter(L.counters, base + (tid & ~63u));
        int slot = wave_append(pred, c);
        if (slot >= 0) {
            slot += (int)((uint32_t)((c - L.counters) / ZT_SHARD_STRIDE) * L.cap5);
            L.meta5[slot] = make_uint4(g, (uint32_t)err, (uint32_t)j, (uint32_t)opaque);
            uint4 *d = L.px5 + (uint64_t)slot * 4;
            d[0] = make_uint4(px[0], px[1], px[2], px[3]); d[1] = make_uint4(px[4], px[5], px[6], px[7]);
            d[2] = make_uint4(px[8], px[9], px[10], px[11]); d[3] = make_uint4(px[12], px[13], px[14], px[15]);
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256, 4) zt_bc7_stage6(zt_jobs jobs, uint32_t total, zt_bc7_lists L) {
    stage6_body(jobs, total, L, blockIdx.x, gridDim.x);
}

// Stage 6 with the refinement steps queued (mode6_first / mode6_refit).  A step whose endpoints quantise to the ones
// the block has ends the reference's loop without a fit, so on textures a block needs its second fit about 40% of
// the time and its third about 10%, where a wave of 64 blocks one lane each runs all three for every lane.  Here a
// wave (one per workgroup) runs the first pass over chunks of 64 blocks; the blocks that go on wait in the wave's
// LDS queue, and whenever 64 are queued the wave runs the next pass over those 64 at once (a block that goes on
// again waits in the second queue).  The waves take `kc` chunks at a time from a cursor until none are left, so they
// all run to the end of the stage (a fixed share per wave left the chip part idle while the last waves ran), then
// run what is left in their queues.  There is a cursor per shard (chunk c is in shard c mod ZT_SHARDS, wave b takes
// from shard b mod ZT_SHARDS), each on its own cache line: atomics on one address serialise, and one cursor for the
// whole stage cost more than the idle tail.
namespace {
constexpr uint32_t kRefitQ = 128;   // queue slots: < 64 waiting + 64 more from a chunk or a pass
struct RefitQueue { uint32_t g[kRefitQ], j[kRefitQ], err[kRefitQ], lo[kRefitQ], hi[kRefitQ]; };

ZT_DEV void list5_put(const zt_bc7_lists &L, int slot, uint32_t g, int err, int j, const uint32_t px[16]) {
    uint32_t amin = 255;
#pragma unroll
    for (int i = 0; i < 16; i++) amin = min(amin, px[i] >> 24);
    L.meta5[slot] = make_uint4(g, (uint32_t)err, (uint32_t)j, amin == 255 ? 1u : 0u);
    uint4 *d = L.px5 + (uint64_t)slot * 4;
    d[0] = make_uint4(px[0], px[1], px[2], px[3]); d[1] = make_uint4(px[4], px[5], px[6], px[7]);
    d[2] = make_uint4(px[8], px[9], px[10], px[11]); d[3] = make_uint4(px[12], px[13], px[14], px[15]);
}
// queue the lanes with pred at tail (wave-uniform); returns how many
ZT_DEV uint32_t refit_push(RefitQueue &Q, uint32_t tail, bool pred, uint32_t g, int j, int err, uint64_t nqp) {
    const uint64_t m = __ballot(pred);
    if (pred) {
        const uint32_t k = (tail + lanes_below(m)) % kRefitQ;
        Q.g[k] = g; Q.j[k] = (uint32_t)j; Q.err[k] = (uint32_t)err; Q.lo[k] = (uint32_t)nqp; Q.hi[k] = (uint32_t)(nqp >> 32);
    }
    __builtin_amdgcn_wave_barrier();
    return __builtin_amdgcn_readfirstlane((uint32_t)__popcll(m));
}
// the next pass over count (<= 64) queued blocks from head; blocks that go on are queued in N at ntail.  A block
// enters list 5 in the shard of its own index (g >> 6), as in the first pass.
template <bool LAST>
ZT_DEV void refit_round(const zt_jobs &jobs, const zt_bc7_lists &L, const RefitQueue &Q, uint32_t head, uint32_t count, RefitQueue &N,
                        uint32_t &ntail) {
    const uint32_t lane = (uint32_t)__lane_id(), k = (head + lane) % kRefitQ;
    const bool act = lane < count;
    uint32_t g = 0, px[16];
    int j = 0, err = 0;
    uint64_t nqp = 0;
    bool more = false, improved = false;
    if (act) {
        g = Q.g[k]; j = (int)Q.j[k]; err = (int)Q.err[k]; nqp = (uint64_t)Q.lo[k] | (uint64_t)Q.hi[k] << 32;
        const zt_job J = job_at(jobs, j);
        const uint32_t b = g - J.first;
        gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
        uint64_t lo, hi;
        err = mode6_refit<LAST>(px, err, nqp, lo, hi, improved, more);
        if (improved) store16(J.dst + (uint64_t)b * 16, lo, hi);
    }
    __builtin_amdgcn_wave_barrier();
    const int slot = wave_append_sharded(act && !more && err > 64, L.counters, L.cap5, g);
    if (slot >= 0) list5_put(L, slot, g, err, j, px);
    if (!LAST) ntail += refit_push(N, ntail, more, g, j, err, nqp);
}
// blockDim.x == 64, gridDim.x >= min(chunks, ZT_SHARDS) (every shard with chunks has a wave); the cursors
// L.counters[ZT_BC7_S6_CURSOR + s * ZT_SHARD_STRIDE] start at 0
ZT_DEV void stage6q_body(const zt_jobs &jobs, uint32_t total, const zt_bc7_lists &L, uint32_t bid, uint32_t kc) {
    __shared__ RefitQueue Q2, Q3;
    const uint32_t lane = zt_tid() & 63u, nch = (total + 63) / 64, sh = bid & (ZT_SHARDS - 1);
    // the shard's chunks are sh, sh + ZT_SHARDS, ...: n of them
    const uint32_t n = nch > sh ? (nch - sh + ZT_SHARDS - 1) / ZT_SHARDS : 0u;
    unsigned *cursor = L.counters + ZT_BC7_S6_CURSOR + sh * ZT_SHARD_STRIDE;
    uint32_t h2 = 0, t2 = 0, h3 = 0, t3 = 0;
    for (;;) {
        uint32_t i0 = 0;
        if (lane == 0) i0 = __hip_atomic_fetch_add(cursor, kc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i0 = __builtin_amdgcn_readfirstlane(__shfl(i0, 0));
        if (i0 >= n) break;
        for (uint32_t i = i0, i1 = min(n, i0 + kc); i < i1; i++) {
            const uint32_t c = i * ZT_SHARDS + sh;
            const uint32_t g = c * 64 + lane;
            const bool valid = g < total;
            bool more = false;
            int err = 0, j = 0;
            uint64_t nqp = 0;
            uint32_t px[16];
            if (valid) {
                j = zt_find_job(jobs, g);
                const zt_job J = job_at(jobs, j);
                const uint32_t b = g - J.first;
                gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
                uint64_t lo, hi;
                err = mode6_first(px, lo, hi, more, nqp);
                store16(J.dst + (uint64_t)b * 16, lo, hi);
            }
            const int slot = wave_append(valid && !more && err > 64, L.counters + sh * ZT_SHARD_STRIDE);
            if (slot >= 0) list5_put(L, slot + (int)(sh * L.cap5), g, err, j, px);
            t2 += refit_push(Q2, t2, more, g, j, err, nqp);
            if (t2 - h2 >= 64) { refit_round<false>(jobs, L, Q2, h2, 64, Q3, t3); h2 += 64; }
            if (t3 - h3 >= 64) { refit_round<true>(jobs, L, Q3, h3, 64, Q3, t3); h3 += 64; }
        }
    }
    while (t2 != h2) {
        const uint32_t n = min(64u, t2 - h2);
        refit_round<false>(jobs, L, Q2, h2, n, Q3, t3);
        h2 += n;
    }
    while (t3 != h3) {
        const uint32_t n = min(64u, t3 - h3);
        refit_round<true>(jobs, L, Q3, h3, n, Q3, t3);
        h3 += n;
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(64, 4) zt_bc7_stage6q(zt_jobs jobs, uint32_t total, zt_bc7_lists L, uint32_t kc) {
    stage6q_body(jobs, total, L, blockIdx.x, kc);
}

// Stage 5: lane = (list-5 entry, rotation). Keeps mode 5 when it beats mode 6; opaque blocks go to list 1.
namespace {
ZT_DEV void stage5_body(const zt_jobs &jobs, const zt_bc7_lists &L, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t n5 = shard_total(L.counters), items = n5 * 4, step = nblk * blockDim.x;
    for (uint32_t base = bid * blockDim.x; base < items; base += step) {
        uint32_t item = base + tid, e = item >> 2;
        int rot = (int)(item & 3);
        bool valid = item < items;
        uint32_t px[16];
        int cq[2][3] = { { 0 } }, aq[2][3] = { { 0 } };
        uint32_t ci = 0, ai = 0;
        int err = 0x3FFFFFF;
        uint4 meta = make_uint4(0, 0, 0, 0);
        if (valid) {
            e = shard_slot(L.counters, L.cap5, e);
            meta = L.meta5[e];
            const uint4 *s = L.px5 + (uint64_t)e * 4;
#pragma unroll
            for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
            err = mode5_rotation(px, rot, cq, aq, ci, ai);
        }
        // best rotation of the 4-lane group: smallest SSE, ties to the lowest rotation
        int key = err * 4 + rot, o;
        o = __shfl_xor(key, 1); key = min(key, o);
        o = __shfl_xor(key, 2); key = min(key, o);
        bool winner = valid && (key & 3) == rot;
        int e5 = key >> 2;
        bool improved = winner && e5 < (int)meta.y;
        if (improved) {
            uint64_t lo, hi;
            pack_mode5(rot, cq, aq, ci, ai, lo, hi);
            const zt_job J = job_at(jobs, (int)meta.z);
            store16(J.dst + (uint64_t)(meta.x - J.first) * 16, lo, hi);
        }
        unsigned *c = shard_counter(L.counters + ZT_BC7_LIST1, base + (tid & ~63u));   // <= 16 appends per wave
        int slot = wave_append(winner && meta.w, c);
        if (slot >= 0)
            L.meta1[(uint32_t)((c - L.counters - ZT_BC7_LIST1) / ZT_SHARD_STRIDE) * L.cap1 + (uint32_t)slot] =
                make_uint4(e, (uint32_t)(improved ? e5 : (int)meta.y), 0, 0);
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_stage5(zt_jobs jobs, zt_bc7_lists L) {
    stage5_body(jobs, L, blockIdx.x, gridDim.x);
}

// Stage 1: one wave per list-1 entry, lane = partition.
namespace {
ZT_DEV void stage1_body(const zt_jobs &jobs, const zt_bc7_lists &L, int m1k, uint32_t bid, uint32_t nblk) {
    const uint32_t n1 = shard_total(L.counters + ZT_BC7_LIST1);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t nw = nblk * (blockDim.x >> 6);
    for (uint32_t e = bid * (blockDim.x >> 6) + wv; e < n1; e += nw) {
        uint4 m = L.meta1[shard_slot(L.counters + ZT_BC7_LIST1, L.cap1, e)];
        uint4 meta = L.meta5[m.x];
        int cur = (int)m.y;
        uint32_t px[16];
        const uint4 *s = L.px5 + (uint64_t)m.x * 4;
#pragma unroll
        for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
        float tot[9];
#pragma unroll
        for (int j = 0; j < 9; j++) tot[j] = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float r = chanf(px[i], 0), g = chanf(px[i], 1), b = chanf(px[i], 2);
            float v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
#pragma unroll
            for (int j = 0; j < 9; j++) tot[j] += v[j];
        }
        float est = mode1_estimate(px, tot, lane);
        if (est == 0.0f) est = 0.0f;  // -0 ranks like +0 (the reference compares with <)
        uint32_t fb = __float_as_uint(est);
        fb = (fb & 0x80000000u) ? ~fb : (fb | 0x80000000u);  // order-preserving float -> uint
        // rank of (estimate, partition) among the 64 lanes
        int rank = 0;
        for (int k = 0; k < 64; k++) {
            uint32_t ofb = __shfl(fb, k);
            rank += ofb < fb || (ofb == fb && k < lane);
        }
        int fe = 0x1FFFFFF;
        if (rank < m1k) {
            uint64_t lo, hi;
            fe = mode1(px, lane, false, lo, hi);
        }
        // first candidate (in rank order) with the smallest fit SSE
        int k2 = fe * 64 + rank;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) k2 = min(k2, __shfl_xor(k2, off));
        if ((k2 >> 6) < cur && (k2 & 63) == rank) {
            uint64_t lo, hi;
            mode1(px, lane, true, lo, hi);
            const zt_job J = job_at(jobs, (int)meta.z);
            store16(J.dst + (uint64_t)(meta.x - J.first) * 16, lo, hi);
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_stage1(zt_jobs jobs, zt_bc7_lists L, int m1k) {
    stage1_body(jobs, L, m1k, blockIdx.x, gridDim.x);
}

extern "C" __global__ void zt_zero_u32(unsigned *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" __global__ void zt_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}

// ---- stage 1 split over three kernels for M1K <= 8 (every lane busy):
//   rank  one wave per list-1 entry: partition estimates, the top-M1K partitions in rank order by M1K wave minima
//   fit   one lane per (entry, candidate, subset): the exact subset fit; SSE and the subset's quantised state
//   pick  one lane per entry: first candidate in rank order with the smallest SSE below the current one, packed
// Capacity: entry e owns cand[8e..8e+7] and items [16e, 16e + 16).
namespace {
ZT_DEV void s1_rank_body(const zt_bc7_lists &L, int m1k, uint8_t *cand, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t n1 = shard_total(L.counters + ZT_BC7_LIST1);
    const int lane = tid & 63, wv = tid >> 6;
    const uint32_t nw = nblk * (blockDim.x >> 6);
    for (uint32_t e = bid * (blockDim.x >> 6) + wv; e < n1; e += nw) {
        const uint4 m = L.meta1[shard_slot(L.counters + ZT_BC7_LIST1, L.cap1, e)];
        uint32_t px[16];
        const uint4 *s = L.px5 + (uint64_t)m.x * 4;
#pragma unroll
        for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
        float tot[9];
#pragma unroll
        for (int j = 0; j < 9; j++) tot[j] = 0.0f;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float r = chanf(px[i], 0), g = chanf(px[i], 1), b = chanf(px[i], 2);
            float v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
#pragma unroll
            for (int j = 0; j < 9; j++) tot[j] += v[j];
        }
        float est = mode1_estimate(px, tot, lane);
        if (est == 0.0f) est = 0.0f;
        uint32_t fb = __float_as_uint(est);
        fb = (fb & 0x80000000u) ? ~fb : (fb | 0x80000000u);
        // the m1k smallest (estimate, partition) pairs in ascending order, one wave minimum each: the partition of
        // rank k is the lowest lane holding the smallest key among the lanes not yet taken
        uint64_t taken = 0;
        for (int k = 0; k < m1k; k++) {
            const bool free = !((taken >> lane) & 1ull);
            const uint32_t key = free ? fb : 0xFFFFFFFFu;
            uint32_t mn = key;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
            const int sel = __ffsll((unsigned long long)__ballot(free && key == mn)) - 1;
            taken |= 1ull << sel;
            if (lane == 0) cand[(uint64_t)e * 8 + k] = (uint8_t)sel;
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_s1_rank(zt_bc7_lists L, int m1k, uint8_t *cand) {
    s1_rank_body(L, m1k, cand, blockIdx.x, gridDim.x);
}

namespace {
ZT_DEV void s1_fit_body(const zt_bc7_lists &L, int m1k, const uint8_t *cand, int *sub_err, uint4 *sub_state, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t n1 = shard_total(L.counters + ZT_BC7_LIST1), items = n1 * (uint32_t)m1k * 2, step = nblk * blockDim.x;
    for (uint32_t it = bid * blockDim.x + tid; it < items; it += step) {
        const uint32_t e = it / (2u * m1k), k = (it >> 1) % (uint32_t)m1k, sb = it & 1;
        const uint4 m = L.meta1[shard_slot(L.counters + ZT_BC7_LIST1, L.cap1, e)];
        uint32_t px[16];
        const uint4 *s = L.px5 + (uint64_t)m.x * 4;
#pragma unroll
        for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
        const int part = cand[(uint64_t)e * 8 + k];
        const uint32_t m1 = kP2[part], mask = sb ? m1 : (~m1 & 0xFFFFu);
        int q[2][4], pb[2];
        uint64_t idx;
        ChanWords T;
        chan_words(px, T);
        const int err = encode_subset<3, 7, 3, true, false>(px, T, mask, __popc(mask), q, pb, idx);
        const uint64_t o = (uint64_t)e * 16 + k * 2 + sb;
        sub_err[o] = err;
        // the 6 endpoint values (6 bits each: endpoint e, channel c at bit 18e + 6c) and the shared p-bit at bit 36
        const uint64_t qs = (uint64_t)q[0][0] | (uint64_t)q[0][1] << 6 | (uint64_t)q[0][2] << 12 | (uint64_t)q[1][0] << 18 |
                            (uint64_t)q[1][1] << 24 | (uint64_t)q[1][2] << 30 | (uint64_t)pb[0] << 36;
        sub_state[o] = make_uint4((uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)qs, (uint32_t)(qs >> 32));
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_s1_fit(zt_bc7_lists L, int m1k, const uint8_t *cand, int *sub_err, uint4 *sub_state) {
    s1_fit_body(L, m1k, cand, sub_err, sub_state, blockIdx.x, gridDim.x);
}

namespace {
// the mode-1 block of partition `part` from its two subsets' fits: indices (4 bits per pixel, the subset's pixels
// only) and endpoint words (6 bits per value: endpoint e, channel c at bit 18e + 6c; the shared p-bit at bit 36),
// anchors' index MSBs cleared by swapping endpoints (zap_tex.h:745-748 and the pack that follows)
ZT_DEV void mode1_pack(int part, uint64_t idx0, uint64_t qs0, uint64_t idx1, uint64_t qs1, uint64_t &lo, uint64_t &hi) {
    const int anc = kA2[part];
    int q[2][2][3], pb[2] = { (int)((qs0 >> 36) & 1), (int)((qs1 >> 36) & 1) };
    uint64_t full = idx0 | idx1;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            q[0][x][c] = (int)((qs0 >> (18 * x + 6 * c)) & 63);
            q[1][x][c] = (int)((qs1 >> (18 * x + 6 * c)) & 63);
        }
    const uint32_t m1 = kP2[part], m0 = ~m1 & 0xFFFFu;
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
            for (int x = 0; x < 2; x++) b.put(8 + 6 * (4 * c + 2 * s + x), (uint64_t)q[s][x][c], 6);
    b.put(80, (uint64_t)pb[0], 1); b.put(81, (uint64_t)pb[1], 1);
    int pos = 82;
    for (int i = 0; i < 16; i++) {
        const int nb = 3 - (i == 0 || i == anc);
        b.put(pos, (full >> (4 * i)) & 7, nb);
        pos += nb;
    }
    lo = b.lo; hi = b.hi;
}
ZT_DEV void s1_pick_body(const zt_jobs &jobs, const zt_bc7_lists &L, int m1k, const uint8_t *cand, const int *sub_err, const uint4 *sub_state,
                         uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t n1 = shard_total(L.counters + ZT_BC7_LIST1), step = nblk * blockDim.x;
    for (uint32_t e = bid * blockDim.x + tid; e < n1; e += step) {
        const uint4 m = L.meta1[shard_slot(L.counters + ZT_BC7_LIST1, L.cap1, e)];
        int best = (int)m.y, bk = -1;
        for (int k = 0; k < m1k; k++) {
            const int err = sub_err[(uint64_t)e * 16 + 2 * k] + sub_err[(uint64_t)e * 16 + 2 * k + 1];
            if (err < best) { best = err; bk = k; }
        }
        if (bk < 0) continue;
        const int part = cand[(uint64_t)e * 8 + bk];
        const uint4 s0 = sub_state[(uint64_t)e * 16 + 2 * bk], s1 = sub_state[(uint64_t)e * 16 + 2 * bk + 1];
        uint64_t lo, hi;
        mode1_pack(part, (uint64_t)s0.x | (uint64_t)s0.y << 32, (uint64_t)s0.z | (uint64_t)s0.w << 32,
                   (uint64_t)s1.x | (uint64_t)s1.y << 32, (uint64_t)s1.z | (uint64_t)s1.w << 32, lo, hi);
        const uint4 meta = L.meta5[m.x];
        const zt_job J = job_at(jobs, (int)meta.z);
        store16(J.dst + (uint64_t)(meta.x - J.first) * 16, lo, hi);
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_s1_pick(zt_jobs jobs, zt_bc7_lists L, int m1k, const uint8_t *cand, const int *sub_err,
                                                                   const uint4 *sub_state) {
    s1_pick_body(jobs, L, m1k, cand, sub_err, sub_state, blockIdx.x, gridDim.x);
}

// The fits and the pick in one launch, when a block's 2 * M1K fits make whole lane groups of a wave (M1K = 1, 2, 4 or
// 8): the group's lanes hand their SSEs and states to the winner by shuffles instead of a second kernel reading them
// back.  The winner is the reference's: the first candidate in rank order whose two subsets' SSE sum is the smallest,
// if that is below the block's best so far (zap_tex.h:771-783).
namespace {
ZT_DEV void s1_fitpick_body(const zt_jobs &jobs, const zt_bc7_lists &L, int m1k, const uint8_t *cand, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t n1 = shard_total(L.counters + ZT_BC7_LIST1), items = n1 * (uint32_t)m1k * 2, step = nblk * blockDim.x;
    const int G = 2 * m1k;
    // a lane group (2 M1K lanes, aligned in the wave) is one block's items, all in or all past `items`
    for (uint32_t it = bid * blockDim.x + tid; it < items; it += step) {
        const uint32_t e = it / (2u * m1k), k = (it >> 1) % (uint32_t)m1k, sb = it & 1;
        const uint4 m = L.meta1[shard_slot(L.counters + ZT_BC7_LIST1, L.cap1, e)];
        uint32_t px[16];
        const uint4 *s = L.px5 + (uint64_t)m.x * 4;
#pragma unroll
        for (int r = 0; r < 4; r++) { uint4 v = s[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
        // the block's destination, loaded now so the two dependent loads overlap the fit (this launch is about one
        // wave per SIMD: nothing else hides their latency)
        const uint4 meta = L.meta5[m.x];
        const zt_job J = job_at(jobs, (int)meta.z);
        uint8_t *const dst = J.dst + (uint64_t)(meta.x - J.first) * 16;
        const int part = cand[(uint64_t)e * 8 + k];
        const uint32_t m1 = kP2[part], mask = sb ? m1 : (~m1 & 0xFFFFu);
        int q[2][4], pb[2];
        uint64_t idx;
        ChanWords T;
        chan_words(px, T);
        const int err = encode_subset<3, 7, 3, true, false>(px, T, mask, __popc(mask), q, pb, idx);
        const uint64_t qs = (uint64_t)q[0][0] | (uint64_t)q[0][1] << 6 | (uint64_t)q[0][2] << 12 | (uint64_t)q[1][0] << 18 |
                            (uint64_t)q[1][1] << 24 | (uint64_t)q[1][2] << 30 | (uint64_t)pb[0] << 36;
        // candidate k's SSE (its subsets sit on lanes 2k and 2k + 1), then (SSE, k) smallest over the group
        const int tot = err + __shfl_xor(err, 1);
        uint32_t key = tot < (int)m.y ? (uint32_t)tot << 4 | k : 0xFFFFFFFFu;   // SSE < 2^22
        for (int off = 2; off < G; off <<= 1) key = min(key, (uint32_t)__shfl_xor((int)key, off));
        // the other subset's state, for the winner's first lane
        const uint32_t o0 = (uint32_t)__shfl_down((int)(uint32_t)idx, 1), o1 = (uint32_t)__shfl_down((int)(uint32_t)(idx >> 32), 1);
        const uint32_t o2 = (uint32_t)__shfl_down((int)(uint32_t)qs, 1), o3 = (uint32_t)__shfl_down((int)(uint32_t)(qs >> 32), 1);
        if (key != 0xFFFFFFFFu && (key & 15u) == k && sb == 0) {
            uint64_t lo, hi;
            mode1_pack(part, idx, qs, (uint64_t)o0 | (uint64_t)o1 << 32, (uint64_t)o2 | (uint64_t)o3 << 32, lo, hi);
            store16(dst, lo, hi);
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc7_s1_fitpick(zt_jobs jobs, zt_bc7_lists L, int m1k, const uint8_t *cand) {
    s1_fitpick_body(jobs, L, m1k, cand, blockIdx.x, gridDim.x);
}
