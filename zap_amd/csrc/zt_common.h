// zt_common.h — device helpers shared by the gfx950 texture kernels.
//
// Exactness rules every kernel follows (the output bytes must equal the reference encoder's):
//   * all TUs are compiled with -ffp-contract=off and HIP's correctly rounded fp32 division, so each
//     float expression rounds exactly as the reference's C does (zap_tex.h is pinned contraction-off);
//   * float accumulations run lane-serially in the reference's pixel order; integer SSE sums and
//     argmins are order-free, with ties going to the lowest index (the reference's strict `<`);
//   * (int)float casts of unbounded values use x86 cvttss2si semantics (cvt_i32), not the AMDGPU
//     saturating convert.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZT_DEV static __device__ __forceinline__

// One unit of encode work: a block-compressed image (or one mip level of one).  Blocks of all jobs
// are numbered consecutively; job j owns global blocks [first, first + bw*bh).
struct zt_job {
    const uint8_t *src;   // RGBA8 (or RGBA float for BC6H) rows
    uint8_t *dst;         // blocks, row-major by block
    uint64_t stride;      // bytes per source row
    uint32_t first;       // global index of this job's first block
    int32_t w, h, bw;     // pixels, pixels, blocks per row
    int32_t fmt;          // zap_bc_format (0 BC1, 1 BC3, 2 BC4, 3 BC5, 4 BC7); 5 BC6H, 6 ASTC
    int32_t pad;
};

// Job tables travel by value in the kernel arguments (up to ZT_INLINE_JOBS) or as a device array (`ext`).
#define ZT_INLINE_JOBS 16
struct zt_jobs {
    const zt_job *ext;    // device table, or null to use `inl`
    int32_t n, pad;
    zt_job inl[ZT_INLINE_JOBS];
};
ZT_DEV zt_job job_at(const zt_jobs &J, int i) { return J.ext ? J.ext[i] : J.inl[i]; }
ZT_DEV uint32_t job_first(const zt_jobs &J, int i) { return J.ext ? J.ext[i].first : J.inl[i].first; }

// job owning global block g (jobs sorted by `first`); binary search
ZT_DEV int zt_find_job(const zt_jobs &J, uint32_t g) {
    int lo = 0, hi = J.n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (job_first(J, mid) <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// x86 (int)f: out-of-range or NaN gives INT_MIN (zap_tex.h:46, :842, :1187 cast unclamped values)
ZT_DEV int cvt_i32(float f) {
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000u;
}

// threadIdx.x as a value the optimiser cannot move.  The persistent batch kernel runs many kernels' bodies in one
// loop; with threadIdx.x itself, loop-invariant code motion hoists each body's thread-index arithmetic to kernel
// entry, where it stays live through every other body (their registers would add up instead of overlapping).
ZT_DEV uint32_t zt_tid() {
    uint32_t v = threadIdx.x;
    asm volatile("" : "+v"(v));
    return v;
}
// RN(v / 255) from one product and one FMA correction (the correction of a product by the correctly rounded
// reciprocal gives the correctly rounded quotient): equal to the division for every finite float except v = -0,
// which gives +0 (every caller adds 0.5 or subtracts a p-bit next, which erases the sign of a zero).  Checked
// exhaustively against v / 255.0f by tests/quant_exhaustive.c.
ZT_DEV float div255(float v) {
    const float R = 1.0f / 255.0f, q0 = v * R;
    return __builtin_fmaf(__builtin_fmaf(-q0, 255.0f, v), R, q0);
}
ZT_DEV int chan(uint32_t p, int c) { return (int)((p >> (8 * c)) & 255u); }
ZT_DEV float chanf(uint32_t p, int c) { return (float)chan(p, c); }
ZT_DEV int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
ZT_DEV uint32_t dot4(uint32_t a, uint32_t b) { return __builtin_amdgcn_udot4(a, b, 0u, false); }
ZT_DEV uint32_t dot4acc(uint32_t a, uint32_t b, uint32_t acc) { return __builtin_amdgcn_udot4(a, b, acc, false); }
// v_perm_b32: byte i of the result = byte sel.byte[i] of the 8 bytes {hi:lo} (0-3 from lo, 4-7 from hi; 0x0C = 0)
ZT_DEV uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) { return __builtin_amdgcn_perm(hi, lo, sel); }
// v as an opaque register value: the optimiser cannot fold it into surrounding arithmetic.  Palette keys are
// (dot4 << s) + K; left visible, K's constant part is re-associated into a separate OR and SUB, three
// instructions per key instead of one v_lshl_add_u32.
ZT_DEV int no_fold(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    int r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
#else
    return v;  // the host pass (and CPU builds of these sources) only need the value
#endif
}

// 16 pixels of block (bx, by), edge-clamped like zap_tex.h:1482-1488, packed RGBA8 (R in the low byte).  Rows are
// read as 16-byte vectors when the block lies inside the image and the base and stride are 16-byte aligned (one
// test for the block, so no row pointer stays live across a branch), else byte by byte.
ZT_DEV void gather_block(const uint8_t *src, int w, int h, uint64_t stride, int bx, int by, uint32_t px[16]) {
    int x0 = bx * 4;
    const bool fast = x0 + 3 < w && (((uintptr_t)src | stride) & 15) == 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        int y = by * 4 + r;
        if (y >= h) y = h - 1;
        const uint8_t *row = src + (uint64_t)y * stride;
        if (fast) {
            uint4 v = *(const uint4 *)(row + x0 * 4);
            px[r * 4 + 0] = v.x; px[r * 4 + 1] = v.y; px[r * 4 + 2] = v.z; px[r * 4 + 3] = v.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                int x = x0 + c;
                if (x >= w) x = w - 1;
                const uint8_t *p = row + x * 4;
                px[r * 4 + c] = (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
            }
        }
    }
}

// 128-bit little-endian bit writer whose positions fold to constants when fully unrolled
struct Bits128 {
    uint64_t lo, hi;
    __device__ __forceinline__ void put(int pos, uint64_t v, int n) {  // v < 2^n, n <= 32
        if (pos >= 64) hi |= v << (pos - 64);
        else if (pos + n <= 64) lo |= v << pos;
        else { lo |= v << pos; hi |= v >> (64 - pos); }
    }
};
ZT_DEV uint64_t get_bits(uint64_t lo, uint64_t hi, int pos, int n) {
    uint64_t v;
    if (pos >= 64) v = hi >> (pos - 64);
    else if (pos + n <= 64) v = lo >> pos;
    else v = (lo >> pos) | (hi << (64 - pos));
    return n >= 64 ? v : v & ((1ull << n) - 1);
}
ZT_DEV void store16(uint8_t *dst, uint64_t lo, uint64_t hi) {
    uint4 v;
    v.x = (uint32_t)lo; v.y = (uint32_t)(lo >> 32); v.z = (uint32_t)hi; v.w = (uint32_t)(hi >> 32);
    *(uint4 *)dst = v;
}
ZT_DEV void store8(uint8_t *dst, uint64_t v) {
    uint2 t; t.x = (uint32_t)v; t.y = (uint32_t)(v >> 32);
    *(uint2 *)dst = t;
}

// Lists that many waves append to keep one counter per shard, each counter on its own 256-byte line (atomics on
// one line complete one after another, about 10 ns apart, so a single counter paces a short kernel): the wave
// holding items [64k, 64k + 64) appends to shard k mod ZT_SHARDS.  Shard s owns slots [s * cap, s * cap + count_s);
// with one append per lane a shard needs cap = 64 * ceil(items / 512).
#define ZT_SHARDS 8
#define ZT_SHARD_STRIDE 64   // counters (uint32) between consecutive shards
ZT_DEV unsigned *shard_counter(unsigned *counts, uint32_t item) { return counts + ((item >> 6) & (ZT_SHARDS - 1)) * ZT_SHARD_STRIDE; }
ZT_DEV uint32_t shard_total(const unsigned *counts) {
    uint32_t n = 0;
#pragma unroll
    for (int s = 0; s < ZT_SHARDS; s++) n += counts[s * ZT_SHARD_STRIDE];
    return n;
}
// slot of flat entry e (shards in order)
ZT_DEV uint32_t shard_slot(const unsigned *counts, uint32_t cap, uint32_t e) {
#pragma unroll
    for (int s = 0; s < ZT_SHARDS; s++) {
        const uint32_t c = counts[s * ZT_SHARD_STRIDE];
        if (e < c) return (uint32_t)s * cap + e;
        e -= c;
    }
    return 0xFFFFFFFFu;
}

// set bits of m below this lane (v_mbcnt: no lane mask is materialised)
ZT_DEV uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// wave-level stream compaction: returns this lane's slot in the list (or -1), one atomic per wave
ZT_DEV int wave_append(bool pred, unsigned *counter) {
    uint64_t mask = __ballot(pred);
    if (!mask) return -1;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (!pred) return -1;
    return (int)(base + lanes_below(mask));
}

// the same for lanes whose items lie in different shards (item i in shard (i >> 6) mod ZT_SHARDS): one atomic per
// shard present in the wave; returns the lane's slot in the sharded list (shard * cap + position) or -1
ZT_DEV int wave_append_sharded(bool pred, unsigned *counts, uint32_t cap, uint32_t item) {
    const uint32_t sh = (item >> 6) & (ZT_SHARDS - 1);
    const int lane = __lane_id();
    int slot = -1;
#pragma unroll
    for (uint32_t k = 0; k < ZT_SHARDS; k++) {
        const bool mine = pred && sh == k;
        const uint64_t mask = __ballot(mine);
        if (!mask) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(counts + k * ZT_SHARD_STRIDE, (unsigned)__popcll(mask));
        base = __shfl(base, leader);
        if (mine) slot = (int)(k * cap + base + lanes_below(mask));
    }
    return slot;
}

// LSB-first reader over a 128-bit block (zap_tex.h:367-371)
struct BitReader {
    uint64_t lo, hi;
    int pos;
    __device__ __forceinline__ int take(int n) {
        if (n == 0) return 0;
        int v = (int)get_bits(lo, hi, pos, n);
        pos += n;
        return v;
    }
};
ZT_DEV BitReader reader_at(const uint8_t *b) {
    uint4 raw = *(const uint4 *)b;
    return BitReader{ (uint64_t)raw.x | (uint64_t)raw.y << 32, (uint64_t)raw.z | (uint64_t)raw.w << 32, 0 };
}
