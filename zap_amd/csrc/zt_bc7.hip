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

// Stage 6: every block. Writes the mode-6 block; blocks with SSE > 64 go to list 5 with their pixels.
extern "C" __global__ void __launch_bounds__(256) zt_bc7_stage6(zt_jobs jobs, uint32_t total, zt_bc7_lists L) {
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += step) {
        uint32_t g = base + threadIdx.x;
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
        int slot = wave_append(pred, &L.counters[0]);
        if (slot >= 0) {
            L.meta5[slot] = make_uint4(g, (uint32_t)err, (uint32_t)j, (uint32_t)opaque);
            uint4 *d = L.px5 + (uint64_t)slot * 4;
            d[0] = make_uint4(px[0], px[1], px[2], px[3]); d[1] = make_uint4(px[4], px[5], px[6], px[7]);
            d[2] = make_uint4(px[8], px[9], px[10], px[11]); d[3] = make_uint4(px[12], px[13], px[14], px[15]);
        }
    }
}

// Stage 5: lane = (list-5 entry, rotation). Keeps mode 5 when it beats mode 6; opaque blocks go to list 1.
extern "C" __global__ void __launch_bounds__(256) zt_bc7_stage5(zt_jobs jobs, zt_bc7_lists L) {
    const uint32_t n5 = L.counters[0], items = n5 * 4, step = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < items; base += step) {
        uint32_t item = base + threadIdx.x, e = item >> 2;
        int rot = (int)(item & 3);
        bool valid = item < items;
        uint32_t px[16];
        int cq[2][3] = { { 0 } }, aq[2][3] = { { 0 } };
        uint32_t ci = 0, ai = 0;
        int err = 0x3FFFFFF;
        uint4 meta = make_uint4(0, 0, 0, 0);
        if (valid) {
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
        int slot = wave_append(winner && meta.w, &L.counters[1]);
        if (slot >= 0) L.meta1[slot] = make_uint4(e, (uint32_t)(improved ? e5 : (int)meta.y), 0, 0);
    }
}

// Stage 1: one wave per list-1 entry, lane = partition.
extern "C" __global__ void __launch_bounds__(256) zt_bc7_stage1(zt_jobs jobs, zt_bc7_lists L, int m1k) {
    const uint32_t n1 = L.counters[1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t e = blockIdx.x * (blockDim.x >> 6) + wv; e < n1; e += nw) {
        uint4 m = L.meta1[e];
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

extern "C" __global__ void zt_zero_u32(unsigned *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" __global__ void zt_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) p[i] = v;
}
