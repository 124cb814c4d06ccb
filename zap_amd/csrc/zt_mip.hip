// zt_mip.hip — 2x2 box-filter mip levels for gfx950 (zap_tex.h:1442-1469).  HBM-bound: each lane makes
// one destination pixel from four source pixels; sRGB levels average in linear light through a 256-entry
// table in LDS and re-encode as the reference's first k with v < thr[k] (the number of thresholds <= v): a
// 4097-entry bucket table in LDS gives the count below v's 1/4096-wide bucket, and one compare adds the bucket's
// own threshold (no bucket of the reference's table holds two; the host checks that when it builds the table).
#include "zt_common.h"
#include "zt_kernels.h"

namespace {
__device__ float g_srgb_lin[256];
__device__ float g_srgb_thr[255];
constexpr int kSrgbBuckets = 4096;
__device__ uint32_t g_srgb_bucket[(kSrgbBuckets + 4) / 4];   // byte b: thresholds below b / 4096, b = 0..4096

// the sRGB tables of one workgroup in LDS: thr[255] = +inf, so the compare past the last threshold is false
struct SrgbLds {
    float lin[256], thr[256];
    uint32_t bkt[(kSrgbBuckets + 4) / 4];
};
ZT_DEV void srgb_load(SrgbLds &S, int tid, int nt) {
    for (int i = tid; i < 256; i += nt) { S.lin[i] = g_srgb_lin[i]; S.thr[i] = i < 255 ? g_srgb_thr[i] : __int_as_float(0x7f800000); }
    for (int i = tid; i < (kSrgbBuckets + 4) / 4; i += nt) S.bkt[i] = g_srgb_bucket[i];
}
// the reference's encode of a linear value f in [0, 1]: the first k with f < thr[k] (zap_tex.h:1429-1433)
ZT_DEV int srgb_encode(float f, const SrgbLds &S) {
    const int b = (int)(f * (float)kSrgbBuckets);   // exact scaling by a power of two, then floor
    const int c = (int)((S.bkt[b >> 2] >> (8 * (b & 3))) & 255u);
    return c + (S.thr[c] <= f);
}

ZT_DEV int src_x0(int x, int w) { int v = 2 * x; return v < w ? v : w - 1; }
ZT_DEV int src_x1(int x, int w) { int v = 2 * x + 1; return v < w ? v : w - 1; }
}  // namespace

hipError_t zt_mip_upload_tables(const float lin[256], const float thr[255]) {
    // bucket b covers [b / 4096, (b + 1) / 4096): its byte counts the thresholds below b / 4096.  The encode is exact
    // only if no bucket holds two thresholds, which the reference's ascending table satisfies (refused otherwise).
    uint8_t bk[kSrgbBuckets + 4] = {};
    int per[kSrgbBuckets + 1] = {};
    for (int k = 0; k < 255; k++) {
        if (!(thr[k] >= 0.0f && thr[k] < 1.0f) || (k && !(thr[k] > thr[k - 1]))) return hipErrorInvalidValue;
        per[(int)(thr[k] * (float)kSrgbBuckets)]++;
    }
    for (int b = 0, below = 0; b <= kSrgbBuckets; b++) {
        if (per[b] > 1) return hipErrorInvalidValue;
        bk[b] = (uint8_t)below;
        below += per[b];
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_srgb_lin), lin, sizeof(float) * 256);
    if (e != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(g_srgb_bucket), bk, sizeof bk)) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_srgb_thr), thr, sizeof(float) * 255);
}

// grid: (ceil(nw/64), up to ceil(nh/4)), block 64x4; the rows repeat with the grid's height (any image height)
extern "C" __global__ void __launch_bounds__(256) zt_mip_level(const uint8_t *src, int w, int h, uint64_t stride, int srgb,
                                                                 uint8_t *dst, uint64_t dst_stride) {
    __shared__ SrgbLds S;
    const int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (srgb) {
        srgb_load(S, tid, blockDim.x * blockDim.y);
        __syncthreads();
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nw) return;
    for (int y = blockIdx.y * blockDim.y + threadIdx.y; y < nh; y += gridDim.y * blockDim.y) {
        const int x0 = src_x0(x, w), x1 = src_x1(x, w), y0 = y * 2 < h ? y * 2 : h - 1, y1 = y * 2 + 1 < h ? y * 2 + 1 : h - 1;
        const uint8_t *r0 = src + (uint64_t)y0 * stride, *r1 = src + (uint64_t)y1 * stride;
        uint32_t a, b, c, d;
        if (x1 == x0 + 1 && ((((uintptr_t)r0 | (uintptr_t)r1) + (uint64_t)x0 * 4) & 7) == 0) {
            uint2 t0 = *(const uint2 *)(r0 + x0 * 4), t1 = *(const uint2 *)(r1 + x0 * 4);
            a = t0.x; b = t0.y; c = t1.x; d = t1.y;
        } else {
            a = *(const uint32_t *)(r0 + x0 * 4); b = *(const uint32_t *)(r0 + x1 * 4);
            c = *(const uint32_t *)(r1 + x0 * 4); d = *(const uint32_t *)(r1 + x1 * 4);
        }
        uint32_t o = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int pa = chan(a, k), pb = chan(b, k), pc = chan(c, k), pd = chan(d, k), v;
            if (srgb && k < 3) v = srgb_encode((S.lin[pa] + S.lin[pb] + S.lin[pc] + S.lin[pd]) * 0.25f, S);
            else v = (pa + pb + pc + pd + 2) >> 2;
            o |= (uint32_t)v << (8 * k);
        }
        *(uint32_t *)(dst + (uint64_t)y * dst_stride + (uint64_t)x * 4) = o;
    }
}

// float RGBA level; strides in floats.  grid as above
extern "C" __global__ void __launch_bounds__(256) zt_mip_level_f(const float *src, int w, int h, uint64_t stride, float *dst,
                                                                   uint64_t dst_stride) {
    const int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nw) return;
    for (int y = blockIdx.y * blockDim.y + threadIdx.y; y < nh; y += gridDim.y * blockDim.y) {
        const int x0 = src_x0(x, w), x1 = src_x1(x, w), y0 = y * 2 < h ? y * 2 : h - 1, y1 = y * 2 + 1 < h ? y * 2 + 1 : h - 1;
        const float *r0 = src + (uint64_t)y0 * stride, *r1 = src + (uint64_t)y1 * stride;
        float4 a = *(const float4 *)(r0 + x0 * 4), b = *(const float4 *)(r0 + x1 * 4);
        float4 c = *(const float4 *)(r1 + x0 * 4), d = *(const float4 *)(r1 + x1 * 4);
        float4 o;
        o.x = 0.25f * (a.x + b.x + c.x + d.x);
        o.y = 0.25f * (a.y + b.y + c.y + d.y);
        o.z = 0.25f * (a.z + b.z + c.z + d.z);
        o.w = 0.25f * (a.w + b.w + c.w + d.w);
        *(float4 *)(dst + (uint64_t)y * dst_stride + (uint64_t)x * 4) = o;
    }
}

// ---------------------------------------------------------------- fused chain
// Level l pixel (x, y) reads level l-1 pixels (2x..2x+1, 2y..2y+1), which exist without clamping while the
// source level is at least 2 wide and 2 high; so a 2^k x 2^k tile of level 1 yields the matching tiles of
// levels 2..k+1 from LDS alone.  One launch of zt_mip_tiles makes up to 7 levels; the tail (a source of at
// most 128 x 128, and every level where a dimension reaches 1) runs in one workgroup (zt_mip_tail).
namespace {
ZT_DEV uint32_t box4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, int srgb, const SrgbLds &S) {
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int pa = chan(a, k), pb = chan(b, k), pc = chan(c, k), pd = chan(d, k), v;
        if (srgb && k < 3) v = srgb_encode((S.lin[pa] + S.lin[pb] + S.lin[pc] + S.lin[pd]) * 0.25f, S);
        else v = (pa + pb + pc + pd + 2) >> 2;
        o |= (uint32_t)v << (8 * k);
    }
    return o;
}
}  // namespace

// One T x T tile (tx, ty) of level 1 (T = 64 or 32) and the matching tiles of levels 2..nlev (at most
// log2(T) + 1 levels).  L.p[i] / L.w[i], L.h[i]: level i+1 buffers (packed rows).  Called by every thread of a
// workgroup (it synchronises).
namespace {
template <int T>
ZT_DEV void tiles_body(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev, const zt_mip_levels &L, int tx, int ty) {
    __shared__ SrgbLds S;
    __shared__ uint32_t t0[T * T], t1[T / 2 * T / 2];
    constexpr uint32_t H = T / 2;                    // pixel pairs per row of the tile
    constexpr int NP = T * T / 2 / 256;               // pairs per thread (256 threads)
    const int tid = (int)zt_tid(), nt = blockDim.x;
    if (srgb) srgb_load(S, tid, nt);
    __syncthreads();
    // level 1: T x T pixels of the tile.  A tile inside both level-1 sides with 16-byte aligned rows and 256
    // threads loads straight through: NP pixel pairs per thread, each two 16-byte reads (4 source pixels of two
    // rows), all issued before the first is used, so a workgroup keeps 4 T^2 bytes in flight.  Other tiles take
    // their pixels one at a time with 8-byte reads.
    const int w1 = L.w[0], h1 = L.h[0];
    if (nt == 256 && (tx + 1) * T <= w1 && (ty + 1) * T <= h1 && ((((uintptr_t)src) | stride) & 15) == 0) {
        uint4 r0[NP], r1[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const uint32_t pp = k * 256u + (uint32_t)tid;
            const int x = tx * T + (int)(pp % H) * 2, y = ty * T + (int)(pp / H);
            const uint8_t *a = src + (uint64_t)(2 * y) * stride + (uint64_t)x * 8;
            r0[k] = *(const uint4 *)a;
            r1[k] = *(const uint4 *)(a + stride);
        }
#pragma unroll
        for (int k = 0; k < NP; k++) {
            const uint32_t pp = k * 256u + (uint32_t)tid;
            const int lx = (int)(pp % H) * 2, ly = (int)(pp / H), x = tx * T + lx, y = ty * T + ly;
            const uint32_t o0 = box4(r0[k].x, r0[k].y, r1[k].x, r1[k].y, srgb, S);
            const uint32_t o1 = box4(r0[k].z, r0[k].w, r1[k].z, r1[k].w, srgb, S);
            t0[ly * T + lx] = o0;
            t0[ly * T + lx + 1] = o1;
            uint8_t *d = L.p[0] + ((uint64_t)y * w1 + x) * 4;
            if ((((uintptr_t)d) & 7) == 0) *(uint2 *)d = make_uint2(o0, o1);
            else { *(uint32_t *)d = o0; *(uint32_t *)(d + 4) = o1; }
        }
    } else {
        for (int p = tid; p < T * T; p += nt) {
            const int lx = p % T, ly = p / T, x = tx * T + lx, y = ty * T + ly;
            uint32_t o = 0;
            if (x < w1 && y < h1) {
                const uint8_t *r0 = src + (uint64_t)(2 * y) * stride + (uint64_t)(2 * x) * 4, *r1 = r0 + stride;
                uint32_t a, b, c, d;
                if ((((uintptr_t)r0 | (uintptr_t)r1) & 7) == 0) {
                    uint2 u0 = *(const uint2 *)r0, u1 = *(const uint2 *)r1;
                    a = u0.x; b = u0.y; c = u1.x; d = u1.y;
                } else {
                    a = *(const uint32_t *)r0; b = *(const uint32_t *)(r0 + 4); c = *(const uint32_t *)r1; d = *(const uint32_t *)(r1 + 4);
                }
                o = box4(a, b, c, d, srgb, S);
                *(uint32_t *)(L.p[0] + ((uint64_t)y * w1 + x) * 4) = o;
            }
            t0[p] = o;
        }
    }
    __syncthreads();
    // levels 2..nlev inside the tile, ping-ponging between t0 (T wide) and t1 (T/2 wide)
    int side = (int)H;
    for (int l = 1; l < nlev; l++, side >>= 1) {
        uint32_t *srcT = (l & 1) ? t0 : t1, *dstT = (l & 1) ? t1 : t0;
        const int sw = side * 2;
        const int wl = L.w[l], hl = L.h[l];
        for (int p = tid; p < side * side; p += nt) {
            const int lx = p % side, ly = p / side, x = tx * side + lx, y = ty * side + ly;
            uint32_t o = box4(srcT[(2 * ly) * sw + 2 * lx], srcT[(2 * ly) * sw + 2 * lx + 1], srcT[(2 * ly + 1) * sw + 2 * lx],
                              srcT[(2 * ly + 1) * sw + 2 * lx + 1], srgb, S);
            if (x < wl && y < hl) *(uint32_t *)(L.p[l] + ((uint64_t)y * wl + x) * 4) = o;
            dstT[p] = o;
        }
        __syncthreads();
    }
}
}  // namespace

// grid: tiles of 64 x 64 (zt_mip_tiles) or 32 x 32 (zt_mip_tiles32) level-1 pixels; block 256
extern "C" __global__ void __launch_bounds__(256) zt_mip_tiles(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev,
                                                                 zt_mip_levels L) {
    tiles_body<64>(src, w, h, stride, srgb, nlev, L, blockIdx.x, blockIdx.y);
}
extern "C" __global__ void __launch_bounds__(256) zt_mip_tiles32(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev,
                                                                   zt_mip_levels L) {
    tiles_body<32>(src, w, h, stride, srgb, nlev, L, blockIdx.x, blockIdx.y);
}

// one workgroup: levels from a source of at most 128 x 128, with the reference's clamping when a side is 1
namespace {
ZT_DEV void tail_body(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev, const zt_mip_levels &L) {
    __shared__ SrgbLds S;
    __shared__ uint32_t a[64 * 64], b[32 * 64];
    const int tid = (int)zt_tid();
    if (srgb) srgb_load(S, tid, blockDim.x);
    __syncthreads();
    int cw = w, ch = h;
    for (int l = 0; l < nlev; l++) {
        const int nw = cw > 1 ? cw / 2 : 1, nh = ch > 1 ? ch / 2 : 1;
        uint32_t *dstT = (l & 1) ? b : a;
        const uint32_t *srcT = (l & 1) ? a : b;
        for (int p = tid; p < nw * nh; p += blockDim.x) {
            const int x = p % nw, y = p / nw;
            const int x0 = 2 * x < cw ? 2 * x : cw - 1, x1 = 2 * x + 1 < cw ? 2 * x + 1 : cw - 1;
            const int y0 = 2 * y < ch ? 2 * y : ch - 1, y1 = 2 * y + 1 < ch ? 2 * y + 1 : ch - 1;
            uint32_t q0, q1, q2, q3;
            if (l == 0) {
                q0 = *(const uint32_t *)(src + (uint64_t)y0 * stride + x0 * 4); q1 = *(const uint32_t *)(src + (uint64_t)y0 * stride + x1 * 4);
                q2 = *(const uint32_t *)(src + (uint64_t)y1 * stride + x0 * 4); q3 = *(const uint32_t *)(src + (uint64_t)y1 * stride + x1 * 4);
            } else {
                q0 = srcT[y0 * cw + x0]; q1 = srcT[y0 * cw + x1]; q2 = srcT[y1 * cw + x0]; q3 = srcT[y1 * cw + x1];
            }
            const uint32_t o = box4(q0, q1, q2, q3, srgb, S);
            dstT[p] = o;
            *(uint32_t *)(L.p[l] + (uint64_t)p * 4) = o;
        }
        __syncthreads();
        cw = nw; ch = nh;
    }
}
}  // namespace

extern "C" __global__ void __launch_bounds__(1024) zt_mip_tail(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev,
                                                                 zt_mip_levels L) {
    tail_body(src, w, h, stride, srgb, nlev, L);
}

// ---------------------------------------------------------------- batched fused steps
// Many textures' fused steps in one launch: zt_mip_tiles_batch runs one tile item per workgroup (items name a
// step and a tile), zt_mip_tail_batch one step per workgroup (any workgroup size).  The bodies take the item or
// step index (bid), so the persistent batch kernel runs them as tasks.
namespace {
template <int T> ZT_DEV void tiles_item_body(const zt_mipstep *steps, const uint2 *items, uint32_t bid) {
    const uint2 it = items[bid];  // (step, tile index in row-major tile order of the step's level 1)
    const zt_mipstep &S = steps[it.x];
    const int tiles_x = (S.L.w[0] + T - 1) / T;
    tiles_body<T>(S.src, S.w, S.h, S.stride, S.srgb, S.nlev, S.L, (int)(it.y % (uint32_t)tiles_x), (int)(it.y / (uint32_t)tiles_x));
}
ZT_DEV void tail_step_body(const zt_mipstep *steps, uint32_t bid) {
    const zt_mipstep &S = steps[bid];
    tail_body(S.src, S.w, S.h, S.stride, S.srgb, S.nlev, S.L);
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_mip_tiles_batch(const zt_mipstep *steps, const uint2 *items) {
    tiles_item_body<64>(steps, items, blockIdx.x);
}
extern "C" __global__ void __launch_bounds__(256) zt_mip_tiles32_batch(const zt_mipstep *steps, const uint2 *items) {
    tiles_item_body<32>(steps, items, blockIdx.x);
}
extern "C" __global__ void __launch_bounds__(256) zt_mip_tail_batch(const zt_mipstep *steps) { tail_step_body(steps, blockIdx.x); }

// ---------------------------------------------------------------- batched level
// One level of many textures in one launch: lane per destination pixel over all jobs (grid-stride); a job is one
// texture's level l from its level l-1, destination rows packed (width * 4 bytes).  Jobs sorted by `first`.
namespace {
ZT_DEV int find_mip_job(const zt_mipjob *J, int n, uint64_t p) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J[mid].first <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}
}  // namespace

namespace {
ZT_DEV void mip_batch_body(const zt_mipjob *jobs, int njobs, uint64_t total, int any_srgb, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    __shared__ SrgbLds S;
    if (any_srgb) srgb_load(S, tid, blockDim.x);
    __syncthreads();
    const uint64_t step = (uint64_t)nblk * blockDim.x;
    for (uint64_t p = (uint64_t)bid * blockDim.x + tid; p < total; p += step) {
        const zt_mipjob J = jobs[find_mip_job(jobs, njobs, p)];
        const int nw = J.w > 1 ? J.w / 2 : 1;
        const uint64_t local = p - J.first;
        const int x = (int)(local % (uint64_t)nw), y = (int)(local / (uint64_t)nw);
        const int x0 = src_x0(x, J.w), x1 = src_x1(x, J.w), y0 = y * 2 < J.h ? y * 2 : J.h - 1, y1 = y * 2 + 1 < J.h ? y * 2 + 1 : J.h - 1;
        const uint8_t *r0 = J.src + (uint64_t)y0 * J.stride, *r1 = J.src + (uint64_t)y1 * J.stride;
        const uint32_t a = *(const uint32_t *)(r0 + x0 * 4), b = *(const uint32_t *)(r0 + x1 * 4);
        const uint32_t c = *(const uint32_t *)(r1 + x0 * 4), d = *(const uint32_t *)(r1 + x1 * 4);
        *(uint32_t *)(J.dst + local * 4) = box4(a, b, c, d, J.srgb, S);
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_mip_batch(const zt_mipjob *jobs, int njobs, uint64_t total, int any_srgb) {
    mip_batch_body(jobs, njobs, total, any_srgb, blockIdx.x, gridDim.x);
}
