// zt_mip.hip — 2x2 box-filter mip levels for gfx950 (zap_tex.h:1442-1469).  HBM-bound: each lane makes
// one destination pixel from four source pixels; sRGB levels average in linear light through a 256-entry
// table in LDS and re-encode by binary search over 255 thresholds (first k with v < thr[k]).
#include "zt_common.h"
#include "zt_kernels.h"

namespace {
__device__ float g_srgb_lin[256];
__device__ float g_srgb_thr[255];

ZT_DEV int src_x0(int x, int w) { int v = 2 * x; return v < w ? v : w - 1; }
ZT_DEV int src_x1(int x, int w) { int v = 2 * x + 1; return v < w ? v : w - 1; }
}  // namespace

hipError_t zt_mip_upload_tables(const float lin[256], const float thr[255]) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_srgb_lin), lin, sizeof(float) * 256);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(g_srgb_thr), thr, sizeof(float) * 255);
}

// grid: (ceil(nw/64), ceil(nh/4)), block 64x4
extern "C" __global__ void __launch_bounds__(256) zt_mip_level(const uint8_t *src, int w, int h, uint64_t stride, int srgb,
                                                                 uint8_t *dst, uint64_t dst_stride) {
    __shared__ float lin[256], thr[256];
    const int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (srgb) {
        lin[tid] = g_srgb_lin[tid];
        thr[tid] = tid < 255 ? g_srgb_thr[tid] : 0.0f;
        __syncthreads();
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= nw || y >= nh) return;
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
        if (srgb && k < 3) {
            float f = (lin[pa] + lin[pb] + lin[pc] + lin[pd]) * 0.25f;
            int lo = 0, hi = 255;
            while (lo < hi) { int mid = (lo + hi) >> 1; if (f < thr[mid]) hi = mid; else lo = mid + 1; }
            v = lo;
        } else v = (pa + pb + pc + pd + 2) >> 2;
        o |= (uint32_t)v << (8 * k);
    }
    *(uint32_t *)(dst + (uint64_t)y * dst_stride + (uint64_t)x * 4) = o;
}

// float RGBA level; strides in floats.  grid as above
extern "C" __global__ void __launch_bounds__(256) zt_mip_level_f(const float *src, int w, int h, uint64_t stride, float *dst,
                                                                   uint64_t dst_stride) {
    const int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= nw || y >= nh) return;
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
