// zt_kernels.h — kernels and device tables shared between the HIP translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct zt_job;
struct zt_jobs;

// BC7 compaction lists (device memory, capacity = total blocks of the launch)
struct zt_bc7_lists {
    unsigned *counters;  // [0] list-5 length, [1] list-1 length
    uint4 *meta5;        // (global block, mode-6 SSE, job, opaque)
    uint4 *px5;          // 4 x uint4 = the block's 16 RGBA8 pixels
    uint4 *meta1;        // (list-5 entry, current best SSE, 0, 0)
};

extern "C" {
__global__ void zt_bc7_stage6(zt_jobs jobs, uint32_t total, zt_bc7_lists L);
__global__ void zt_bc7_stage5(zt_jobs jobs, zt_bc7_lists L);
__global__ void zt_bc7_stage1(zt_jobs jobs, zt_bc7_lists L, int m1k);
__global__ void zt_zero_u32(unsigned *p, int n);
// exhaustive all-mode BC7 search (optionally restricted to listed global blocks); best_cand may be null
__global__ void zt_bc7_full(zt_jobs jobs, uint32_t total, const uint32_t *list, const unsigned *nlist, uint32_t *best_cand);

__global__ void zt_bc14_encode(zt_jobs jobs, uint32_t total);
__global__ void zt_bc6h_encode(zt_jobs jobs, uint32_t total);
__global__ void zt_astc_encode(zt_jobs jobs, uint32_t total);

__global__ void zt_bc_decode(const uint8_t *bc, int w, int h, int fmt, uint8_t *rgba, uint64_t stride);
__global__ void zt_bc6h_decode(const uint8_t *bc, int w, int h, uint16_t *rgba, uint64_t stride);
__global__ void zt_astc_decode(const uint8_t *bc, int w, int h, uint8_t *rgba, uint64_t stride);
__global__ void zt_bc_shuffle(const uint8_t *in, uint8_t *out, uint64_t nb, int fmt, int split);

// mips: one level (any size); tile chain (levels 1..k inside a workgroup tile); float level
__global__ void zt_mip_level(const uint8_t *src, int w, int h, uint64_t stride, int srgb, uint8_t *dst, uint64_t dst_stride);
__global__ void zt_mip_level_f(const float *src, int w, int h, uint64_t stride, float *dst, uint64_t dst_stride);
}

// host-side table uploads (called once per device by the C-ABI layer); return hipError_t
hipError_t zt_mip_upload_tables(const float lin[256], const float thr[255]);
hipError_t zt_astc_upload_tables(const void *setup, size_t bytes);
