// zt_probe.hip — measurement kernels for the roofline denominators (libztprobe.so, loaded by bench.py only).
//
// ztprobe_valu runs `iters` rounds over 8 independent dependency chains of one VALU instruction per lane, written
// in inline asm so the compiler can neither fold nor reorder them: every round issues exactly 32 wave64 VALU
// instructions of the chosen kind per wave (4 steps of each chain; the loop counter is scalar).  Launched with
// many waves per SIMD it shows the chip's VALU issue ceiling for that instruction (wave instructions per second),
// which the encoders' counter rates are set against.
#include <hip/hip_runtime.h>
#include <stdint.h>

template <int KIND> __global__ void __launch_bounds__(256) zt_valu_probe(uint32_t *out, int iters, uint32_t seed) {
    uint32_t a[8];
    uint64_t w[8];   // the 64-bit chains of v_mad_u64_u32
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = 0; k < 8; k++) { a[k] = t * 2654435761u + seed * (k + 1); w[k] = a[k]; }
    const uint32_t b = t ^ seed, sel = 0x05040100u ^ (seed & 0x03030303u);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 32; j++) {
            const int k = j & 7;
            if (KIND == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            else if (KIND == 1) asm volatile("v_dot4_u32_u8 %0, %1, %1, %0" : "+v"(a[k]) : "v"(b));
            else if (KIND == 2) asm volatile("v_fma_f32 %0, %0, %1, %1" : "+v"(a[k]) : "v"(b));
            else if (KIND == 3) asm volatile("v_max_i32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            else if (KIND == 4) asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(a[k]) : "v"(b));
            else if (KIND == 5) asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(a[k]) : "v"(b), "v"(sel));
            else if (KIND == 6) asm volatile("v_mad_u64_u32 %0, vcc, %1, %1, %0" : "+v"(w[k]) : "v"(b) : "vcc");
            else if (KIND == 7) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            else asm volatile("v_mad_u32_u24 %0, %0, %1, %0" : "+v"(a[k]) : "v"(b));
        }
    }
    uint32_t r = 0;
    for (int k = 0; k < 8; k++) r ^= a[k] ^ (uint32_t)w[k] ^ (uint32_t)(w[k] >> 32);
    out[t] = r;
}

// a plain device copy of known size (16-byte vectors, grid-stride): what the HBM byte counters read for it
// calibrates FETCH_SIZE / WRITE_SIZE, and its rate is the practical bandwidth bar for the HBM-bound kernels
__global__ void __launch_bounds__(256) zt_copy_probe(const uint4 *src, uint4 *dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull) dst[i] = src[i];
}

extern "C" {
// dst <- src, `bytes` a multiple of 16, both 16-byte aligned device pointers
int ztprobe_copy(const void *src, void *dst, uint64_t bytes, void *stream) {
    if (!src || !dst || bytes % 16 || ((uintptr_t)src | (uintptr_t)dst) % 16) return 1;
    const uint64_t n = bytes / 16;
    const uint64_t groups = n / 256 < 256 * 16 ? (n + 255) / 256 : 256 * 16;
    hipLaunchKernelGGL(zt_copy_probe, dim3((unsigned)(groups ? groups : 1)), dim3(256), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst, n);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// kind 0 v_add_u32, 1 v_dot4_u32_u8, 2 v_fma_f32, 3 v_max_i32, 4 v_lshl_add_u32, 5 v_perm_b32, 6 v_mad_u64_u32,
// 7 v_mul_lo_u32, 8 v_mad_u32_u24;
// out: groups * 256 uint32 of device memory
int ztprobe_valu(int kind, int groups, int iters, uint32_t *out, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (groups < 1 || iters < 1 || !out) return 1;
    switch (kind) {
    case 0: hipLaunchKernelGGL(zt_valu_probe<0>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 1: hipLaunchKernelGGL(zt_valu_probe<1>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 2: hipLaunchKernelGGL(zt_valu_probe<2>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 3: hipLaunchKernelGGL(zt_valu_probe<3>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 4: hipLaunchKernelGGL(zt_valu_probe<4>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 5: hipLaunchKernelGGL(zt_valu_probe<5>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 6: hipLaunchKernelGGL(zt_valu_probe<6>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 7: hipLaunchKernelGGL(zt_valu_probe<7>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    case 8: hipLaunchKernelGGL(zt_valu_probe<8>, dim3(groups), dim3(256), 0, s, out, iters, 12345u); break;
    default: return 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
}
