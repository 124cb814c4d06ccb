// zt_capi.cpp — the C ABI of libzaptex (include/zaptex.h): argument checks, device set-up, table uploads,
// staging of host buffers, and kernel launches.  No CPU fallback: every path runs the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/zaptex.h"
#include "zt_astc_setup.h"
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

thread_local char g_err[512];
std::mutex g_mu;     // serialises host entry points (shared staging buffers)
int g_m1k = 4;

// optional per-kernel timing: HIP events recorded on the launch stream around every kernel
bool g_profile = false;
std::mutex g_prof_mu;
struct ProfPending { const char *name; hipEvent_t a, b; };
struct ProfTotal { const char *name; double ms; long n; };
std::vector<ProfPending> g_prof_pending;
std::vector<ProfTotal> g_prof_total;

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
#define HIP_TRY(x)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return fail(ZAPTEX_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// launch `fn` on stream s; with profiling on, bracket it with events
template <typename F> int launch(const char *name, hipStream_t s, F &&fn) {
    hipEvent_t a = nullptr, b = nullptr;
    if (g_profile) {
        HIP_TRY(hipEventCreate(&a)); HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ZAPTEX_E_DEVICE, "launch of %s: %s", name, hipGetErrorString(e));
    if (g_profile) {
        HIP_TRY(hipEventRecord(b, s));
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_pending.push_back({ name, a, b });
    }
    return ZAPTEX_OK;
}

size_t block_bytes(int fmt) { return fmt == ZAPTEX_BC1 || fmt == ZAPTEX_BC4 ? 8 : 16; }
size_t out_bytes(int w, int h, int fmt) { return (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4) * block_bytes(fmt); }
size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ---- sRGB tables: float32 of the sRGB transfer function, except five entries one ulp away, exactly as the
// reference's literal tables (zap_tex.h:1360-1427); checked by hash before upload.
void srgb_tables(float lin[256], float thr[255]) {
    auto f = [](double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); };
    for (int v = 0; v < 256; v++) lin[v] = (float)f(v / 255.0);
    for (int k = 0; k < 255; k++) thr[k] = (float)f((k + 0.5) / 255.0);
    auto nudge = [](float &x, int d) { uint32_t b; memcpy(&b, &x, 4); b += (uint32_t)d; memcpy(&x, &b, 4); };
    nudge(lin[98], -1); nudge(lin[122], -1); nudge(lin[164], -1); nudge(thr[34], -1); nudge(thr[133], +1);
}
uint32_t fnv1a(const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    uint32_t h = 2166136261u;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 16777619u; }
    return h;
}
const uint32_t kSrgbLinHash = 0xab9a3c1bu, kSrgbThrHash = 0x4f5c9c4fu;

struct DeviceState {
    bool ready = false;
    hipStream_t stream = nullptr;
    // staging for the host entry points
    void *in = nullptr, *out = nullptr, *ws = nullptr;
    size_t in_cap = 0, out_cap = 0, ws_cap = 0;
};
DeviceState g_dev[64];

int ensure_buf(void *&p, size_t &cap, size_t need) {
    if (need <= cap) return ZAPTEX_OK;
    if (p) { hipFree(p); p = nullptr; cap = 0; }
    size_t n = std::max(align_up(need, 1 << 20), (size_t)1 << 20);
    if (hipMalloc(&p, n) != hipSuccess) { p = nullptr; return fail(ZAPTEX_E_NOMEM, "hipMalloc(%zu) failed", n); }
    cap = n;
    return ZAPTEX_OK;
}

int current_device(int &dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(ZAPTEX_E_DEVICE, "no HIP device available");
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(ZAPTEX_E_DEVICE, "device ordinal %d out of range", dev);
    return ZAPTEX_OK;
}

// per-device one-time set-up: tables into device memory, a stream for the host entry points
int ensure_ready(DeviceState *&st) {
    int dev;
    if (int rc = current_device(dev)) return rc;
    st = &g_dev[dev];
    if (st->ready) return ZAPTEX_OK;
    float lin[256], thr[255];
    srgb_tables(lin, thr);
    if (fnv1a(lin, sizeof lin) != kSrgbLinHash || fnv1a(thr, sizeof thr) != kSrgbThrHash)
        return fail(ZAPTEX_E_DEVICE, "sRGB tables differ from the reference values (libm pow mismatch)");
    HIP_TRY(zt_mip_upload_tables(lin, thr));
    zt_astc_setup setup;
    zt_astc_setup_build(&setup);
    HIP_TRY(zt_astc_upload_tables(&setup, sizeof setup));
    HIP_TRY(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    st->ready = true;
    return ZAPTEX_OK;
}

bool is_bcn(int fmt) { return fmt >= ZAPTEX_BC1 && fmt <= ZAPTEX_BC7; }
bool valid_fmt(int fmt) { return fmt >= ZAPTEX_BC1 && fmt <= ZAPTEX_ASTC; }
size_t image_out_bytes(const zaptex_image &im) { return out_bytes(im.w, im.h, im.fmt == ZAPTEX_BC6H || im.fmt == ZAPTEX_ASTC ? ZAPTEX_BC7 : im.fmt); }
uint32_t nblocks(const zaptex_image &im) { return (uint32_t)(((im.w + 3) / 4) * ((im.h + 3) / 4)); }

int check_image(const zaptex_image &im) {
    if (!valid_fmt(im.fmt)) return fail(ZAPTEX_E_ARG, "unknown format %d", im.fmt);
    if (im.w < 1 || im.h < 1 || im.w > (1 << 16) || im.h > (1 << 16)) return fail(ZAPTEX_E_ARG, "bad size %dx%d", im.w, im.h);
    if (!im.src || !im.dst) return fail(ZAPTEX_E_ARG, "null image pointer");
    size_t px = im.fmt == ZAPTEX_BC6H ? 16 : 4;
    if (im.stride < (size_t)im.w * px) return fail(ZAPTEX_E_ARG, "stride %zu below row size", im.stride);
    if ((((uintptr_t)im.src) | im.stride) % px) return fail(ZAPTEX_E_ARG, "source rows must be %zu-byte aligned", px);
    if (((uintptr_t)im.dst) % block_bytes(im.fmt == ZAPTEX_BC6H || im.fmt == ZAPTEX_ASTC ? ZAPTEX_BC7 : im.fmt))
        return fail(ZAPTEX_E_ARG, "destination must be block-aligned");
    return ZAPTEX_OK;
}

// workspace layout: [counters 256 B][meta5 16 B/blk][px5 64 B/blk][meta1 16 B/blk] over the BC7 blocks
size_t ws_bytes_for(uint64_t bc7_blocks) { return bc7_blocks ? 256 + bc7_blocks * 96 : 0; }

zt_bc7_lists lists_in(void *ws, uint64_t n) {
    uint8_t *p = (uint8_t *)ws;
    zt_bc7_lists L;
    L.counters = (unsigned *)p;
    L.meta5 = (uint4 *)(p + 256);
    L.px5 = (uint4 *)(p + 256 + n * 16);
    L.meta1 = (uint4 *)(p + 256 + n * 80);
    return L;
}

int grid_for(uint64_t items, int per) {
    uint64_t g = (items + per - 1) / per;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 8192));
}

// enqueue the encoders for up to ZT_INLINE_JOBS images (already validated)
int enqueue(const zaptex_image *imgs, int n, void *ws, size_t ws_bytes, hipStream_t s) {
    zt_jobs fam[3];  // 0: BC1-5, 1: BC7, 2: BC6H, ASTC handled per kernel below
    zt_jobs astc, b6;
    memset(fam, 0, sizeof fam); memset(&astc, 0, sizeof astc); memset(&b6, 0, sizeof b6);
    uint32_t tot[3] = { 0, 0, 0 }, tot_astc = 0, tot_b6 = 0;
    for (int i = 0; i < n; i++) {
        const zaptex_image &im = imgs[i];
        zt_jobs *J;
        uint32_t *t;
        if (im.fmt == ZAPTEX_BC7) { J = &fam[1]; t = &tot[1]; }
        else if (im.fmt == ZAPTEX_BC6H) { J = &b6; t = &tot_b6; }
        else if (im.fmt == ZAPTEX_ASTC) { J = &astc; t = &tot_astc; }
        else { J = &fam[0]; t = &tot[0]; }
        zt_job &j = J->inl[J->n++];
        j.src = (const uint8_t *)im.src; j.dst = (uint8_t *)im.dst; j.stride = im.stride;
        j.first = *t; j.w = im.w; j.h = im.h; j.bw = (im.w + 3) / 4; j.fmt = im.fmt; j.pad = 0;
        *t += nblocks(im);
    }
    int rc;
    if (tot[0] && (rc = launch("zt_bc14_encode", s, [&] {
            hipLaunchKernelGGL(zt_bc14_encode, dim3(grid_for(tot[0], 256)), dim3(256), 0, s, fam[0], tot[0]); })))
        return rc;
    if (tot_b6 && (rc = launch("zt_bc6h_encode", s, [&] {
            hipLaunchKernelGGL(zt_bc6h_encode, dim3(grid_for(tot_b6, 256)), dim3(256), 0, s, b6, tot_b6); })))
        return rc;
    if (tot_astc && (rc = launch("zt_astc_encode", s, [&] {
            hipLaunchKernelGGL(zt_astc_encode, dim3(grid_for(tot_astc, 256)), dim3(256), 0, s, astc, tot_astc); })))
        return rc;
    if (tot[1]) {
        if (ws_bytes < ws_bytes_for(tot[1]) || !ws) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu bytes", ws_bytes, ws_bytes_for(tot[1]));
        zt_bc7_lists L = lists_in(ws, tot[1]);
        if ((rc = launch("zt_zero_u32", s, [&] { hipLaunchKernelGGL(zt_zero_u32, dim3(1), dim3(64), 0, s, L.counters, 64); })))
            return rc;
        if ((rc = launch("zt_bc7_stage6", s, [&] {
                hipLaunchKernelGGL(zt_bc7_stage6, dim3(grid_for(tot[1], 256)), dim3(256), 0, s, fam[1], tot[1], L); })))
            return rc;
        if ((rc = launch("zt_bc7_stage5", s, [&] {
                hipLaunchKernelGGL(zt_bc7_stage5, dim3(grid_for((uint64_t)tot[1] * 4, 256)), dim3(256), 0, s, fam[1], L); })))
            return rc;
        if ((rc = launch("zt_bc7_stage1", s, [&] {
                hipLaunchKernelGGL(zt_bc7_stage1, dim3(grid_for((uint64_t)tot[1], 4)), dim3(256), 0, s, fam[1], L, g_m1k); })))
            return rc;
    }
    return ZAPTEX_OK;
}

int enqueue_mip(const void *src, int w, int h, size_t stride, int srgb, void *dst, size_t dst_stride, hipStream_t s) {
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    dim3 grid((nw + 63) / 64, (nh + 3) / 4), block(64, 4);
    return launch("zt_mip_level", s, [&] {
        hipLaunchKernelGGL(zt_mip_level, grid, block, 0, s, (const uint8_t *)src, w, h, (uint64_t)stride, srgb, (uint8_t *)dst, (uint64_t)dst_stride);
    });
}

// run one host-buffer image through the device: stage in, encode, copy out
int host_encode(const void *src, int w, int h, size_t stride, size_t px_bytes, int fmt, void *out) {
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t row = (size_t)w * px_bytes, inb = row * (size_t)h;
    zaptex_image im{ nullptr, nullptr, row, w, h, fmt };
    size_t ob = image_out_bytes(im);
    uint64_t n7 = fmt == ZAPTEX_BC7 ? nblocks(im) : 0;
    if (int rc = ensure_buf(st->in, st->in_cap, inb)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, ob)) return rc;
    if (n7) if (int rc = ensure_buf(st->ws, st->ws_cap, ws_bytes_for(n7))) return rc;
    im.src = st->in; im.dst = st->out;
    HIP_TRY(hipMemcpy2DAsync(st->in, row, src, stride, row, (size_t)h, hipMemcpyHostToDevice, st->stream));
    if (int rc = enqueue(&im, 1, st->ws, st->ws_cap, st->stream)) return rc;
    HIP_TRY(hipMemcpyAsync(out, st->out, ob, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

}  // namespace

extern "C" {

int zaptex_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, void *out) {
    if (!is_bcn(fmt)) return fail(ZAPTEX_E_ARG, "unknown zap_bc_format %d", fmt);
    if (rdo > 0) return fail(ZAPTEX_E_UNSUPPORTED, "rdo > 0 (rate-distortion optimisation) is not implemented by libzaptex");
    if (w < 1 || h < 1 || !rgba || !out || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    return host_encode(rgba, w, h, stride, 4, fmt, out);
}

int zaptex_bc6h_encode(const float *rgba, int w, int h, size_t stride, void *out) {
    if (w < 1 || h < 1 || !rgba || !out || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    return host_encode(rgba, w, h, stride * sizeof(float), 16, ZAPTEX_BC6H, out);
}

int zaptex_astc_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out) {
    if (w < 1 || h < 1 || !rgba || !out || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    return host_encode(rgba, w, h, stride, 4, ZAPTEX_ASTC, out);
}

int zaptex_mip_next(const uint8_t *src, int w, int h, size_t stride, int srgb, uint8_t *dst, size_t dst_stride) {
    if (w < 1 || h < 1 || !src || !dst || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    if (dst_stride < (size_t)nw * 4) return fail(ZAPTEX_E_ARG, "dst_stride below row size");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t row = (size_t)w * 4, orow = (size_t)nw * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, row * h)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, orow * nh)) return rc;
    HIP_TRY(hipMemcpy2DAsync(st->in, row, src, stride, row, (size_t)h, hipMemcpyHostToDevice, st->stream));
    if (int rc = enqueue_mip(st->in, w, h, row, srgb, st->out, orow, st->stream)) return rc;
    HIP_TRY(hipMemcpy2DAsync(dst, dst_stride, st->out, orow, orow, (size_t)nh, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

int zaptex_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t dst_stride) {
    if (w < 1 || h < 1 || !src || !dst || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    if (dst_stride < (size_t)nw * 4) return fail(ZAPTEX_E_ARG, "dst_stride below row size");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t row = (size_t)w * 16, orow = (size_t)nw * 16;
    if (int rc = ensure_buf(st->in, st->in_cap, row * h)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, orow * nh)) return rc;
    HIP_TRY(hipMemcpy2DAsync(st->in, row, src, stride * 4, row, (size_t)h, hipMemcpyHostToDevice, st->stream));
    dim3 grid((nw + 63) / 64, (nh + 3) / 4), block(64, 4);
    hipLaunchKernelGGL(zt_mip_level_f, grid, block, 0, st->stream, (const float *)st->in, w, h, (uint64_t)w * 4, (float *)st->out, (uint64_t)nw * 4);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy2DAsync(dst, dst_stride * 4, st->out, orow, orow, (size_t)nh, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

int zaptex_bc_decode(const void *bc, int w, int h, int fmt, uint8_t *rgba, size_t stride) {
    if (!is_bcn(fmt) || w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t ib = out_bytes(w, h, fmt), row = (size_t)w * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_bc_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, fmt, (uint8_t *)st->out, (uint64_t)row);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy2DAsync(rgba, stride, st->out, row, row, (size_t)h, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

int zaptex_bc6h_decode(const void *bc, int w, int h, uint16_t *rgba, size_t stride) {
    if (w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t ib = out_bytes(w, h, ZAPTEX_BC7), row = (size_t)w * 8;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_bc6h_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, (uint16_t *)st->out, (uint64_t)w * 4);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy2DAsync(rgba, stride * 2, st->out, row, row, (size_t)h, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

int zaptex_astc_decode(const void *bc, int w, int h, uint8_t *rgba, size_t stride) {
    if (w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    size_t ib = out_bytes(w, h, ZAPTEX_BC7), row = (size_t)w * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_astc_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, (uint8_t *)st->out, (uint64_t)row);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy2DAsync(rgba, stride, st->out, row, row, (size_t)h, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

static int shuffle(const void *in, size_t n, int fmt, void *out, int split) {
    if (!is_bcn(fmt) || (!in && n) || (!out && n)) return fail(ZAPTEX_E_ARG, "bad arguments");
    if (!n) return ZAPTEX_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState *st;
    if (int rc = ensure_ready(st)) return rc;
    if (int rc = ensure_buf(st->in, st->in_cap, n)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, n)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, in, n, hipMemcpyHostToDevice, st->stream));
    uint64_t nb = n / block_bytes(fmt);
    size_t tail = n - nb * block_bytes(fmt);
    if (nb) {
        hipLaunchKernelGGL(zt_bc_shuffle, dim3(grid_for(nb, 256)), dim3(256), 0, st->stream, (const uint8_t *)st->in, (uint8_t *)st->out, nb, fmt, split);
        HIP_TRY(hipGetLastError());
    }
    if (tail) HIP_TRY(hipMemcpyAsync((uint8_t *)st->out + nb * block_bytes(fmt), (const uint8_t *)st->in + nb * block_bytes(fmt), tail, hipMemcpyDeviceToDevice, st->stream));
    HIP_TRY(hipMemcpyAsync(out, st->out, n, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}
int zaptex_bc_split(const void *bc, size_t n, int fmt, void *out) { return shuffle(bc, n, fmt, out, 1); }
int zaptex_bc_merge(const void *planes, size_t n, int fmt, void *out) { return shuffle(planes, n, fmt, out, 0); }

size_t zaptex_workspace_size(const zaptex_image *imgs, int n) {
    uint64_t n7 = 0;
    for (int i = 0; i < n; i++) if (imgs[i].fmt == ZAPTEX_BC7) n7 += nblocks(imgs[i]);
    return ws_bytes_for(n7);
}

int zaptex_encode_dev(const zaptex_image *imgs, int n, void *workspace, size_t ws_bytes, void *stream) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS) return fail(ZAPTEX_E_ARG, "image count %d outside 1..%d", n, ZT_INLINE_JOBS);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        if (int rc = check_image(imgs[i])) return rc;
        total += nblocks(imgs[i]);
    }
    if (total >= (1ull << 31)) return fail(ZAPTEX_E_ARG, "too many blocks in one call");
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    return enqueue(imgs, n, workspace, ws_bytes, (hipStream_t)stream);
}

int zaptex_bc7_full_dev(const zaptex_image *imgs, int n, uint32_t *best_cand, void *stream) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS) return fail(ZAPTEX_E_ARG, "image count %d outside 1..%d", n, ZT_INLINE_JOBS);
    zt_jobs J;
    memset(&J, 0, sizeof J);
    uint32_t total = 0;
    for (int i = 0; i < n; i++) {
        if (int rc = check_image(imgs[i])) return rc;
        if (imgs[i].fmt != ZAPTEX_BC7) return fail(ZAPTEX_E_ARG, "the exhaustive search encodes BC7 only");
        zt_job &j = J.inl[J.n++];
        j.src = (const uint8_t *)imgs[i].src; j.dst = (uint8_t *)imgs[i].dst; j.stride = imgs[i].stride;
        j.first = total; j.w = imgs[i].w; j.h = imgs[i].h; j.bw = (imgs[i].w + 3) / 4; j.fmt = ZAPTEX_BC7; j.pad = 0;
        total += nblocks(imgs[i]);
    }
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    return launch("zt_bc7_full", s, [&] {
        hipLaunchKernelGGL(zt_bc7_full, dim3(grid_for(total, 4)), dim3(256), 0, s, J, total, (const uint32_t *)nullptr, (const unsigned *)nullptr, best_cand);
    });
}

int zaptex_mip_next_dev(const void *src, int w, int h, size_t stride, int srgb, void *dst, size_t dst_stride, void *stream) {
    if (w < 1 || h < 1 || !src || !dst || stride < (size_t)w * 4 || ((uintptr_t)src | stride | (uintptr_t)dst | dst_stride) % 4)
        return fail(ZAPTEX_E_ARG, "bad arguments (rows must be 4-byte aligned)");
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    return enqueue_mip(src, w, h, stride, srgb, dst, dst_stride, (hipStream_t)stream);
}

int zaptex_mip_next_f_dev(const void *src, int w, int h, size_t stride, void *dst, size_t dst_stride, void *stream) {
    if (w < 1 || h < 1 || !src || !dst || stride < (size_t)w * 4 || ((uintptr_t)src | (uintptr_t)dst) % 16 || stride % 4 || dst_stride % 4)
        return fail(ZAPTEX_E_ARG, "bad arguments (rows must be 16-byte aligned)");
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    dim3 grid((nw + 63) / 64, (nh + 3) / 4), block(64, 4);
    hipLaunchKernelGGL(zt_mip_level_f, grid, block, 0, (hipStream_t)stream, (const float *)src, w, h, (uint64_t)stride, (float *)dst, (uint64_t)dst_stride);
    HIP_TRY(hipGetLastError());
    return ZAPTEX_OK;
}

int zaptex_mip_chain_dev(const void *base, int w, int h, size_t stride, int srgb, void *const *levels, int nlevels, void *stream) {
    if (w < 1 || h < 1 || !base || !levels || stride < (size_t)w * 4 || ((uintptr_t)base | stride) % 4)
        return fail(ZAPTEX_E_ARG, "bad arguments");
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    const void *src = base;
    size_t ss = stride;
    for (int i = 0; i < nlevels && (w > 1 || h > 1); i++) {
        int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
        if (!levels[i] || ((uintptr_t)levels[i]) % 4) return fail(ZAPTEX_E_ARG, "bad level pointer %d", i);
        if (int rc = enqueue_mip(src, w, h, ss, srgb, levels[i], (size_t)nw * 4, (hipStream_t)stream)) return rc;
        src = levels[i]; ss = (size_t)nw * 4; w = nw; h = nh;
    }
    return ZAPTEX_OK;
}

int zaptex_bc_decode_dev(const void *bc, int w, int h, int fmt, void *rgba, size_t stride, void *stream) {
    if (!is_bcn(fmt) || w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4 || ((uintptr_t)rgba | stride) % 4)
        return fail(ZAPTEX_E_ARG, "bad arguments");
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_bc_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, (hipStream_t)stream, (const uint8_t *)bc, w, h, fmt, (uint8_t *)rgba, (uint64_t)stride);
    HIP_TRY(hipGetLastError());
    return ZAPTEX_OK;
}

int zaptex_set_option(int key, int value) {
    if (key == ZAPTEX_OPT_PROFILE) {
        g_profile = value != 0;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_BC7_M1K) {
        if (value < 1 || value > 64) return fail(ZAPTEX_E_ARG, "BC7 M1K must be 1..64");
        g_m1k = value;
        return ZAPTEX_OK;
    }
    return fail(ZAPTEX_E_ARG, "unknown option %d", key);
}
int zaptex_get_option(int key) { return key == ZAPTEX_OPT_BC7_M1K ? g_m1k : key == ZAPTEX_OPT_PROFILE ? (int)g_profile : -1; }

int zaptex_profile_collect(void) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (const ProfPending &p : g_prof_pending) {
        HIP_TRY(hipEventSynchronize(p.b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        auto it = std::find_if(g_prof_total.begin(), g_prof_total.end(), [&](const ProfTotal &t) { return !strcmp(t.name, p.name); });
        if (it == g_prof_total.end()) g_prof_total.push_back({ p.name, ms, 1 });
        else { it->ms += ms; it->n++; }
        hipEventDestroy(p.a); hipEventDestroy(p.b);
    }
    g_prof_pending.clear();
    return ZAPTEX_OK;
}
int zaptex_profile_get(const char *kernel, double *ms_total, int *launches) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (const ProfTotal &t : g_prof_total)
        if (!strcmp(t.name, kernel)) { *ms_total = t.ms; *launches = (int)t.n; return ZAPTEX_OK; }
    *ms_total = 0; *launches = 0;
    return ZAPTEX_OK;
}
void zaptex_profile_reset(void) {
    zaptex_profile_collect();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_total.clear();
}
const char *zaptex_last_error(void) { return g_err; }
int zaptex_device_count(void) {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess ? n : 0;
}
const char *zaptex_version(void) { return "zaptex 0.1 (gfx950; reference zap_tex.h 1.3.0)"; }

}  // extern "C"
