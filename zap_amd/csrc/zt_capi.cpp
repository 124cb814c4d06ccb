// zt_capi.cpp — the C ABI of libzaptex (include/zaptex.h): argument checks, device set-up, table uploads,
// staging of host buffers, and kernel launches.  No CPU fallback: every path runs the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/zaptex.h"
#include "zt_astc_setup.h"
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

thread_local char g_err[512];
std::mutex g_mu;     // per-device set-up, and the host slots' busy flags
int g_m1k = 4;
int g_climb_refill = 16;  // zt_bc1_climb refills once this many lanes of a wave are idle
int g_climb4_max = 1 << 20;   // BC1-family blocks up to which a launch starts with two lanes per block (zt_bc14_encode2)
                              // and climbs with four (zt_bc1_climb4)
int g_climb4_range = 0;       // zt_bc1_climb4's blocks per wave (0: automatic)
int g_climb_group = 4;        // lanes per block of the grouped climb (4 or 8)
int g_s6_chunks = 2;          // BC7 mode-6 stage: 0 = one pass per block (zt_bc7_stage6), k > 0 = refits queued per
                              // wave, which takes k 64-block chunks at a time (zt_bc7_stage6q)

// launch shapes of the grid-stride kernels: workgroup size (64, 128 or 256) and the cap on workgroups;
// zaptex_set_launch changes them (the autotuner's results), a batch captures those in force when it is built
struct LaunchCfg { const char *name; int wg, max_groups; };
LaunchCfg g_cfg[] = { { "zt_bc14_encode", 256, 8192 }, { "zt_bc7_stage6", 256, 8192 }, { "zt_bc7_stage5", 256, 2048 },
                      { "zt_bc7_s1", 256, 1024 }, { "zt_mip_batch", 256, 8192 }, { "zt_bc1_climb", 256, 8192 } };
enum { CFG_BC14, CFG_S6, CFG_S5, CFG_S1, CFG_MIPB, CFG_CLIMB, CFG_N };
// workgroups for `items` work items of `per_group` each under config c (at least 1)
int groups_for(int c, uint64_t items, uint64_t per_group) {
    uint64_t g = (items + per_group - 1) / per_group;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)g_cfg[c].max_groups));
}

// optional per-kernel timing: HIP events recorded on the launch stream around every kernel (not while a
// batch graph is being captured)
bool g_profile = false;
bool g_capturing = false;
// set while a no-alloc call enqueues, so that kernel timing (which creates HIP events) stays off
thread_local bool t_noalloc = false;
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
    const bool prof = g_profile && !g_capturing && !t_noalloc;
    if (prof) {
        HIP_TRY(hipEventCreate(&a)); HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(a, s));
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ZAPTEX_E_DEVICE, "launch of %s: %s", name, hipGetErrorString(e));
    if (prof) {
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

// The encoders' workspace for zaptex_encode_dev_noalloc: a file-scope device array, one copy per GPU, held until
// the process ends.  Host code never passes its name to a kernel (on the host the name is a stand-in, not a
// device address); it uses the per-device address that ensure_ready reads with hipGetSymbolAddress.
constexpr size_t kScratchBytes = ZAPTEX_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kScratchBytes];
// one host entry point call's staging: its stream, device buffers, and pinned host memory for rows that are not
// packed.  Each call holds one slot from start to finish (HostLease).
struct HostSlot {
    hipStream_t stream = nullptr;
    void *in = nullptr, *out = nullptr, *ws = nullptr, *pin = nullptr, *mip = nullptr, *rdo = nullptr;
    size_t in_cap = 0, out_cap = 0, ws_cap = 0, pin_cap = 0, mip_cap = 0, rdo_cap = 0;
};
constexpr int kHostSlots = 8;   // host calls that run side by side on one device
struct DeviceState {
    bool ready = false;
    hipStream_t stream = nullptr;   // the library's own set-up work (batch tables, network initialisation)
    void *scratch = nullptr;        // this device's address of `scratch`
    HostSlot slots[kHostSlots];
    bool busy[kHostSlots] = {};
};
DeviceState g_dev[64];

int ensure_buf(void *&p, size_t &cap, size_t need) {
    if (need <= cap) return ZAPTEX_OK;
    if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
    size_t n = std::max(align_up(need, 1 << 20), (size_t)1 << 20);
    if (hipMalloc(&p, n) != hipSuccess) { p = nullptr; return fail(ZAPTEX_E_NOMEM, "hipMalloc(%zu) failed", n); }
    cap = n;
    return ZAPTEX_OK;
}

// Host <-> device rows for the host entry points.  Packed rows move in one copy; rows with gaps between them are
// packed through pinned staging memory by the CPU, because a 2D copy from pageable memory is staged one row at a
// time (a 4096-row image took 60 ms that way instead of 5).
int ensure_pinned(HostSlot *st, size_t need) {
    if (need <= st->pin_cap) return ZAPTEX_OK;
    HIP_TRY(hipStreamSynchronize(st->stream));   // the old buffer may still feed a copy
    if (st->pin) { (void)hipHostFree(st->pin); st->pin = nullptr; st->pin_cap = 0; }
    size_t n = std::max(align_up(need, 1 << 20), (size_t)1 << 20);
    if (hipHostMalloc(&st->pin, n, hipHostMallocDefault) != hipSuccess) { st->pin = nullptr; return fail(ZAPTEX_E_NOMEM, "hipHostMalloc(%zu) failed", n); }
    st->pin_cap = n;
    return ZAPTEX_OK;
}
int rows_to_device(HostSlot *st, void *dst, const void *src, size_t spitch, size_t row, size_t h) {
    if (spitch == row) { HIP_TRY(hipMemcpyAsync(dst, src, row * h, hipMemcpyHostToDevice, st->stream)); return ZAPTEX_OK; }
    if (int rc = ensure_pinned(st, row * h)) return rc;
    HIP_TRY(hipStreamSynchronize(st->stream));
    for (size_t y = 0; y < h; y++) memcpy((uint8_t *)st->pin + y * row, (const uint8_t *)src + y * spitch, row);
    HIP_TRY(hipMemcpyAsync(dst, st->pin, row * h, hipMemcpyHostToDevice, st->stream));
    return ZAPTEX_OK;
}
// ends with the stream synchronised
int rows_to_host(HostSlot *st, void *dst, size_t dpitch, const void *src, size_t row, size_t h) {
    if (dpitch == row) {
        HIP_TRY(hipMemcpyAsync(dst, src, row * h, hipMemcpyDeviceToHost, st->stream));
        HIP_TRY(hipStreamSynchronize(st->stream));
        return ZAPTEX_OK;
    }
    if (int rc = ensure_pinned(st, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->pin, src, row * h, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    for (size_t y = 0; y < h; y++) memcpy((uint8_t *)dst + y * dpitch, (const uint8_t *)st->pin + y * row, row);
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
    HIP_TRY(hipGetSymbolAddress(&st->scratch, HIP_SYMBOL(scratch)));
    if (!st->scratch || ((uintptr_t)st->scratch) % 256) return fail(ZAPTEX_E_DEVICE, "scratch array of device %d is not 256-byte aligned", dev);
    HIP_TRY(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    st->ready = true;
    return ZAPTEX_OK;
}

// A host entry point's lease on one of its device's host slots: calls from different threads run side by side on
// their own streams and staging buffers, up to kHostSlots per device; further calls wait for a slot.  (The
// reference functions share no state, so concurrent callers are the norm, e.g. one thread per band.)
std::condition_variable g_slot_cv;
struct HostLease {
    DeviceState *dev = nullptr;
    HostSlot *h = nullptr;
    int idx = -1;
    int acquire() {
        std::unique_lock<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(dev)) return rc;
        g_slot_cv.wait(lk, [&] {
            for (int i = 0; i < kHostSlots; i++) if (!dev->busy[i]) return true;
            return false;
        });
        for (idx = 0; dev->busy[idx]; idx++) {}
        HostSlot &slot = dev->slots[idx];
        if (!slot.stream) HIP_TRY(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking));
        dev->busy[idx] = true;
        h = &slot;
        return ZAPTEX_OK;
    }
    ~HostLease() {
        if (!h) return;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            dev->busy[idx] = false;
        }
        g_slot_cv.notify_one();
    }
};

bool is_bcn(int fmt) { return fmt >= ZAPTEX_BC1 && fmt <= ZAPTEX_BC7; }
bool valid_fmt(int fmt) { return fmt >= ZAPTEX_BC1 && fmt <= ZAPTEX_ASTC; }
size_t image_out_bytes(const zaptex_image &im) { return out_bytes(im.w, im.h, im.fmt == ZAPTEX_BC6H || im.fmt == ZAPTEX_ASTC ? ZAPTEX_BC7 : im.fmt); }
uint32_t nblocks(const zaptex_image &im) { return (uint32_t)((uint64_t)((im.w + 3) / 4) * (uint64_t)((im.h + 3) / 4)); }
// Sizes the kernels can index: sides up to 2^30 and fewer than 2^31 blocks per image (block indices are 32-bit;
// a launch's total is checked the same way).  The reference takes any size; these bounds are far beyond memory.
constexpr int kMaxSide = 1 << 30;
bool size_ok(int w, int h) {
    return w >= 1 && h >= 1 && w <= kMaxSide && h <= kMaxSide && (uint64_t)((w + 3) / 4) * (uint64_t)((h + 3) / 4) < (1ull << 31);
}

int check_image(const zaptex_image &im) {
    if (!valid_fmt(im.fmt)) return fail(ZAPTEX_E_ARG, "unknown format %d", im.fmt);
    if (!size_ok(im.w, im.h)) return fail(ZAPTEX_E_ARG, "bad size %dx%d", im.w, im.h);
    if (!im.src || !im.dst) return fail(ZAPTEX_E_ARG, "null image pointer");
    size_t px = im.fmt == ZAPTEX_BC6H ? 16 : 4;
    if (im.stride < (size_t)im.w * px) return fail(ZAPTEX_E_ARG, "stride %zu below row size", im.stride);
    if ((((uintptr_t)im.src) | im.stride) % px) return fail(ZAPTEX_E_ARG, "source rows must be %zu-byte aligned", px);
    if (((uintptr_t)im.dst) % block_bytes(im.fmt == ZAPTEX_BC6H || im.fmt == ZAPTEX_ASTC ? ZAPTEX_BC7 : im.fmt))
        return fail(ZAPTEX_E_ARG, "destination must be block-aligned");
    return ZAPTEX_OK;
}

// workspace layout over the BC7 blocks: [sharded counters][meta5 16][px5 64][meta1 16][cand1 8][sub_err 64][sub_state 256] B
// per block, for the block count rounded up to 512: list 5 takes one append per stage-6 lane (64 * ceil(n / 512)
// slots per shard), list 1 at most 16 per stage-5 wave of 64 (entry, rotation) lanes (16 * ceil(n / 128) per shard)
uint64_t bc7_rows(uint64_t n) { return (n + 511) / 512 * 512; }
size_t ws_bytes_for(uint64_t bc7_blocks) { return bc7_blocks ? ZT_BC7_COUNTERS * 4 + bc7_rows(bc7_blocks) * 424 : 0; }
// an encode workspace: the BC7 lists, then 16 B of BC1 hill-climb state per BC1/BC3/BC4/BC5 block
size_t ws_climb_offset(uint64_t bc7_blocks) { return (ws_bytes_for(bc7_blocks) + 255) / 256 * 256; }
size_t ws_total(uint64_t bc7_blocks, uint64_t bc14_blocks) { return ws_climb_offset(bc7_blocks) + bc14_blocks * 16; }
// then the ASTC stages' regions (zt_astc_ws): the sharded counters of three lists (ZT_SHARDS lines of 256 B each),
// 300 B per ASTC block, and the three lists at 4 B per block rounded up to 512 blocks
constexpr size_t kAstcCounters = 3 * ZT_SHARDS * ZT_SHARD_STRIDE;   // uint32 counters
size_t ws_astc_offset(uint64_t bc7_blocks, uint64_t bc14_blocks) { return (ws_total(bc7_blocks, bc14_blocks) + 255) / 256 * 256; }
uint64_t astc_list_rows(uint64_t n) { return (n + 511) / 512 * 512; }
size_t ws_astc_bytes(uint64_t n) { return n ? kAstcCounters * 4 + n * 300 + astc_list_rows(n) * 12 : 0; }
size_t ws_total(uint64_t bc7_blocks, uint64_t bc14_blocks, uint64_t astc_blocks) {
    return astc_blocks ? ws_astc_offset(bc7_blocks, bc14_blocks) + ws_astc_bytes(astc_blocks) : ws_total(bc7_blocks, bc14_blocks);
}
zt_astc_ws astc_ws_in(void *ws, uint64_t n) {
    uint8_t *p = (uint8_t *)ws;
    zt_astc_ws W;
    W.counters = (unsigned *)p; p += kAstcCounters * 4;
    W.px = (uint4 *)p; p += n * 64;
    W.blk1 = (uint4 *)p; p += n * 48;
    W.blk2 = (uint4 *)p; p += n * 128;
    W.err1 = (int *)p; p += n * 12;
    W.err2 = (int *)p; p += n * 32;
    W.cand = (uint32_t *)p; p += n * 16;
    const uint64_t r = astc_list_rows(n);
    W.cap = (uint32_t)(r / ZT_SHARDS);
    W.n = (uint32_t)n;
    W.lo = (uint32_t *)p; p += r * 4;
    W.lt = (uint32_t *)p; p += r * 4;
    W.lp = (uint32_t *)p;
    return W;
}

zt_bc7_lists lists_in(void *ws, uint64_t blocks) {
    uint8_t *p = (uint8_t *)ws + ZT_BC7_COUNTERS * 4;
    const uint64_t n = bc7_rows(blocks);
    zt_bc7_lists L;
    L.cap5 = (uint32_t)((blocks + 511) / 512 * 64);
    L.cap1 = (uint32_t)((blocks + 127) / 128 * 16);
    L.counters = (unsigned *)ws;
    L.meta5 = (uint4 *)p;
    L.px5 = (uint4 *)(p + n * 16);
    L.meta1 = (uint4 *)(p + n * 80);
    L.cand1 = p + n * 96;
    L.sub_err = (int *)(p + n * 104);
    L.sub_state = (uint4 *)(p + n * 168);
    return L;
}

int grid_for(uint64_t items, int per) {
    uint64_t g = (items + per - 1) / per;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 8192));
}

// BC1/BC3/BC4/BC5 over one job table: everything but the BC1 hill climb, then the hill climb (climb: 16 B/block)
int enqueue_bc14(const zt_jobs &J, uint32_t tot, uint4 *climb, hipStream_t s) {
    const int wg = g_cfg[CFG_BC14].wg, wc = g_cfg[CFG_CLIMB].wg;
    const bool small = tot <= (uint64_t)g_climb4_max;   // too few blocks to fill the GPU one lane per block
    if (int rc = launch("zt_bc14_encode", s, [&] {
            if (small) hipLaunchKernelGGL(zt_bc14_encode2, dim3((unsigned)((2ull * tot + 255) / 256)), dim3(256), 0, s, J, tot, climb);
            else hipLaunchKernelGGL(zt_bc14_encode, dim3(groups_for(CFG_BC14, tot, wg)), dim3(wg), 0, s, J, tot, climb); }))
        return rc;
    if (small) {
        // G lanes per block: ranges of g_climb4_range blocks (automatic: 64 / G blocks per wave up to 4096 waves)
        const uint32_t G = (uint32_t)g_climb_group, per = 64 / G;
        const uint32_t range = g_climb4_range ? (uint32_t)g_climb4_range
                                              : (uint32_t)std::max<uint64_t>(per, ((tot + 4095) / 4096 + per - 1) / per * per);
        const uint64_t waves = (tot + range - 1) / range;
        if (int rc = launch("zt_bc1_climb", s, [&] {
                if (G == 8) hipLaunchKernelGGL(zt_bc1_climb8, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, J, tot, climb, range);
                else hipLaunchKernelGGL(zt_bc1_climb4, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, J, tot, climb, range); }))
            return rc;
    } else {
        // ranges of up to ZT_CLIMB_RANGE blocks, smaller for small jobs so that about 4096 waves share the work
        const uint32_t range = (uint32_t)std::min<uint64_t>(ZT_CLIMB_RANGE, std::max<uint64_t>(64, ((tot + 4095) / 4096 + 63) / 64 * 64));
        const uint64_t waves = (tot + range - 1) / range;
        if (int rc = launch("zt_bc1_climb", s, [&] {
                hipLaunchKernelGGL(zt_bc1_climb, dim3(groups_for(CFG_CLIMB, waves, wc / 64)), dim3(wc), 0, s, J, tot, climb, g_climb_refill, range); }))
            return rc;
    }
    if (small) return ZAPTEX_OK;   // the grouped climb gives its moved blocks their indices itself
    return launch("zt_bc1_finish", s, [&] {
        hipLaunchKernelGGL(zt_bc1_finish, dim3(groups_for(CFG_BC14, tot, wg)), dim3(wg), 0, s, J, tot, (const uint4 *)climb); });
}

// the reference BC7 search (mode 6, then 5, then 1) over one job table; ws >= ws_bytes_for(tot)
int enqueue_bc7(const zt_jobs &J, uint32_t tot, void *ws, size_t ws_bytes, hipStream_t s) {
    if (ws_bytes < ws_bytes_for(tot) || !ws) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu bytes", ws_bytes, ws_bytes_for(tot));
    zt_bc7_lists L = lists_in(ws, tot);
    int rc;
    if ((rc = launch("zt_zero_u32", s, [&] {
            hipLaunchKernelGGL(zt_zero_u32, dim3(ZT_BC7_COUNTERS / 256), dim3(256), 0, s, L.counters, ZT_BC7_COUNTERS); })))
        return rc;
    const int w6 = g_cfg[CFG_S6].wg, w5 = g_cfg[CFG_S5].wg, w1 = g_cfg[CFG_S1].wg;
    if (g_s6_chunks > 0) {
        // the waves take chunks until none are left: as many as the chip holds at once, no more than there are grabs
        int dev = 0, cus = 256, per_cu = 16;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)zt_bc7_stage6q, 64, 0) != hipSuccess || per_cu < 1) per_cu = 16;
        const uint32_t kc = (uint32_t)g_s6_chunks;
        const uint64_t nch = ((uint64_t)tot + 63) / 64;   // (a wave for every shard that has chunks)
        const uint32_t waves = (uint32_t)std::max<uint64_t>(std::min<uint64_t>((nch + kc - 1) / kc, (uint64_t)cus * (uint64_t)per_cu),
                                                            std::min<uint64_t>(nch, ZT_SHARDS));
        if ((rc = launch("zt_bc7_stage6", s, [&] { hipLaunchKernelGGL(zt_bc7_stage6q, dim3(waves), dim3(64), 0, s, J, tot, L, kc); })))
            return rc;
    } else if ((rc = launch("zt_bc7_stage6", s, [&] {
                   hipLaunchKernelGGL(zt_bc7_stage6, dim3(groups_for(CFG_S6, tot, w6)), dim3(w6), 0, s, J, tot, L); })))
        return rc;
    if ((rc = launch("zt_bc7_stage5", s, [&] {
            hipLaunchKernelGGL(zt_bc7_stage5, dim3(groups_for(CFG_S5, (uint64_t)tot * 4, w5)), dim3(w5), 0, s, J, L); })))
        return rc;
    if (g_m1k <= 8) {
        const int gsz = groups_for(CFG_S1, (uint64_t)tot / 16 + 1, w1 / 64);
        if ((rc = launch("zt_bc7_s1_rank", s, [&] { hipLaunchKernelGGL(zt_bc7_s1_rank, dim3(gsz), dim3(w1), 0, s, L, g_m1k, L.cand1); })))
            return rc;
        if (64 % (2 * g_m1k) == 0)   // a block's fits make whole lane groups: they pick the winner themselves
            return launch("zt_bc7_s1_fitpick", s, [&] {
                hipLaunchKernelGGL(zt_bc7_s1_fitpick, dim3(gsz), dim3(w1), 0, s, J, L, g_m1k, (const uint8_t *)L.cand1); });
        if ((rc = launch("zt_bc7_s1_fit", s, [&] {
                hipLaunchKernelGGL(zt_bc7_s1_fit, dim3(gsz), dim3(w1), 0, s, L, g_m1k, (const uint8_t *)L.cand1, L.sub_err, L.sub_state); })))
            return rc;
        return launch("zt_bc7_s1_pick", s, [&] {
            hipLaunchKernelGGL(zt_bc7_s1_pick, dim3(gsz), dim3(w1), 0, s, J, L, g_m1k, (const uint8_t *)L.cand1, (const int *)L.sub_err,
                               (const uint4 *)L.sub_state); });
    }
    return launch("zt_bc7_stage1", s, [&] {
        hipLaunchKernelGGL(zt_bc7_stage1, dim3(grid_for((uint64_t)tot, 4)), dim3(256), 0, s, J, L, g_m1k); });
}

// ASTC 4x4 in stages: classify, the 1-partition configurations, pick, pattern search, 2-partition trials, pick
int enqueue_astc(const zt_jobs &J, uint32_t tot, const zt_astc_ws &W, hipStream_t s) {
    const int g = grid_for(tot, 256);
    int rc;
    if ((rc = launch("zt_zero_u32", s, [&] {
            hipLaunchKernelGGL(zt_zero_u32, dim3((kAstcCounters + 255) / 256), dim3(256), 0, s, W.counters, (int)kAstcCounters); })))
        return rc;
    if ((rc = launch("zt_astc_classify", s, [&] { hipLaunchKernelGGL(zt_astc_classify, dim3(g), dim3(256), 0, s, J, tot, W); }))) return rc;
    if ((rc = launch("zt_astc_single", s, [&] { hipLaunchKernelGGL(zt_astc_single, dim3(grid_for((uint64_t)tot * 3, 256)), dim3(256), 0, s, W); })))
        return rc;
    if ((rc = launch("zt_astc_pick1", s, [&] { hipLaunchKernelGGL(zt_astc_pick1, dim3(g), dim3(256), 0, s, J, tot, W); }))) return rc;
    if ((rc = launch("zt_astc_search", s, [&] { hipLaunchKernelGGL(zt_astc_search, dim3(g), dim3(256), 0, s, W); }))) return rc;
    if ((rc = launch("zt_astc_pair", s, [&] { hipLaunchKernelGGL(zt_astc_pair, dim3(grid_for((uint64_t)tot * 8, 256)), dim3(256), 0, s, W); })))
        return rc;
    return launch("zt_astc_pick2", s, [&] { hipLaunchKernelGGL(zt_astc_pick2, dim3(g), dim3(256), 0, s, J, W); });
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
    const size_t need = ws_total(tot[1], tot[0], tot_astc);
    if (ws_bytes < need || (!ws && need)) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu bytes", ws_bytes, need);
    int rc;
    if (tot[0] && (rc = enqueue_bc14(fam[0], tot[0], (uint4 *)((uint8_t *)ws + ws_climb_offset(tot[1])), s))) return rc;
    if (tot_b6 && (rc = launch("zt_bc6h_encode", s, [&] {
            hipLaunchKernelGGL(zt_bc6h_encode, dim3(grid_for(tot_b6, 256)), dim3(256), 0, s, b6, tot_b6); })))
        return rc;
    if (tot_astc && (rc = enqueue_astc(astc, tot_astc, astc_ws_in((uint8_t *)ws + ws_astc_offset(tot[1], tot[0]), tot_astc), s))) return rc;
    if (tot[1]) return enqueue_bc7(fam[1], tot[1], ws, ws_bytes, s);
    return ZAPTEX_OK;
}

// ---- rate-distortion passes (rdo > 0) over images whose plain encoding is already enqueued into their dst: one
// stream per BC1, BC4 or BC7 image, two per BC3 (alpha BC4 half, 4-colour BC1 half) or BC5 (R and G halves).  Each
// stream is cut into segments walked speculatively, one wave each, then checked and repaired in order.
int g_rdo_seg = 0;      // segment length in blocks (0: automatic)
int g_rdo_warm = 512;   // warm-up blocks before a segment
uint32_t rdo_seg_len(uint64_t nb) {
    if (g_rdo_seg) return (uint32_t)g_rdo_seg;
    return (uint32_t)std::max<uint64_t>(64, ((nb + 1023) / 1024 + 15) / 16 * 16);   // about 1,024 waves per stream
}
int rdo_streams_of(int fmt) { return fmt == ZAPTEX_BC3 || fmt == ZAPTEX_BC5 ? 2 : 1; }
// scratch of one image: the plain copy, then per stream the per-block distances, the segment states and ends
size_t rdo_scratch_image(const zaptex_image &im) {
    const uint64_t nb = nblocks(im), nseg = (nb + rdo_seg_len(nb) - 1) / rdo_seg_len(nb);
    const size_t cache = nseg > 1 ? align_up(nb * ZT_RDO_LANES * 16, 256) + align_up(nb * ZT_RDO_LANES * 4, 256) : 0;
    return align_up(image_out_bytes(im), 256) +
           (size_t)rdo_streams_of(im.fmt) * (align_up(nb * 2, 256) + align_up(nseg * sizeof(zt_rdo_state), 256) + align_up(nseg * 4, 256) + cache);
}
size_t rdo_scratch_bytes(const zaptex_image *imgs, int n) {
    size_t t = 0;
    for (int i = 0; i < n; i++) t += rdo_scratch_image(imgs[i]);
    return t;
}
int enqueue_rdo(const zaptex_image *imgs, int n, float rdo, void *scratch, size_t scratch_bytes, hipStream_t s) {
    if (!(rdo > 0)) return ZAPTEX_OK;   // zap_tex.h:249, :284, :794: rdo <= 0 (or NaN) is the plain encoding
    for (int i = 0; i < n; i++)
        if (!is_bcn(imgs[i].fmt)) return fail(ZAPTEX_E_ARG, "rdo applies to BC1/BC3/BC4/BC5/BC7 only (image %d)", i);
    if (scratch_bytes < rdo_scratch_bytes(imgs, n) || !scratch) return fail(ZAPTEX_E_ARG, "rdo scratch %zu < %zu bytes", scratch_bytes, rdo_scratch_bytes(imgs, n));
    zt_rdo_streams S;
    memset(&S, 0, sizeof S);
    int k = 0;
    uint64_t segs = 0;
    auto flush = [&]() -> int {
        if (!k) return ZAPTEX_OK;
        const int cnt = k;
        const uint64_t g = segs;
        k = 0; segs = 0;
        int rc = launch("zt_rdo_spec", s, [&] { hipLaunchKernelGGL(zt_rdo_spec, dim3((unsigned)g), dim3(64), 0, s, S, rdo); });
        if (rc) return rc;
        return launch("zt_rdo_fix", s, [&] { hipLaunchKernelGGL(zt_rdo_fix, dim3(cnt), dim3(64), 0, s, S, rdo); });
    };
    uint8_t *p = (uint8_t *)scratch;
    for (int i = 0; i < n; i++) {
        const zaptex_image &im = imgs[i];
        const uint64_t nb = nblocks(im);
        const uint32_t seg = rdo_seg_len(nb), nseg = (uint32_t)((nb + seg - 1) / seg);
        uint8_t *plain = p;
        p += align_up(image_out_bytes(im), 256);
        HIP_TRY(hipMemcpyAsync(plain, im.dst, image_out_bytes(im), hipMemcpyDeviceToDevice, s));
        const int ns = rdo_streams_of(im.fmt);
        for (int t = 0; t < ns; t++) {
            if (k == ZT_RDO_STREAMS) if (int rc = flush()) return rc;
            zt_rdo_stream &r = S.s[k++];
            r.src = (const uint8_t *)im.src; r.out = (uint8_t *)im.dst; r.plain = plain; r.stride = im.stride;
            r.dists = (uint16_t *)p; p += align_up(nb * 2, 256);
            r.states = (zt_rdo_state *)p; p += align_up((uint64_t)nseg * sizeof(zt_rdo_state), 256);
            r.seg_end = (uint32_t *)p; p += align_up((uint64_t)nseg * 4, 256);
            if (nseg > 1) {   // the speculative walk's candidates, which the ordered repair reuses
                r.cand = (uint4 *)p; p += align_up(nb * ZT_RDO_LANES * 16, 256);
                r.cand_err = (int *)p; p += align_up(nb * ZT_RDO_LANES * 4, 256);
            }
            r.w = im.w; r.h = im.h; r.bw = (im.w + 3) / 4; r.bh = (im.h + 3) / 4;
            r.bs = (int)block_bytes(im.fmt); r.seg = seg; r.warm = (uint32_t)g_rdo_warm; r.nseg = nseg;
            switch (im.fmt) {   // zap_tex.h:1494-1503
            case ZAPTEX_BC1: r.kind = ZT_RDO_BC1; break;
            case ZAPTEX_BC4: r.kind = ZT_RDO_BC4; break;
            case ZAPTEX_BC3: if (t == 0) { r.kind = ZT_RDO_BC4; r.chan = 3; } else { r.kind = ZT_RDO_BC1; r.off = 8; r.force4 = 1; } break;
            case ZAPTEX_BC5: r.kind = ZT_RDO_BC4; r.chan = t; r.off = 8 * t; break;
            default: r.kind = ZT_RDO_BC7; break;
            }
            segs += nseg;
        }
    }
    return flush();
}

int enqueue_mip(const void *src, int w, int h, size_t stride, int srgb, void *dst, size_t dst_stride, hipStream_t s) {
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    dim3 grid((nw + 63) / 64, std::min((nh + 3) / 4, 65535)), block(64, 4);
    return launch("zt_mip_level", s, [&] {
        hipLaunchKernelGGL(zt_mip_level, grid, block, 0, s, (const uint8_t *)src, w, h, (uint64_t)stride, srgb, (uint8_t *)dst, (uint64_t)dst_stride);
    });
}

// level-1 tile side of the fused mip passes (ZAPTEX_OPT_MIP_TILE): 32 (up to 6 levels per pass; the default: four
// times the workgroups of 64, 1 us faster on the 4096^2 chain and 0.7% on the 512-texture batch) or 64 (up to 7)
int g_mip_tile = 32;
int tile_levels(int tile) { return tile == 32 ? 6 : 7; }

// whole chain: fused tile passes (up to 7 or 6 levels each) while the source is larger than 128 on a side, then
// one workgroup for the rest; a level with a side of 1 and the other side above 128 goes level by level
int enqueue_mip_chain(const void *base, int w, int h, size_t stride, int srgb, void *const *levels, int nlevels, hipStream_t s) {
    std::vector<int> lw(1, w), lh(1, h);
    for (int i = 0; i < nlevels; i++) { lw.push_back(lw.back() > 1 ? lw.back() / 2 : 1); lh.push_back(lh.back() > 1 ? lh.back() / 2 : 1); }
    const void *src = base;
    size_t ss = stride;
    int l = 0, rc;
    while (l < nlevels && (lw[l] > 1 || lh[l] > 1)) {
        const int cw = lw[l], ch = lh[l];
        zt_mip_levels L;
        memset(&L, 0, sizeof L);
        const int T = g_mip_tile;
        if ((cw > 128 || ch > 128) && cw >= 2 && ch >= 2 && (lh[l + 1] + T - 1) / T <= 65535) {
            int k = 0;
            while (k < tile_levels(T) && l + k < nlevels && lw[l + k] >= 2 && lh[l + k] >= 2) {
                L.p[k] = (uint8_t *)levels[l + k]; L.w[k] = lw[l + k + 1]; L.h[k] = lh[l + k + 1];
                k++;
            }
            dim3 grid((lw[l + 1] + T - 1) / T, (lh[l + 1] + T - 1) / T);
            if (T == 32) {
                if ((rc = launch("zt_mip_tiles", s, [&] {
                        hipLaunchKernelGGL(zt_mip_tiles32, grid, dim3(256), 0, s, (const uint8_t *)src, cw, ch, (uint64_t)ss, srgb, k, L); })))
                    return rc;
            } else if ((rc = launch("zt_mip_tiles", s, [&] {
                           hipLaunchKernelGGL(zt_mip_tiles, grid, dim3(256), 0, s, (const uint8_t *)src, cw, ch, (uint64_t)ss, srgb, k, L); })))
                return rc;
            src = levels[l + k - 1]; ss = (size_t)lw[l + k] * 4; l += k;
        } else if (cw <= 128 && ch <= 128) {
            const int k = std::min(16, nlevels - l);
            for (int i = 0; i < k; i++) { L.p[i] = (uint8_t *)levels[l + i]; L.w[i] = lw[l + i + 1]; L.h[i] = lh[l + i + 1]; }
            if ((rc = launch("zt_mip_tail", s, [&] {
                    hipLaunchKernelGGL(zt_mip_tail, dim3(1), dim3(1024), 0, s, (const uint8_t *)src, cw, ch, (uint64_t)ss, srgb, k, L); })))
                return rc;
            l += k;
            src = levels[l - 1]; ss = (size_t)lw[l] * 4;
        } else {
            if ((rc = enqueue_mip(src, cw, ch, ss, srgb, levels[l], (size_t)lw[l + 1] * 4, s))) return rc;
            src = levels[l]; ss = (size_t)lw[l + 1] * 4; l++;
        }
    }
    return ZAPTEX_OK;
}

// run one host-buffer image through the device: stage in, encode, copy out
int host_encode(const void *src, int w, int h, size_t stride, size_t px_bytes, int fmt, void *out, float rdo = 0.0f) {
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t row = (size_t)w * px_bytes, inb = row * (size_t)h;
    zaptex_image im{ nullptr, nullptr, row, w, h, fmt };
    size_t ob = image_out_bytes(im);
    uint64_t n7 = fmt == ZAPTEX_BC7 ? nblocks(im) : 0, n14 = fmt <= ZAPTEX_BC5 ? nblocks(im) : 0, na = fmt == ZAPTEX_ASTC ? nblocks(im) : 0;
    if (int rc = ensure_buf(st->in, st->in_cap, inb)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, ob)) return rc;
    if (n7 || n14 || na) if (int rc = ensure_buf(st->ws, st->ws_cap, ws_total(n7, n14, na))) return rc;
    im.src = st->in; im.dst = st->out;
    if (int rc = rows_to_device(st, st->in, src, stride, row, (size_t)h)) return rc;
    if (int rc = enqueue(&im, 1, st->ws, st->ws_cap, st->stream)) return rc;
    if (rdo > 0) {
        if (int rc = ensure_buf(st->rdo, st->rdo_cap, rdo_scratch_bytes(&im, 1))) return rc;
        if (int rc = enqueue_rdo(&im, 1, rdo, st->rdo, st->rdo_cap, st->stream)) return rc;
    }
    HIP_TRY(hipMemcpyAsync(out, st->out, ob, hipMemcpyDeviceToHost, st->stream));
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

// ---- texture files as `zap tex` writes them (zap.c:309-355)
void put32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }
int level_count(int w, int h, int mips) {
    int n = 1;
    while (mips && (w > 1 || h > 1)) { w = w > 1 ? w / 2 : 1; h = h > 1 ? h / 2 : 1; n++; }
    return n;
}
// DDS file header for BC1/3/4/5/7: "DDS " + DDS_HEADER (124 B), + DDS_HEADER_DXT10 (20 B) for BC7 or an sRGB
// format; returns its length (128 or 148).  BC4/BC5 are never sRGB (the caller clears the flag).
size_t dds_header(uint8_t hdr[148], int w, int h, int fmt, int levels, int mips, int srgb) {
    static const char kFourCC[5][5] = { "DXT1", "DXT5", "BC4U", "BC5U", "DX10" };
    static const uint32_t kDxgi[5] = { 71, 77, 80, 83, 98 };   // DXGI_FORMAT_BC1/3/4/5/7_UNORM; +1 is the _SRGB one
    const bool dx10 = fmt == ZAPTEX_BC7 || srgb;
    memset(hdr, 0, 148);
    memcpy(hdr, "DDS ", 4);
    put32(hdr + 4, 124);
    put32(hdr + 8, 0x1 | 0x2 | 0x4 | 0x1000 | 0x80000 | (mips ? 0x20000 : 0));  // CAPS HEIGHT WIDTH PIXELFORMAT LINEARSIZE (MIPMAPCOUNT)
    put32(hdr + 12, (uint32_t)h);
    put32(hdr + 16, (uint32_t)w);
    put32(hdr + 20, (uint32_t)out_bytes(w, h, fmt));                              // level 0's bytes
    put32(hdr + 28, (uint32_t)levels);
    put32(hdr + 76, 32);                                                          // DDS_PIXELFORMAT: size, FOURCC flag
    put32(hdr + 80, 0x4);
    memcpy(hdr + 84, dx10 ? "DX10" : kFourCC[fmt], 4);
    put32(hdr + 108, 0x1000 | (mips ? 0x400008 : 0));                             // TEXTURE (+ COMPLEX | MIPMAP)
    if (!dx10) return 128;
    put32(hdr + 128, kDxgi[fmt] + (srgb ? 1u : 0u));
    put32(hdr + 132, 3);                                                          // TEXTURE2D
    put32(hdr + 140, 1);                                                          // array size
    return 148;
}
// .astc file header: magic, 4x4x1 block, then width, height and depth as 24-bit little-endian values
void astc_header(uint8_t hdr[16], int w, int h) {
    static const uint8_t kMagic[4] = { 0x13, 0xAB, 0xA1, 0x5C };
    memcpy(hdr, kMagic, 4);
    hdr[4] = 4; hdr[5] = 4; hdr[6] = 1;
    for (int k = 0; k < 3; k++) { hdr[7 + k] = (uint8_t)(w >> (8 * k)); hdr[10 + k] = (uint8_t)(h >> (8 * k)); hdr[13 + k] = (uint8_t)(1 >> (8 * k)); }
}
size_t tex_file_bytes(int w, int h, int fmt, int mips, int srgb) {
    if (fmt == ZAPTEX_ASTC) return 16 + out_bytes(w, h, ZAPTEX_BC7);
    if (fmt == ZAPTEX_BC4 || fmt == ZAPTEX_BC5) srgb = 0;
    size_t total = fmt == ZAPTEX_BC7 || srgb ? 148 : 128;
    for (int l = 0, n = level_count(w, h, mips), lw = w, lh = h; l < n; l++, lw = lw > 1 ? lw / 2 : 1, lh = lh > 1 ? lh / 2 : 1)
        total += out_bytes(lw, lh, fmt);
    return total;
}

}  // namespace

extern "C" {

size_t zaptex_tex_file_size(int w, int h, int fmt, int mips, int srgb) {
    if (!size_ok(w, h) || (!is_bcn(fmt) && fmt != ZAPTEX_ASTC)) return 0;
    return tex_file_bytes(w, h, fmt, mips, srgb);
}

int zaptex_tex_file(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, int mips, int srgb, void *out, size_t out_size) {
    if ((!is_bcn(fmt) && fmt != ZAPTEX_ASTC) || !size_ok(w, h) || !rgba || !out || stride < (size_t)w * 4)
        return fail(ZAPTEX_E_ARG, "bad arguments");
    const size_t need = tex_file_bytes(w, h, fmt, mips, srgb);
    if (out_size < need) return fail(ZAPTEX_E_ARG, "output buffer %zu < %zu bytes", out_size, need);
    uint8_t *o = (uint8_t *)out;
    if (fmt == ZAPTEX_ASTC) {   // one level, like the reference (its -m is ignored)
        astc_header(o, w, h);
        return host_encode(rgba, w, h, stride, 4, ZAPTEX_ASTC, o + 16);
    }
    if (fmt == ZAPTEX_BC4 || fmt == ZAPTEX_BC5) srgb = 0;   // channel data, never sRGB
    const int levels = level_count(w, h, mips);
    uint8_t hdr[148];   // the file can be shorter than 148 bytes (a 1x1 BC1 file is 136)
    const size_t hn = dds_header(hdr, w, h, fmt, levels, mips, srgb);
    memcpy(o, hdr, hn);
    // the whole chain on the device: base in once, levels 1.. generated in HBM, every level encoded, blocks out once
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    std::vector<int> lw(1, w), lh(1, h);
    for (int l = 1; l < levels; l++) {
        lw.push_back(lw.back() > 1 ? lw.back() / 2 : 1);
        lh.push_back(lh.back() > 1 ? lh.back() / 2 : 1);
    }
    // device offsets, 256-byte aligned: generated level l (>= 1) at moff[l - 1], level l's blocks at ooff[l]
    std::vector<size_t> moff, ooff;
    size_t mip_bytes = 0, blk_bytes = 0;
    for (int l = 0; l < levels; l++) {
        if (l) { moff.push_back(mip_bytes); mip_bytes += align_up((size_t)lw[l] * lh[l] * 4, 256); }
        ooff.push_back(blk_bytes);
        blk_bytes += align_up(out_bytes(lw[l], lh[l], fmt), 256);
    }
    const size_t row = (size_t)w * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, row * (size_t)h)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, blk_bytes)) return rc;
    if (levels > 1) if (int rc = ensure_buf(st->mip, st->mip_cap, mip_bytes)) return rc;
    uint64_t ws_need = 0;
    for (int l0 = 0; l0 < levels; l0 += ZT_INLINE_JOBS) {
        uint64_t n = 0;
        for (int l = l0; l < std::min(levels, l0 + ZT_INLINE_JOBS); l++) n += (uint64_t)((lw[l] + 3) / 4) * ((lh[l] + 3) / 4);
        ws_need = std::max<uint64_t>(ws_need, fmt == ZAPTEX_BC7 ? ws_total(n, 0) : ws_total(0, n));
    }
    if (int rc = ensure_buf(st->ws, st->ws_cap, ws_need)) return rc;
    if (rdo > 0) {   // the rate-distortion scratch of the largest group of levels, before anything is enqueued
        size_t rdo_need = 0;
        for (int l0 = 0; l0 < levels; l0 += ZT_INLINE_JOBS) {
            size_t b = 0;
            for (int l = l0; l < std::min(levels, l0 + ZT_INLINE_JOBS); l++) {
                const zaptex_image im{ nullptr, nullptr, (size_t)lw[l] * 4, lw[l], lh[l], fmt };
                b += rdo_scratch_image(im);
            }
            rdo_need = std::max(rdo_need, b);
        }
        if (int rc = ensure_buf(st->rdo, st->rdo_cap, rdo_need)) return rc;
    }
    if (int rc = rows_to_device(st, st->in, rgba, stride, row, (size_t)h)) return rc;
    std::vector<void *> lvl;
    for (int l = 1; l < levels; l++) lvl.push_back((uint8_t *)st->mip + moff[l - 1]);
    if (levels > 1) if (int rc = enqueue_mip_chain(st->in, w, h, row, srgb, lvl.data(), levels - 1, st->stream)) return rc;
    for (int l0 = 0; l0 < levels; l0 += ZT_INLINE_JOBS) {
        zaptex_image im[ZT_INLINE_JOBS];
        int n = 0;
        for (int l = l0; l < std::min(levels, l0 + ZT_INLINE_JOBS); l++, n++)
            im[n] = zaptex_image{ l ? lvl[l - 1] : st->in, (uint8_t *)st->out + ooff[l], (size_t)lw[l] * 4, lw[l], lh[l], fmt };
        if (int rc = enqueue(im, n, st->ws, st->ws_cap, st->stream)) return rc;
        if (rdo > 0) if (int rc = enqueue_rdo(im, n, rdo, st->rdo, st->rdo_cap, st->stream)) return rc;
    }
    size_t pos = hn;
    for (int l = 0; l < levels; l++) {
        const size_t b = out_bytes(lw[l], lh[l], fmt);
        HIP_TRY(hipMemcpyAsync(o + pos, (uint8_t *)st->out + ooff[l], b, hipMemcpyDeviceToHost, st->stream));
        pos += b;
    }
    HIP_TRY(hipStreamSynchronize(st->stream));
    return ZAPTEX_OK;
}

int zaptex_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, void *out) {
    if (!is_bcn(fmt)) return fail(ZAPTEX_E_ARG, "unknown zap_bc_format %d", fmt);
    if (w < 1 || h < 1 || !rgba || !out || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    return host_encode(rgba, w, h, stride, 4, fmt, out, rdo);
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
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t row = (size_t)w * 4, orow = (size_t)nw * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, row * h)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, orow * nh)) return rc;
    if (int rc = rows_to_device(st, st->in, src, stride, row, (size_t)h)) return rc;
    if (int rc = enqueue_mip(st->in, w, h, row, srgb, st->out, orow, st->stream)) return rc;
    return rows_to_host(st, dst, dst_stride, st->out, orow, (size_t)nh);
}

int zaptex_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t dst_stride) {
    if (w < 1 || h < 1 || !src || !dst || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    if (dst_stride < (size_t)nw * 4) return fail(ZAPTEX_E_ARG, "dst_stride below row size");
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t row = (size_t)w * 16, orow = (size_t)nw * 16;
    if (int rc = ensure_buf(st->in, st->in_cap, row * h)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, orow * nh)) return rc;
    if (int rc = rows_to_device(st, st->in, src, stride * 4, row, (size_t)h)) return rc;
    dim3 grid((nw + 63) / 64, std::min((nh + 3) / 4, 65535)), block(64, 4);
    hipLaunchKernelGGL(zt_mip_level_f, grid, block, 0, st->stream, (const float *)st->in, w, h, (uint64_t)w * 4, (float *)st->out, (uint64_t)nw * 4);
    HIP_TRY(hipGetLastError());
    return rows_to_host(st, dst, dst_stride * 4, st->out, orow, (size_t)nh);
}

int zaptex_bc_decode(const void *bc, int w, int h, int fmt, uint8_t *rgba, size_t stride) {
    if (!is_bcn(fmt) || w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t ib = out_bytes(w, h, fmt), row = (size_t)w * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_bc_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, fmt, (uint8_t *)st->out, (uint64_t)row);
    HIP_TRY(hipGetLastError());
    return rows_to_host(st, rgba, stride, st->out, row, (size_t)h);
}

int zaptex_bc6h_decode(const void *bc, int w, int h, uint16_t *rgba, size_t stride) {
    if (w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t ib = out_bytes(w, h, ZAPTEX_BC7), row = (size_t)w * 8;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_bc6h_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, (uint16_t *)st->out, (uint64_t)w * 4);
    HIP_TRY(hipGetLastError());
    return rows_to_host(st, rgba, stride * 2, st->out, row, (size_t)h);
}

int zaptex_astc_decode(const void *bc, int w, int h, uint8_t *rgba, size_t stride) {
    if (w < 1 || h < 1 || !bc || !rgba || stride < (size_t)w * 4) return fail(ZAPTEX_E_ARG, "bad arguments");
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
    size_t ib = out_bytes(w, h, ZAPTEX_BC7), row = (size_t)w * 4;
    if (int rc = ensure_buf(st->in, st->in_cap, ib)) return rc;
    if (int rc = ensure_buf(st->out, st->out_cap, row * h)) return rc;
    HIP_TRY(hipMemcpyAsync(st->in, bc, ib, hipMemcpyHostToDevice, st->stream));
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    hipLaunchKernelGGL(zt_astc_decode, dim3((bw * bh + 255) / 256), dim3(256), 0, st->stream, (const uint8_t *)st->in, w, h, (uint8_t *)st->out, (uint64_t)row);
    HIP_TRY(hipGetLastError());
    return rows_to_host(st, rgba, stride, st->out, row, (size_t)h);
}

static int shuffle(const void *in, size_t n, int fmt, void *out, int split) {
    if (!is_bcn(fmt) || (!in && n) || (!out && n)) return fail(ZAPTEX_E_ARG, "bad arguments");
    if (!n) return ZAPTEX_OK;
    HostLease lease;
    if (int rc = lease.acquire()) return rc;
    HostSlot *st = lease.h;
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
    uint64_t n7 = 0, n14 = 0, na = 0;
    for (int i = 0; i < n; i++) {
        if (imgs[i].fmt == ZAPTEX_BC7) n7 += nblocks(imgs[i]);
        else if (imgs[i].fmt == ZAPTEX_ASTC) na += nblocks(imgs[i]);
        else if (imgs[i].fmt <= ZAPTEX_BC5) n14 += nblocks(imgs[i]);
    }
    return ws_total(n7, n14, na);
}

int zaptex_encode_dev(const zaptex_image *imgs, int n, void *workspace, size_t ws_bytes, void *stream) {
    if (((uintptr_t)workspace) % 256) return fail(ZAPTEX_E_ARG, "workspace must be 256-byte aligned");
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

size_t zaptex_rdo_workspace_size(const zaptex_image *imgs, int n) {
    return align_up(zaptex_workspace_size(imgs, n), 256) + rdo_scratch_bytes(imgs, n);
}

int zaptex_encode_rdo_dev(const zaptex_image *imgs, int n, float rdo, void *workspace, size_t ws_bytes, void *stream) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS) return fail(ZAPTEX_E_ARG, "image count %d outside 1..%d", n, ZT_INLINE_JOBS);
    for (int i = 0; i < n; i++)
        if (!is_bcn(imgs[i].fmt)) return fail(ZAPTEX_E_ARG, "rdo applies to BC1/BC3/BC4/BC5/BC7 only (image %d)", i);
    const size_t plain_ws = align_up(zaptex_workspace_size(imgs, n), 256);
    if (rdo > 0 && ws_bytes < zaptex_rdo_workspace_size(imgs, n))
        return fail(ZAPTEX_E_ARG, "workspace %zu < %zu bytes (zaptex_rdo_workspace_size)", ws_bytes, zaptex_rdo_workspace_size(imgs, n));
    if (int rc = zaptex_encode_dev(imgs, n, workspace, ws_bytes, stream)) return rc;
    return enqueue_rdo(imgs, n, rdo, (uint8_t *)workspace + plain_ws, ws_bytes - plain_ws, (hipStream_t)stream);
}

// zaptex_encode_dev with the library's per-device scratch array as its workspace: no allocation of any kind
// (set-up on a device's first call reads the array's address; profiling events are not recorded)
int zaptex_encode_dev_noalloc(const zaptex_image *imgs, int n, void *stream) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS) return fail(ZAPTEX_E_ARG, "image count %d outside 1..%d", n, ZT_INLINE_JOBS);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) {
        if (int rc = check_image(imgs[i])) return rc;
        total += nblocks(imgs[i]);
    }
    if (total >= (1ull << 31)) return fail(ZAPTEX_E_ARG, "too many blocks in one call");
    const size_t need = zaptex_workspace_size(imgs, n);
    if (need > kScratchBytes)
        return fail(ZAPTEX_E_ARG, "these images need %zu bytes of workspace, above the %zu-byte scratch of zaptex_encode_dev_noalloc "
                    "(encode them with zaptex_encode_dev and a workspace of zaptex_workspace_size bytes)", need, kScratchBytes);
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    t_noalloc = true;
    const int rc = enqueue(imgs, n, st->scratch, kScratchBytes, (hipStream_t)stream);
    t_noalloc = false;
    return rc;
}

namespace {
// job table of BC7 images (all blocks numbered in image order); total blocks out
int bc7_jobs(const zaptex_image *imgs, int n, zt_jobs &J, uint32_t &total, const char *what) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS) return fail(ZAPTEX_E_ARG, "image count %d outside 1..%d", n, ZT_INLINE_JOBS);
    memset(&J, 0, sizeof J);
    uint64_t t = 0;
    for (int i = 0; i < n; i++) {
        if (int rc = check_image(imgs[i])) return rc;
        if (imgs[i].fmt != ZAPTEX_BC7) return fail(ZAPTEX_E_ARG, "%s encodes BC7 only", what);
        zt_job &j = J.inl[J.n++];
        j.src = (const uint8_t *)imgs[i].src; j.dst = (uint8_t *)imgs[i].dst; j.stride = imgs[i].stride;
        j.first = (uint32_t)t; j.w = imgs[i].w; j.h = imgs[i].h; j.bw = (imgs[i].w + 3) / 4; j.fmt = ZAPTEX_BC7; j.pad = 0;
        t += nblocks(imgs[i]);
    }
    if (t >= (1ull << 26)) return fail(ZAPTEX_E_ARG, "too many blocks in one call (%llu)", (unsigned long long)t);
    total = (uint32_t)t;
    return ZAPTEX_OK;
}
// the nine per-mode exhaustive launches over `total` dense blocks, or over the device list (list, *nlist <= total)
int enqueue_exhaustive(const uint4 *blocks, uint32_t total, const uint32_t *list, const unsigned *nlist, uint32_t *key, uint32_t *errs,
                       hipStream_t s) {
    typedef void (*Exh)(const uint4 *, uint32_t, const uint32_t *, const unsigned *, uint32_t *, uint32_t *);
    struct K { const char *name; Exh fn; int slots; } ks[9] = {
        { "zt_bc7_exh_m0", zt_bc7_exh_m0, 16 }, { "zt_bc7_exh_m1", zt_bc7_exh_m1, 64 }, { "zt_bc7_exh_m2", zt_bc7_exh_m2, 64 },
        { "zt_bc7_exh_m3", zt_bc7_exh_m3, 64 }, { "zt_bc7_exh_m4i0", zt_bc7_exh_m4i0, 4 }, { "zt_bc7_exh_m4i1", zt_bc7_exh_m4i1, 4 },
        { "zt_bc7_exh_m5", zt_bc7_exh_m5, 4 }, { "zt_bc7_exh_m6", zt_bc7_exh_m6, 1 }, { "zt_bc7_exh_m7", zt_bc7_exh_m7, 64 } };
    for (const K &k : ks)
        if (int rc = launch(k.name, s, [&] {
                hipLaunchKernelGGL(k.fn, dim3(grid_for((uint64_t)total * k.slots, 256)), dim3(256), 0, s, blocks, total, list, nlist, key, errs); }))
            return rc;
    return ZAPTEX_OK;
}
size_t full_ws_bytes(uint64_t total, int with_deltas) { return 256 + total * (64 + 4) + (with_deltas ? total * 2048 + 256 : 0); }
}  // namespace

size_t zaptex_bc7_full_workspace_size(const zaptex_image *imgs, int n, int with_deltas) {
    uint64_t t = 0;
    for (int i = 0; i < n; i++) t += nblocks(imgs[i]);
    return full_ws_bytes(t, with_deltas);
}

int zaptex_bc7_full_dev(const zaptex_image *imgs, int n, uint32_t *best_cand, uint16_t *deltas, void *ws, size_t ws_bytes, void *stream) {
    zt_jobs J;
    uint32_t total;
    if (int rc = bc7_jobs(imgs, n, J, total, "the exhaustive search")) return rc;
    if (!ws || ws_bytes < full_ws_bytes(total, deltas != nullptr))
        return fail(ZAPTEX_E_ARG, "workspace %zu < %zu", ws_bytes, full_ws_bytes(total, deltas != nullptr));
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    uint8_t *b = (uint8_t *)align_up((uintptr_t)ws, 256);
    uint4 *blocks = (uint4 *)b; b += (uint64_t)total * 64;
    uint32_t *key = (uint32_t *)b; b += (uint64_t)total * 4;
    uint32_t *errs = deltas ? (uint32_t *)align_up((uintptr_t)b, 256) : nullptr;
    int rc;
    if ((rc = launch("zt_bc7_gather", s, [&] { hipLaunchKernelGGL(zt_bc7_gather, dim3(grid_for(total, 256)), dim3(256), 0, s, J, total, blocks); })))
        return rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(grid_for(total, 256)), dim3(256), 0, s, key, (uint64_t)total, 0xFFFFFFFFu); })))
        return rc;
    if (errs && (rc = launch("zt_fill_u32", s, [&] {
                     hipLaunchKernelGGL(zt_fill_u32, dim3(grid_for((uint64_t)total * 512, 256)), dim3(256), 0, s, errs, (uint64_t)total * 512, 0xFFFFFFFFu); })))
        return rc;
    if ((rc = enqueue_exhaustive(blocks, total, nullptr, nullptr, key, errs, s))) return rc;
    if ((rc = launch("zt_bc7_finish", s, [&] {
            hipLaunchKernelGGL(zt_bc7_finish, dim3(grid_for(total, 256)), dim3(256), 0, s, J, (const uint4 *)blocks, total, (const uint32_t *)key, best_cand,
                               (const uint2 *)nullptr, (const unsigned *)nullptr); })))
        return rc;
    if (!deltas) return ZAPTEX_OK;
    return launch("zt_bc7_deltas", s, [&] {
        hipLaunchKernelGGL(zt_bc7_deltas, dim3(grid_for((uint64_t)total * 512, 256)), dim3(256), 0, s, (const uint32_t *)errs, (const uint32_t *)key, total,
                           deltas); });
}

// ===================================================================================== proposal network
namespace {
// parameters (fp32 master, weights stored [in x out]): W1 64x128, b1, W2 128x128, b2, W3 128x512, b3
constexpr int kIn = 64, kH = 128, kOut = 512;
constexpr int64_t oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH * kOut, kP = oB3 + kOut;
// bf16 weight copies, stored transposed [out x in] so forward GEMM operands stream along k
constexpr int64_t bW1 = 0, bW2 = bW1 + kIn * kH, bW3 = bW2 + kH * kH, kPb = bW3 + kH * kOut;
struct MlpState {  // views into the caller's state buffer
    float *p, *g, *m1, *m2, *scale;
    int *bfi;
    uint16_t *w;
    int *cfg;  // [0] feature normalisation (zt_mlp_features)
};
MlpState mlp_view(void *state) {
    uint8_t *b = (uint8_t *)state;
    MlpState v;
    v.p = (float *)b; v.g = v.p + kP; v.m1 = v.g + kP; v.m2 = v.m1 + kP; v.scale = v.m2 + kP;
    v.bfi = (int *)(v.scale + kP);
    v.w = (uint16_t *)(v.bfi + kP);
    v.cfg = (int *)(v.w + kPb);
    return v;
}
size_t mlp_state_bytes() { return (size_t)kP * 4 * 6 + (size_t)kPb * 2 + 256; }
struct MlpWs { uint16_t *X, *H1, *H2, *dZ, *dH2, *dH1; float *Z, *stats; };
size_t mlp_ws_bytes(int B) { return align_up((size_t)B * (kIn + 2 * kH + kOut + 2 * kH) * 2 + (size_t)B * kOut * 4 + 256, 256); }
MlpWs mlp_ws(void *ws, int B) {
    uint8_t *b = (uint8_t *)ws;
    MlpWs v;
    v.stats = (float *)b; b += 256;
    v.Z = (float *)b; b += (size_t)B * kOut * 4;
    v.X = (uint16_t *)b; b += (size_t)B * kIn * 2;
    v.H1 = (uint16_t *)b; b += (size_t)B * kH * 2;
    v.H2 = (uint16_t *)b; b += (size_t)B * kH * 2;
    v.dZ = (uint16_t *)b; b += (size_t)B * kOut * 2;
    v.dH2 = (uint16_t *)b; b += (size_t)B * kH * 2;
    v.dH1 = (uint16_t *)b;
    return v;
}
zt_gemm_args gemm(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn, int M, int N, int K) {
    zt_gemm_args g;
    memset(&g, 0, sizeof g);
    g.A = A; g.sam = sam; g.sak = sak; g.B = B; g.sbk = sbk; g.sbn = sbn; g.M = M; g.N = N; g.K = K;
    return g;
}
int run_gemm(const zt_gemm_args &g, int ksplit, hipStream_t s) {
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, ksplit);
    return launch("zt_gemm_bf16", s, [&] { hipLaunchKernelGGL(zt_gemm_bf16, grid, dim3(256), 0, s, g); });
}
// forward pass on B blocks (16 RGBA8 pixels each) -> Z logits (and X, H1, H2 kept for backward)
int mlp_forward(const MlpState &st, const void *blocks, int B, const MlpWs &w, hipStream_t s) {
    int rc;
    if ((rc = launch("zt_mlp_features", s, [&] {
            hipLaunchKernelGGL(zt_mlp_features, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, (const uint8_t *)blocks, B, (const int *)st.cfg, w.X); })))
        return rc;
    zt_gemm_args g = gemm(w.X, kIn, 1, st.w + bW1, 1, kIn, B, kH, kIn);
    g.bias = st.p + oB1; g.relu = 1; g.Cb = w.H1; g.ldc = kH;
    if ((rc = run_gemm(g, 1, s))) return rc;
    g = gemm(w.H1, kH, 1, st.w + bW2, 1, kH, B, kH, kH);
    g.bias = st.p + oB2; g.relu = 1; g.Cb = w.H2; g.ldc = kH;
    if ((rc = run_gemm(g, 1, s))) return rc;
    g = gemm(w.H2, kH, 1, st.w + bW3, 1, kH, B, kOut, kH);
    g.bias = st.p + oB3; g.Cf = w.Z; g.ldc = kOut;
    return run_gemm(g, 1, s);
}
// loss and gradients into st.g (zeroed first); stats = (loss, top-1 hits).  Targets: the one-hot labels, or with
// deltas and a temperature the softmax of -delta / temperature over the candidate slots; with deltas and risk_k,
// also the risk slot's regression.
int mlp_backward(const MlpState &st, const zaptex_mlp_targets &t, int B, const MlpWs &w, hipStream_t s) {
    int rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(256), dim3(256), 0, s, (uint32_t *)st.g, (uint64_t)kP, 0u); })))
        return rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)w.stats, (uint64_t)64, 0u); })))
        return rc;
    if ((rc = launch("zt_mlp_xent", s, [&] { hipLaunchKernelGGL(zt_mlp_xent, dim3((B + 3) / 4), dim3(256), 0, s, w.Z, t.labels, t.deltas,
                               t.temperature > 0.0f ? 1.0f / t.temperature : 0.0f, t.deltas ? t.risk_k : 0, t.risk_weight, B, w.dZ, w.stats); })))
        return rc;
    const int ks = std::max(1, std::min(64, B / 256));
    auto wgrad = [&](const uint16_t *act, int in, const uint16_t *d, int out, int64_t off) {  // dW = act^T d
        zt_gemm_args g = gemm(act, 1, in, d, out, 1, in, out, B);
        g.Cf = st.g + off; g.ldc = out; g.accumulate = 1;
        return run_gemm(g, ks, s);
    };
    auto bgrad = [&](const uint16_t *d, int out, int64_t off) {
        return launch("zt_mlp_colsum", s, [&] {
            hipLaunchKernelGGL(zt_mlp_colsum, dim3((out + 127) / 128, std::max(1, std::min(256, B / 64))), dim3(128), 0, s, d, B, out, st.g + off); });
    };
    auto dgrad = [&](const uint16_t *d, int out, const uint16_t *wt, int in, const uint16_t *act, uint16_t *dst) {  // (d W^T) * (act > 0)
        zt_gemm_args g = gemm(d, out, 1, wt, in, 1, B, in, out);
        g.Cb = dst; g.ldc = in; g.mask = act; g.ldm = in;
        return run_gemm(g, 1, s);
    };
    if ((rc = wgrad(w.H2, kH, w.dZ, kOut, oW3))) return rc;
    if ((rc = bgrad(w.dZ, kOut, oB3))) return rc;
    if ((rc = dgrad(w.dZ, kOut, st.w + bW3, kH, w.H2, w.dH2))) return rc;
    if ((rc = wgrad(w.H1, kH, w.dH2, kH, oW2))) return rc;
    if ((rc = bgrad(w.dH2, kH, oB2))) return rc;
    if ((rc = dgrad(w.dH2, kH, st.w + bW2, kH, w.H1, w.dH1))) return rc;
    if ((rc = wgrad(w.X, kIn, w.dH1, kH, oW1))) return rc;
    return bgrad(w.dH1, kH, oB1);
}
}  // namespace

size_t zaptex_mlp_state_size(void) { return mlp_state_bytes(); }
size_t zaptex_mlp_workspace_size(int batch) { return batch > 0 ? mlp_ws_bytes(batch) : 0; }

int zaptex_mlp_init(void *state, uint32_t seed, void *stream) {
    if (!state) return fail(ZAPTEX_E_ARG, "null state");
    DeviceState *dst;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(dst)) return rc;
    }
    MlpState st = mlp_view(state);
    std::vector<int> bfi(kP, -1);
    std::vector<float> scale(kP, 0.0f);
    struct L { int64_t off, boff; int in, out; } layers[3] = { { oW1, bW1, kIn, kH }, { oW2, bW2, kH, kH }, { oW3, bW3, kH, kOut } };
    for (const L &l : layers)
        for (int i = 0; i < l.in; i++)
            for (int o = 0; o < l.out; o++) {
                bfi[l.off + (int64_t)i * l.out + o] = (int)(l.boff + (int64_t)o * l.in + i);
                scale[l.off + (int64_t)i * l.out + o] = std::sqrt(6.0f / (float)l.in);
            }
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(st.bfi, bfi.data(), kP * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(st.scale, scale.data(), kP * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    int rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(512), dim3(256), 0, s, (uint32_t *)st.g, (uint64_t)kP * 3, 0u); })))
        return rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)st.cfg, (uint64_t)64, 0u); })))
        return rc;
    return launch("zt_mlp_init", s, [&] {
        hipLaunchKernelGGL(zt_mlp_init, dim3((unsigned)((kP + 255) / 256)), dim3(256), 0, s, st.p, (int)kP, seed, st.w, (const int *)st.bfi, (const float *)st.scale); });
}

int zaptex_mlp_set_params(void *state, const float *host_params, void *stream) {  // load fp32 master weights; refreshes bf16 copies
    if (!state || !host_params) return fail(ZAPTEX_E_ARG, "null argument");
    MlpState st = mlp_view(state);
    {  // index tables (as zaptex_mlp_init writes them)
        std::vector<int> bfi(kP, -1);
        struct L { int64_t off, boff; int in, out; } layers[3] = { { oW1, bW1, kIn, kH }, { oW2, bW2, kH, kH }, { oW3, bW3, kH, kOut } };
        for (const L &l : layers)
            for (int i = 0; i < l.in; i++)
                for (int o = 0; o < l.out; o++) bfi[l.off + (int64_t)i * l.out + o] = (int)(l.boff + (int64_t)o * l.in + i);
        HIP_TRY(hipMemcpy(st.bfi, bfi.data(), kP * 4, hipMemcpyHostToDevice));
    }
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(st.p, host_params, kP * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return launch("zt_mlp_refresh", s, [&] {
        hipLaunchKernelGGL(zt_mlp_refresh, dim3((unsigned)((kP + 255) / 256)), dim3(256), 0, s, (const float *)st.p, (int)kP, st.w, (const int *)st.bfi);
    });
}

int zaptex_mlp_set_features(void *state, int mode, void *stream) {
    if (!state || mode < 0 || mode > 4) return fail(ZAPTEX_E_ARG, "feature mode must be 0..4");
    MlpState st = mlp_view(state);
    hipStream_t s = (hipStream_t)stream;
    return launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(1), dim3(64), 0, s, (uint32_t *)st.cfg, (uint64_t)1, (uint32_t)mode); });
}

int zaptex_mlp_forward(const void *state, const void *blocks, int batch, float *logits, void *ws, size_t ws_bytes, void *stream) {
    if (!state || !blocks || batch < 1) return fail(ZAPTEX_E_ARG, "bad arguments");
    if (!ws || ws_bytes < mlp_ws_bytes(batch)) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu", ws_bytes, mlp_ws_bytes(batch));
    MlpState st = mlp_view((void *)state);
    MlpWs w = mlp_ws(ws, batch);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = mlp_forward(st, blocks, batch, w, s)) return rc;
    if (logits) HIP_TRY(hipMemcpyAsync(logits, w.Z, (size_t)batch * kOut * 4, hipMemcpyDeviceToDevice, s));
    return ZAPTEX_OK;
}

namespace {
// the network's top-K proposals and risk (and, for tests, the kept logits) for n dense blocks, one persistent launch
// of 8 waves per CU; a wave takes 32 blocks per pass
int enqueue_propose(const MlpState &st, const uint4 *blocks, uint32_t n, int topk, uint32_t *cand, float *risk, float *logits, hipStream_t s) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const uint64_t groups = ((uint64_t)n + 255) / 256;   // 8 waves x 32 blocks per workgroup pass
    const dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(groups, (uint64_t)cus)));
    zt_propose_args P{ blocks, n, topk, st.w, st.p + oB1, st.p + oB2, st.p + oB3, st.cfg, cand, risk, logits };
    if (topk <= 4) return launch("zt_mlp_propose4", s, [&] { hipLaunchKernelGGL(zt_mlp_propose4, grid, dim3(512), 0, s, P); });
    return launch("zt_mlp_propose16", s, [&] { hipLaunchKernelGGL(zt_mlp_propose16, grid, dim3(512), 0, s, P); });
}
}  // namespace

int zaptex_mlp_propose(const void *state, const void *blocks, int n, int topk, uint32_t *cand, float *risk, void *stream) {
    if (!state || !blocks || !cand || !risk || n < 1) return fail(ZAPTEX_E_ARG, "bad arguments");
    if (topk < 1 || topk > 16) return fail(ZAPTEX_E_ARG, "topk must be 1..16");
    if (((uintptr_t)blocks) % 16) return fail(ZAPTEX_E_ARG, "blocks must be 16-byte aligned");
    return enqueue_propose(mlp_view((void *)state), (const uint4 *)blocks, (uint32_t)n, topk, cand, risk, nullptr, (hipStream_t)stream);
}

int zaptex_mlp_propose_logits(const void *state, const void *blocks, int n, float *logits, void *stream) {
    if (!state || !blocks || !logits || n < 1) return fail(ZAPTEX_E_ARG, "bad arguments");
    if (((uintptr_t)blocks) % 16 || ((uintptr_t)logits) % 16) return fail(ZAPTEX_E_ARG, "blocks and logits must be 16-byte aligned");
    return enqueue_propose(mlp_view((void *)state), (const uint4 *)blocks, (uint32_t)n, 1, nullptr, nullptr, logits, (hipStream_t)stream);
}

int zaptex_mlp_gradients(void *state, const void *blocks, const zaptex_mlp_targets *targets, int batch, void *ws, size_t ws_bytes, float *stats,
                         void *stream) {
    if (!state || !blocks || !targets || batch < 1) return fail(ZAPTEX_E_ARG, "bad arguments");
    const zaptex_mlp_targets &t = *targets;
    if (!t.labels && !(t.deltas && t.temperature > 0.0f)) return fail(ZAPTEX_E_ARG, "need labels, or deltas with a positive temperature");
    if (t.risk_k < 0 || t.risk_k > 16 || (t.risk_k > 0 && !t.deltas)) return fail(ZAPTEX_E_ARG, "risk_k must be 0..16 and needs deltas");
    if (!ws || ws_bytes < mlp_ws_bytes(batch)) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu", ws_bytes, mlp_ws_bytes(batch));
    MlpState st = mlp_view(state);
    MlpWs w = mlp_ws(ws, batch);
    hipStream_t s = (hipStream_t)stream;
    if (int rc = mlp_forward(st, blocks, batch, w, s)) return rc;
    if (int rc = mlp_backward(st, t, batch, w, s)) return rc;
    if (stats) HIP_TRY(hipMemcpyAsync(stats, w.stats, 8, hipMemcpyDeviceToDevice, s));
    return ZAPTEX_OK;
}

int zaptex_mlp_train_step(void *state, const void *blocks, const zaptex_mlp_targets *targets, int batch, float lr, int step, void *ws,
                          size_t ws_bytes, float *stats, void *stream) {
    if (int rc = zaptex_mlp_gradients(state, blocks, targets, batch, ws, ws_bytes, stats, stream)) return rc;
    MlpState st = mlp_view(state);
    hipStream_t s = (hipStream_t)stream;
    const float b1 = 0.9f, b2 = 0.999f, c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    return launch("zt_mlp_adam", s, [&] {
        hipLaunchKernelGGL(zt_mlp_adam, dim3((unsigned)((kP + 255) / 256)), dim3(256), 0, s, st.p, (const float *)st.g, st.m1, st.m2, (int)kP, lr,
                           b1, b2, c1, c2, 1e-8f, st.w, (const int *)st.bfi);
    });
}

int zaptex_bc7_gather_dev(const zaptex_image *imgs, int n, void *blocks, void *stream) {
    if (!imgs || n < 1 || n > ZT_INLINE_JOBS || !blocks) return fail(ZAPTEX_E_ARG, "bad arguments");
    zt_jobs J;
    memset(&J, 0, sizeof J);
    uint32_t total = 0;
    for (int i = 0; i < n; i++) {
        if (int rc = check_image(imgs[i])) return rc;
        zt_job &j = J.inl[J.n++];
        j.src = (const uint8_t *)imgs[i].src; j.dst = (uint8_t *)imgs[i].dst; j.stride = imgs[i].stride;
        j.first = total; j.w = imgs[i].w; j.h = imgs[i].h; j.bw = (imgs[i].w + 3) / 4; j.fmt = imgs[i].fmt; j.pad = 0;
        total += nblocks(imgs[i]);
    }
    hipStream_t s = (hipStream_t)stream;
    return launch("zt_bc7_gather", s, [&] {
        hipLaunchKernelGGL(zt_bc7_gather, dim3(grid_for(total, 256)), dim3(256), 0, s, J, total, (uint4 *)blocks); });
}

// ---- learned-proposal BC7 encode
namespace {
// per block: its pixels (64 B), K proposals (item 8 B, encoding 16 B, candidate 4 B, item index 4 B), key, risk and
// a fallback-list entry and its mode-sorted finish item (8 B); with a wide fallback (K2 > 0) also K2 wide candidates
// (4 B) and items (8 B)
struct LearnedWs {
    uint4 *blocks, *enc;
    uint2 *items, *items2, *items3;
    uint32_t *cand, *slot, *fallback, *key, *cand2;
    float *risk;
    unsigned *counters, *counters2, *counters3;
};
constexpr uint64_t kBucketBytes = ZT_BUCKET_WORDS * 4;   // one bucketing counter array (a multiple of 256 B)
// 14 regions, each starting on a 256-byte boundary
size_t learned_ws_bytes(uint64_t total, int K, int K2) {
    return 256 + 3 * kBucketBytes + total * (64 + 32 * (uint64_t)K + 4 + 4 + 4 + 8) + total * 12 * (uint64_t)K2 + 15 * 256;
}
LearnedWs learned_ws(void *ws, uint64_t total, int K, int K2) {
    uint8_t *b = (uint8_t *)align_up((uintptr_t)ws, 256);
    auto take = [&b](uint64_t bytes) { uint8_t *p = b; b = (uint8_t *)align_up((uintptr_t)(b + bytes), 256); return (void *)p; };
    LearnedWs v;
    v.counters = (unsigned *)take(kBucketBytes);
    v.counters2 = (unsigned *)take(kBucketBytes);
    v.counters3 = (unsigned *)take(kBucketBytes);
    v.blocks = (uint4 *)take(total * 64);
    v.items3 = (uint2 *)take(total * 8);
    v.enc = (uint4 *)take(total * 16 * K);
    v.items = (uint2 *)take(total * 8 * K);
    v.items2 = (uint2 *)take(total * 8 * (uint64_t)K2);
    v.cand = (uint32_t *)take(total * 4 * K);
    v.slot = (uint32_t *)take(total * 4 * K);
    v.cand2 = (uint32_t *)take(total * 4 * (uint64_t)K2);
    v.key = (uint32_t *)take(total * 4);
    v.risk = (float *)take(total * 4);
    v.fallback = (uint32_t *)take(total * 4);
    return v;
}
typedef void (*PropFn)(const uint4 *, const uint2 *, const unsigned *, uint32_t *, uint4 *);
// one launch per mode over the bucketed items (counters: per-mode counts at [0..7], run starts at [24..31])
int enqueue_props(const uint4 *blocks, const uint2 *items, const unsigned *counters, uint32_t *key, uint4 *enc, uint64_t max_items,
                  hipStream_t s) {
    const PropFn props[8] = { zt_bc7_prop_m0, zt_bc7_prop_m1, zt_bc7_prop_m2, zt_bc7_prop_m3, zt_bc7_prop_m4, zt_bc7_prop_m5, zt_bc7_prop_m6, zt_bc7_prop_m7 };
    static const char *names[8] = { "zt_bc7_prop_m0", "zt_bc7_prop_m1", "zt_bc7_prop_m2", "zt_bc7_prop_m3",
                                    "zt_bc7_prop_m4", "zt_bc7_prop_m5", "zt_bc7_prop_m6", "zt_bc7_prop_m7" };
    const dim3 grid((unsigned)std::min<uint64_t>((uint64_t)grid_for(max_items, 256), 65536));
    for (int m = 0; m < 8; m++)
        if (int rc = launch(names[m], s, [&] { hipLaunchKernelGGL(props[m], grid, dim3(256), 0, s, blocks, items, counters, key, enc); })) return rc;
    return ZAPTEX_OK;
}
}  // namespace

size_t zaptex_bc7_learned_workspace_size(const zaptex_image *imgs, int n, int topk, int fallback_k) {
    uint64_t t = 0;
    for (int i = 0; i < n; i++) t += nblocks(imgs[i]);
    return learned_ws_bytes(t, topk, fallback_k > 0 ? fallback_k : 0);
}

int zaptex_bc7_learned_dev(const zaptex_image *imgs, int n, const void *state, int topk, float threshold, int fallback_k, void *ws,
                           size_t ws_bytes, void *stream) {
    if (!state) return fail(ZAPTEX_E_ARG, "null network state");
    if (topk < 1 || topk > 16) return fail(ZAPTEX_E_ARG, "topk must be 1..16");
    if (fallback_k < 0 || fallback_k > 285) return fail(ZAPTEX_E_ARG, "fallback_k must be 0 (exhaustive) or 1..285");
    zt_jobs J;
    uint32_t total;
    if (int rc = bc7_jobs(imgs, n, J, total, "the learned proposal")) return rc;
    const size_t need = learned_ws_bytes(total, topk, fallback_k);
    if (!ws || ws_bytes < need) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu", ws_bytes, need);
    DeviceState *dst;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(dst)) return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    LearnedWs w = learned_ws(ws, total, topk, fallback_k);
    MlpState st = mlp_view((void *)state);
    int rc;
    if ((rc = launch("zt_zero_u32", s, [&] {
            hipLaunchKernelGGL(zt_zero_u32, dim3((3 * ZT_BUCKET_WORDS + 255) / 256), dim3(256), 0, s, w.counters, 3 * ZT_BUCKET_WORDS); })))
        return rc;
    if ((rc = launch("zt_fill_u32", s, [&] { hipLaunchKernelGGL(zt_fill_u32, dim3(grid_for(total, 256)), dim3(256), 0, s, w.key, (uint64_t)total, 0xFFFFFFFFu); })))
        return rc;
    if ((rc = launch("zt_bc7_gather", s, [&] { hipLaunchKernelGGL(zt_bc7_gather, dim3(grid_for(total, 256)), dim3(256), 0, s, J, total, w.blocks); })))
        return rc;
    if ((rc = enqueue_propose(st, w.blocks, total, topk, w.cand, w.risk, nullptr, s))) return rc;
    for (uint32_t pass = 0; pass < 2; pass++) {
        if ((rc = launch("zt_bc7_bucket", s, [&] {
                hipLaunchKernelGGL(zt_bc7_bucket, dim3(grid_for(total, 256)), dim3(256), 0, s, (const uint32_t *)w.cand, (const float *)w.risk, total, topk,
                                   threshold, w.counters, w.items, pass, w.fallback, w.slot); })))
            return rc;
        if (pass == 0 && (rc = launch("zt_bc7_bucket_scan", s, [&] { hipLaunchKernelGGL(zt_bc7_bucket_scan, dim3(1), dim3(64), 0, s, w.counters); })))
            return rc;
    }
    if ((rc = enqueue_props(w.blocks, w.items, w.counters, w.key, w.enc, (uint64_t)total * topk, s))) return rc;
    if (fallback_k == 0) {
        if ((rc = enqueue_exhaustive(w.blocks, total, w.fallback, w.counters + 8, w.key, nullptr, s))) return rc;
    } else {
        // the fallback list's wide proposals: the network's top fallback_k, bucketed by mode and encoded exactly
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
        const dim3 pgrid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)total + 63) / 64, (uint64_t)cus)));
        zt_wide_args P{ w.blocks, w.fallback, w.counters + 8, fallback_k, st.w, st.p + oB1, st.p + oB2, st.p + oB3, st.cfg, w.cand2 };
        if ((rc = launch("zt_mlp_propose_wide", s, [&] { hipLaunchKernelGGL(zt_mlp_propose_wide, pgrid, dim3(256), 0, s, P); }))) return rc;
        const uint64_t max_items = (uint64_t)total * fallback_k;
        const dim3 bgrid((unsigned)std::min<uint64_t>((uint64_t)grid_for(max_items, 256), 4096));
        for (uint32_t pass = 0; pass < 2; pass++) {
            if ((rc = launch("zt_bc7_bucket_wide", s, [&] {
                    hipLaunchKernelGGL(zt_bc7_bucket_wide, bgrid, dim3(256), 0, s, (const uint32_t *)w.cand2, (const uint32_t *)nullptr,
                                       (const uint32_t *)w.fallback, (const unsigned *)(w.counters + 8), fallback_k, w.counters2, w.items2, pass); })))
                return rc;
            if (pass == 0 && (rc = launch("zt_bc7_bucket_scan", s, [&] { hipLaunchKernelGGL(zt_bc7_bucket_scan, dim3(1), dim3(64), 0, s, w.counters2); })))
                return rc;
        }
        if ((rc = enqueue_props(w.blocks, w.items2, w.counters2, w.key, nullptr, max_items, s))) return rc;
    }
    // fallback blocks re-encode their winner, sorted by its mode; the others copy their winning proposal's stored encoding
    const dim3 fgrid((unsigned)grid_for(total / 16 + 1, 256));
    for (uint32_t pass = 0; pass < 2; pass++) {
        if ((rc = launch("zt_bc7_bucket_wide", s, [&] {
                hipLaunchKernelGGL(zt_bc7_bucket_wide, fgrid, dim3(256), 0, s, (const uint32_t *)nullptr, (const uint32_t *)w.key,
                                   (const uint32_t *)w.fallback, (const unsigned *)(w.counters + 8), 1, w.counters3, w.items3, pass); })))
            return rc;
        if (pass == 0 && (rc = launch("zt_bc7_bucket_scan", s, [&] { hipLaunchKernelGGL(zt_bc7_bucket_scan, dim3(1), dim3(64), 0, s, w.counters3); })))
            return rc;
    }
    if ((rc = launch("zt_bc7_finish", s, [&] {
            hipLaunchKernelGGL(zt_bc7_finish, fgrid, dim3(256), 0, s, J, (const uint4 *)w.blocks, total, (const uint32_t *)w.key, (uint32_t *)nullptr,
                               (const uint2 *)w.items3, (const unsigned *)(w.counters + 8)); })))
        return rc;
    return launch("zt_bc7_finish_learned", s, [&] {
        hipLaunchKernelGGL(zt_bc7_finish_learned, dim3(grid_for(total, 256)), dim3(256), 0, s, J, total, topk, (const uint32_t *)w.cand,
                           (const uint32_t *)w.slot, (const uint4 *)w.enc, (const uint32_t *)w.key); });
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
    dim3 grid((nw + 63) / 64, std::min((nh + 3) / 4, 65535)), block(64, 4);
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
    for (int i = 0; i < nlevels; i++)
        if (!levels[i] || ((uintptr_t)levels[i]) % 4) return fail(ZAPTEX_E_ARG, "bad level pointer %d", i);
    return enqueue_mip_chain(base, w, h, stride, srgb, levels, nlevels, (hipStream_t)stream);
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

// ===================================================================================== batch
}  // extern "C"
namespace {
// byte offsets of levels 1..levels-1 in a texture's mip arena (tight rows, each level 256-B aligned) and of
// levels 0..levels-1 in its block arena
void batch_layout(int w, int h, int fmt, int levels, std::vector<size_t> *mip_off, std::vector<size_t> *out_off, size_t *mip_total,
                  size_t *out_total) {
    size_t mo = 0, oo = 0;
    int cw = w, ch = h;
    for (int l = 0; l < levels; l++) {
        if (out_off) out_off->push_back(oo);
        oo += align_up(out_bytes(cw, ch, fmt), 256);
        if (l + 1 < levels) {
            cw = cw > 1 ? cw / 2 : 1; ch = ch > 1 ? ch / 2 : 1;
            if (mip_off) mip_off->push_back(mo);
            mo += align_up((size_t)cw * ch * 4, 256);
        }
    }
    if (mip_total) *mip_total = mo;
    if (out_total) *out_total = oo;
}
int max_levels(int w, int h) { int n = 1; while (w > 1 || h > 1) { w = w > 1 ? w / 2 : 1; h = h > 1 ? h / 2 : 1; n++; } return n; }
int check_texture(const zaptex_texture &t) {
    if (t.fmt < ZAPTEX_BC1 || t.fmt > ZAPTEX_BC7) return fail(ZAPTEX_E_ARG, "batch formats are BC1, BC3, BC4, BC5 and BC7 (got %d)", t.fmt);
    if (!size_ok(t.w, t.h)) return fail(ZAPTEX_E_ARG, "bad size %dx%d", t.w, t.h);
    if (t.levels < 1 || t.levels > max_levels(t.w, t.h)) return fail(ZAPTEX_E_ARG, "levels %d outside 1..%d", t.levels, max_levels(t.w, t.h));
    if (!t.base || !t.out || (t.levels > 1 && !t.mips)) return fail(ZAPTEX_E_ARG, "null texture pointer");
    if (t.stride < (size_t)t.w * 4 || (((uintptr_t)t.base) | t.stride) % 4) return fail(ZAPTEX_E_ARG, "base rows must be 4-byte aligned, stride >= 4 w");
    if (((uintptr_t)t.out | (uintptr_t)t.mips) % 256) return fail(ZAPTEX_E_ARG, "mips and out arenas must be 256-byte aligned");
    return ZAPTEX_OK;
}
// The mip chains of a batch as launches: each texture's chain is cut into steps exactly as a single chain is
// (a fused tile pass of up to 7 levels while the source exceeds 128 on a side, single levels while a side is 1
// and the other exceeds 128, then one workgroup for a source of at most 128 x 128); step s of every texture goes
// into the launches of round s (one per step kind), so each texture's steps run in order.
struct MipLaunch { int kind; size_t first; int count; uint64_t px; int srgb; int round; };  // kind 0 tiles, 1 single level, 2 tail
struct MipPlan {
    std::vector<zt_mipstep> steps;   // tile and tail launches' steps, launch by launch
    std::vector<uint2> items;        // tile items (step within its launch, tile)
    std::vector<zt_mipjob> jobs;     // single-level jobs, launch by launch
    std::vector<MipLaunch> launches; // first = offset into items (kind 0), jobs (kind 1) or steps (kind 2)
    std::vector<size_t> tile_steps;  // kind-0 launches: offset of their steps
    std::vector<std::vector<int>> level_round;  // [texture][level]: the round whose step writes it (-1: level 0)
    int rounds = 0;
    int tile = 64;                   // level-1 tile side of the kind-0 launches
};
MipPlan plan_mips(const zaptex_texture *t, int n, int tile) {
    struct TexStep { int kind, l, k; };
    std::vector<std::vector<TexStep>> ts(n);
    std::vector<std::vector<int>> LW(n), LH(n);
    std::vector<std::vector<size_t>> MO(n);
    size_t rounds = 0;
    for (int i = 0; i < n; i++) {
        const int L = t[i].levels;
        int cw = t[i].w, ch = t[i].h;
        for (int l = 0; l < L; l++) { LW[i].push_back(cw); LH[i].push_back(ch); cw = cw > 1 ? cw / 2 : 1; ch = ch > 1 ? ch / 2 : 1; }
        batch_layout(t[i].w, t[i].h, t[i].fmt, L, &MO[i], nullptr, nullptr, nullptr);
        for (int l = 0; l < L - 1;) {
            const int w = LW[i][l], h = LH[i][l];
            TexStep st;
            if ((w > 128 || h > 128) && w >= 2 && h >= 2) {
                int k = 0;
                while (k < tile_levels(tile) && l + k < L - 1 && LW[i][l + k] >= 2 && LH[i][l + k] >= 2) k++;
                st = { 0, l, k };
            } else if (w <= 128 && h <= 128) {
                st = { 2, l, L - 1 - l };
            } else {
                st = { 1, l, 1 };
            }
            ts[i].push_back(st);
            l += st.k;
        }
        rounds = std::max(rounds, ts[i].size());
    }
    auto level_ptr = [&](int i, int l) -> uint8_t * {
        return l == 0 ? (uint8_t *)t[i].base : (uint8_t *)t[i].mips + MO[i][l - 1];
    };
    auto make_step = [&](int i, const TexStep &st) {
        zt_mipstep m;
        memset(&m, 0, sizeof m);
        m.src = level_ptr(i, st.l);
        m.stride = st.l == 0 ? t[i].stride : (uint64_t)LW[i][st.l] * 4;
        m.w = LW[i][st.l]; m.h = LH[i][st.l]; m.srgb = t[i].srgb ? 1 : 0; m.nlev = st.k;
        for (int j = 0; j < st.k; j++) { m.L.p[j] = level_ptr(i, st.l + 1 + j); m.L.w[j] = LW[i][st.l + 1 + j]; m.L.h[j] = LH[i][st.l + 1 + j]; }
        return m;
    };
    MipPlan P;
    P.tile = tile;
    P.rounds = (int)rounds;
    P.level_round.resize(n);
    for (int i = 0; i < n; i++) {
        P.level_round[i].assign(t[i].levels, -1);
        for (size_t r = 0; r < ts[i].size(); r++)
            for (int j = 1; j <= ts[i][r].k; j++) P.level_round[i][ts[i][r].l + j] = (int)r;
    }
    for (size_t r = 0; r < rounds; r++)
        for (int kind = 0; kind < 3; kind++) {
            MipLaunch ml{ kind, 0, 0, 0, 0, (int)r };
            ml.first = kind == 0 ? P.items.size() : kind == 1 ? P.jobs.size() : P.steps.size();
            const size_t step0 = P.steps.size();
            for (int i = 0; i < n; i++) {
                if (ts[i].size() <= r || ts[i][r].kind != kind) continue;
                const TexStep &st = ts[i][r];
                if (kind == 1) {
                    zt_mipjob j{};
                    j.src = level_ptr(i, st.l); j.dst = level_ptr(i, st.l + 1);
                    j.stride = st.l == 0 ? t[i].stride : (uint64_t)LW[i][st.l] * 4;
                    j.first = ml.px; j.w = LW[i][st.l]; j.h = LH[i][st.l]; j.srgb = t[i].srgb ? 1 : 0;
                    ml.px += (uint64_t)LW[i][st.l + 1] * LH[i][st.l + 1];
                    ml.srgb |= j.srgb;
                    P.jobs.push_back(j);
                } else {
                    const uint32_t si = (uint32_t)(P.steps.size() - step0);
                    P.steps.push_back(make_step(i, st));
                    if (kind == 0) {
                        const uint32_t tiles = (uint32_t)(((LW[i][st.l + 1] + tile - 1) / tile) * ((LH[i][st.l + 1] + tile - 1) / tile));
                        for (uint32_t k = 0; k < tiles; k++) P.items.push_back(make_uint2(si, k));
                    }
                }
                ml.count++;
            }
            if (!ml.count) continue;
            if (kind == 0) { P.tile_steps.push_back(step0); ml.count = (int)(P.items.size() - ml.first); }
            P.launches.push_back(ml);
        }
    return P;
}

// ---- the persistent batch kernel's queue (zt_batch.hip)
// Encode jobs of a family in queue order: level 0 of every texture first, then the levels written by mip round 0,
// round 1, ... (texture and level order within a round), so the level-0 chunks need no mips and every later chunk
// waits for one round's counter.
struct FamJob { int round, tex, lvl; uint32_t blocks; };
std::vector<FamJob> family_jobs(const zaptex_texture *t, int n, const MipPlan &P, bool bc7) {
    std::vector<FamJob> v;
    for (int i = 0; i < n; i++) {
        if ((t[i].fmt == ZAPTEX_BC7) != bc7) continue;
        int cw = t[i].w, ch = t[i].h;
        for (int l = 0; l < t[i].levels; l++) {
            v.push_back({ P.level_round[i][l], i, l, (uint32_t)((uint64_t)((cw + 3) / 4) * (uint64_t)((ch + 3) / 4)) });
            cw = cw > 1 ? cw / 2 : 1; ch = ch > 1 ? ch / 2 : 1;
        }
    }
    std::stable_sort(v.begin(), v.end(), [](const FamJob &a, const FamJob &b) { return a.round < b.round; });
    return v;
}

int g_batch_persistent = 0;  // 1: batches run as one persistent kernel; 0 (default): a graph of the per-stage kernels
int g_batch_mix = 2;         // level-0 encode tasks per round-0 mip task at the head of the queue
int g_batch_per = 1;         // 256-block chunks per encode task (BC1-family start, BC7 stage 6)
int g_batch_groups = 0;      // persistent workgroups per CU (0: as many as fit)
// the batch's mip tiles: the persistent kernel runs the 64-pixel tile body only
int batch_mip_tile() { return g_batch_persistent ? 64 : g_mip_tile; }

// counters after the mip rounds' (round r: counter r)
enum { PC_BC14, PC_CLIMB, PC_FINISH, PC_S6, PC_S5, PC_RANK, PC_FIT, PC_PICK, PC_N };
struct PersistentQueue {
    std::vector<zt_seg> segs;
    std::vector<uint32_t> queue;   // segment << ZT_TASK_SEG_SHIFT | task within the segment
    uint32_t nctl = 0;
};
// device tables the mip segments point into (null while only counting tasks)
struct PqTables { const zt_mipstep *steps = nullptr; const uint2 *items = nullptr; const zt_mipjob *mipjobs = nullptr; };

// x and y interleaved evenly (each prefix holds the two in proportion to their lengths)
std::vector<uint32_t> even_merge(const std::vector<uint32_t> &x, const std::vector<uint32_t> &y) {
    std::vector<uint32_t> o;
    o.reserve(x.size() + y.size());
    size_t i = 0, j = 0;
    while (i < x.size() || j < y.size()) {
        if (j == y.size() || (i < x.size() && i * y.size() <= j * x.size())) o.push_back(x[i++]);
        else o.push_back(y[j++]);
    }
    return o;
}
// one task of a, then `rb` of b, until a runs out; then the rest of b
std::vector<uint32_t> lead_merge(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b, int rb) {
    std::vector<uint32_t> o;
    o.reserve(a.size() + b.size());
    size_t j = 0;
    for (size_t i = 0; i < a.size(); i++) {
        o.push_back(a[i]);
        for (int k = 0; k < rb && j < b.size(); k++) o.push_back(b[j++]);
    }
    while (j < b.size()) o.push_back(b[j++]);
    return o;
}

// The queue: every task waits only for tasks placed before it (mip round r after round r - 1; a chunk of levels
// from round r after every mip task; each encode stage after the whole stage it reads), so it drains.
int plan_queue(const MipPlan &P, const std::vector<FamJob> &f14, const std::vector<FamJob> &f7, uint32_t tot14, uint32_t tot7,
               const PqTables &tb, PersistentQueue &Q) {
    Q = PersistentQueue();
    const uint32_t R = (uint32_t)P.rounds, C = R;   // stage counter s: C + s
    std::vector<std::vector<uint32_t>> tasks;       // per segment, its queue entries
    auto add = [&](uint32_t kind, uint32_t vbase, uint32_t count, uint32_t per, uint32_t vgrid, uint32_t wait, uint32_t need, uint32_t done) {
        zt_seg sg;
        memset(&sg, 0, sizeof sg);
        sg.kind = kind; sg.vbase = vbase; sg.count = count; sg.per = per; sg.vgrid = vgrid; sg.wait = wait; sg.need = need; sg.done = done;
        Q.segs.push_back(sg);
        const uint32_t id = (uint32_t)(Q.segs.size() - 1), nt = (count + per - 1) / per;
        std::vector<uint32_t> e(nt);
        for (uint32_t j = 0; j < nt; j++) e[j] = id << ZT_TASK_SEG_SHIFT | j;
        tasks.push_back(std::move(e));
        return id;
    };
    auto cat = [&](const std::vector<uint32_t> &ids) {
        std::vector<uint32_t> o;
        for (uint32_t id : ids) o.insert(o.end(), tasks[id].begin(), tasks[id].end());
        return o;
    };
    // mip rounds (P.launches come in round order)
    std::vector<uint32_t> round_tasks(R, 0);
    std::vector<std::vector<uint32_t>> round_segs(R);
    size_t tl = 0;
    for (const MipLaunch &ml : P.launches) {
        const uint32_t r = (uint32_t)ml.round, wait = r ? r - 1 : ZT_NO_WAIT, need = r ? round_tasks[r - 1] : 0;
        uint32_t id;
        if (ml.kind == 0) {
            id = add(ZT_T_MIP_TILE, 0, (uint32_t)ml.count, 1, (uint32_t)ml.count, wait, need, r);
            Q.segs[id].a = tb.steps ? (const void *)(tb.steps + P.tile_steps[tl]) : nullptr;
            Q.segs[id].b = tb.items ? (const void *)(tb.items + ml.first) : nullptr;
            tl++;
        } else if (ml.kind == 1) {
            const uint32_t g = (uint32_t)groups_for(CFG_MIPB, ml.px, 256);
            id = add(ZT_T_MIP_LEVEL, 0, g, 1, g, wait, need, r);
            Q.segs[id].a = tb.mipjobs ? (const void *)(tb.mipjobs + ml.first) : nullptr;
            Q.segs[id].n = (uint32_t)ml.count; Q.segs[id].total = ml.px; Q.segs[id].srgb = (uint32_t)ml.srgb;
        } else {
            id = add(ZT_T_MIP_TAIL, 0, (uint32_t)ml.count, 1, (uint32_t)ml.count, wait, need, r);
            Q.segs[id].a = tb.steps ? (const void *)(tb.steps + ml.first) : nullptr;
        }
        round_segs[r].push_back(id);
        round_tasks[r] += (uint32_t)tasks[id].size();
    }
    // encode chunks of 256 blocks by the round of their levels; a chunk waits for the round of its last block's job
    // (the jobs are sorted by round, so that is the latest round it reads)
    auto chunks = [&](const std::vector<FamJob> &f, uint32_t tot, uint32_t kind, uint32_t done, std::vector<std::vector<uint32_t>> &by_round) {
        by_round.assign(R + 1, {});
        const uint32_t nch = (tot + 255) / 256;
        std::vector<int> cr(nch, -1);
        size_t j = 0;
        uint64_t end = f[0].blocks;   // job j's blocks end here (the jobs' blocks sum to tot)
        for (uint32_t c = 0; c < nch; c++) {
            const uint64_t last = std::min<uint64_t>(256ull * c + 255, (uint64_t)tot - 1);
            while (last >= end) end += f[++j].blocks;
            cr[c] = f[j].round;
        }
        uint32_t ntasks = 0;
        for (uint32_t c0 = 0; c0 < nch;) {
            uint32_t c1 = c0;
            while (c1 < nch && cr[c1] == cr[c0]) c1++;
            const int r = cr[c0];
            const uint32_t id = add(kind, c0, c1 - c0, (uint32_t)g_batch_per, nch, r < 0 ? ZT_NO_WAIT : (uint32_t)r, r < 0 ? 0 : round_tasks[r], done);
            by_round[r + 1].push_back(id);
            ntasks += (uint32_t)tasks[id].size();
            c0 = c1;
        }
        return ntasks;
    };
    std::vector<std::vector<uint32_t>> c14, c7;
    const uint32_t n14 = tot14 ? chunks(f14, tot14, ZT_T_BC14, C + PC_BC14, c14) : 0;
    const uint32_t n7 = tot7 ? chunks(f7, tot7, ZT_T_S6, C + PC_S6, c7) : 0;
    if (!tot14) c14.assign(R + 1, {});
    if (!tot7) c7.assign(R + 1, {});
    // the later stages, each after the whole stage it reads
    std::vector<uint32_t> climb, finish, s5, rank, fit, pick;
    if (tot14) {
        const uint32_t range = (uint32_t)std::min<uint64_t>(ZT_CLIMB_RANGE, std::max<uint64_t>(64, ((tot14 + 4095) / 4096 + 63) / 64 * 64));
        const uint32_t gc = (uint32_t)groups_for(CFG_CLIMB, (tot14 + range - 1) / range, 4), gf = (uint32_t)groups_for(CFG_BC14, tot14, 256);
        const uint32_t ic = add(ZT_T_CLIMB, 0, gc, 1, gc, C + PC_BC14, n14, C + PC_CLIMB);
        const uint32_t ifn = add(ZT_T_FINISH, 0, gf, 1, gf, C + PC_CLIMB, gc, C + PC_FINISH);
        climb = tasks[ic]; finish = tasks[ifn];
    }
    if (tot7) {
        const uint32_t g5 = (uint32_t)groups_for(CFG_S5, (uint64_t)tot7 * 4, 256), g1 = (uint32_t)groups_for(CFG_S1, (uint64_t)tot7 / 16 + 1, 4);
        s5 = tasks[add(ZT_T_S5, 0, g5, 1, g5, C + PC_S6, n7, C + PC_S5)];
        rank = tasks[add(ZT_T_RANK, 0, g1, 1, g1, C + PC_S5, g5, C + PC_RANK)];
        fit = tasks[add(ZT_T_FIT, 0, g1, 1, g1, C + PC_RANK, g1, C + PC_FIT)];
        pick = tasks[add(ZT_T_PICK, 0, g1, 1, g1, C + PC_FIT, g1, C + PC_PICK)];
    }
    if (Q.segs.size() > ZT_MAX_SEGS) return fail(ZAPTEX_E_ARG, "batch plan needs %zu queue segments (at most %d)", Q.segs.size(), ZT_MAX_SEGS);
    for (const auto &e : tasks)
        if (e.size() >= (1u << ZT_TASK_SEG_SHIFT)) return fail(ZAPTEX_E_ARG, "batch plan has a segment of %zu tasks", e.size());
    // the order: round-0 mip tasks led through the level-0 encode chunks, the later rounds, the chunks of their
    // levels, then the stages
    std::vector<uint32_t> &q = Q.queue;
    const std::vector<uint32_t> lvl0 = even_merge(cat(c14[0]), cat(c7[0]));
    q = R ? lead_merge(cat(round_segs[0]), lvl0, g_batch_mix) : lvl0;
    for (uint32_t r = 1; r < R; r++) { const auto v = cat(round_segs[r]); q.insert(q.end(), v.begin(), v.end()); }
    for (uint32_t r = 0; r < R; r++) { const auto v = even_merge(cat(c14[r + 1]), cat(c7[r + 1])); q.insert(q.end(), v.begin(), v.end()); }
    for (const auto &v : { even_merge(climb, s5), even_merge(finish, rank), fit, pick }) q.insert(q.end(), v.begin(), v.end());
    Q.nctl = ZT_CTL_COUNTERS + R + PC_N;
    // what keeps the kernel from waiting forever, checked before anything runs: every task appears once, and when a
    // task that waits comes up, every task counting towards its counter is already before it and `need` is their
    // number
    std::vector<uint32_t> total(Q.nctl, 0), seen(Q.nctl, 0), uses(Q.segs.size(), 0);
    size_t ntasks = 0;
    for (size_t i = 0; i < tasks.size(); i++) { total[Q.segs[i].done] += (uint32_t)tasks[i].size(); ntasks += tasks[i].size(); }
    if (q.size() != ntasks) return fail(ZAPTEX_E_ARG, "batch plan: %zu queue entries for %zu tasks", q.size(), ntasks);
    for (uint32_t e : q) {
        const uint32_t sg = e >> ZT_TASK_SEG_SHIFT, j = e & ((1u << ZT_TASK_SEG_SHIFT) - 1u);
        if (sg >= Q.segs.size() || j >= tasks[sg].size()) return fail(ZAPTEX_E_ARG, "batch plan: bad queue entry %08x", e);
        const zt_seg &S = Q.segs[sg];
        if (S.wait != ZT_NO_WAIT && (S.wait >= Q.nctl || S.need != total[S.wait] || seen[S.wait] != total[S.wait]))
            return fail(ZAPTEX_E_ARG, "batch plan: a task of segment %u waits for tasks after it", sg);
        seen[S.done]++;
        uses[sg]++;
    }
    for (size_t i = 0; i < tasks.size(); i++)
        if (uses[i] != tasks[i].size()) return fail(ZAPTEX_E_ARG, "batch plan: segment %zu queued %u of %zu tasks", i, uses[i], tasks[i].size());
    return ZAPTEX_OK;
}

struct BatchSizes { size_t steps, items, mipjobs, jobs14, jobs7; uint64_t blocks14, blocks7, blocks14_0, blocks7_0; int n14, n7; };
BatchSizes batch_sizes(const zaptex_texture *t, int n, const MipPlan &P) {
    BatchSizes z{};
    for (int i = 0; i < n; i++) {
        int cw = t[i].w, ch = t[i].h;
        for (int l = 0; l < t[i].levels; l++) {
            const uint64_t nb = (uint64_t)((cw + 3) / 4) * ((ch + 3) / 4);
            if (t[i].fmt == ZAPTEX_BC7) { z.n7++; z.blocks7 += nb; z.blocks7_0 += l ? 0 : nb; }
            else { z.n14++; z.blocks14 += nb; z.blocks14_0 += l ? 0 : nb; }
            cw = cw > 1 ? cw / 2 : 1; ch = ch > 1 ? ch / 2 : 1;
        }
    }
    z.steps = align_up(P.steps.size() * sizeof(zt_mipstep), 256);
    z.items = align_up(P.items.size() * sizeof(uint2), 256);
    z.mipjobs = align_up(P.jobs.size() * sizeof(zt_mipjob), 256);
    z.jobs14 = align_up((size_t)z.n14 * sizeof(zt_job), 256);
    z.jobs7 = align_up((size_t)z.n7 * sizeof(zt_job), 256);
    return z;   // (each family's table is held twice: whole, and its levels >= 1 rebased to block 0)
}
// The encoders' workspace of a batch: the BC7 lists of the level-0 blocks, those of the later levels (the two run
// at once, on their own streams), then 16 B of BC1 climb state per BC1-family block (level 0 first).  The two BC7
// regions together hold at least ws_bytes_for(all BC7 blocks), what one list over every block needs.
size_t batch_bc7_a_bytes(const BatchSizes &z) { return align_up(ws_bytes_for(z.blocks7_0), 256); }
size_t batch_bc7_bytes(const BatchSizes &z) { return batch_bc7_a_bytes(z) + ws_bytes_for(z.blocks7 - z.blocks7_0); }
size_t batch_enc_bytes(const BatchSizes &z) { return align_up(batch_bc7_bytes(z), 256) + z.blocks14 * 16;
}
}  // namespace

struct zaptex_batch {
    int device = 0;
    std::vector<MipLaunch> mip_launches;
    std::vector<size_t> tile_steps;
    int mip_tile = 64;
    zt_mipstep *steps = nullptr;
    uint2 *items = nullptr;
    zt_mipjob *mipjobs = nullptr;
    zt_jobs j14{}, j7{};
    uint32_t tot14 = 0, tot7 = 0;
    // the graph's branches: each family's level-0 jobs (they read the bases only, so they run beside the mips) and
    // its later levels (rebased to block 0, after the mips and the family's level-0 encode, sharing its workspace)
    zt_jobs j14a{}, j14b{}, j7a{}, j7b{};
    uint32_t tot14a = 0, tot14b = 0, tot7a = 0, tot7b = 0;
    void *bc7ws = nullptr;     // the BC7 lists: the whole region (one list over every block) ...
    size_t bc7ws_bytes = 0;
    void *bc7ws_b = nullptr;   // ... or the level-0 lists at its start and the later levels' lists from here
    size_t bc7ws_a_bytes = 0, bc7ws_b_bytes = 0;
    uint4 *climb = nullptr;
    bool persistent = false;   // one zt_batch_persistent launch (else the per-stage kernels)
    zt_batch_args args;
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

size_t zaptex_batch_mip_bytes(int w, int h, int levels) {
    size_t m = 0;
    if (w < 1 || h < 1 || levels < 1) return 0;
    batch_layout(w, h, ZAPTEX_BC1, levels, nullptr, nullptr, &m, nullptr);
    return m;
}
size_t zaptex_batch_out_bytes(int w, int h, int fmt, int levels) {
    size_t o = 0;
    if (w < 1 || h < 1 || levels < 1) return 0;
    batch_layout(w, h, fmt, levels, nullptr, nullptr, nullptr, &o);
    return o;
}
}  // extern "C"
namespace {
// the persistent queue's regions: task list, segments, control words (counted with null table addresses)
struct QueueSizes { size_t queue, segs, ctl; };
QueueSizes queue_sizes(const PersistentQueue &Q) {
    return { align_up(Q.queue.size() * 4, 256), align_up(Q.segs.size() * sizeof(zt_seg), 256), align_up((size_t)Q.nctl * 4, 256) };
}
}  // namespace
extern "C" {

size_t zaptex_batch_workspace_size(const zaptex_texture *t, int n) {
    if (!t || n < 1) return 0;
    for (int i = 0; i < n; i++)
        if (!size_ok(t[i].w, t[i].h) || t[i].levels < 1 || t[i].levels > max_levels(t[i].w, t[i].h)) return 0;
    MipPlan P = plan_mips(t, n, batch_mip_tile());
    BatchSizes z = batch_sizes(t, n, P);
    PersistentQueue Q;
    if (z.blocks14 >= (1ull << 31) || z.blocks7 >= (1ull << 31)) return 0;
    if (plan_queue(P, family_jobs(t, n, P, false), family_jobs(t, n, P, true), (uint32_t)z.blocks14, (uint32_t)z.blocks7, PqTables(), Q)) return 0;
    const QueueSizes qz = queue_sizes(Q);
    return 256 + z.steps + z.items + z.mipjobs + 2 * (z.jobs14 + z.jobs7) + qz.queue + qz.segs + qz.ctl + batch_enc_bytes(z);
}

int zaptex_batch_create(const zaptex_texture *t, int n, void *ws, size_t ws_bytes, zaptex_batch **out) {
    if (!t || n < 1 || !out) return fail(ZAPTEX_E_ARG, "bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (int rc = check_texture(t[i])) return rc;
    MipPlan P = plan_mips(t, n, batch_mip_tile());
    BatchSizes z = batch_sizes(t, n, P);
    if (z.blocks7 >= (1ull << 31) || z.blocks14 >= (1ull << 31)) return fail(ZAPTEX_E_ARG, "too many blocks in one batch");
    if (!ws || ws_bytes < zaptex_batch_workspace_size(t, n)) return fail(ZAPTEX_E_ARG, "workspace %zu < %zu", ws_bytes, zaptex_batch_workspace_size(t, n));
    DeviceState *st;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (int rc = ensure_ready(st)) return rc;
    }
    zaptex_batch *b = new zaptex_batch();
    HIP_TRY(hipGetDevice(&b->device));
    const std::vector<FamJob> f14 = family_jobs(t, n, P, false), f7 = family_jobs(t, n, P, true);
    PersistentQueue Q;
    if (int rc = plan_queue(P, f14, f7, (uint32_t)z.blocks14, (uint32_t)z.blocks7, PqTables(), Q)) { delete b; return rc; }
    const QueueSizes qz = queue_sizes(Q);
    uint8_t *p = (uint8_t *)align_up((uintptr_t)ws, 256);
    b->steps = (zt_mipstep *)p; p += z.steps;
    b->items = (uint2 *)p; p += z.items;
    b->mipjobs = (zt_mipjob *)p; p += z.mipjobs;
    zt_job *d14 = (zt_job *)p; p += z.jobs14;
    zt_job *d7 = (zt_job *)p; p += z.jobs7;
    zt_job *d14b = (zt_job *)p; p += z.jobs14;
    zt_job *d7b = (zt_job *)p; p += z.jobs7;
    uint32_t *dqueue = (uint32_t *)p; p += qz.queue;
    zt_seg *dsegs = (zt_seg *)p; p += qz.segs;
    unsigned *dctl = (unsigned *)p; p += qz.ctl;
    b->bc7ws = p;
    b->bc7ws_bytes = batch_bc7_bytes(z);
    b->bc7ws_a_bytes = batch_bc7_a_bytes(z);
    b->bc7ws_b = p + b->bc7ws_a_bytes;
    b->bc7ws_b_bytes = b->bc7ws_bytes - b->bc7ws_a_bytes;
    b->climb = (uint4 *)(p + align_up(b->bc7ws_bytes, 256));
    b->mip_launches = P.launches;
    b->tile_steps = P.tile_steps;
    b->mip_tile = P.tile;
    // encode jobs per family in queue order (family_jobs: level 0 of every texture, then by mip round); level
    // l >= 1 is read from the texture's mip arena
    std::vector<zt_job> h14, h7;
    for (int fam = 0; fam < 2; fam++)
        for (const FamJob &fj : fam ? f7 : f14) {
            const zaptex_texture &tx = t[fj.tex];
            std::vector<size_t> mo, oo;
            batch_layout(tx.w, tx.h, tx.fmt, tx.levels, &mo, &oo, nullptr, nullptr);
            int cw = tx.w, ch = tx.h;
            for (int l = 0; l < fj.lvl; l++) { cw = cw > 1 ? cw / 2 : 1; ch = ch > 1 ? ch / 2 : 1; }
            zt_job j{};
            j.src = fj.lvl ? (const uint8_t *)tx.mips + mo[fj.lvl - 1] : (const uint8_t *)tx.base;
            j.stride = fj.lvl ? (uint64_t)cw * 4 : tx.stride;
            j.dst = (uint8_t *)tx.out + oo[fj.lvl]; j.w = cw; j.h = ch; j.bw = (cw + 3) / 4; j.fmt = tx.fmt;
            if (fam) { j.first = b->tot7; b->tot7 += fj.blocks; h7.push_back(j); }
            else { j.first = b->tot14; b->tot14 += fj.blocks; h14.push_back(j); }
        }
    hipStream_t s = st->stream;
    if (!P.steps.empty()) HIP_TRY(hipMemcpyAsync(b->steps, P.steps.data(), P.steps.size() * sizeof(zt_mipstep), hipMemcpyHostToDevice, s));
    if (!P.items.empty()) HIP_TRY(hipMemcpyAsync(b->items, P.items.data(), P.items.size() * sizeof(uint2), hipMemcpyHostToDevice, s));
    if (!P.jobs.empty()) HIP_TRY(hipMemcpyAsync(b->mipjobs, P.jobs.data(), P.jobs.size() * sizeof(zt_mipjob), hipMemcpyHostToDevice, s));
    if (!h14.empty()) HIP_TRY(hipMemcpyAsync(d14, h14.data(), h14.size() * sizeof(zt_job), hipMemcpyHostToDevice, s));
    if (!h7.empty()) HIP_TRY(hipMemcpyAsync(d7, h7.data(), h7.size() * sizeof(zt_job), hipMemcpyHostToDevice, s));
    b->j14.ext = d14; b->j14.n = (int)h14.size();
    b->j7.ext = d7; b->j7.n = (int)h7.size();
    // the level-0 prefix of each family (family_jobs puts level 0 of every texture first) and the rest, rebased
    std::vector<zt_job> r14, r7;
    auto split = [&](const std::vector<FamJob> &f, const std::vector<zt_job> &h, zt_job *dall, zt_job *drest, std::vector<zt_job> &rest,
                     zt_jobs &ja, zt_jobs &jb, uint32_t &ta, uint32_t &tb, uint32_t tot) {
        size_t k = 0;
        while (k < f.size() && f[k].lvl == 0) k++;
        ta = k < h.size() ? h[k].first : tot;
        tb = tot - ta;
        for (size_t i = k; i < h.size(); i++) { zt_job j = h[i]; j.first -= ta; rest.push_back(j); }
        ja.ext = dall; ja.n = (int)k;
        jb.ext = drest; jb.n = (int)rest.size();
    };
    split(f14, h14, d14, d14b, r14, b->j14a, b->j14b, b->tot14a, b->tot14b, b->tot14);
    split(f7, h7, d7, d7b, r7, b->j7a, b->j7b, b->tot7a, b->tot7b, b->tot7);
    if (b->tot7a != z.blocks7_0 || b->tot14a != z.blocks14_0) { delete b; return fail(ZAPTEX_E_ARG, "batch plan: level-0 block counts disagree"); }
    if (!r14.empty()) HIP_TRY(hipMemcpyAsync(d14b, r14.data(), r14.size() * sizeof(zt_job), hipMemcpyHostToDevice, s));
    if (!r7.empty()) HIP_TRY(hipMemcpyAsync(d7b, r7.data(), r7.size() * sizeof(zt_job), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    // the persistent kernel (the mode-1 split it runs covers M1K <= 8, the default 4): queue, segments, counters
    b->persistent = g_batch_persistent && g_m1k <= 8;
    if (b->persistent) {
        PqTables tb;
        tb.steps = b->steps; tb.items = b->items; tb.mipjobs = b->mipjobs;
        if (int rc = plan_queue(P, f14, f7, b->tot14, b->tot7, tb, Q)) { delete b; return rc; }
        if (queue_sizes(Q).queue > qz.queue || queue_sizes(Q).segs > qz.segs) { delete b; return fail(ZAPTEX_E_ARG, "batch queue outgrew its workspace"); }
        HIP_TRY(hipMemcpyAsync(dqueue, Q.queue.data(), Q.queue.size() * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(dsegs, Q.segs.data(), Q.segs.size() * sizeof(zt_seg), hipMemcpyHostToDevice, s));
        zt_batch_args &A = b->args;
        memset(&A, 0, sizeof A);
        A.queue = dqueue; A.ntasks = (uint32_t)Q.queue.size(); A.nctl = Q.nctl; A.segs = dsegs; A.ctl = dctl;
        A.j14 = b->j14; A.j7 = b->j7; A.tot14 = b->tot14; A.tot7 = b->tot7; A.climb = b->climb;
        A.refill_min = g_climb_refill;
        A.range = (uint32_t)std::min<uint64_t>(ZT_CLIMB_RANGE, std::max<uint64_t>(64, ((b->tot14 + 4095) / 4096 + 63) / 64 * 64));
        if (b->tot7) A.L = lists_in(b->bc7ws, b->tot7);   // (no BC7 blocks: no lists, null counters)
        A.m1k = g_m1k;
        // the queue state starts cleared; the kernel's last workgroup clears it again for the next launch
        HIP_TRY(hipMemsetAsync(dctl, 0, (size_t)Q.nctl * 4, s));
        if (b->tot7) HIP_TRY(hipMemsetAsync(A.L.counters, 0, ZT_BC7_COUNTERS * 4, s));
        HIP_TRY(hipStreamSynchronize(s));
        int per_cu = 0, cus = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)zt_batch_persistent, 256, 0));
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
        b->grid = (g_batch_groups > 0 ? g_batch_groups : std::max(1, per_cu)) * std::max(1, cus);
    }
    // capture as one graph: the mip levels in order on the capturing stream; beside them, on two more streams, each
    // family's level-0 encode (BC1-5 on one, BC7 on the other); then each family's later levels once the mips are done
    // streams: each family's level-0 encode (side 0: BC1-5, side 1: BC7) and its later levels (sides 2 and 3), the
    // later levels in their own workspace (BC7 lists of their own, the BC1 climb state after level 0's), so they start
    // as soon as the mips are done
    hipStream_t side[4] = { nullptr, nullptr, nullptr, nullptr };
    hipEvent_t ev[6] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };   // fork, mips done, sides 0..3 done
    auto cleanup = [&] {
        for (hipEvent_t &e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
        for (hipStream_t &x : side) if (x) { (void)hipStreamDestroy(x); x = nullptr; }
    };
    const bool branches = !b->persistent && !P.launches.empty();
    if (branches) {
        for (hipStream_t &x : side) if (hipStreamCreateWithFlags(&x, hipStreamNonBlocking) != hipSuccess) { cleanup(); delete b; return fail(ZAPTEX_E_DEVICE, "stream create"); }
        for (hipEvent_t &e : ev) if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { cleanup(); delete b; return fail(ZAPTEX_E_DEVICE, "event create"); }
    }
    auto mips_body = [&](hipStream_t cs) -> int {
        int rc;
        size_t tl = 0;  // kind-0 launches seen
        for (const MipLaunch &ml : b->mip_launches) {
            if (ml.kind == 0) {
                const zt_mipstep *steps = b->steps + b->tile_steps[tl++];
                const uint2 *items = b->items + ml.first;
                if ((rc = launch("zt_mip_tiles_batch", cs, [&] {
                         if (b->mip_tile == 32) hipLaunchKernelGGL(zt_mip_tiles32_batch, dim3(ml.count), dim3(256), 0, cs, steps, items);
                         else hipLaunchKernelGGL(zt_mip_tiles_batch, dim3(ml.count), dim3(256), 0, cs, steps, items); })))
                    return rc;
            } else if (ml.kind == 1) {
                const zt_mipjob *jobs = b->mipjobs + ml.first;
                const int wg = g_cfg[CFG_MIPB].wg;
                if ((rc = launch("zt_mip_batch", cs, [&] {
                         hipLaunchKernelGGL(zt_mip_batch, dim3(groups_for(CFG_MIPB, ml.px, wg)), dim3(wg), 0, cs, jobs, ml.count, ml.px, ml.srgb); })))
                    return rc;
            } else {
                const zt_mipstep *steps = b->steps + ml.first;
                // 256 lanes: in the graph the tail steps run beside the level-0 encoders, where a 1024-lane workgroup
                // waits for a CU with 16 free wave slots (a round took 5 ms instead of 22 us alone), a 256-lane one for 4
                if ((rc = launch("zt_mip_tail_batch", cs, [&] { hipLaunchKernelGGL(zt_mip_tail_batch, dim3(ml.count), dim3(256), 0, cs, steps); })))
                    return rc;
            }
        }
        return ZAPTEX_OK;
    };
    auto body = [&](hipStream_t cs) -> int {
        int rc;
        if (b->persistent)
            return launch("zt_batch_persistent", cs, [&] { hipLaunchKernelGGL(zt_batch_persistent, dim3(b->grid), dim3(256), 0, cs, b->args); });
        if (!branches) {   // no mips: the encoders in order
            if (b->tot14 && (rc = enqueue_bc14(b->j14, b->tot14, b->climb, cs))) return rc;
            if (b->tot7 && (rc = enqueue_bc7(b->j7, b->tot7, b->bc7ws, b->bc7ws_bytes, cs))) return rc;
            return ZAPTEX_OK;
        }
        HIP_TRY(hipEventRecord(ev[0], cs));
        for (int k = 0; k < 2; k++) HIP_TRY(hipStreamWaitEvent(side[k], ev[0], 0));
        if (b->tot14a && (rc = enqueue_bc14(b->j14a, b->tot14a, b->climb, side[0]))) return rc;
        if (b->tot7a && (rc = enqueue_bc7(b->j7a, b->tot7a, b->bc7ws, b->bc7ws_a_bytes, side[1]))) return rc;
        if ((rc = mips_body(cs))) return rc;
        HIP_TRY(hipEventRecord(ev[1], cs));
        for (int k = 2; k < 4; k++) HIP_TRY(hipStreamWaitEvent(side[k], ev[1], 0));
        if (b->tot14b && (rc = enqueue_bc14(b->j14b, b->tot14b, b->climb + b->tot14a, side[2]))) return rc;
        if (b->tot7b && (rc = enqueue_bc7(b->j7b, b->tot7b, b->bc7ws_b, b->bc7ws_b_bytes, side[3]))) return rc;
        for (int k = 0; k < 4; k++) {
            HIP_TRY(hipEventRecord(ev[2 + k], side[k]));
            HIP_TRY(hipStreamWaitEvent(cs, ev[2 + k], 0));
        }
        return ZAPTEX_OK;
    };
    {
        std::lock_guard<std::mutex> lk(g_mu);
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        g_capturing = true;
        int rc = body(s);
        g_capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        cleanup();
        if (rc) { if (g) (void)hipGraphDestroy(g); delete b; return rc; }
        if (e != hipSuccess) { delete b; return fail(ZAPTEX_E_DEVICE, "graph capture: %s", hipGetErrorString(e)); }
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { (void)hipGraphDestroy(g); delete b; return fail(ZAPTEX_E_DEVICE, "graph instantiate: %s", hipGetErrorString(e)); }
    }
    *out = b;
    return ZAPTEX_OK;
}

int zaptex_batch_run(zaptex_batch *b, void *stream) {
    if (!b || !b->exec) return fail(ZAPTEX_E_ARG, "no batch");
    int dev;
    if (int rc = current_device(dev)) return rc;
    if (dev != b->device) return fail(ZAPTEX_E_ARG, "batch was built on device %d, current device is %d", b->device, dev);
    HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return ZAPTEX_OK;
}

void zaptex_batch_destroy(zaptex_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    delete b;
}

int zaptex_set_launch(const char *kernel, int wg, int max_groups) {
    if (!kernel || (wg != 64 && wg != 128 && wg != 256) || max_groups < 1)
        return fail(ZAPTEX_E_ARG, "workgroup size must be 64, 128 or 256 and the group cap >= 1");
    for (LaunchCfg &c : g_cfg)
        if (!strcmp(c.name, kernel)) { c.wg = wg; c.max_groups = max_groups; return ZAPTEX_OK; }
    return fail(ZAPTEX_E_ARG, "no tunable kernel named %s", kernel);
}
int zaptex_get_launch(const char *kernel, int *wg, int *max_groups) {
    for (LaunchCfg &c : g_cfg)
        if (kernel && !strcmp(c.name, kernel)) { if (wg) *wg = c.wg; if (max_groups) *max_groups = c.max_groups; return ZAPTEX_OK; }
    return fail(ZAPTEX_E_ARG, "no tunable kernel named %s", kernel ? kernel : "(null)");
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
    if (key == ZAPTEX_OPT_RDO_SEGMENT) {
        if (value != 0 && (value < 16 || value > (1 << 24))) return fail(ZAPTEX_E_ARG, "rdo segment must be 0 (automatic) or 16..2^24 blocks");
        g_rdo_seg = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_RDO_WARMUP) {
        if (value < 0 || value > (1 << 24)) return fail(ZAPTEX_E_ARG, "rdo warm-up must be 0..2^24 blocks");
        g_rdo_warm = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_CLIMB_REFILL) {
        if (value < 1 || value > 64) return fail(ZAPTEX_E_ARG, "refill threshold must be 1..64");
        g_climb_refill = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_CLIMB4_MAX) {
        if (value < 0) return fail(ZAPTEX_E_ARG, "grouped-climb threshold must be >= 0");
        g_climb4_max = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_CLIMB4_RANGE) {
        if (value != 0 && (value < 8 || value > 65536 || value % 8)) return fail(ZAPTEX_E_ARG, "grouped-climb range must be 0 or a multiple of 8 in 8..65536");
        g_climb4_range = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_CLIMB_GROUP) {
        if (value != 4 && value != 8) return fail(ZAPTEX_E_ARG, "grouped-climb lanes per block must be 4 or 8");
        g_climb_group = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_S6_CHUNKS) {
        if (value < 0 || value > 64) return fail(ZAPTEX_E_ARG, "mode-6 chunks per wave must be 0..64");
        g_s6_chunks = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_MIP_TILE) {
        if (value != 32 && value != 64) return fail(ZAPTEX_E_ARG, "mip tile side must be 32 or 64");
        g_mip_tile = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_BATCH_PERSISTENT) {
        g_batch_persistent = value != 0;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_BATCH_MIX) {
        if (value < 0 || value > 64) return fail(ZAPTEX_E_ARG, "batch mix must be 0..64");
        g_batch_mix = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_BATCH_CHUNKS) {
        if (value < 1 || value > 64) return fail(ZAPTEX_E_ARG, "batch chunks per task must be 1..64");
        g_batch_per = value;
        return ZAPTEX_OK;
    }
    if (key == ZAPTEX_OPT_BATCH_GROUPS) {
        if (value < 0 || value > 32) return fail(ZAPTEX_E_ARG, "batch workgroups per CU must be 0..32");
        g_batch_groups = value;
        return ZAPTEX_OK;
    }
    return fail(ZAPTEX_E_ARG, "unknown option %d", key);
}
int zaptex_get_option(int key) {
    return key == ZAPTEX_OPT_BC7_M1K ? g_m1k : key == ZAPTEX_OPT_PROFILE ? (int)g_profile : key == ZAPTEX_OPT_CLIMB_REFILL ? g_climb_refill
         : key == ZAPTEX_OPT_RDO_SEGMENT ? g_rdo_seg : key == ZAPTEX_OPT_RDO_WARMUP ? g_rdo_warm
         : key == ZAPTEX_OPT_BATCH_PERSISTENT ? g_batch_persistent : key == ZAPTEX_OPT_BATCH_MIX ? g_batch_mix
         : key == ZAPTEX_OPT_BATCH_CHUNKS ? g_batch_per : key == ZAPTEX_OPT_BATCH_GROUPS ? g_batch_groups
         : key == ZAPTEX_OPT_MIP_TILE ? g_mip_tile : key == ZAPTEX_OPT_CLIMB4_MAX ? g_climb4_max
         : key == ZAPTEX_OPT_CLIMB4_RANGE ? g_climb4_range : key == ZAPTEX_OPT_CLIMB_GROUP ? g_climb_group
         : key == ZAPTEX_OPT_S6_CHUNKS ? g_s6_chunks : -1;
}

int zaptex_profile_collect(void) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (const ProfPending &p : g_prof_pending) {
        HIP_TRY(hipEventSynchronize(p.b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        auto it = std::find_if(g_prof_total.begin(), g_prof_total.end(), [&](const ProfTotal &t) { return !strcmp(t.name, p.name); });
        if (it == g_prof_total.end()) g_prof_total.push_back({ p.name, ms, 1 });
        else { it->ms += ms; it->n++; }
        (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b);
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
