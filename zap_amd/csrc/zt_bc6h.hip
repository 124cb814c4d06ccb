// zt_bc6h.hip — BC6H unsigned (mode 11 only) encoder and decoder for gfx950, one lane per block
// (zap_tex.h:817-975).  The reference scores palettes with a double-precision squared error of integer half
// bits; every term is an integer below 2^32 and every sum below 2^53, so uint32/uint64 arithmetic gives the
// same comparisons exactly.
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

ZT_DEV uint32_t half_of(float f) {  // zap_tex.h:820-827
    if (!(f > 0)) return 0;
    uint32_t x = __float_as_uint(f);
    if (x >= 0x477FE000u) return 0x7BFF;
    if (x < 0x38800000u) return (uint32_t)(f * 16777216.0f + 0.5f);
    return ((x - 0x38000000u) + 0x1000u) >> 13;
}
ZT_DEV int unquant(int e) { return e == 0 ? 0 : e == 1023 ? 0xFFFF : ((e << 16) + 0x8000) >> 10; }
ZT_DEV int finish(int u) { return (u * 31) >> 6; }
ZT_DEV int w4(int j) { return (j * 64 + 7) / 15; }  // BC7/BC6H 4-bit weights

ZT_DEV int quant(float u) {  // zap_tex.h:841-851
    int g = cvt_i32((u - 32.0f) / 64.0f + 0.5f), best = 0;
    float bd = 1e30f;
#pragma unroll
    for (int dt = -1; dt <= 1; dt++) {
        long long t = (long long)g + dt;
        int e = t < 0 ? 0 : t > 1023 ? 1023 : (int)t;
        float d = (float)unquant(e) - u;
        d = d < 0 ? -d : d;
        if (d < bd) { bd = d; best = e; }
    }
    return best;
}

// v_dot2_u32_u16: a.lo * b.lo + a.hi * b.hi + c, modulo 2^32
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
ZT_DEV uint32_t dot2(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2_t, a), __builtin_bit_cast(u16x2_t, b), c, false);
}
// the block's half values in LDS (a lane's pixel i at word i * ZT_B6_WG, so a wave's reads hit consecutive banks),
// two channels per word: T01(i) = h0 | h1 << 16, T2(i) = h2 (each h <= 0x7BFF)
#define ZT_B6_WG 256
struct HalfPx {
    uint32_t *t01, *t2;
    __device__ __forceinline__ uint32_t T01(int i) const { return t01[i * ZT_B6_WG]; }
    __device__ __forceinline__ uint32_t T2(int i) const { return t2[i * ZT_B6_WG]; }
};
ZT_DEV int hch(const HalfPx &H, int i, int c) { return c == 0 ? (int)(H.T01(i) & 0xFFFFu) : c == 1 ? (int)(H.T01(i) >> 16) : (int)H.T2(i); }
ZT_DEV float uval(const HalfPx &H, int i, int c) { return (float)hch(H, i, c) * 64.0f / 31.0f; }   // zap_tex.h:878

// zap_tex.h:854-871: best 4-bit index per pixel for endpoints e, SSE in half-bit space.  Palette values p <= 0x7BFF,
// so 2p fits 16 bits and |t - p|^2 = |t|^2 + |p|^2 - t . 2p is formed with two v_dot2_u32_u16; every true squared
// error is below 3 * 0x7BFF^2 < 2^32, so the uint32 result is exact whatever the intermediate wraps.
ZT_DEV uint64_t fit(const HalfPx &H, const int e[2][3], uint64_t &idx) {
    int u0[3], u1[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { u0[c] = unquant(e[0][c]); u1[c] = unquant(e[1][c]); }
    uint32_t P01[16], P2[16], PP[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int w = w4(j);
        uint32_t p[3];
#pragma unroll
        for (int c = 0; c < 3; c++) p[c] = (uint32_t)finish(((64 - w) * u0[c] + w * u1[c] + 32) >> 6);
        P01[j] = 2 * p[0] | (2 * p[1]) << 16;
        P2[j] = 2 * p[2];
        PP[j] = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    }
    uint64_t err = 0, bits = 0;
#pragma unroll 2
    for (int i = 0; i < 16; i++) {
        const uint32_t t01 = H.T01(i), t2 = H.T2(i), tt = dot2(t01, t01, dot2(t2, t2, 0u));
        uint32_t be = 0xFFFFFFFFu;
        int bi = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint32_t d = tt + PP[j] - dot2(t01, P01[j], dot2(t2, P2[j], 0u));
            if (d < be) { be = d; bi = j; }
        }
        bits |= (uint64_t)bi << (4 * i);
        err += be;
    }
    idx = bits;
    return err;
}

ZT_DEV void encode_block(const HalfPx &H, uint64_t &olo, uint64_t &ohi) {  // zap_tex.h:873-931
    float mean[3] = { 0.0f, 0.0f, 0.0f }, cv[6] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f }, ax[3] = { 1.0f, 1.0f, 1.0f };
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int c = 0; c < 3; c++) mean[c] += uval(H, i, c) / 16.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float r = uval(H, i, 0) - mean[0], g = uval(H, i, 1) - mean[1], b = uval(H, i, 2) - mean[2];
        cv[0] += r * r; cv[1] += r * g; cv[2] += r * b; cv[3] += g * g; cv[4] += g * b; cv[5] += b * b;
    }
    for (int it = 0; it < 6; it++) {
        float x = cv[0] * ax[0] + cv[1] * ax[1] + cv[2] * ax[2], y = cv[1] * ax[0] + cv[3] * ax[1] + cv[4] * ax[2];
        float z = cv[2] * ax[0] + cv[4] * ax[1] + cv[5] * ax[2], m = x < 0 ? -x : x;
        if ((y < 0 ? -y : y) > m) m = y < 0 ? -y : y;
        if ((z < 0 ? -z : z) > m) m = z < 0 ? -z : z;
        if (m < 1e-6f) break;
        ax[0] = x / m; ax[1] = y / m; ax[2] = z / m;
    }
    float n2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], tl = 1e30f, th = -1e30f;
    float bl[3] = { 1e30f, 1e30f, 1e30f }, bh[3] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float d = (uval(H, i, 0) - mean[0]) * ax[0] + (uval(H, i, 1) - mean[1]) * ax[1] + (uval(H, i, 2) - mean[2]) * ax[2];
        if (d < tl) tl = d;
        if (d > th) th = d;
#pragma unroll
        for (int c = 0; c < 3; c++) { const float u = uval(H, i, c); if (u < bl[c]) bl[c] = u; if (u > bh[c]) bh[c] = u; }
    }
    uint64_t best = ~0ull, bidx = 0;
    int be[2][3] = { { 0, 0, 0 }, { 0, 0, 0 } };
    for (int s = 0; s < 2; s++) {  // principal axis, then the per-channel bounding box
        int e[2][3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float lo = s ? bl[c] : mean[c] + ax[c] * tl / n2, hi = s ? bh[c] : mean[c] + ax[c] * th / n2;
            e[0][c] = quant(lo); e[1][c] = quant(hi);
        }
        uint64_t idx;
        uint64_t err = fit(H, e, idx);
        for (int it = 0; it < 3 && err > 0; it++) {
            float aa = 0.0f, ab = 0.0f, bb = 0.0f, xa[3] = { 0.0f, 0.0f, 0.0f }, xb[3] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int i = 0; i < 16; i++) {
                float w = (float)w4((int)((idx >> (4 * i)) & 15)) / 64.0f, v = 1 - w;
                aa += v * v; ab += v * w; bb += w * w;
#pragma unroll
                for (int c = 0; c < 3; c++) { const float u = uval(H, i, c); xa[c] += v * u; xb[c] += w * u; }
            }
            float det = aa * bb - ab * ab;
            if (det < 1e-6f && det > -1e-6f) break;
            int ne[2][3];
#pragma unroll
            for (int c = 0; c < 3; c++) { ne[0][c] = quant((xa[c] * bb - xb[c] * ab) / det); ne[1][c] = quant((xb[c] * aa - xa[c] * ab) / det); }
            uint64_t ni;
            uint64_t e2 = fit(H, ne, ni);
            if (e2 >= err) break;
            err = e2; idx = ni;
#pragma unroll
            for (int c = 0; c < 3; c++) { e[0][c] = ne[0][c]; e[1][c] = ne[1][c]; }
        }
        if (err < best) {
            best = err; bidx = idx;
#pragma unroll
            for (int c = 0; c < 3; c++) { be[0][c] = e[0][c]; be[1][c] = e[1][c]; }
        }
    }
    if (bidx & 8) {
#pragma unroll
        for (int c = 0; c < 3; c++) { int x = be[0][c]; be[0][c] = be[1][c]; be[1][c] = x; }
        bidx ^= ~0ull;
    }
    Bits128 b{ 3, 0 };  // mode 11: 5-bit code 00011
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) b.put(5 + 10 * (3 * k + c), (uint64_t)be[k][c], 10);
    uint64_t field = (bidx & 7) | ((bidx >> 1) & ~7ull);
    b.hi |= field << 1;
    olo = b.lo; ohi = b.hi;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(ZT_B6_WG) zt_bc6h_encode(zt_jobs jobs, uint32_t total) {
    __shared__ uint32_t s_t01[16 * ZT_B6_WG], s_t2[16 * ZT_B6_WG];   // each lane's own pixels (no sharing, no barrier)
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const zt_job J = job_at(jobs, zt_find_job(jobs, g));
        uint32_t b = g - J.first;
        int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw);
        const HalfPx H{ &s_t01[threadIdx.x], &s_t2[threadIdx.x] };
#pragma unroll
        for (int i = 0; i < 16; i++) {
            int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
            if (x >= J.w) x = J.w - 1;
            if (y >= J.h) y = J.h - 1;
            float4 p = *(const float4 *)(J.src + (uint64_t)y * J.stride + (uint64_t)x * 16);
            H.t01[i * ZT_B6_WG] = half_of(p.x) | half_of(p.y) << 16;
            H.t2[i * ZT_B6_WG] = half_of(p.z);
        }
        uint64_t lo, hi;
        encode_block(H, lo, hi);
        store16(J.dst + (uint64_t)b * 16, lo, hi);
    }
}

// zap_tex.h:933-943 / :963-975: mode 11 only, anything else decodes to 0; alpha is half 1.0
extern "C" __global__ void __launch_bounds__(256) zt_bc6h_decode(const uint8_t *bc, int w, int h, uint16_t *rgba, uint64_t stride) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bw * bh) return;
    const int bx = g % bw, by = g / bw;
    uint4 raw = *(const uint4 *)(bc + (uint64_t)g * 16);
    uint64_t lo = (uint64_t)raw.x | (uint64_t)raw.y << 32, hi = (uint64_t)raw.z | (uint64_t)raw.w << 32;
    bool ok = (lo & 31) == 3;
    int e[2][3];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) e[k][c] = unquant((int)get_bits(lo, hi, 5 + 10 * (3 * k + c), 10));
    for (int i = 0; i < 16; i++) {
        int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
        if (x >= w || y >= h) continue;
        uint16_t *o = rgba + (uint64_t)y * stride + (uint64_t)x * 4;
        if (!ok) { o[0] = o[1] = o[2] = o[3] = 0; continue; }
        int pos = i ? 64 + 4 * i : 65, nb = i ? 4 : 3;
        int wt = w4((int)get_bits(lo, hi, pos, nb));
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = (uint16_t)finish(((64 - wt) * e[0][c] + wt * e[1][c] + 32) >> 6);
        o[3] = 0x3C00;
    }
}
