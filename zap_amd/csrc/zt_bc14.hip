// zt_bc14.hip — BC1 / BC3 / BC4 / BC5 encoders for gfx950, one lane per block.
//
// BC1 (zap_tex.h:97-168): PCA by power iteration, two starts (inset or not) with <= 3 least-squares
// refinements, then a 565 hill climb (<= 8 passes of 12 one-step moves).  Every trial palette argmin is an
// exact integer max over 4 keys built with v_dot4_u32_u8.  BC4 (zap_tex.h:198-213): 9 inset endpoint pairs,
// 8-entry palette argmin, integer only.  BC3 = BC4(alpha) @0 + BC1 @8; BC5 = BC4(R) @0 + BC4(G) @8
// (zap_tex.h:1490-1501 at rdo = 0).
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

ZT_DEV uint32_t rgb565_expand(uint32_t c) {  // zap_tex.h:41-44, packed R | G << 8 | B << 16
    uint32_t R = (c >> 11) & 31, G = (c >> 5) & 63, B = c & 31;
    return ((R << 3) | (R >> 2)) | ((G << 2) | (G >> 4)) << 8 | ((B << 3) | (B >> 2)) << 16;
}
ZT_DEV uint32_t to565(float r, float g, float b) {  // zap_tex.h:45-49
    int R = cvt_i32(r * 31.0f / 255.0f + 0.5f), G = cvt_i32(g * 63.0f / 255.0f + 0.5f), B = cvt_i32(b * 31.0f / 255.0f + 0.5f);
    return (uint32_t)(clampi(R, 0, 31) << 11 | clampi(G, 0, 63) << 5 | clampi(B, 0, 31));
}

// SSE and indices of a canonical (c0 >= c1) 4-colour endpoint pair; c0 == c1 is solid with all indices 0
// (zap_tex.h:52-94 with force4 = 1).  idx: 2 bits per pixel.
ZT_DEV int bc1_trial(const uint32_t px[16], uint32_t c0, uint32_t c1, uint32_t &idx) {
    uint32_t a = rgb565_expand(c0), b = rgb565_expand(c1);
    uint32_t p2 = 0, p3 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int x = chan(a, k), y = chan(b, k);
        p2 |= (uint32_t)((2 * x + y + 1) / 3) << (8 * k);
        p3 |= (uint32_t)((x + 2 * y + 1) / 3) << (8 * k);
    }
    if (c0 == c1) {
        int err = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            uint32_t p = px[i] & 0xFFFFFFu;
            err += (int)dot4(p, p) - 2 * (int)dot4(p, a) + (int)dot4(a, a);
        }
        idx = 0;
        return err;
    }
    const uint32_t pal[4] = { a, b, p2, p3 };
    int K[4];
#pragma unroll
    for (int j = 0; j < 4; j++) K[j] = 15 - j - 16 * (int)dot4(pal[j], pal[j]);
    int err = 0;
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t p = px[i] & 0xFFFFFFu;
        int best = (int)(dot4(p, pal[0]) << 5) + K[0];
#pragma unroll
        for (int j = 1; j < 4; j++) best = max(best, (int)(dot4(p, pal[j]) << 5) + K[j]);
        bits |= (uint32_t)(15 - (best & 15)) << (2 * i);
        err += (int)dot4(p, p) - (best >> 4);
    }
    idx = bits;
    return err;
}
ZT_DEV int bc1_attempt(const uint32_t px[16], uint32_t &c0, uint32_t &c1, uint32_t &idx) {  // zap_tex.h:90-94
    if (c0 < c1) { uint32_t t = c0; c0 = c1; c1 = t; }
    return bc1_trial(px, c0, c1, idx);
}

ZT_DEV int bc1_block(const uint32_t px[16], uint32_t &oc0, uint32_t &oc1, uint32_t &oidx) {
    float mean[3] = { 0.0f, 0.0f, 0.0f }, cv[6] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) mean[k] += chanf(px[i], k) / 16.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float r = chanf(px[i], 0) - mean[0], g = chanf(px[i], 1) - mean[1], b = chanf(px[i], 2) - mean[2];
        cv[0] += r * r; cv[1] += r * g; cv[2] += r * b; cv[3] += g * g; cv[4] += g * b; cv[5] += b * b;
    }
    float ax[3] = { 1.0f, 1.0f, 1.0f };
    for (int it = 0; it < 6; it++) {
        float x = cv[0] * ax[0] + cv[1] * ax[1] + cv[2] * ax[2];
        float y = cv[1] * ax[0] + cv[3] * ax[1] + cv[4] * ax[2];
        float z = cv[2] * ax[0] + cv[4] * ax[1] + cv[5] * ax[2];
        float m = x < 0 ? -x : x, my = y < 0 ? -y : y, mz = z < 0 ? -z : z;
        if (my > m) m = my;
        if (mz > m) m = mz;
        if (m < 1e-6f) break;
        ax[0] = x / m; ax[1] = y / m; ax[2] = z / m;
    }
    float n2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], lo = 1e9f, hi = -1e9f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float t = (chanf(px[i], 0) - mean[0]) * ax[0] + (chanf(px[i], 1) - mean[1]) * ax[1] + (chanf(px[i], 2) - mean[2]) * ax[2];
        if (t < lo) lo = t;
        if (t > hi) hi = t;
    }
    uint32_t b0 = 0, b1 = 0, bidx = 0;
    int best = 1 << 30;
    for (int s = 0; s < 2; s++) {
        float in = s ? 0.0f : (hi - lo) / 16.0f, l = (lo + in) / n2, u = (hi - in) / n2;
        uint32_t c0 = to565(mean[0] + ax[0] * u, mean[1] + ax[1] * u, mean[2] + ax[2] * u);
        uint32_t c1 = to565(mean[0] + ax[0] * l, mean[1] + ax[1] * l, mean[2] + ax[2] * l);
        uint32_t idx = 0;
        int e = bc1_attempt(px, c0, c1, idx);
        for (int it = 0; it < 3 && e > 0; it++) {  // least squares on the chosen indices
            const float W1 = 1.0f, W2 = 0.0f, W3 = 2.0f / 3.0f, W4 = 1.0f / 3.0f;
            float aa = 0.0f, ab = 0.0f, bb = 0.0f, sa[3] = { 0.0f, 0.0f, 0.0f }, sb[3] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int i = 0; i < 16; i++) {
                uint32_t k = (idx >> (2 * i)) & 3;
                float w = k == 0 ? W1 : k == 1 ? W2 : k == 2 ? W3 : W4, v = 1.0f - w;
                aa += w * w; ab += w * v; bb += v * v;
#pragma unroll
                for (int c = 0; c < 3; c++) { sa[c] += w * chanf(px[i], c); sb[c] += v * chanf(px[i], c); }
            }
            float det = aa * bb - ab * ab;
            if (det < 1e-6f && det > -1e-6f) break;
            float A[3], B[3];
#pragma unroll
            for (int c = 0; c < 3; c++) { A[c] = (sa[c] * bb - sb[c] * ab) / det; B[c] = (sb[c] * aa - sa[c] * ab) / det; }
            uint32_t n0 = to565(A[0], A[1], A[2]), n1 = to565(B[0], B[1], B[2]), ni = 0;
            int ne = bc1_attempt(px, n0, n1, ni);
            if (ne >= e) break;
            e = ne; c0 = n0; c1 = n1; idx = ni;
        }
        if (e < best) { best = e; b0 = c0; b1 = c1; bidx = idx; }
    }
    // hill climb: +-1 on each 565 channel of each endpoint while it helps (order: ep, channel, -1 then +1)
    for (int pass = 0, improved = 1; improved && pass < 8 && best > 0; pass++) {
        improved = 0;
        for (int mv = 0; mv < 12; mv++) {
            int ep = mv / 6, ch = (mv >> 1) % 3, d = (mv & 1) ? 1 : -1;
            int sh = ch == 0 ? 11 : ch == 1 ? 5 : 0, mx = ch == 1 ? 63 : 31;
            uint32_t c = ep ? b1 : b0;
            int v = (int)((c >> sh) & (uint32_t)mx);
            if (v + d < 0 || v + d > mx) continue;
            c = (c & ~((uint32_t)mx << sh)) | ((uint32_t)(v + d) << sh);
            uint32_t n0 = ep ? b0 : c, n1 = ep ? c : b1, ni = 0;
            int ne = bc1_attempt(px, n0, n1, ni);
            if (ne < best) { best = ne; b0 = n0; b1 = n1; bidx = ni; improved = 1; }
        }
    }
    oc0 = b0; oc1 = b1; oidx = bidx;
    return best;
}

// BC4 on channel `ch` of the block; returns the 8-byte block (a0, a1, 48 index bits)
ZT_DEV uint64_t bc4_block(const uint32_t px[16], int ch) {
    int v[16], lo = 255, hi = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) { v[i] = chan(px[i], ch); lo = min(lo, v[i]); hi = max(hi, v[i]); }
    int best = 1 << 30, b0 = hi, b1 = lo;
    uint64_t bi = 0;
    for (int d0 = 0; d0 <= 2; d0++)
        for (int d1 = 0; d1 <= 2; d1++) {
            int a0 = hi - d0, a1 = lo + d1;
            if (a0 <= a1) { a0 = hi; a1 = lo; }
            int pal[8];
            pal[0] = a0; pal[1] = a1;
            if (a0 > a1) {
#pragma unroll
                for (int i = 2; i < 8; i++) pal[i] = ((8 - i) * a0 + (i - 1) * a1 + 3) / 7;
            } else {
#pragma unroll
                for (int i = 2; i < 6; i++) pal[i] = ((6 - i) * a0 + (i - 1) * a1 + 2) / 5;
                pal[6] = 0; pal[7] = 255;
            }
            int e = 0;
            uint64_t idx = 0;
            if (a0 == a1) {
#pragma unroll
                for (int i = 0; i < 16; i++) e += (v[i] - a0) * (v[i] - a0);
            } else {
                int K[8];
#pragma unroll
                for (int j = 0; j < 8; j++) K[j] = 15 - j - 16 * pal[j] * pal[j];
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    int bk = ((v[i] * pal[0]) << 5) + K[0];
#pragma unroll
                    for (int j = 1; j < 8; j++) bk = max(bk, ((v[i] * pal[j]) << 5) + K[j]);
                    idx |= (uint64_t)(15 - (bk & 15)) << (3 * i);
                    e += v[i] * v[i] - (bk >> 4);
                }
            }
            if (e < best) { best = e; b0 = a0; b1 = a1; bi = idx; }
        }
    return (uint64_t)b0 | (uint64_t)b1 << 8 | bi << 16;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256) zt_bc14_encode(zt_jobs jobs, uint32_t total) {
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += step) {
        const zt_job J = job_at(jobs, zt_find_job(jobs, g));
        uint32_t b = g - J.first;
        int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw);
        uint32_t px[16];
        gather_block(J.src, J.w, J.h, J.stride, bx, by, px);
        if (J.fmt == 0) {
            uint32_t c0, c1, ix;
            bc1_block(px, c0, c1, ix);
            store8(J.dst + (uint64_t)b * 8, (uint64_t)c0 | (uint64_t)c1 << 16 | (uint64_t)ix << 32);
        } else if (J.fmt == 2) {
            store8(J.dst + (uint64_t)b * 8, bc4_block(px, 0));
        } else if (J.fmt == 1) {
            uint64_t a = bc4_block(px, 3);
            uint32_t c0, c1, ix;
            bc1_block(px, c0, c1, ix);
            store16(J.dst + (uint64_t)b * 16, a, (uint64_t)c0 | (uint64_t)c1 << 16 | (uint64_t)ix << 32);
        } else {
            store16(J.dst + (uint64_t)b * 16, bc4_block(px, 0), bc4_block(px, 1));
        }
    }
}
