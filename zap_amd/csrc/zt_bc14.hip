// zt_bc14.hip — BC1 / BC3 / BC4 / BC5 encoders for gfx950, one lane per block.
//
// BC1 (zap_tex.h:97-168): PCA by power iteration, two starts (inset or not) with <= 3 least-squares
// refinements, then a 565 hill climb (<= 8 passes of 12 one-step moves).  Every trial palette argmin is an
// exact integer max over 4 keys built with v_dot4_u32_u8.  The hill climb's length varies per block (a
// wave would wait for its slowest lane, about twice the mean), so it runs in its own kernel where every lane
// that finishes takes the next block of its wave's range: each step is one trial per lane.  BC4 (zap_tex.h:198-213): 9 inset endpoint pairs,
// 8-entry palette argmin, integer only.  BC3 = BC4(alpha) @0 + BC1 @8; BC5 = BC4(R) @0 + BC4(G) @8
// (zap_tex.h:1490-1501 at rdo = 0).
#include "zt_common.h"
#include "zt_kernels.h"

namespace {

ZT_DEV uint32_t rgb565_expand(uint32_t c) {  // zap_tex.h:41-44, packed R | G << 8 | B << 16
    uint32_t R = (c >> 11) & 31, G = (c >> 5) & 63, B = c & 31;
    return ((R << 3) | (R >> 2)) | ((G << 2) | (G >> 4)) << 8 | ((B << 3) | (B >> 2)) << 16;
}
ZT_DEV uint32_t to565(float r, float g, float b) {  // zap_tex.h:45-49; x / 255.0f as div255 (+0.5 erases a -0 difference)
    int R = cvt_i32(div255(r * 31.0f) + 0.5f), G = cvt_i32(div255(g * 63.0f) + 0.5f), B = cvt_i32(div255(b * 31.0f) + 0.5f);
    return (uint32_t)(clampi(R, 0, 31) << 11 | clampi(G, 0, 63) << 5 | clampi(B, 0, 31));
}

// SSE and indices of a canonical (c0 >= c1) 4-colour endpoint pair; c0 == c1 is solid with all indices 0
// (zap_tex.h:52-94 with force4 = 1).  idx: 2 bits per pixel.  pm = the block's pixels with alpha cleared,
// spp = sum of |pm|^2 (both per block).  key_j = 128 p.pal_j - 64 |pal_j|^2 + (3 - j) = 64 S_j + (3 - j) with
// S_j = 2 p.pal_j - |pal_j|^2 = |p|^2 - SSE_j: the largest key is the smallest SSE, lowest index on ties, and its
// low 2 bits are 3 - j.  The 16 keys' tie terms add up to at most 48 < 64, so (sum of keys) >> 6 = sum of S_j
// exactly (|key| < 2^25, the sum < 2^31).
ZT_DEV int bc1_trial(const uint32_t pm[16], int spp, uint32_t c0, uint32_t c1, uint32_t &idx) {
    const uint32_t a = rgb565_expand(c0), b = rgb565_expand(c1);
    if (c0 == c1) {
        uint32_t pa = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) pa = __builtin_amdgcn_udot4(pm[i], a, pa, false);
        idx = 0;
        return spp - 2 * (int)pa + 16 * (int)dot4(a, a);
    }
    uint32_t p2 = 0, p3 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int x = chan(a, k), y = chan(b, k);
        p2 |= (uint32_t)((2 * x + y + 1) / 3) << (8 * k);
        p3 |= (uint32_t)((x + 2 * y + 1) / 3) << (8 * k);
    }
    const int K0 = no_fold(3 - 64 * (int)dot4(a, a)), K1 = no_fold(2 - 64 * (int)dot4(b, b)), K2 = no_fold(1 - 64 * (int)dot4(p2, p2)),
              K3 = no_fold(-64 * (int)dot4(p3, p3));
    int sum = 0;
    uint32_t inv = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int k0 = (int)(dot4(pm[i], a) << 7) + K0, k1 = (int)(dot4(pm[i], b) << 7) + K1;
        const int k2 = (int)(dot4(pm[i], p2) << 7) + K2, k3 = (int)(dot4(pm[i], p3) << 7) + K3;
        const int best = max(max(k0, k1), max(k2, k3));
        inv |= (uint32_t)(best & 3) << (2 * i);
        sum += best;
    }
    idx = ~inv;
    return spp - (sum >> 6);
}
ZT_DEV int bc1_attempt(const uint32_t pm[16], int spp, uint32_t &c0, uint32_t &c1, uint32_t &idx) {  // zap_tex.h:90-94
    if (c0 < c1) { uint32_t t = c0; c0 = c1; c1 = t; }
    return bc1_trial(pm, spp, c0, c1, idx);
}
// bc1_attempt's SSE alone (the hill climb compares errors; the winner's indices are recomputed once, in
// zt_bc1_finish or at the end of the grouped climb's range): the same keys and maxima without the index bits
ZT_DEV int bc1_attempt_err(const uint32_t pm[16], int spp, uint32_t &c0, uint32_t &c1) {
    if (c0 < c1) { uint32_t t = c0; c0 = c1; c1 = t; }
    const uint32_t a = rgb565_expand(c0), b = rgb565_expand(c1);
    if (c0 == c1) {
        uint32_t pa = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) pa = __builtin_amdgcn_udot4(pm[i], a, pa, false);
        return spp - 2 * (int)pa + 16 * (int)dot4(a, a);
    }
    uint32_t p2 = 0, p3 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int x = chan(a, k), y = chan(b, k);
        p2 |= (uint32_t)((2 * x + y + 1) / 3) << (8 * k);
        p3 |= (uint32_t)((x + 2 * y + 1) / 3) << (8 * k);
    }
    const int K0 = no_fold(3 - 64 * (int)dot4(a, a)), K1 = no_fold(2 - 64 * (int)dot4(b, b)), K2 = no_fold(1 - 64 * (int)dot4(p2, p2)),
              K3 = no_fold(-64 * (int)dot4(p3, p3));
    int sum = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int k0 = (int)(dot4(pm[i], a) << 7) + K0, k1 = (int)(dot4(pm[i], b) << 7) + K1;
        const int k2 = (int)(dot4(pm[i], p2) << 7) + K2, k3 = (int)(dot4(pm[i], p3) << 7) + K3;
        sum += max(max(k0, k1), max(k2, k3));
    }
    return spp - (sum >> 6);
}
// alpha-cleared pixels and their sum of squares
ZT_DEV int bc1_prepare(const uint32_t px[16], uint32_t pm[16]) {
    uint32_t spp = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) { pm[i] = px[i] & 0xFFFFFFu; spp = __builtin_amdgcn_udot4(pm[i], pm[i], spp, false); }
    return (int)spp;
}

// the block's principal axis (zap_tex.h:99-124): mean, axis, |axis|^2 and the projections' range
struct Bc1Axis { float mean[3], ax[3], n2, lo, hi; };
ZT_DEV void bc1_axis(const uint32_t px[16], Bc1Axis &A) {
    float cv[6] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int k = 0; k < 3; k++) A.mean[k] = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) A.mean[k] += chanf(px[i], k) / 16.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float r = chanf(px[i], 0) - A.mean[0], g = chanf(px[i], 1) - A.mean[1], b = chanf(px[i], 2) - A.mean[2];
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
    float lo = 1e9f, hi = -1e9f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float t = (chanf(px[i], 0) - A.mean[0]) * ax[0] + (chanf(px[i], 1) - A.mean[1]) * ax[1] + (chanf(px[i], 2) - A.mean[2]) * ax[2];
        if (t < lo) lo = t;
        if (t > hi) hi = t;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) A.ax[k] = ax[k];
    A.n2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]; A.lo = lo; A.hi = hi;
}
// start s (0: endpoints inset by 1/16 of the range, 1: not) with at most 3 least-squares refits (zap_tex.h:129-148)
ZT_DEV int bc1_one_start(const uint32_t px[16], const uint32_t pm[16], int spp, const Bc1Axis &A, int s, uint32_t &oc0, uint32_t &oc1,
                         uint32_t &oidx) {
    float in = s ? 0.0f : (A.hi - A.lo) / 16.0f, l = (A.lo + in) / A.n2, u = (A.hi - in) / A.n2;
    uint32_t c0 = to565(A.mean[0] + A.ax[0] * u, A.mean[1] + A.ax[1] * u, A.mean[2] + A.ax[2] * u);
    uint32_t c1 = to565(A.mean[0] + A.ax[0] * l, A.mean[1] + A.ax[1] * l, A.mean[2] + A.ax[2] * l);
    uint32_t idx = 0;
    int e = bc1_attempt(pm, spp, c0, c1, idx);
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
        float Av[3], Bv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) { Av[c] = (sa[c] * bb - sb[c] * ab) / det; Bv[c] = (sb[c] * aa - sa[c] * ab) / det; }
        uint32_t n0 = to565(Av[0], Av[1], Av[2]), n1 = to565(Bv[0], Bv[1], Bv[2]), ni = 0;
        int ne = bc1_attempt(pm, spp, n0, n1, ni);
        if (ne >= e) break;
        e = ne; c0 = n0; c1 = n1; idx = ni;
    }
    oc0 = c0; oc1 = c1; oidx = idx;
    return e;
}
// everything before the hill climb: PCA, two starts, least squares; returns the best SSE (start 0 on ties)
ZT_DEV int bc1_start(const uint32_t px[16], uint32_t &oc0, uint32_t &oc1, uint32_t &oidx) {
    uint32_t pm[16];
    const int spp = bc1_prepare(px, pm);
    Bc1Axis A;
    bc1_axis(px, A);
    uint32_t b0 = 0, b1 = 0, bidx = 0;
    int best = 1 << 30;
    for (int s = 0; s < 2; s++) {
        uint32_t c0, c1, idx;
        const int e = bc1_one_start(px, pm, spp, A, s, c0, c1, idx);
        if (e < best) { best = e; b0 = c0; b1 = c1; bidx = idx; }
    }
    oc0 = b0; oc1 = b1; oidx = bidx;
    return best;
}

// hill-climb state of one block: +-1 on each 565 channel of each endpoint (order: endpoint, channel, -1 then +1),
// passes of 12 moves while a pass improved, at most 8 (zap_tex.h:152-166).  Within a pass the moves after the pass's
// last improvement were tried from the pass's final endpoints and failed; the next pass starts from those endpoints,
// so until something improves they fail again.  A pass that reaches them without an improvement therefore ends the
// climb there: `stop` is the previous pass's last improving move (11 in the first pass), `last` this pass's.  The
// result is the reference's; on textures it skips 16-21% of the trials.
struct Climb {
    uint32_t b0, b1, bidx;   // bidx: the start's indices (valid while moved == 0)
    int best, pass, mv, improved, moved, stop, last;
};
// the next valid move's endpoint pair (cheap; no trial); false when the climb is over.  Stopping as soon as the
// SSE is 0 gives the reference's result: no later move can improve on 0.
ZT_DEV bool climb_next(Climb &s, uint32_t &n0, uint32_t &n1) {
    for (int guard = 0; guard < 16; guard++) {  // >= 6 of any pass's 12 moves are valid, so this always ends early
        if (s.best == 0) return false;
        if (!s.improved && s.mv > s.stop) return false;   // the rest of the pass fails again (with s.mv == 12 as well)
        if (s.mv == 12) {
            if (s.pass == 7) return false;
            s.pass++; s.mv = 0; s.improved = 0; s.stop = s.last;
        }
        const int mv = s.mv, ep = mv / 6, ch = (mv >> 1) % 3, d = (mv & 1) ? 1 : -1;
        const int sh = ch == 0 ? 11 : ch == 1 ? 5 : 0, mx = ch == 1 ? 63 : 31;
        uint32_t c = ep ? s.b1 : s.b0;
        const int v = (int)((c >> sh) & (uint32_t)mx);
        s.mv++;
        if (v + d < 0 || v + d > mx) continue;
        c = (c & ~((uint32_t)mx << sh)) | ((uint32_t)(v + d) << sh);
        n0 = ep ? s.b0 : c; n1 = ep ? c : s.b1;
        return true;
    }
    return false;
}
// fold a trial (endpoints as bc1_attempt_err left them, canonical order; the move climb_next just gave) into the state
ZT_DEV void climb_apply(Climb &s, uint32_t n0, uint32_t n1, int ne) {
    if (ne < s.best) { s.best = ne; s.b0 = n0; s.b1 = n1; s.improved = 1; s.moved = 1; s.last = s.mv - 1; }
}
ZT_DEV uint64_t bc1_pack(uint32_t c0, uint32_t c1, uint32_t idx) { return (uint64_t)c0 | (uint64_t)c1 << 16 | (uint64_t)idx << 32; }
// the 2-bit indices of row `row` of block b (pixels 4 row .. 4 row + 3, edge-clamped like gather_block) for the
// canonical pair (c0 >= c1), at their places in the block's 32 index bits: bc1_trial's keys for four pixels
ZT_DEV uint32_t bc1_row_indices(const zt_job &J, uint32_t b, int row, uint32_t c0, uint32_t c1) {
    if (c0 == c1) return 0;
    const int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw), y = min(by * 4 + row, J.h - 1);
    const uint32_t a = rgb565_expand(c0), e = rgb565_expand(c1);
    uint32_t p2 = 0, p3 = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int x = chan(a, k), z = chan(e, k);
        p2 |= (uint32_t)((2 * x + z + 1) / 3) << (8 * k);
        p3 |= (uint32_t)((x + 2 * z + 1) / 3) << (8 * k);
    }
    const int K0 = no_fold(3 - 64 * (int)dot4(a, a)), K1 = no_fold(2 - 64 * (int)dot4(e, e)), K2 = no_fold(1 - 64 * (int)dot4(p2, p2)),
              K3 = no_fold(-64 * (int)dot4(p3, p3));
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int x = min(bx * 4 + i, J.w - 1);
        // (source rows are 4-byte aligned: the entry points refuse other bases and strides)
        const uint32_t pm = *(const uint32_t *)(J.src + (uint64_t)y * J.stride + (uint64_t)x * 4) & 0xFFFFFFu;
        const int k0 = (int)(dot4(pm, a) << 7) + K0, k1 = (int)(dot4(pm, e) << 7) + K1;
        const int k2 = (int)(dot4(pm, p2) << 7) + K2, k3 = (int)(dot4(pm, p3) << 7) + K3;
        const int best = max(max(k0, k1), max(k2, k3));
        bits |= (uint32_t)(3 - (best & 3)) << (2 * (4 * row + i));
    }
    return bits;
}

// BC4 palette of an endpoint pair with a0 > a1 (zap_tex.h:172-176, first branch): a0, a1 and the six interpolants
// ((8 - i) a0 + (i - 1) a1 + 3) / 7.  The encoder never meets a0 <= a1 with a0 != a1 (a pair that would be one is
// replaced by (hi, lo), and hi == lo is a flat block), so the other branch is the decoder's only.  The numerators are
// at most 7 * 255 + 3 = 1788, below the 13109 up to which (n * 9363) >> 16 = floor(n / 7) exactly (checked for every
// n by tests/test_kernel_identities.py).
ZT_DEV void bc4_palette8(int a0, int a1, int pal[8]) {
    pal[0] = a0; pal[1] = a1;
    int n = 6 * a0 + a1 + 3;
#pragma unroll
    for (int i = 2; i < 8; i++, n += a1 - a0) pal[i] = (int)(__umul24((uint32_t)n, 9363u) >> 16);
}

// BC4 on channel `ch` of the block; returns the 8-byte block (a0, a1, 48 index bits).  The 9 inset pairs
// (zap_tex.h:198-213) are compared by SSE alone, which does not depend on how ties between palette entries are
// broken: key_j = v * 2 pal_j - pal_j^2 (one 24-bit multiply-add: the constant goes through no_fold, or the compiler
// splits it into a multiply and a subtract) and SSE = sum (v^2 - max_j key_j).  Only the chosen pair's indices are
// then computed, with the reference's lowest-index rule.  A flat block (hi == lo) is (hi, hi) with every index 0,
// which is what all nine of its pairs give.
ZT_DEV uint64_t bc4_block(const uint32_t px[16], int ch) {
    int v[16], lo = 255, hi = 0, svv = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) { v[i] = chan(px[i], ch); lo = min(lo, v[i]); hi = max(hi, v[i]); svv += v[i] * v[i]; }
    if (hi == lo) return (uint64_t)hi | (uint64_t)hi << 8;
    int best = 1 << 30, b0 = hi, b1 = lo;
#pragma unroll 1
    for (int pr = 0; pr < 9; pr++) {
        const int d0 = pr / 3, d1 = pr % 3;
        int a0 = hi - d0, a1 = lo + d1;
        if (a0 <= a1) { a0 = hi; a1 = lo; }   // hi > lo here
        int pal[8], P2[8], K[8];
        bc4_palette8(a0, a1, pal);
#pragma unroll
        for (int j = 0; j < 8; j++) { P2[j] = 2 * pal[j]; K[j] = no_fold(-pal[j] * pal[j]); }
        int sk = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            int m = __mul24(v[i], P2[0]) + K[0];
#pragma unroll
            for (int j = 1; j < 8; j++) m = max(m, __mul24(v[i], P2[j]) + K[j]);
            sk += m;
        }
        const int e = svv - sk;
        if (e < best) { best = e; b0 = a0; b1 = a1; }
    }
    int pal[8], P32[8], K[8];
    bc4_palette8(b0, b1, pal);
#pragma unroll
    for (int j = 0; j < 8; j++) { P32[j] = 32 * pal[j]; K[j] = no_fold(15 - j - 16 * pal[j] * pal[j]); }
    uint64_t bi = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int bk = __mul24(v[i], P32[0]) + K[0];
#pragma unroll
        for (int j = 1; j < 8; j++) bk = max(bk, __mul24(v[i], P32[j]) + K[j]);
        bi |= (uint64_t)(15 - (bk & 15)) << (3 * i);
    }
    return (uint64_t)b0 | (uint64_t)b1 << 8 | bi << 16;
}

}  // namespace

// The kernels' bodies take the workgroup index and grid size as arguments (bid, nblk), so the persistent batch
// kernel can run any workgroup of them as a task.

// BC4 / BC5 blocks and the BC3 alpha half complete here; BC1 blocks and the BC3 colour half stop before the hill
// climb and leave their state in climb[g] = (c0 | c1 << 16, SSE, indices, job) for zt_bc1_climb; the other
// blocks leave (0, 0, 0, job).
namespace {
ZT_DEV void bc14_encode_body(const zt_jobs &jobs, uint32_t total, uint4 *climb, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t step = nblk * blockDim.x;
    for (uint32_t g = bid * blockDim.x + tid; g < total; g += step) {
        const int ji = zt_find_job(jobs, g);
        const zt_job J = job_at(jobs, ji);
        uint32_t b = g - J.first;
        int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw);
        uint32_t px[16];
        gather_block(J.src, J.w, J.h, J.stride, bx, by, px);
        if (J.fmt == 0 || J.fmt == 1) {
            if (J.fmt == 1) store8(J.dst + (uint64_t)b * 16, bc4_block(px, 3));
            uint32_t c0, c1, ix;
            const int e = bc1_start(px, c0, c1, ix);
            climb[g] = make_uint4(c0 | c1 << 16, (uint32_t)e, ix, (uint32_t)ji);
        } else {
            climb[g] = make_uint4(0, 0, 0, (uint32_t)ji);  // every entry names its job: zt_bc1_climb reads it first
            if (J.fmt == 2) store8(J.dst + (uint64_t)b * 8, bc4_block(px, 0));
            else store16(J.dst + (uint64_t)b * 16, bc4_block(px, 0), bc4_block(px, 1));
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc14_encode(zt_jobs jobs, uint32_t total, uint4 *climb) {
    bc14_encode_body(jobs, total, climb, blockIdx.x, gridDim.x);
}

// The same with two lanes per block, for launches too small to fill the GPU one lane per block: lane 0 of a pair
// runs start 0 (and BC5's R half), lane 1 start 1 (and BC3's BC4 alpha half, BC5's G half); the pair keeps start 1
// when its SSE is strictly below start 0's, as the reference's loop over the starts does.
namespace {
ZT_DEV void bc14_encode2_body(const zt_jobs &jobs, uint32_t total, uint4 *climb, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid(), q = tid & 1u;
    const uint64_t step = (uint64_t)nblk * blockDim.x, n = 2ull * total;
    for (uint64_t t = (uint64_t)bid * blockDim.x + tid; t < n; t += step) {   // both lanes of a pair take the same path
        const uint32_t g = (uint32_t)(t >> 1);
        const int ji = zt_find_job(jobs, g);
        const zt_job J = job_at(jobs, ji);
        uint32_t b = g - J.first;
        int bx = (int)(b % (uint32_t)J.bw), by = (int)(b / (uint32_t)J.bw);
        uint32_t px[16];
        gather_block(J.src, J.w, J.h, J.stride, bx, by, px);
        if (J.fmt == 0 || J.fmt == 1) {
            if (J.fmt == 1 && q) store8(J.dst + (uint64_t)b * 16, bc4_block(px, 3));
            uint32_t pm[16], c0, c1, ix;
            const int spp = bc1_prepare(px, pm);
            Bc1Axis A;
            bc1_axis(px, A);
            const int e = bc1_one_start(px, pm, spp, A, (int)q, c0, c1, ix);
            const uint32_t oc = (uint32_t)__shfl_xor((int)(c0 | c1 << 16), 1), oix = (uint32_t)__shfl_xor((int)ix, 1);
            const int oe = __shfl_xor(e, 1);
            if (!q) climb[g] = oe < e ? make_uint4(oc, (uint32_t)oe, oix, (uint32_t)ji) : make_uint4(c0 | c1 << 16, (uint32_t)e, ix, (uint32_t)ji);
        } else {
            if (!q) climb[g] = make_uint4(0, 0, 0, (uint32_t)ji);  // every entry names its job: the climb reads it first
            if (J.fmt == 2) { if (!q) store8(J.dst + (uint64_t)b * 8, bc4_block(px, 0)); }
            else store8(J.dst + (uint64_t)b * 16 + 8 * q, bc4_block(px, (int)q));
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc14_encode2(zt_jobs jobs, uint32_t total, uint4 *climb) {
    bc14_encode2_body(jobs, total, climb, blockIdx.x, gridDim.x);
}

// The BC1 hill climb with lane refill.  Each wave owns ranges of `range` blocks (wave-strided; a multiple of 64, at
// most ZT_CLIMB_RANGE, chosen by the host so that small jobs still spread over enough waves); a lane that has no
// block (or has finished one) takes the next block of the range, numbered by a ballot prefix count, so every step
// runs one trial on (nearly) every lane.  Non-BC1 blocks are skipped at refill.
// Refills wait until at least refill_min lanes are idle (or none is busy): a refill costs every lane of the wave
// a job lookup and a 16-pixel gather, so doing it for one lane per step would cost more than the idle lanes.
// Trials compute errors only.  A block the climb never moved is stored with its start's indices; a moved block
// leaves (c0 | c1 << 16, SSE, 0, job | 2^31) in climb[] for zt_bc1_finish, which computes its indices once.
namespace {
ZT_DEV void bc1_climb_body(const zt_jobs &jobs, uint32_t total, uint4 *climb, int refill_min, uint32_t range, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const int lane = tid & 63;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t wave = (bid * blockDim.x + tid) >> 6, nwaves = (nblk * blockDim.x) >> 6;
    for (uint64_t r0 = (uint64_t)wave * range; r0 < total; r0 += (uint64_t)nwaves * range) {
        const uint32_t end = (uint32_t)min<uint64_t>(total, r0 + range);
        uint32_t next = (uint32_t)r0;  // same on every lane
        bool active = false;
        uint32_t px[16];  // alpha cleared (bc1_prepare)
        int spp = 0;
        Climb s = { 0, 0, 0, 0, 0, 0, 0, 0, 11, -1 };
        uint8_t *dst = nullptr;
        uint32_t cur = 0, job = 0;
        for (;;) {
            const uint64_t need = __ballot(!active);
            const int idle = __popcll(need);
            const bool refill = idle >= refill_min || idle == 64;
            const uint32_t mine = next + (uint32_t)__popcll(need & below);
            if (refill) next += (uint32_t)idle;
            if (refill && !active && mine < end) {
                const uint4 st = climb[mine];
                const zt_job J = job_at(jobs, (int)st.w);
                if (J.fmt == 0 || J.fmt == 1) {
                    const uint32_t b = mine - J.first;
                    dst = J.dst + (J.fmt == 0 ? (uint64_t)b * 8 : (uint64_t)b * 16 + 8);
                    s.b0 = st.x & 0xFFFFu; s.b1 = st.x >> 16; s.best = (int)st.y; s.bidx = st.z; s.pass = 0; s.mv = 0; s.improved = 0; s.moved = 0;
                    s.stop = 11; s.last = -1;
                    cur = mine; job = st.w;
                    gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
                    spp = bc1_prepare(px, px);
                    active = true;
                }
            }
            if (__ballot(active) == 0 && next >= end) break;
            uint32_t n0 = 0, n1 = 0;
            const bool has = active && climb_next(s, n0, n1);
            if (active && !has) {
                if (s.moved) climb[cur] = make_uint4(s.b0 | s.b1 << 16, (uint32_t)s.best, 0u, job | 0x80000000u);
                else store8(dst, bc1_pack(s.b0, s.b1, s.bidx));
                active = false;
            }
            if (has) {
                const int ne = bc1_attempt_err(px, spp, n0, n1);   // puts (n0, n1) in canonical order first
                climb_apply(s, n0, n1, ne);
            }
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc1_climb(zt_jobs jobs, uint32_t total, uint4 *climb, int refill_min, uint32_t range) {
    bc1_climb_body(jobs, total, climb, refill_min, range, blockIdx.x, gridDim.x);
}

// The same hill climb with four lanes per block, for images too small to fill the GPU one lane per block (a wave's
// time is then its slowest block's chain of trials).  Lane q of a block's group tries move mv + q of the current
// pass from the group's endpoints; the group commits the first of them in the reference's move order that improves
// and goes on from the move after it, or from mv + 4 when none improves.  A move tried from the endpoints the
// reference would try it from gives the reference's result: the moves before the committed one saw no improvement
// from those same endpoints, and the moves after it are tried again from the new ones.  Groups refill from their
// wave's range like the lanes of zt_bc1_climb.  When its range is done the wave writes the moved blocks' indices
// itself (zt_bc1_finish's work, with four lanes per block), so a small launch needs no finish kernel.
namespace {
template <int G>
ZT_DEV void bc1_climb_group_body(const zt_jobs &jobs, uint32_t total, uint4 *climb, uint32_t range, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const int lane = tid & 63, q = lane & (G - 1), gbase = lane & ~(G - 1);
    const uint64_t below = gbase ? (~0ull >> (64 - gbase)) : 0ull;
    const uint32_t wave = (bid * blockDim.x + tid) >> 6, nwaves = (nblk * blockDim.x) >> 6;
    for (uint64_t r0 = (uint64_t)wave * range; r0 < total; r0 += (uint64_t)nwaves * range) {
        const uint32_t end = (uint32_t)min<uint64_t>(total, r0 + range);
        uint32_t next = (uint32_t)r0;  // same on every lane
        bool active = false;
        uint32_t px[16];  // alpha cleared (bc1_prepare)
        int spp = 0, best = 0, pass = 0, mv = 0, stop = 11, last = -1;   // stop, last: as in Climb
        bool improved = false, moved = false;
        uint32_t b0 = 0, b1 = 0, bidx = 0, cur = 0, job = 0;
        uint8_t *dst = nullptr;
        for (;;) {
            // idle groups take the next blocks of the range in group order (one ballot over the groups' first lanes)
            const uint64_t need = __ballot(!active && q == 0);
            const uint32_t mine = next + (uint32_t)__popcll(need & below);
            next += (uint32_t)__popcll(need);
            if (!active && mine < end) {
                const uint4 st = climb[mine];
                const zt_job J = job_at(jobs, (int)st.w);
                if (J.fmt == 0 || J.fmt == 1) {
                    const uint32_t b = mine - J.first;
                    dst = J.dst + (J.fmt == 0 ? (uint64_t)b * 8 : (uint64_t)b * 16 + 8);
                    b0 = st.x & 0xFFFFu; b1 = st.x >> 16; best = (int)st.y; bidx = st.z; pass = 0; mv = 0; improved = false; moved = false;
                    stop = 11; last = -1;
                    cur = mine; job = st.w;
                    gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
                    spp = bc1_prepare(px, px);
                    active = true;
                }
            }
            if (__ballot(active) == 0 && next >= end) break;
            // this lane's move from the group's endpoints
            const int m = mv + q;
            uint32_t n0 = 0, n1 = 0;
            int ne = 0;
            bool better = false;
            if (active && best > 0 && m < 12) {
                const int ep = m / 6, ch = (m >> 1) % 3, d = (m & 1) ? 1 : -1;
                const int sh = ch == 0 ? 11 : ch == 1 ? 5 : 0, mx = ch == 1 ? 63 : 31;
                uint32_t c = ep ? b1 : b0;
                const int v = (int)((c >> sh) & (uint32_t)mx);
                if (v + d >= 0 && v + d <= mx) {
                    c = (c & ~((uint32_t)mx << sh)) | ((uint32_t)(v + d) << sh);
                    n0 = ep ? b0 : c; n1 = ep ? c : b1;
                    ne = bc1_attempt_err(px, spp, n0, n1);   // puts (n0, n1) in canonical order first
                    better = ne < best;
                }
            }
            // the group's first improving move, taken by every lane of the group
            const uint32_t gb = (uint32_t)(__ballot(better) >> gbase) & ((1u << G) - 1u);
            const int first = gb ? __ffs(gb) - 1 : 0;
            const uint32_t pair = (uint32_t)__shfl((int)(n0 | n1 << 16), gbase + first);
            const int pbest = __shfl(ne, gbase + first);
            if (active) {
                bool done = best == 0;
                if (!done) {
                    if (gb) { best = pbest; b0 = pair & 0xFFFFu; b1 = pair >> 16; improved = true; moved = true; last = mv + first; mv += first + 1; }
                    else mv += G;
                    if (best == 0) done = true;   // no later move can improve on 0
                    else if (!improved && mv > stop) done = true;   // the rest of the pass fails again (mv >= 12 too)
                    else if (mv >= 12) {
                        if (pass == 7) done = true;
                        else { pass++; mv = 0; improved = false; stop = last; }
                    }
                }
                if (done) {
                    if (q == 0) {
                        if (moved) climb[cur] = make_uint4(b0 | b1 << 16, (uint32_t)best, 0u, job | 0x80000000u);
                        else store8(dst, bc1_pack(b0, b1, bidx));
                    }
                    active = false;
                }
            }
        }
        // the range's moved blocks take their indices here, as zt_bc1_finish would: four lanes per block, one row of
        // four pixels each, the rows' index bits OR-ed over the quad.  The climb state they read was stored by this wave.
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        for (uint32_t c = (uint32_t)r0; c < end; c += 16) {
            const uint32_t g = c + (uint32_t)(lane >> 2);
            const int row = lane & 3;
            const uint4 st = g < end ? climb[g] : make_uint4(0u, 0u, 0u, 0u);
            const bool fin = g < end && (st.w >> 31);
            uint32_t bits = 0, c0 = st.x & 0xFFFFu, c1 = st.x >> 16;
            const zt_job J = job_at(jobs, fin ? (int)(st.w & 0x7FFFFFFFu) : 0);
            const uint32_t b = g - (fin ? J.first : 0u);
            if (fin) bits = bc1_row_indices(J, b, row, c0, c1);
            bits |= (uint32_t)__shfl_xor((int)bits, 1);
            bits |= (uint32_t)__shfl_xor((int)bits, 2);
            if (fin && row == 0) store8(J.dst + (J.fmt == 0 ? (uint64_t)b * 8 : (uint64_t)b * 16 + 8), bc1_pack(c0, c1, bits));
        }
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc1_climb4(zt_jobs jobs, uint32_t total, uint4 *climb, uint32_t range) {
    bc1_climb_group_body<4>(jobs, total, climb, range, blockIdx.x, gridDim.x);
}
extern "C" __global__ void __launch_bounds__(256) zt_bc1_climb8(zt_jobs jobs, uint32_t total, uint4 *climb, uint32_t range) {
    bc1_climb_group_body<8>(jobs, total, climb, range, blockIdx.x, gridDim.x);
}

// the indices of every block the hill climb moved (climb[g].w has bit 31 set): one full trial of its final pair
namespace {
ZT_DEV void bc1_finish_body(const zt_jobs &jobs, uint32_t total, const uint4 *climb, uint32_t bid, uint32_t nblk) {
    const uint32_t tid = zt_tid();
    const uint32_t step = nblk * blockDim.x;
    for (uint32_t g = bid * blockDim.x + tid; g < total; g += step) {
        const uint4 st = climb[g];
        if (!(st.w >> 31)) continue;
        const zt_job J = job_at(jobs, (int)(st.w & 0x7FFFFFFFu));
        const uint32_t b = g - J.first;
        uint32_t px[16];
        gather_block(J.src, J.w, J.h, J.stride, (int)(b % (uint32_t)J.bw), (int)(b / (uint32_t)J.bw), px);
        const int spp = bc1_prepare(px, px);
        uint32_t c0 = st.x & 0xFFFFu, c1 = st.x >> 16, idx = 0;
        (void)bc1_trial(px, spp, c0, c1, idx);   // (c0, c1) is canonical already
        store8(J.dst + (J.fmt == 0 ? (uint64_t)b * 8 : (uint64_t)b * 16 + 8), bc1_pack(c0, c1, idx));
    }
}
}  // namespace
extern "C" __global__ void __launch_bounds__(256) zt_bc1_finish(zt_jobs jobs, uint32_t total, const uint4 *climb) {
    bc1_finish_body(jobs, total, climb, blockIdx.x, gridDim.x);
}
