// zt_pca.h — principal-axis endpoints of a pixel subset (zap_tex.h:441-462), shared by the BC7 and ASTC
// kernels.  Lane-serial in pixel order; FULL = all 16 pixels (the reference divides by (float)16).
#pragma once
#include "zt_common.h"

namespace {
// Marks the 16 pixel registers as rewritten (an empty asm that claims to modify them).  The optimiser then cannot
// reuse a byte -> float conversion made before this point after it, so each pass over the pixels converts again (one
// VALU op each) instead of keeping 16 x NC floats live across passes.
ZT_DEV void refresh16(uint32_t *p) {
    asm volatile("" : "+v"(p[0]), "+v"(p[1]), "+v"(p[2]), "+v"(p[3]), "+v"(p[4]), "+v"(p[5]), "+v"(p[6]), "+v"(p[7]),
                 "+v"(p[8]), "+v"(p[9]), "+v"(p[10]), "+v"(p[11]), "+v"(p[12]), "+v"(p[13]), "+v"(p[14]), "+v"(p[15]));
}

// `sums` (optional, FULL only): the integer channel sums.  Each term p/16 and each partial sum of the reference's
// mean is then exact (multiples of 1/16 below 256), so the mean is exactly sum * (1/16).
// REFRESH: refresh16 the (then non-const) pixels before the covariance and projection passes.
template <int NC, bool FULL, bool REFRESH = false, typename PX>
ZT_DEV void pca(PX *px, uint32_t mask, int cnt, float lo[4], float hi[4], const int *sums = nullptr) {
    float mean[NC], cv[NC][NC], ax[NC];
    const float fc = (float)cnt;
#pragma unroll
    for (int c = 0; c < NC; c++) { mean[c] = 0.0f; ax[c] = 1.0f; }
#pragma unroll
    for (int a = 0; a < NC; a++)
#pragma unroll
        for (int b = 0; b < NC; b++) cv[a][b] = 0.0f;
    if (FULL && sums) {
#pragma unroll
        for (int c = 0; c < NC; c++) mean[c] = (float)sums[c] * 0.0625f;
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (!FULL && !((mask >> k) & 1)) continue;
#pragma unroll
            for (int c = 0; c < NC; c++) mean[c] += FULL ? chanf(px[k], c) / 16.0f : chanf(px[k], c) / fc;
        }
    }
    if constexpr (REFRESH) refresh16(px);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (!FULL && !((mask >> k) & 1)) continue;
        float d[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) d[c] = chanf(px[k], c) - mean[c];
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
            for (int b = a; b < NC; b++) cv[a][b] += d[a] * d[b];
    }
#pragma unroll
    for (int a = 0; a < NC; a++)
#pragma unroll
        for (int b = 0; b < a; b++) cv[a][b] = cv[b][a];
    for (int it = 0; it < 6; it++) {
        float nv[NC], m = 0.0f;
#pragma unroll
        for (int a = 0; a < NC; a++) {
            nv[a] = 0.0f;
#pragma unroll
            for (int b = 0; b < NC; b++) nv[a] += cv[a][b] * ax[b];
            float av = nv[a] < 0 ? -nv[a] : nv[a];
            if (av > m) m = av;
        }
        if (m < 1e-6f) break;
#pragma unroll
        for (int a = 0; a < NC; a++) ax[a] = nv[a] / m;
    }
    if constexpr (REFRESH) refresh16(px);
    float n2 = 0.0f, tl = 1e9f, th = -1e9f;
#pragma unroll
    for (int a = 0; a < NC; a++) n2 += ax[a] * ax[a];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (!FULL && !((mask >> k) & 1)) continue;
        float t = 0.0f;
#pragma unroll
        for (int c = 0; c < NC; c++) t += (chanf(px[k], c) - mean[c]) * ax[c];
        if (t < tl) tl = t;
        if (t > th) th = t;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c < NC) { lo[c] = mean[c] + ax[c] * tl / n2; hi[c] = mean[c] + ax[c] * th / n2; }
        else { lo[c] = 255.0f; hi[c] = 255.0f; }
    }
}

}  // namespace
