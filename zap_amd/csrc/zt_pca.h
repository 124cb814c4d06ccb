// zt_pca.h — principal-axis endpoints of a pixel subset (zap_tex.h:441-462), shared by the BC7 and ASTC
// kernels.  Lane-serial in pixel order; FULL = all 16 pixels (the reference divides by (float)16).
#pragma once
#include "zt_common.h"

namespace {
// `sums` (optional, FULL only): the integer channel sums.  Each term p/16 and each partial sum of the reference's
// mean is then exact (multiples of 1/16 below 256), so the mean is exactly sum * (1/16).
template <int NC, bool FULL>
ZT_DEV void pca(const uint32_t px[16], uint32_t mask, int cnt, float lo[4], float hi[4], const int *sums = nullptr) {
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
