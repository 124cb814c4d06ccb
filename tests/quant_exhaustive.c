/* Exhaustive check of the kernels' BC7 endpoint quantisers (zap_amd/csrc/zt_common.h: div255, also used by the
 * BC1 565 quantiser; zt_bc7_core.h: quant_pbit1,
 * quant_plain) against the reference's arithmetic (zap_tex.h:426-438 zap__bc7q, :613-624 zap__bc7_qn).
 *   ./quant_exhaustive [stride] [limit]
 * checks every stride-th float bit pattern: div255 for all finite floats, the quantisers for |v| <= limit
 * (default 2048; the encoders' endpoint values stay within [-1020, 1275]).  Exit status 1 on any mismatch.
 * Build: gcc -O2 -ffp-contract=off [-mfma] [-fopenmp] quant_exhaustive.c -lm */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static int x86i(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000u; }
static int bc7x(int v, int n) { v <<= 8 - n; return v | v >> n; }
static int clampi(int a, int lo, int hi) { return a < lo ? lo : a > hi ? hi : a; }
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

/* ---- the reference's arithmetic */
static int ref_q(float v, int n, int p, int *q) {
    int mx = (1 << (n - 1)) - 1, best = 0, be = 1 << 30, g = x86i((v / 255.0f * (float)((1 << n) - 1) - (float)p) / 2.0f + 0.5f);
    for (long t = (long)g - 1; t <= (long)g + 1; t++) {
        int tq = t < 0 ? 0 : t > mx ? mx : (int)t, x = bc7x(tq << 1 | p, n);
        float d = (float)x - v;
        if (x86i(d * d) < be) { be = x86i(d * d); best = tq; }
    }
    *q = best;
    return be;
}
static int ref_qn(float v, int n, int *q) {
    int mx = (1 << n) - 1, g = x86i(v * (float)mx / 255.0f + 0.5f), bq = 0, bd = 1 << 30;
    for (long t = (long)g - 1; t <= (long)g + 1; t++) {
        int tq = t < 0 ? 0 : t > mx ? mx : (int)t, x = bc7x(tq, n);
        int d = x86i((float)x - v < 0 ? v - (float)x : (float)x - v);
        if (d < bd) { bd = d; bq = tq; }
    }
    *q = bq;
    return bd;
}

/* ---- the kernels' formulation, restated operation for operation */
static float div255(float v) { const float R = 1.0f / 255.0f, q0 = v * R; return fmaf(fmaf(-q0, 255.0f, v), R, q0); }
static int ker_q(float v, float t, int n, int p, int *q) {
    const int mx = (1 << (n - 1)) - 1, g = clampi((int)((t - (float)p) * 0.5f + 0.5f), 1, mx - 1), vm = g << 1 | p;
    uint32_t k[3];
    if (n == 8) {
        const float xm = (float)vm, xs[3] = { xm - 2.0f, xm, xm + 2.0f };
        for (int dt = 0; dt < 3; dt++) { const float d = xs[dt] - v; k[dt] = (uint32_t)(int)(d * d) << 2 | (uint32_t)dt; }
    } else {
        for (int dt = 0; dt < 3; dt++) { const float d = (float)bc7x(vm + 2 * (dt - 1), n) - v; k[dt] = (uint32_t)(int)(d * d) << 2 | (uint32_t)dt; }
    }
    const uint32_t best = umin(umin(k[0], k[1]), k[2]);
    *q = g + (int)(best & 3u) - 1;
    return (int)(best >> 2);
}
static int ker_qn(float v, int n, int *q) {
    const int mx = (1 << n) - 1, g = clampi((int)(div255(v * (float)mx) + 0.5f), 1, mx - 1);
    uint32_t k[3];
    if (n == 8) {
        const float xm = (float)g, xs[3] = { xm - 1.0f, xm, xm + 1.0f };
        for (int dt = 0; dt < 3; dt++) k[dt] = (uint32_t)(int)fabsf(xs[dt] - v) << 2 | (uint32_t)dt;
    } else {
        for (int dt = 0; dt < 3; dt++) k[dt] = (uint32_t)(int)fabsf((float)bc7x(g + dt - 1, n) - v) << 2 | (uint32_t)dt;
    }
    const uint32_t best = umin(umin(k[0], k[1]), k[2]);
    *q = g + (int)(best & 3u) - 1;
    return (int)(best >> 2);
}

int main(int argc, char **argv) {
    const long stride = argc > 1 ? atol(argv[1]) : 1;
    const float lim = argc > 2 ? (float)atof(argv[2]) : 2048.0f;
    long bad_div = 0, bad_q = 0, bad_qn = 0, n_div = 0, n_q = 0;
#pragma omp parallel for reduction(+ : bad_div, bad_q, bad_qn, n_div, n_q) schedule(dynamic, 4096)
    for (long i = 0; i < (1L << 32); i += stride) {
        const float v = f_of((uint32_t)i);
        if (!isfinite(v)) continue;
        n_div++;
        /* -0 gives +0; the quantisers add 0.5 or subtract the p-bit next, which erases the sign of a zero */
        if (v != 0.0f && u_of(div255(v)) != u_of(v / 255.0f)) bad_div++;
        if (!(fabsf(v) <= lim)) continue;
        n_q++;
        for (int n = 5; n <= 8; n++) {
            const float t = div255(v) * (float)((1 << n) - 1);
            for (int p = 0; p < 2; p++) {
                int q0, q1, e0 = ref_q(v, n, p, &q0), e1 = ker_q(v, t, n, p, &q1);
                if (e0 != e1 || q0 != q1) bad_q++;
            }
        }
        for (int n = 4; n <= 8; n++) {
            int q0, q1, e0 = ref_qn(v, n, &q0), e1 = ker_qn(v, n, &q1);
            if (e0 != e1 || q0 != q1) bad_qn++;
        }
    }
    printf("div255: %ld floats, %ld mismatches; quantisers: %ld floats, p-bit %ld, plain %ld mismatches\n", n_div, bad_div, n_q, bad_q, bad_qn);
    return (bad_div || bad_q || bad_qn) ? 1 : 0;
}
