/* zt_oracle.c — CPU restatement of zap_tex.h (WatchDogStudios/zap 1.3.0): the parity oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is compared against and the timed
 * CPU baseline; it is never linked into, loaded by or called from the product library (zap_amd/).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function restates the reference algorithm in its exact floating-point operation order (the
 * reference is pinned with FMA contraction off; build this file with -std=c11 -O2 -ffp-contract=off).
 * Citations are zap_tex.h:line of the reference.  Pinning of each part:
 *   - whole-encoder outputs vs the compiled reference (oracle/_ref) on seeded and adversarial inputs, and vs
 *     committed SHA-256 golden fixtures generated from that build (tests/golden/);
 *   - reference known-answer tests (tex_bench.c:45-179) re-run against this file (tests/test_oracle_kats.py);
 *   - closed forms: mip (sum+2)>>2, BC4 exhaustive search, decoder/encoder SSE identity for BC7 modes 5/1.
 * With -DZTO_COUNT every stage adds its algorithmic op count (arithmetic, compare/select and convert ops,
 * one each; loads, stores and branches zero) to a thread-local counter, for the VALU roofline denominator.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ZTO_COUNT
static _Thread_local unsigned long long g_ops, g_stage[4];
#define OPS(n) (g_ops += (unsigned long long)(n))
#define STAGE_BEGIN() unsigned long long stage_t0_ = g_ops
#define STAGE_END(k) (g_stage[k] += g_ops - stage_t0_, stage_t0_ = g_ops)
unsigned long long zto_ops_get(void) { return g_ops; }
void zto_ops_reset(void) { g_ops = 0; g_stage[0] = g_stage[1] = g_stage[2] = g_stage[3] = 0; }
/* BC7 ops by search stage: [0] mode 6 + opaque test, [1] mode 5, [2] mode-1 ranking + fits, [3] blocks reaching mode 1 */
void zto_stage_ops(unsigned long long out[4]) { for (int k = 0; k < 4; k++) out[k] = g_stage[k]; }
/* Unclamped (int) casts, per site (0: BC1 zap__to565, 1: BC6H zap__bc6_q, 2: ASTC zap__astc_quant): the largest
 * |argument| that fits int32 and the number that do not (NaN included); and least-squares steps that pass the
 * reference's |det| >= 1e-6 guard with a nearly singular system (0: BC1 |det| < 1e-2, 1: ASTC |det| < 1e-4). */
static _Thread_local double g_cast_max[3];
static _Thread_local unsigned long long g_cast_bad[3], g_ls_near[2];
#define CAST_NOTE(site, f)                                                                 \
    do {                                                                                   \
        double a_ = fabs((double)(f));                                                     \
        if (a_ != a_ || a_ >= 2147483648.0) g_cast_bad[site]++;                            \
        else if (a_ > g_cast_max[site]) g_cast_max[site] = a_;                             \
    } while (0)
#define LS_NEAR(k, det, lim) ((det) < (lim) && (det) > -(lim) ? (void)g_ls_near[k]++ : (void)0)
void zto_cast_stats(double maxabs[3], unsigned long long bad[3], unsigned long long near[2]) {
    for (int k = 0; k < 3; k++) { maxabs[k] = g_cast_max[k]; bad[k] = g_cast_bad[k]; }
    near[0] = g_ls_near[0]; near[1] = g_ls_near[1];
}
void zto_cast_reset(void) { for (int k = 0; k < 3; k++) { g_cast_max[k] = 0; g_cast_bad[k] = 0; } g_ls_near[0] = g_ls_near[1] = 0; }
#else
#define CAST_NOTE(site, f) ((void)0)
#define LS_NEAR(k, det, lim) ((void)0)
#define OPS(n) ((void)0)
#define STAGE_BEGIN() ((void)0)
#define STAGE_END(k) ((void)0)
#endif

typedef uint8_t px4[4];

/* x86 cvttss2si semantics for (int)f: a value whose truncation leaves int32 (or NaN) gives INT_MIN.
 * The reference casts unclamped least-squares results (zap_tex.h:46, :842, :1187); this pins that case. */
static int cvt_i32(float f, int site) {
    CAST_NOTE(site, f);
    (void)site;
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int)f;
    return INT_MIN;
}

static uint64_t ld_le(const uint8_t *p, int n) { uint64_t v = 0; for (int i = n - 1; i >= 0; i--) v = v << 8 | p[i]; return v; }
static void st_le(uint8_t *p, uint64_t v, int n) { for (int i = 0; i < n; i++) { p[i] = (uint8_t)v; v >>= 8; } }

static unsigned getbits(const uint8_t *b, int *pos, int n) { /* zap_tex.h:367-371, LSB first */
    unsigned v = 0;
    for (int i = 0; i < n; i++, (*pos)++) v |= (unsigned)(b[*pos >> 3] >> (*pos & 7) & 1) << i;
    return v;
}
static void putbits(uint8_t *b, int *pos, unsigned v, int n) { /* zap_tex.h:372-374 */
    for (int i = 0; i < n; i++, (*pos)++) if (v >> i & 1) b[*pos >> 3] |= (uint8_t)(1 << (*pos & 7));
}

/* ======================================================================= BC1 (zap_tex.h:39-168) */

static void unpack565(unsigned c, int rgb[3]) { /* zap_tex.h:41-44 */
    int R = (c >> 11) & 31, G = (c >> 5) & 63, B = c & 31;
    rgb[0] = (R << 3) | (R >> 2); rgb[1] = (G << 2) | (G >> 4); rgb[2] = (B << 3) | (B >> 2);
}
static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }
static uint16_t pack565(float r, float g, float b) { /* zap_tex.h:45-49 */
    int R = cvt_i32(r * 31.0f / 255.0f + 0.5f, 0), G = cvt_i32(g * 63.0f / 255.0f + 0.5f, 0), B = cvt_i32(b * 31.0f / 255.0f + 0.5f, 0);
    OPS(3 * 4 + 3 * 2 + 2);
    return (uint16_t)(clampi(R, 0, 31) << 11 | clampi(G, 0, 63) << 5 | clampi(B, 0, 31));
}

/* palette (zap_tex.h:52-62); returns the entry count */
static int bc1_palette(uint16_t c0, uint16_t c1, int force4, int pal[4][3]) {
    unpack565(c0, pal[0]); unpack565(c1, pal[1]);
    int four = force4 || c0 > c1;
    for (int k = 0; k < 3; k++) {
        int a = pal[0][k], b = pal[1][k];
        pal[2][k] = four ? (2 * a + b + 1) / 3 : (a + b + 1) / 2;
        pal[3][k] = four ? (a + 2 * b + 1) / 3 : 0;
    }
    OPS(24);
    return four ? 4 : 3;
}
static int sse3(const uint8_t *p, const int *q) { int a = p[0] - q[0], b = p[1] - q[1], c = p[2] - q[2]; return a * a + b * b + c * c; }

static int bc1_indices(const px4 *px, uint16_t c0, uint16_t c1, int force4, uint32_t *idx) { /* zap_tex.h:70-80 */
    int pal[4][3], n = bc1_palette(c0, c1, force4, pal), err = 0;
    uint32_t bits = 0;
    for (int i = 0; i < 16; i++) {
        int bi = 0, be = sse3(px[i], pal[0]);
        for (int k = 1; k < n; k++) { int e = sse3(px[i], pal[k]); if (e < be) { be = e; bi = k; } }
        bits |= (uint32_t)bi << (2 * i); err += be;
    }
    OPS(16 * (n * 8 + (n - 1) * 3 + 2));
    *idx = bits;
    return err;
}
static int bc1_score(const px4 *px, uint16_t c0, uint16_t c1, int force4, uint32_t idx) { /* zap_tex.h:82-87 */
    int pal[4][3], err = 0;
    bc1_palette(c0, c1, force4, pal);
    for (int i = 0; i < 16; i++) err += sse3(px[i], pal[(idx >> (2 * i)) & 3]);
    OPS(16 * 9);
    return err;
}
static int bc1_attempt(const px4 *px, uint16_t *c0, uint16_t *c1, uint32_t *idx) { /* zap_tex.h:90-94 */
    if (*c0 < *c1) { uint16_t t = *c0; *c0 = *c1; *c1 = t; }
    OPS(3);
    if (*c0 == *c1) { *idx = 0; return bc1_score(px, *c0, *c1, 1, 0); }
    return bc1_indices(px, *c0, *c1, 1, idx);
}

/* zap_tex.h:97-168: PCA by 6 power iterations, 2 starts with LS refinement, 565 hill climb */
static int bc1_block(const px4 *px, uint16_t *oc0, uint16_t *oc1, uint32_t *oidx) {
    float mean[3] = { 0, 0, 0 }, cv[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < 16; i++) for (int k = 0; k < 3; k++) mean[k] += px[i][k] / 16.0f;
    for (int i = 0; i < 16; i++) {
        float r = px[i][0] - mean[0], g = px[i][1] - mean[1], b = px[i][2] - mean[2];
        cv[0] += r * r; cv[1] += r * g; cv[2] += r * b; cv[3] += g * g; cv[4] += g * b; cv[5] += b * b;
    }
    OPS(16 * 3 * 3 + 16 * (3 + 12));
    float ax[3] = { 1, 1, 1 };
    for (int it = 0; it < 6; it++) {
        float x = cv[0] * ax[0] + cv[1] * ax[1] + cv[2] * ax[2];
        float y = cv[1] * ax[0] + cv[3] * ax[1] + cv[4] * ax[2];
        float z = cv[2] * ax[0] + cv[4] * ax[1] + cv[5] * ax[2];
        float m = x < 0 ? -x : x, my = y < 0 ? -y : y, mz = z < 0 ? -z : z;
        if (my > m) m = my;
        if (mz > m) m = mz;
        OPS(15 + 9 + 1);
        if (m < 1e-6f) break;
        ax[0] = x / m; ax[1] = y / m; ax[2] = z / m;
        OPS(3);
    }
    float n2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], lo = 1e9f, hi = -1e9f;
    for (int i = 0; i < 16; i++) {
        float t = (px[i][0] - mean[0]) * ax[0] + (px[i][1] - mean[1]) * ax[1] + (px[i][2] - mean[2]) * ax[2];
        if (t < lo) lo = t;
        if (t > hi) hi = t;
    }
    OPS(5 + 16 * (3 + 3 + 3 + 2 + 4));
    uint16_t b0 = 0, b1 = 0;
    uint32_t bidx = 0;
    int best = 1 << 30;
    for (int s = 0; s < 2; s++) {
        float in = s ? 0.0f : (hi - lo) / 16.0f, l = (lo + in) / n2, u = (hi - in) / n2;
        uint16_t c0 = pack565(mean[0] + ax[0] * u, mean[1] + ax[1] * u, mean[2] + ax[2] * u);
        uint16_t c1 = pack565(mean[0] + ax[0] * l, mean[1] + ax[1] * l, mean[2] + ax[2] * l);
        OPS(6 + 12);
        uint32_t idx = 0;
        int e = bc1_attempt(px, &c0, &c1, &idx);
        for (int it = 0; it < 3 && e > 0; it++) {
            static const float W[4] = { 1.0f, 0.0f, 2.0f / 3.0f, 1.0f / 3.0f };
            float aa = 0, ab = 0, bb = 0, sa[3] = { 0, 0, 0 }, sb[3] = { 0, 0, 0 };
            for (int i = 0; i < 16; i++) {
                float w = W[(idx >> (2 * i)) & 3], v = 1.0f - w;
                aa += w * w; ab += w * v; bb += v * v;
                for (int k = 0; k < 3; k++) { sa[k] += w * px[i][k]; sb[k] += v * px[i][k]; }
            }
            float det = aa * bb - ab * ab;
            OPS(16 * (1 + 6 + 3 + 12) + 3 + 2);
            if (det < 1e-6f && det > -1e-6f) break;
            LS_NEAR(0, det, 1e-2f);
            float A[3], B[3];
            for (int k = 0; k < 3; k++) { A[k] = (sa[k] * bb - sb[k] * ab) / det; B[k] = (sb[k] * aa - sa[k] * ab) / det; }
            OPS(3 * 8);
            uint16_t n0 = pack565(A[0], A[1], A[2]), n1 = pack565(B[0], B[1], B[2]);
            uint32_t ni = 0;
            int ne = bc1_attempt(px, &n0, &n1, &ni);
            OPS(1);
            if (ne >= e) break;
            e = ne; c0 = n0; c1 = n1; idx = ni;
        }
        OPS(1);
        if (e < best) { best = e; b0 = c0; b1 = c1; bidx = idx; }
    }
    static const int SH[3] = { 11, 5, 0 }, MX[3] = { 31, 63, 31 };
    for (int pass = 0, improved = 1; improved && pass < 8 && best > 0; pass++) {
        improved = 0;
        for (int ep = 0; ep < 2; ep++) for (int ch = 0; ch < 3; ch++) for (int d = -1; d <= 1; d += 2) {
            uint16_t c = ep ? b1 : b0;
            int v = (c >> SH[ch]) & MX[ch];
            OPS(4);
            if (v + d < 0 || v + d > MX[ch]) continue;
            c = (uint16_t)((c & ~(MX[ch] << SH[ch])) | ((v + d) << SH[ch]));
            uint16_t n0 = ep ? b0 : c, n1 = ep ? c : b1;
            uint32_t ni = 0;
            int ne = bc1_attempt(px, &n0, &n1, &ni);
            OPS(6);
            if (ne < best) { best = ne; b0 = n0; b1 = n1; bidx = ni; improved = 1; }
        }
    }
    *oc0 = b0; *oc1 = b1; *oidx = bidx;
    return best;
}

/* ======================================================================= BC4 (zap_tex.h:170-213) */

static void bc4_palette(int a0, int a1, int pal[8]) {
    pal[0] = a0; pal[1] = a1;
    if (a0 > a1) for (int i = 2; i < 8; i++) pal[i] = ((8 - i) * a0 + (i - 1) * a1 + 3) / 7;
    else { for (int i = 2; i < 6; i++) pal[i] = ((6 - i) * a0 + (i - 1) * a1 + 2) / 5; pal[6] = 0; pal[7] = 255; }
    OPS(6 * 5);
}
static int bc4_indices(const uint8_t v[16], int a0, int a1, uint64_t *idx) {
    int pal[8], err = 0;
    uint64_t bits = 0;
    bc4_palette(a0, a1, pal);
    for (int i = 0; i < 16; i++) {
        int bi = 0, be = (v[i] - pal[0]) * (v[i] - pal[0]);
        for (int k = 1; k < 8; k++) { int e = (v[i] - pal[k]) * (v[i] - pal[k]); if (e < be) { be = e; bi = k; } }
        bits |= (uint64_t)bi << (3 * i); err += be;
    }
    OPS(16 * (8 * 2 + 7 * 3 + 2));
    *idx = bits;
    return err;
}
static int bc4_score(const uint8_t v[16], int a0, int a1, uint64_t idx) {
    int pal[8], err = 0;
    bc4_palette(a0, a1, pal);
    for (int i = 0; i < 16; i++) { int d = v[i] - pal[(idx >> (3 * i)) & 7]; err += d * d; }
    OPS(16 * 3);
    return err;
}
static int bc4_block(const uint8_t v[16], uint8_t *oa0, uint8_t *oa1, uint64_t *oidx) { /* zap_tex.h:198-213 */
    int lo = 255, hi = 0;
    for (int i = 0; i < 16; i++) { if (v[i] < lo) lo = v[i]; if (v[i] > hi) hi = v[i]; }
    OPS(16 * 4);
    int best = 1 << 30, b0 = hi, b1 = lo;
    uint64_t bi = 0;
    for (int d0 = 0; d0 <= 2; d0++) for (int d1 = 0; d1 <= 2; d1++) {
        int a0 = hi - d0, a1 = lo + d1;
        if (a0 <= a1) { a0 = hi; a1 = lo; }
        uint64_t idx = 0;
        int e = a0 == a1 ? bc4_score(v, a0, a1, 0) : bc4_indices(v, a0, a1, &idx);
        if (a0 == a1) idx = 0;
        OPS(6);
        if (e < best) { best = e; b0 = a0; b1 = a1; bi = idx; }
    }
    *oa0 = (uint8_t)b0; *oa1 = (uint8_t)b1; *oidx = bi;
    return best;
}

static void put_bc1(uint8_t *o, uint16_t c0, uint16_t c1, uint32_t idx) { st_le(o, c0, 2); st_le(o + 2, c1, 2); st_le(o + 4, idx, 4); }
static void put_bc4(uint8_t *o, uint8_t a0, uint8_t a1, uint64_t idx) { o[0] = a0; o[1] = a1; st_le(o + 2, idx, 6); }

/* ======================================================================= BC7 (zap_tex.h:308-784) */

/* per-mode layout: subsets, partition bits, rotation bits, index-selector bits, colour bits, alpha bits,
   per-endpoint p-bit, shared p-bit, index bits, secondary index bits (BC7 format specification) */
typedef struct { uint8_t ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2; } bc7_layout;
static const bc7_layout L7[8] = {
    { 3, 4, 0, 0, 4, 0, 1, 0, 3, 0 }, { 2, 6, 0, 0, 6, 0, 0, 1, 3, 0 }, { 3, 6, 0, 0, 5, 0, 0, 0, 2, 0 },
    { 2, 6, 0, 0, 7, 0, 1, 0, 2, 0 }, { 1, 0, 2, 1, 5, 6, 0, 0, 2, 3 }, { 1, 0, 2, 0, 7, 8, 0, 0, 2, 2 },
    { 1, 0, 0, 0, 7, 7, 1, 0, 4, 0 }, { 2, 6, 0, 0, 5, 5, 1, 0, 2, 0 } };

/* BC7 spec partition tables. Two subsets: bit i = pixel i in subset 1. */
static const uint16_t P2[64] = {
    0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8, 0xFF00, 0xFFF0, 0xF000,
    0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110, 0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C,
    0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696, 0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660,
    0x0272, 0x04E4, 0x4E40, 0x2720, 0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22 };
/* Three subsets, one row of 16 subset digits per partition (pixel 0 first). */
static const char *P3S[64] = {
    "0011001102212222", "0001001122112221", "0000200122112211", "0222002200110111", "0000000011221122", "0011001100220022",
    "0022002211111111", "0011001122112211", "0000000011112222", "0000111111112222", "0000111122222222", "0012001200120012",
    "0112011201120112", "0122012201220122", "0011011211221222", "0011200122002220", "0001001101121122", "0111001120012200",
    "0000112211221122", "0022002200221111", "0111011102220222", "0001000122212221", "0000001101220122", "0000110022102210",
    "0122012200110000", "0012001211222222", "0110122112210110", "0000011012211221", "0022110211020022", "0110011020022222",
    "0011012201220011", "0000200022112221", "0000000211221222", "0222002200120011", "0011001200220222", "0120012001200120",
    "0000111122220000", "0120120120120120", "0120201212010120", "0011220011220011", "0011112222000011", "0101010122222222",
    "0000000021212121", "0022112200221122", "0022001100220011", "0220122102201221", "0101222222220101", "0000212121212121",
    "0101010101012222", "0222011102220111", "0002111200021112", "0000211221122112", "0222011101110222", "0002111211120002",
    "0110011001102222", "0000000021122112", "0110011022222222", "0022001100110022", "0022112211220022", "0000000000002112",
    "0002000100020001", "0222122202221222", "0101222222222222", "0111201122012220" };
static uint32_t P3[64];
static const uint8_t A2[64] = {
    15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2, 8, 2, 2, 8, 8, 15, 2, 8, 2, 2, 8, 8, 2, 2,
    15, 15, 6, 8, 2, 8, 15, 15, 2, 8, 2, 2, 2, 15, 15, 6, 6, 2, 6, 8, 15, 15, 2, 2, 15, 15, 15, 15, 15, 2, 2, 15 };
static const uint8_t A3a[64] = {
    3, 3, 15, 15, 8, 3, 15, 15, 8, 8, 6, 6, 6, 5, 3, 3, 3, 3, 8, 15, 3, 3, 6, 10, 5, 8, 8, 6, 8, 5, 15, 15,
    8, 15, 3, 5, 6, 10, 8, 15, 15, 3, 15, 5, 15, 15, 15, 15, 3, 15, 5, 5, 5, 8, 5, 10, 5, 10, 8, 13, 15, 12, 3, 3 };
static const uint8_t A3b[64] = {
    15, 8, 8, 3, 15, 15, 3, 8, 15, 15, 15, 15, 15, 15, 15, 8, 15, 8, 15, 3, 15, 8, 15, 8, 3, 15, 6, 10, 15, 15, 10, 8,
    15, 3, 15, 10, 10, 8, 9, 10, 6, 15, 8, 15, 3, 6, 6, 8, 15, 3, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 3, 15, 15, 8 };
static const uint8_t W2[4] = { 0, 21, 43, 64 }, W3[8] = { 0, 9, 18, 27, 37, 46, 55, 64 },
                     W4[16] = { 0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64 };
static int weight(int bits, int i) { return bits == 2 ? W2[i] : bits == 3 ? W3[i] : W4[i]; }

static void p3_init(void) {
    static int done;
    if (done) return;
    for (int p = 0; p < 64; p++) { uint32_t v = 0; for (int i = 0; i < 16; i++) v |= (uint32_t)(P3S[p][i] - '0') << (2 * i); P3[p] = v; }
    done = 1;
}
static int subset_of(int ns, int part, int i) { return ns == 1 ? 0 : ns == 2 ? (P2[part] >> i) & 1 : (int)(P3[part] >> (2 * i)) & 3; }
static int anchor_of(int ns, int part, int sub) { return sub == 0 ? 0 : ns == 2 ? A2[part] : sub == 1 ? A3a[part] : A3b[part]; }
static int expand_bits(int v, int n) { v <<= 8 - n; return v | v >> n; } /* zap_tex.h:375 */

/* zap_tex.h:378-417: decoder for all 8 modes; reserved mode 8 decodes to zeros */
static void bc7_unpack(const uint8_t *b, px4 *px) {
    int mode = 0;
    p3_init();
    while (mode < 8 && !(b[0] >> mode & 1)) mode++;
    if (mode == 8) { memset(px, 0, 64); return; }
    const bc7_layout *m = &L7[mode];
    int pos = mode + 1, part = (int)getbits(b, &pos, m->pb), rot = (int)getbits(b, &pos, m->rb);
    int isb = (int)getbits(b, &pos, m->isb), ne = 2 * m->ns, ep[6][4], cb = m->cb, ab = m->ab;
    for (int c = 0; c < 3; c++) for (int e = 0; e < ne; e++) ep[e][c] = (int)getbits(b, &pos, m->cb);
    for (int e = 0; e < ne; e++) ep[e][3] = ab ? (int)getbits(b, &pos, ab) : 255;
    if (m->epb || m->spb) {
        for (int e = 0; e < ne; e++) {
            if (m->epb || !(e & 1)) {
                int p = (int)getbits(b, &pos, 1);
                for (int k = e; k <= (m->epb ? e : e + 1); k++) {
                    for (int c = 0; c < 3; c++) ep[k][c] = ep[k][c] << 1 | p;
                    if (ab) ep[k][3] = ep[k][3] << 1 | p;
                }
            }
        }
        cb++; if (ab) ab++;
    }
    for (int e = 0; e < ne; e++) {
        for (int c = 0; c < 3; c++) ep[e][c] = expand_bits(ep[e][c], cb);
        if (ab) ep[e][3] = expand_bits(ep[e][3], ab);
    }
    int i1[16], i2[16];
    for (int i = 0; i < 16; i++) {
        int anc = i == anchor_of(m->ns, part, subset_of(m->ns, part, i));
        i1[i] = (int)getbits(b, &pos, m->ib - anc);
    }
    for (int i = 0; i < 16 && m->ib2; i++) i2[i] = (int)getbits(b, &pos, m->ib2 - (i == 0));
    for (int i = 0; i < 16; i++) {
        int sb = subset_of(m->ns, part, i), *e0 = ep[2 * sb], *e1 = ep[2 * sb + 1];
        int cw = m->ib2 && isb ? weight(m->ib2, i2[i]) : weight(m->ib, i1[i]);
        int aw = !m->ib2 ? cw : isb ? weight(m->ib, i1[i]) : weight(m->ib2, i2[i]);
        for (int c = 0; c < 3; c++) px[i][c] = (uint8_t)(((64 - cw) * e0[c] + cw * e1[c] + 32) >> 6);
        px[i][3] = (uint8_t)(((64 - aw) * e0[3] + aw * e1[3] + 32) >> 6);
        if (rot) { uint8_t t = px[i][3]; px[i][3] = px[i][rot - 1]; px[i][rot - 1] = t; }
    }
}

static int sse_n(const uint8_t *a, const int *b, int nc) { int e = 0; for (int c = 0; c < nc; c++) { int d = a[c] - b[c]; e += d * d; } return e; }

/* zap_tex.h:426-438: quantise to n bits whose low bit is the p-bit; returns summed squared error */
static int quant_pbit(const float *v, int nc, int n, int p, int *q, int *out) {
    int err = 0, mx = (1 << (n - 1)) - 1;
    for (int c = 0; c < nc; c++) {
        int bq = 0, be = 1 << 30, g = (int)((v[c] / 255.0f * (float)((1 << n) - 1) - (float)p) / 2.0f + 0.5f);
        for (int t = g - 1; t <= g + 1; t++) {
            int tq = t < 0 ? 0 : t > mx ? mx : t, x = expand_bits(tq << 1 | p, n);
            float d = (float)x - v[c];
            if ((int)(d * d) < be) { be = (int)(d * d); bq = tq; }
        }
        q[c] = bq; out[c] = expand_bits(bq << 1 | p, n); err += be;
    }
    OPS(nc * (6 + 3 * 10 + 4));
    return err;
}

/* zap_tex.h:441-462: principal-axis endpoints of a pixel subset */
static void pca_endpoints(const px4 *px, const int *sel, int cnt, int nc, float lo[4], float hi[4]) {
    float mean[4] = { 0, 0, 0, 0 }, cv[4][4] = { { 0 } }, ax[4] = { 1, 1, 1, 1 };
    for (int k = 0; k < cnt; k++) for (int c = 0; c < nc; c++) mean[c] += px[sel[k]][c] / (float)cnt;
    for (int k = 0; k < cnt; k++)
        for (int a = 0; a < nc; a++) for (int b = 0; b < nc; b++) cv[a][b] += (px[sel[k]][a] - mean[a]) * (px[sel[k]][b] - mean[b]);
    OPS(cnt * nc * 3 + cnt * nc * nc * 5);
    for (int it = 0; it < 6; it++) {
        float nv[4] = { 0, 0, 0, 0 }, m = 0;
        for (int a = 0; a < nc; a++) {
            for (int b = 0; b < nc; b++) nv[a] += cv[a][b] * ax[b];
            float av = nv[a] < 0 ? -nv[a] : nv[a];
            if (av > m) m = av;
        }
        OPS(nc * nc * 2 + nc * 4 + 1);
        if (m < 1e-6f) break;
        for (int a = 0; a < nc; a++) ax[a] = nv[a] / m;
        OPS(nc);
    }
    float n2 = 0, tl = 1e9f, th = -1e9f;
    for (int a = 0; a < nc; a++) n2 += ax[a] * ax[a];
    for (int k = 0; k < cnt; k++) {
        float t = 0;
        for (int c = 0; c < nc; c++) t += (px[sel[k]][c] - mean[c]) * ax[c];
        if (t < tl) tl = t;
        if (t > th) th = t;
    }
    for (int c = 0; c < nc; c++) { lo[c] = mean[c] + ax[c] * tl / n2; hi[c] = mean[c] + ax[c] * th / n2; }
    for (int c = nc; c < 4; c++) lo[c] = hi[c] = 255;
    OPS(nc * 2 + cnt * (nc * 4 + 4) + nc * 6);
}

/* zap_tex.h:466-486: quantised endpoints (best p-bit pair) and best indices for one subset */
static int fit_subset(const px4 *px, const int *sel, int cnt, int nc, int n, int ib, int shared_p,
                      const float lo[4], const float hi[4], int q[2][4], int pb[2], int *idx) {
    int e[2][4], best = 1 << 30, bq[2][4], bp[2] = { 0, 0 };
    for (int p0 = 0; p0 < 2; p0++) for (int p1 = 0; p1 < 2; p1++) {
        if (shared_p && p0 != p1) continue;
        int tq[2][4];
        int err = quant_pbit(lo, nc, n, p0, tq[0], e[0]) + quant_pbit(hi, nc, n, p1, tq[1], e[1]);
        OPS(2);
        if (err < best) { best = err; memcpy(bq, tq, sizeof bq); bp[0] = p0; bp[1] = p1; }
    }
    memcpy(q, bq, sizeof bq); pb[0] = bp[0]; pb[1] = bp[1];
    for (int k = 0; k < 2; k++) for (int c = 0; c < nc; c++) e[k][c] = expand_bits(q[k][c] << 1 | pb[k], n);
    int pal[16][4], np = 1 << ib, err = 0;
    for (int j = 0; j < np; j++) for (int c = 0; c < nc; c++) { int w = weight(ib, j); pal[j][c] = ((64 - w) * e[0][c] + w * e[1][c] + 32) >> 6; }
    for (int k = 0; k < cnt; k++) {
        int bi = 0, be = 1 << 30;
        for (int j = 0; j < np; j++) { int d = sse_n(px[sel[k]], pal[j], nc); if (d < be) { be = d; bi = j; } }
        idx[k] = bi; err += be;
    }
    OPS(2 * nc * 4 + np * nc * 6 + cnt * (np * (3 * nc + 3) + 1));
    return err;
}

/* zap_tex.h:489-515: PCA start, fit, then at most 2 least-squares refinements */
static int encode_subset(const px4 *px, const int *sel, int cnt, int nc, int n, int ib, int shared_p, int q[2][4], int pb[2], int *idx) {
    float lo[4], hi[4];
    pca_endpoints(px, sel, cnt, nc, lo, hi);
    int err = fit_subset(px, sel, cnt, nc, n, ib, shared_p, lo, hi, q, pb, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float aa = 0, ab = 0, bb = 0, sa[4] = { 0, 0, 0, 0 }, sb[4] = { 0, 0, 0, 0 };
        for (int k = 0; k < cnt; k++) {
            float w = weight(ib, idx[k]) / 64.0f, v = 1 - w;
            aa += v * v; ab += v * w; bb += w * w;
            for (int c = 0; c < nc; c++) { sa[c] += v * px[sel[k]][c]; sb[c] += w * px[sel[k]][c]; }
        }
        float det = aa * bb - ab * ab;
        OPS(cnt * (3 + 6 + nc * 5) + 3 + 2);
        if (det < 1e-6f && det > -1e-6f) break;
        float nlo[4], nhi[4];
        for (int c = 0; c < nc; c++) {
            nlo[c] = (sa[c] * bb - sb[c] * ab) / det; nhi[c] = (sb[c] * aa - sa[c] * ab) / det;
            nlo[c] = nlo[c] < 0 ? 0 : nlo[c] > 255 ? 255 : nlo[c]; nhi[c] = nhi[c] < 0 ? 0 : nhi[c] > 255 ? 255 : nhi[c];
        }
        OPS(nc * 12);
        for (int c = nc; c < 4; c++) nlo[c] = nhi[c] = 255;
        int nq[2][4], np[2], ni[16];
        int ne = fit_subset(px, sel, cnt, nc, n, ib, shared_p, nlo, nhi, nq, np, ni);
        OPS(1);
        if (ne >= err) break;
        err = ne; memcpy(q, nq, sizeof nq); pb[0] = np[0]; pb[1] = np[1]; memcpy(idx, ni, sizeof(int) * (size_t)cnt);
    }
    return err;
}

/* mode 6 (zap_tex.h:520-527, :582-593) */
static int bc7_mode6(const px4 *px, uint8_t out[16]) {
    int sel[16], q[2][4], pb[2], idx[16];
    for (int i = 0; i < 16; i++) sel[i] = i;
    int err = encode_subset(px, sel, 16, 4, 8, 4, 0, q, pb, idx);
    if (idx[0] & 8) {
        for (int c = 0; c < 4; c++) { int t = q[0][c]; q[0][c] = q[1][c]; q[1][c] = t; }
        int t = pb[0]; pb[0] = pb[1]; pb[1] = t;
        for (int i = 0; i < 16; i++) idx[i] = 15 - idx[i];
    }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 1 << 6, 7);
    for (int c = 0; c < 4; c++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)q[e][c], 7);
    putbits(out, &pos, (unsigned)pb[0], 1); putbits(out, &pos, (unsigned)pb[1], 1);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)idx[i], i ? 4 : 3);
    OPS(40);
    return err;
}

/* mode 5 helpers (zap_tex.h:597-657) */
static int fit_2bit(const px4 *r, int c0, int nc, const int *e0, const int *e1, int idx[16]) {
    int pal[4][4], err = 0;
    for (int j = 0; j < 4; j++) for (int c = 0; c < nc; c++) pal[j][c] = ((64 - W2[j]) * e0[c] + W2[j] * e1[c] + 32) >> 6;
    for (int i = 0; i < 16; i++) {
        int bi = 0, be = 1 << 30;
        for (int j = 0; j < 4; j++) {
            int d = 0;
            for (int c = 0; c < nc; c++) { int x = r[i][c0 + c] - pal[j][c]; d += x * x; }
            if (d < be) { be = d; bi = j; }
        }
        idx[i] = bi; err += be;
    }
    OPS(4 * nc * 6 + 16 * (4 * (3 * nc + 3) + 1));
    return err;
}
static void quant_plain(const float *v, int nc, int n, int *q, int *e) {
    int mx = (1 << n) - 1;
    for (int c = 0; c < nc; c++) {
        int g = (int)(v[c] * (float)mx / 255.0f + 0.5f), bq = 0, bd = 1 << 30;
        for (int t = g - 1; t <= g + 1; t++) {
            int tq = t < 0 ? 0 : t > mx ? mx : t, x = expand_bits(tq, n);
            int d = (int)((float)x - v[c] < 0 ? v[c] - (float)x : (float)x - v[c]);
            if (d < bd) { bd = d; bq = tq; }
        }
        q[c] = bq; e[c] = expand_bits(bq, n);
    }
    OPS(nc * (4 + 3 * 10 + 3));
}
static int lsq_2bit(const px4 *r, int c0, int nc, const int idx[16], float lo[4], float hi[4]) {
    float aa = 0, ab = 0, bb = 0, sa[4] = { 0, 0, 0, 0 }, sb[4] = { 0, 0, 0, 0 };
    for (int i = 0; i < 16; i++) {
        float w = W2[idx[i]] / 64.0f, v = 1 - w;
        aa += v * v; ab += v * w; bb += w * w;
        for (int c = 0; c < nc; c++) { sa[c] += v * r[i][c0 + c]; sb[c] += w * r[i][c0 + c]; }
    }
    float det = aa * bb - ab * ab;
    OPS(16 * (3 + 6 + nc * 5) + 5);
    if (det < 1e-6f && det > -1e-6f) return 0;
    for (int c = 0; c < nc; c++) {
        lo[c] = (sa[c] * bb - sb[c] * ab) / det; hi[c] = (sb[c] * aa - sa[c] * ab) / det;
        lo[c] = lo[c] < 0 ? 0 : lo[c] > 255 ? 255 : lo[c]; hi[c] = hi[c] < 0 ? 0 : hi[c] > 255 ? 255 : hi[c];
    }
    OPS(nc * 12);
    return 1;
}
static int part_mode5(const px4 *r, int c0, int nc, int n, float lo[4], float hi[4], int q[2][3], int idx[16]) {
    int e0[3], e1[3], ti[16], tq[2][3], te0[3], te1[3];
    quant_plain(lo, nc, n, q[0], e0); quant_plain(hi, nc, n, q[1], e1);
    int err = fit_2bit(r, c0, nc, e0, e1, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float nlo[4], nhi[4];
        if (!lsq_2bit(r, c0, nc, idx, nlo, nhi)) break;
        quant_plain(nlo, nc, n, tq[0], te0); quant_plain(nhi, nc, n, tq[1], te1);
        int e = fit_2bit(r, c0, nc, te0, te1, ti);
        OPS(1);
        if (e >= err) break;
        err = e; memcpy(q, tq, sizeof tq); memcpy(idx, ti, sizeof ti);
    }
    return err;
}

/* mode 5 (zap_tex.h:659-684): 4 rotations, colour 7 bits + alpha 8 bits, 2-bit index sets */
static int bc7_mode5(const px4 *px, uint8_t out[16]) {
    int best = 1 << 30, brot = 0, bcq[2][3] = { { 0 } }, baq[2][3] = { { 0 } }, bci[16] = { 0 }, bai[16] = { 0 };
    for (int rot = 0; rot < 4; rot++) {
        px4 r[16];
        int sel[16], cq[2][3], aq[2][3], ci[16], ai[16];
        memcpy(r, px, sizeof r);
        for (int i = 0; i < 16; i++) { sel[i] = i; if (rot) { uint8_t t = r[i][3]; r[i][3] = r[i][rot - 1]; r[i][rot - 1] = t; } }
        float lo[4], hi[4], alo[4] = { 255 }, ahi[4] = { 0 };
        pca_endpoints((const px4 *)r, sel, 16, 3, lo, hi);
        int err = part_mode5((const px4 *)r, 0, 3, 7, lo, hi, cq, ci);
        for (int i = 0; i < 16; i++) { if (r[i][3] < alo[0]) alo[0] = r[i][3]; if (r[i][3] > ahi[0]) ahi[0] = r[i][3]; }
        err += part_mode5((const px4 *)r, 3, 1, 8, alo, ahi, aq, ai);
        OPS(16 * 4 + 2);
        if (err < best) { best = err; brot = rot; memcpy(bcq, cq, sizeof cq); memcpy(baq, aq, sizeof aq); memcpy(bci, ci, sizeof ci); memcpy(bai, ai, sizeof ai); }
    }
    if (bci[0] & 2) { for (int c = 0; c < 3; c++) { int t = bcq[0][c]; bcq[0][c] = bcq[1][c]; bcq[1][c] = t; } for (int i = 0; i < 16; i++) bci[i] = 3 - bci[i]; }
    if (bai[0] & 2) { int t = baq[0][0]; baq[0][0] = baq[1][0]; baq[1][0] = t; for (int i = 0; i < 16; i++) bai[i] = 3 - bai[i]; }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 1 << 5, 6);
    putbits(out, &pos, (unsigned)brot, 2);
    for (int c = 0; c < 3; c++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)bcq[e][c], 7);
    for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)baq[e][0], 8);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)bci[i], i ? 2 : 1);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)bai[i], i ? 2 : 1);
    OPS(60);
    return best;
}

/* mode 1 (zap_tex.h:687-708): 2 subsets, RGB 6 bits + shared p-bit, 3-bit indices; out NULL = error only */
static int bc7_mode1(const px4 *px, int part, uint8_t *out) {
    int sel[2][16], cnt[2] = { 0, 0 }, q[2][2][4], pb[2][2], idx[2][16], err = 0, full[16];
    for (int i = 0; i < 16; i++) { int s = (P2[part] >> i) & 1; sel[s][cnt[s]++] = i; }
    for (int s = 0; s < 2; s++) err += encode_subset(px, sel[s], cnt[s], 3, 7, 3, 1, q[s], pb[s], idx[s]);
    if (!out) return err;
    for (int s = 0; s < 2; s++) {
        for (int k = 0; k < cnt[s]; k++) full[sel[s][k]] = idx[s][k];
        int anc = s ? A2[part] : 0;
        if (full[anc] & 4) {
            for (int c = 0; c < 3; c++) { int t = q[s][0][c]; q[s][0][c] = q[s][1][c]; q[s][1][c] = t; }
            for (int k = 0; k < cnt[s]; k++) full[sel[s][k]] = 7 - full[sel[s][k]];
        }
    }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 2, 2);
    putbits(out, &pos, (unsigned)part, 6);
    for (int c = 0; c < 3; c++) for (int s = 0; s < 2; s++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)q[s][e][c], 6);
    putbits(out, &pos, (unsigned)pb[0][0], 1); putbits(out, &pos, (unsigned)pb[1][0], 1);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)full[i], 3 - (i == 0 || i == A2[part]));
    OPS(60);
    return err;
}

static int bc7_decoded_sse(const px4 *px, const uint8_t *blk) { /* zap_tex.h:710-716 */
    px4 d[16];
    int err = 0;
    bc7_unpack(blk, d);
    for (int i = 0; i < 16; i++) for (int c = 0; c < 4; c++) { int x = px[i][c] - d[i][c]; err += x * x; }
    return err;
}

/* zap_tex.h:722-748: partition score from per-pixel moments {r,g,b,rr,gg,bb,rg,rb,gb} */
static float mode1_estimate(const float mo[16][9], const float tot[9], int part) {
    float s[2][9];
    int n1 = 0;
    for (int j = 0; j < 9; j++) s[1][j] = 0;
    for (int i = 0; i < 16; i++) if (P2[part] >> i & 1) { n1++; for (int j = 0; j < 9; j++) s[1][j] += mo[i][j]; }
    for (int j = 0; j < 9; j++) s[0][j] = tot[j] - s[1][j];
    OPS(16 * 9 + 9);
    float err = 0;
    for (int k = 0; k < 2; k++) {
        float n = (float)(k ? n1 : 16 - n1), *v = s[k];
        if (n < 2) continue;
        float S[3][3];
        S[0][0] = v[3] - v[0] * v[0] / n; S[0][1] = v[6] - v[0] * v[1] / n; S[0][2] = v[7] - v[0] * v[2] / n;
        S[1][1] = v[4] - v[1] * v[1] / n; S[1][2] = v[8] - v[1] * v[2] / n; S[2][2] = v[5] - v[2] * v[2] / n;
        S[1][0] = S[0][1]; S[2][0] = S[0][2]; S[2][1] = S[1][2];
        float tr = S[0][0] + S[1][1] + S[2][2], ax[3];
        int r0 = S[0][0] >= S[1][1] && S[0][0] >= S[2][2] ? 0 : S[1][1] >= S[2][2] ? 1 : 2;
        for (int c = 0; c < 3; c++) ax[c] = S[r0][c];
        for (int it = 0; it < 3; it++) {
            float t[3], mx = 1e-12f;
            for (int c = 0; c < 3; c++) { t[c] = S[c][0] * ax[0] + S[c][1] * ax[1] + S[c][2] * ax[2]; mx = t[c] > mx ? t[c] : -t[c] > mx ? -t[c] : mx; }
            for (int c = 0; c < 3; c++) ax[c] = t[c] / mx;
        }
        float vv = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], vsv = 0;
        for (int c = 0; c < 3; c++) vsv += ax[c] * (S[c][0] * ax[0] + S[c][1] * ax[1] + S[c][2] * ax[2]);
        float lam = vv > 1e-12f ? vsv / vv : 0;
        err += tr - lam * (48.0f / 49.0f);
        OPS(6 * 3 + 2 + 5 + 3 * (3 * 5 + 3 * 5 + 3) + 5 + 3 * 7 + 2 + 4);
    }
    return err;
}

static int g_m1k = 4; /* zap_tex.h:750-752 (ZAP__BC7_M1K, compile-time there) */

/* zap_tex.h:755-784 */
static int bc7_block(const px4 *px, uint8_t out[16]) {
    STAGE_BEGIN();
    int err = bc7_mode6(px, out), opaque = 1;
    for (int i = 0; i < 16; i++) opaque &= px[i][3] == 255;
    OPS(32 + 1);
    STAGE_END(0);
    if (err <= 16 * 4) return err;
    uint8_t m5[16];
    bc7_mode5(px, m5);
    int e5 = bc7_decoded_sse(px, m5);
    if (e5 < err) { memcpy(out, m5, 16); err = e5; }
    STAGE_END(1);
    if (!opaque) return err;
#ifdef ZTO_COUNT
    g_stage[3]++;
#endif
    float mo[16][9], tot[9] = { 0 }, ce[64];
    int cand[64], bp = -1, be = err, K = g_m1k;
    for (int i = 0; i < 16; i++) {
        float r = px[i][0], g = px[i][1], b = px[i][2], v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
        for (int j = 0; j < 9; j++) { mo[i][j] = v[j]; tot[j] += v[j]; }
    }
    OPS(16 * (3 + 6 + 9));
    for (int j = 0; j < K; j++) { ce[j] = 1e30f; cand[j] = -1; }
    for (int part = 0; part < 64; part++) {
        float e = mode1_estimate((const float(*)[9])mo, tot, part);
        for (int j = 0; j < K; j++)
            if (e < ce[j]) { for (int t = K - 1; t > j; t--) { ce[t] = ce[t - 1]; cand[t] = cand[t - 1]; } ce[j] = e; cand[j] = part; break; }
        OPS(4);
    }
    for (int j = 0; j < K; j++) { int e = bc7_mode1(px, cand[j], NULL); if (e < be) { be = e; bp = cand[j]; } }
    if (bp >= 0) {
        uint8_t m1[16];
        bc7_mode1(px, bp, m1);
        int e1 = bc7_decoded_sse(px, m1);
        if (e1 < err) { memcpy(out, m1, 16); err = e1; }
    }
    STAGE_END(2);
    return err;
}

static void gather(const uint8_t *rgba, int w, int h, size_t stride, int bx, int by, px4 *px);
/* ======================================================================= exhaustive BC7 search (all 8 modes)
 * Project-defined, with no reference counterpart: the reference encoder only tries modes 6, 5 and 1.  Candidates
 * are numbered mode*64 + slot: slot = partition for modes 0 (0..15), 1, 2, 3, 7 (0..63); rot*2 + isb for mode 4;
 * rot for mode 5; 0 for mode 6 (285 candidates).  Every subset is encoded the reference's way (PCA start, fit,
 * <= 2 least-squares refinements); modes 6, 5 and 1 are exactly the reference's encoders, so the exhaustive
 * result's SSE never exceeds the reference search's.  The SSE counts all four channels of the decoded block. */

static int fit_subset_plain(const px4 *px, const int *sel, int cnt, int nc, int n, int ib, const float lo[4], const float hi[4],
                            int q[2][4], int *idx) {
    int e0[4], e1[4];
    quant_plain(lo, nc, n, q[0], e0); quant_plain(hi, nc, n, q[1], e1);
    int pal[16][4], np = 1 << ib, err = 0;
    for (int j = 0; j < np; j++) for (int c = 0; c < nc; c++) { int w = weight(ib, j); pal[j][c] = ((64 - w) * e0[c] + w * e1[c] + 32) >> 6; }
    for (int k = 0; k < cnt; k++) {
        int bi = 0, be = 1 << 30;
        for (int j = 0; j < np; j++) { int d = sse_n(px[sel[k]], pal[j], nc); if (d < be) { be = d; bi = j; } }
        idx[k] = bi; err += be;
    }
    OPS(np * nc * 6 + cnt * (np * (3 * nc + 3) + 1));
    return err;
}
/* subset without p-bits (modes 2 / 4-5 colour): PCA, plain n-bit quantisation, <= 2 least-squares refinements */
static int encode_subset_plain(const px4 *px, const int *sel, int cnt, int nc, int n, int ib, int q[2][4], int *idx) {
    float lo[4], hi[4];
    pca_endpoints(px, sel, cnt, nc, lo, hi);
    int err = fit_subset_plain(px, sel, cnt, nc, n, ib, lo, hi, q, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float aa = 0, ab = 0, bb = 0, sa[4] = { 0, 0, 0, 0 }, sb[4] = { 0, 0, 0, 0 };
        for (int k = 0; k < cnt; k++) {
            float w = weight(ib, idx[k]) / 64.0f, v = 1 - w;
            aa += v * v; ab += v * w; bb += w * w;
            for (int c = 0; c < nc; c++) { sa[c] += v * px[sel[k]][c]; sb[c] += w * px[sel[k]][c]; }
        }
        float det = aa * bb - ab * ab;
        if (det < 1e-6f && det > -1e-6f) break;
        float nlo[4], nhi[4];
        for (int c = 0; c < nc; c++) {
            nlo[c] = (sa[c] * bb - sb[c] * ab) / det; nhi[c] = (sb[c] * aa - sa[c] * ab) / det;
            nlo[c] = nlo[c] < 0 ? 0 : nlo[c] > 255 ? 255 : nlo[c]; nhi[c] = nhi[c] < 0 ? 0 : nhi[c] > 255 ? 255 : nhi[c];
        }
        for (int c = nc; c < 4; c++) nlo[c] = nhi[c] = 255;
        int nq[2][4], ni[16];
        int ne = fit_subset_plain(px, sel, cnt, nc, n, ib, nlo, nhi, nq, ni);
        if (ne >= err) break;
        err = ne; memcpy(q, nq, sizeof nq); memcpy(idx, ni, sizeof(int) * (size_t)cnt);
    }
    return err;
}

/* modes 0, 1, 2, 3, 6, 7: one encode per subset, anchor fix-ups, pack */
static int p1_subsets(const px4 *px, int mode, int part, uint8_t out[16]) {
    const bc7_layout *m = &L7[mode];
    const int ns = m->ns, nc = m->ab ? 4 : 3, haspb = m->epb || m->spb, n = m->cb + haspb, ib = m->ib;
    int sel[3][16], cnt[3] = { 0, 0, 0 }, q[3][2][4], pb[3][2] = { { 0, 0 }, { 0, 0 }, { 0, 0 } }, idx[3][16], full[16], err = 0;
    for (int i = 0; i < 16; i++) { int sb = subset_of(ns, part, i); sel[sb][cnt[sb]++] = i; }
    for (int sb = 0; sb < ns; sb++)
        err += haspb ? encode_subset(px, sel[sb], cnt[sb], nc, n, ib, m->spb, q[sb], pb[sb], idx[sb])
                     : encode_subset_plain(px, sel[sb], cnt[sb], nc, n, ib, q[sb], idx[sb]);
    if (nc == 3) for (int i = 0; i < 16; i++) { int d = 255 - px[i][3]; err += d * d; }  /* decoded alpha is 255 */
    for (int sb = 0; sb < ns; sb++) {
        for (int k = 0; k < cnt[sb]; k++) full[sel[sb][k]] = idx[sb][k];
        int anc = anchor_of(ns, part, sb), top = 1 << (ib - 1), mx = (1 << ib) - 1;
        if (full[anc] & top) {
            for (int c = 0; c < 4; c++) { int t = q[sb][0][c]; q[sb][0][c] = q[sb][1][c]; q[sb][1][c] = t; }
            if (m->epb) { int t = pb[sb][0]; pb[sb][0] = pb[sb][1]; pb[sb][1] = t; }
            for (int k = 0; k < cnt[sb]; k++) full[sel[sb][k]] = mx - full[sel[sb][k]];
        }
    }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 1u << mode, mode + 1);
    putbits(out, &pos, (unsigned)part, m->pb);
    for (int c = 0; c < 3; c++) for (int sb = 0; sb < ns; sb++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)q[sb][e][c], m->cb);
    if (m->ab) for (int sb = 0; sb < ns; sb++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)q[sb][e][3], m->ab);
    if (m->epb) for (int sb = 0; sb < ns; sb++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)pb[sb][e], 1);
    if (m->spb) for (int sb = 0; sb < ns; sb++) putbits(out, &pos, (unsigned)pb[sb][0], 1);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)full[i], ib - (i == anchor_of(ns, part, subset_of(ns, part, i))));
    return err;
}

/* n-bit / ib-bit part of a rotation mode: plain quantisation, argmin, <= 2 refinements (generalises zap_tex.h:644-657) */
static int fit_nbit(const px4 *r, int c0, int nc, int ib, const int *e0, const int *e1, int idx[16]) {
    int pal[8][4], err = 0, np = 1 << ib;
    for (int j = 0; j < np; j++) for (int c = 0; c < nc; c++) { int w = weight(ib, j); pal[j][c] = ((64 - w) * e0[c] + w * e1[c] + 32) >> 6; }
    for (int i = 0; i < 16; i++) {
        int bi = 0, be = 1 << 30;
        for (int j = 0; j < np; j++) {
            int d = 0;
            for (int c = 0; c < nc; c++) { int x = r[i][c0 + c] - pal[j][c]; d += x * x; }
            if (d < be) { be = d; bi = j; }
        }
        idx[i] = bi; err += be;
    }
    OPS(np * nc * 6 + 16 * (np * (3 * nc + 3) + 1));
    return err;
}
static int lsq_nbit(const px4 *r, int c0, int nc, int ib, const int idx[16], float lo[4], float hi[4]) {
    float aa = 0, ab = 0, bb = 0, sa[4] = { 0, 0, 0, 0 }, sb[4] = { 0, 0, 0, 0 };
    for (int i = 0; i < 16; i++) {
        float w = weight(ib, idx[i]) / 64.0f, v = 1 - w;
        aa += v * v; ab += v * w; bb += w * w;
        for (int c = 0; c < nc; c++) { sa[c] += v * r[i][c0 + c]; sb[c] += w * r[i][c0 + c]; }
    }
    float det = aa * bb - ab * ab;
    if (det < 1e-6f && det > -1e-6f) return 0;
    for (int c = 0; c < nc; c++) {
        lo[c] = (sa[c] * bb - sb[c] * ab) / det; hi[c] = (sb[c] * aa - sa[c] * ab) / det;
        lo[c] = lo[c] < 0 ? 0 : lo[c] > 255 ? 255 : lo[c]; hi[c] = hi[c] < 0 ? 0 : hi[c] > 255 ? 255 : hi[c];
    }
    return 1;
}
static int part_nbit(const px4 *r, int c0, int nc, int n, int ib, const float lo[4], const float hi[4], int q[2][3], int idx[16]) {
    int e0[3], e1[3], ti[16], tq[2][3], te0[3], te1[3];
    quant_plain(lo, nc, n, q[0], e0); quant_plain(hi, nc, n, q[1], e1);
    int err = fit_nbit(r, c0, nc, ib, e0, e1, idx);
    for (int it = 0; it < 2 && err > 0; it++) {
        float nlo[4], nhi[4];
        if (!lsq_nbit(r, c0, nc, ib, idx, nlo, nhi)) break;
        quant_plain(nlo, nc, n, tq[0], te0); quant_plain(nhi, nc, n, tq[1], te1);
        int e = fit_nbit(r, c0, nc, ib, te0, te1, ti);
        if (e >= err) break;
        err = e; memcpy(q, tq, sizeof tq); memcpy(idx, ti, sizeof ti);
    }
    return err;
}
/* modes 4 and 5: rotation rot, index selector isb (mode 4) */
static int p1_rotation(const px4 *px, int mode, int rot, int isb, uint8_t out[16]) {
    const bc7_layout *m = &L7[mode];
    const int ibc = mode == 4 && isb ? 3 : 2, iba = mode == 4 && !isb ? 3 : 2;
    px4 r[16];
    int sel[16], cq[2][3], aq[2][3], ci[16], ai[16];
    memcpy(r, px, sizeof r);
    for (int i = 0; i < 16; i++) { sel[i] = i; if (rot) { uint8_t t = r[i][3]; r[i][3] = r[i][rot - 1]; r[i][rot - 1] = t; } }
    float lo[4], hi[4], alo[4] = { 255 }, ahi[4] = { 0 };
    pca_endpoints((const px4 *)r, sel, 16, 3, lo, hi);
    int err = part_nbit((const px4 *)r, 0, 3, m->cb, ibc, lo, hi, cq, ci);
    for (int i = 0; i < 16; i++) { if (r[i][3] < alo[0]) alo[0] = r[i][3]; if (r[i][3] > ahi[0]) ahi[0] = r[i][3]; }
    err += part_nbit((const px4 *)r, 3, 1, m->ab, iba, alo, ahi, aq, ai);
    if (ci[0] & (1 << (ibc - 1))) { for (int c = 0; c < 3; c++) { int t = cq[0][c]; cq[0][c] = cq[1][c]; cq[1][c] = t; } for (int i = 0; i < 16; i++) ci[i] = (1 << ibc) - 1 - ci[i]; }
    if (ai[0] & (1 << (iba - 1))) { int t = aq[0][0]; aq[0][0] = aq[1][0]; aq[1][0] = t; for (int i = 0; i < 16; i++) ai[i] = (1 << iba) - 1 - ai[i]; }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 1u << mode, mode + 1);
    putbits(out, &pos, (unsigned)rot, 2);
    if (mode == 4) putbits(out, &pos, (unsigned)isb, 1);
    for (int c = 0; c < 3; c++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)cq[e][c], m->cb);
    for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)aq[e][0], m->ab);
    const int *first = isb ? ai : ci, *second = isb ? ci : ai;   /* the 2-bit set is stored first */
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)first[i], m->ib - (i == 0));
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)second[i], m->ib2 - (i == 0));
    return err;
}

static int p1_valid(int cand) {
    int mode = cand >> 6, slot = cand & 63;
    return mode == 0 ? slot < 16 : mode == 4 ? slot < 8 : mode == 5 ? slot < 4 : mode == 6 ? slot == 0 : 1;
}
/* one candidate: SSE of its decoded block, block bytes in out; -1 for an unused slot */
int zto_bc7_candidate(const uint8_t px64[64], int cand, uint8_t out[16]) {
    const px4 *px = (const px4 *)px64;
    p3_init();
    if (cand < 0 || cand >= 512 || !p1_valid(cand)) return -1;
    int mode = cand >> 6, slot = cand & 63;
    if (mode == 4) return p1_rotation(px, 4, slot >> 1, slot & 1, out);
    if (mode == 5) return p1_rotation(px, 5, slot, 0, out);
    return p1_subsets(px, mode, slot, out);
}
/* exhaustive search: best candidate (smallest SSE, then lowest candidate number) */
int zto_bc7_full_block(const uint8_t px64[64], uint8_t out[16], int *best_cand) {
    int best = 1 << 30, bc = -1;
    uint8_t tmp[16];
    for (int c = 0; c < 512; c++) {
        int e = zto_bc7_candidate(px64, c, tmp);
        if (e >= 0 && e < best) { best = e; bc = c; memcpy(out, tmp, 16); }
    }
    if (best_cand) *best_cand = bc;
    return best;
}
void zto_bc7_full_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out_) {
    uint8_t *out = (uint8_t *)out_;
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            px4 px[16];
            gather(rgba, w, h, stride, bx, by, px);
            zto_bc7_full_block((const uint8_t *)px, out + ((size_t)by * (size_t)bw + (size_t)bx) * 16, NULL);
        }
}

/* ======================================================================= BC6H (zap_tex.h:817-975) */

static uint16_t half_of(float f) { /* zap_tex.h:820-827 */
    uint32_t x;
    if (!(f > 0)) return 0;
    memcpy(&x, &f, 4);
    if (x >= 0x477FE000u) return 0x7BFF;
    if (x < 0x38800000u) return (uint16_t)(f * 16777216.0f + 0.5f);
    return (uint16_t)(((x - 0x38000000u) + 0x1000u) >> 13);
}
static float float_of(uint16_t h) { /* zap_tex.h:828-835 */
    uint32_t e = (h >> 10) & 31, m = h & 1023u, x;
    float f;
    if (e == 0) return (float)m * (1.0f / 16777216.0f);
    x = ((e + 112u) << 23) | (m << 13);
    memcpy(&f, &x, 4);
    return f;
}
static int bc6_unquant(int e) { return e == 0 ? 0 : e == 1023 ? 0xFFFF : ((e << 16) + 0x8000) >> 10; }
static int bc6_finish(int u) { return (u * 31) >> 6; }
static int bc6_quant(float u) { /* zap_tex.h:841-851 */
    int g = cvt_i32((u - 32.0f) / 64.0f + 0.5f, 1), best = 0;
    float bd = 1e30f;
    for (long long t = (long long)g - 1; t <= (long long)g + 1; t++) {
        int e = t < 0 ? 0 : t > 1023 ? 1023 : (int)t;
        float d = (float)bc6_unquant(e) - u;
        d = d < 0 ? -d : d;
        if (d < bd) { bd = d; best = e; }
    }
    OPS(4 + 3 * 12);
    return best;
}
static uint64_t bc6_fit(const uint16_t t[16][3], const int e[2][3], int idx[16]) { /* zap_tex.h:854-871 (exact in int64) */
    int pal[16][3];
    uint64_t err = 0;
    for (int j = 0; j < 16; j++)
        for (int c = 0; c < 3; c++)
            pal[j][c] = bc6_finish(((64 - W4[j]) * bc6_unquant(e[0][c]) + W4[j] * bc6_unquant(e[1][c]) + 32) >> 6);
    for (int i = 0; i < 16; i++) {
        uint64_t be = UINT64_MAX;
        int bi = 0;
        for (int j = 0; j < 16; j++) {
            uint64_t d = 0;
            for (int c = 0; c < 3; c++) { int64_t x = (int64_t)t[i][c] - pal[j][c]; d += (uint64_t)(x * x); }
            if (d < be) { be = d; bi = j; }
        }
        idx[i] = bi; err += be;
    }
    OPS(16 * 3 * 8 + 16 * (16 * 12 + 1));
    return err;
}
static uint64_t bc6_block(const uint16_t t[16][3], uint8_t out[16]) { /* zap_tex.h:873-931 */
    float U[16][3], mean[3] = { 0, 0, 0 }, cv[6] = { 0, 0, 0, 0, 0, 0 }, ax[3] = { 1, 1, 1 };
    for (int i = 0; i < 16; i++) for (int c = 0; c < 3; c++) { U[i][c] = t[i][c] * 64.0f / 31.0f; mean[c] += U[i][c] / 16.0f; }
    for (int i = 0; i < 16; i++) {
        float r = U[i][0] - mean[0], g = U[i][1] - mean[1], b = U[i][2] - mean[2];
        cv[0] += r * r; cv[1] += r * g; cv[2] += r * b; cv[3] += g * g; cv[4] += g * b; cv[5] += b * b;
    }
    OPS(16 * 3 * 5 + 16 * 15);
    for (int it = 0; it < 6; it++) {
        float x = cv[0] * ax[0] + cv[1] * ax[1] + cv[2] * ax[2], y = cv[1] * ax[0] + cv[3] * ax[1] + cv[4] * ax[2];
        float z = cv[2] * ax[0] + cv[4] * ax[1] + cv[5] * ax[2], m = x < 0 ? -x : x;
        if ((y < 0 ? -y : y) > m) m = y < 0 ? -y : y;
        if ((z < 0 ? -z : z) > m) m = z < 0 ? -z : z;
        OPS(25);
        if (m < 1e-6f) break;
        ax[0] = x / m; ax[1] = y / m; ax[2] = z / m;
        OPS(3);
    }
    float n2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2], tl = 1e30f, th = -1e30f, bl[3] = { 1e30f, 1e30f, 1e30f }, bh[3] = { 0, 0, 0 };
    for (int i = 0; i < 16; i++) {
        float d = (U[i][0] - mean[0]) * ax[0] + (U[i][1] - mean[1]) * ax[1] + (U[i][2] - mean[2]) * ax[2];
        if (d < tl) tl = d;
        if (d > th) th = d;
        for (int c = 0; c < 3; c++) { if (U[i][c] < bl[c]) bl[c] = U[i][c]; if (U[i][c] > bh[c]) bh[c] = U[i][c]; }
    }
    OPS(5 + 16 * (8 + 4 + 12));
    uint64_t best = UINT64_MAX;
    int be[2][3] = { { 0 } }, bidx[16] = { 0 };
    for (int s = 0; s < 2; s++) {
        int e[2][3], idx[16];
        for (int c = 0; c < 3; c++) {
            float lo = s ? bl[c] : mean[c] + ax[c] * tl / n2, hi = s ? bh[c] : mean[c] + ax[c] * th / n2;
            e[0][c] = bc6_quant(lo); e[1][c] = bc6_quant(hi);
        }
        OPS(s ? 0 : 18);
        uint64_t err = bc6_fit(t, (const int(*)[3])e, idx);
        for (int it = 0; it < 3 && err > 0; it++) {
            float aa = 0, ab = 0, bb = 0, xa[3] = { 0, 0, 0 }, xb[3] = { 0, 0, 0 };
            for (int i = 0; i < 16; i++) {
                float w = W4[idx[i]] / 64.0f, v = 1 - w;
                aa += v * v; ab += v * w; bb += w * w;
                for (int c = 0; c < 3; c++) { xa[c] += v * U[i][c]; xb[c] += w * U[i][c]; }
            }
            float det = aa * bb - ab * ab;
            OPS(16 * (3 + 6 + 12) + 5);
            if (det < 1e-6f && det > -1e-6f) break;
            int ne[2][3], ni[16];
            for (int c = 0; c < 3; c++) { ne[0][c] = bc6_quant((xa[c] * bb - xb[c] * ab) / det); ne[1][c] = bc6_quant((xb[c] * aa - xa[c] * ab) / det); }
            OPS(3 * 8);
            uint64_t e2 = bc6_fit(t, (const int(*)[3])ne, ni);
            if (e2 >= err) break;
            err = e2; memcpy(e, ne, sizeof ne); memcpy(idx, ni, sizeof ni);
        }
        if (err < best) { best = err; memcpy(be, e, sizeof e); memcpy(bidx, idx, sizeof idx); }
    }
    if (bidx[0] & 8) {
        for (int c = 0; c < 3; c++) { int x = be[0][c]; be[0][c] = be[1][c]; be[1][c] = x; }
        for (int i = 0; i < 16; i++) bidx[i] = 15 - bidx[i];
    }
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 3, 5);
    for (int k = 0; k < 2; k++) for (int c = 0; c < 3; c++) putbits(out, &pos, (unsigned)be[k][c], 10);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)bidx[i], i ? 4 : 3);
    OPS(40);
    return best;
}
static void bc6_unpack(const uint8_t *b, uint16_t px[16][4]) { /* zap_tex.h:933-943 */
    int pos = 0, e[2][3];
    memset(px, 0, 16 * 4 * sizeof(uint16_t));
    if (getbits(b, &pos, 5) != 3) return;
    for (int k = 0; k < 2; k++) for (int c = 0; c < 3; c++) e[k][c] = bc6_unquant((int)getbits(b, &pos, 10));
    for (int i = 0; i < 16; i++) {
        int w = W4[getbits(b, &pos, i ? 4 : 3)];
        for (int c = 0; c < 3; c++) px[i][c] = (uint16_t)bc6_finish(((64 - w) * e[0][c] + w * e[1][c] + 32) >> 6);
        px[i][3] = 0x3C00;
    }
}

/* ======================================================================= ASTC 4x4 (zap_tex.h:977-1356) */

/* integer sequence encoding ranges (ASTC specification): levels, trit, quint, bits */
typedef struct { uint16_t levels; uint8_t trit, quint, bits; } ise_range;
static const ise_range ISE[21] = {
    { 2, 0, 0, 1 }, { 3, 1, 0, 0 }, { 4, 0, 0, 2 }, { 5, 0, 1, 0 }, { 6, 1, 0, 1 }, { 8, 0, 0, 3 }, { 10, 0, 1, 1 },
    { 12, 1, 0, 2 }, { 16, 0, 0, 4 }, { 20, 0, 1, 2 }, { 24, 1, 0, 3 }, { 32, 0, 0, 5 }, { 40, 0, 1, 3 }, { 48, 1, 0, 4 },
    { 64, 0, 0, 6 }, { 80, 0, 1, 4 }, { 96, 1, 0, 5 }, { 128, 0, 0, 7 }, { 160, 0, 1, 5 }, { 192, 1, 0, 6 }, { 256, 0, 0, 8 } };
static int ise_bitcount(const ise_range *r, int n) { return n * r->bits + (r->trit ? (8 * n + 4) / 5 : r->quint ? (7 * n + 2) / 3 : 0); }
static const ise_range *colour_range(int nvals, int avail) { /* zap_tex.h:996-999 */
    for (int i = 20; i >= 4; i--) if (ise_bitcount(&ISE[i], nvals) <= avail) return &ISE[i];
    return NULL;
}
static int replicate(int v, int n, int to) { /* zap_tex.h:1001-1005 */
    int x = 0;
    for (int sh = to - n; sh > -n; sh -= n) x |= sh >= 0 ? v << sh : v >> -sh;
    return x & ((1 << to) - 1);
}
static int colour_unquant(const ise_range *r, int v) { /* zap_tex.h:1008-1027 (spec colour unquantisation) */
    if (!r->trit && !r->quint) return replicate(v, r->bits, 8);
    int n = r->bits, m = v & ((1 << n) - 1), D = v >> n;
    int a = m & 1, b = m >> 1 & 1, c = m >> 2 & 1, d = m >> 3 & 1, e = m >> 4 & 1, f = m >> 5 & 1, A = a ? 0x1FF : 0, B = 0, C = 0;
    switch (r->levels) {
    case 6: C = 204; break;
    case 10: C = 113; break;
    case 12: B = b << 8 | b << 4 | b << 2 | b << 1; C = 93; break;
    case 20: B = b << 8 | b << 3 | b << 2; C = 54; break;
    case 24: B = c << 8 | b << 7 | c << 3 | b << 2 | c << 1 | b; C = 44; break;
    case 40: B = c << 8 | b << 7 | c << 2 | b << 1 | c; C = 26; break;
    case 48: B = d << 8 | c << 7 | b << 6 | d << 2 | c << 1 | b; C = 22; break;
    case 80: B = d << 8 | c << 7 | b << 6 | d << 1 | c; C = 13; break;
    case 96: B = e << 8 | d << 7 | c << 6 | b << 5 | e << 1 | d; C = 11; break;
    case 160: B = e << 8 | d << 7 | c << 6 | b << 5 | e; C = 6; break;
    case 192: B = f << 8 | e << 7 | d << 6 | c << 5 | b << 4 | f; C = 5; break;
    }
    int T = (D * C + B) ^ A;
    return (A & 0x80) | (T >> 2);
}
static int weight_unquant(int w, int n) { int x = replicate(w, n, 6); return x > 32 ? x + 1 : x; }

static void trits_of(int T, int t[5]) { /* zap_tex.h:1032-1042 (spec C.2.12) */
    int C;
    if ((T >> 2 & 7) == 7) { C = (T >> 5 & 7) << 2 | (T & 3); t[4] = 2; t[3] = 2; }
    else {
        C = T & 31;
        if ((T >> 5 & 3) == 3) { t[4] = 2; t[3] = T >> 7 & 1; } else { t[4] = T >> 7 & 1; t[3] = T >> 5 & 3; }
    }
    if ((C & 3) == 3) { t[2] = 2; t[1] = C >> 4 & 1; t[0] = (C >> 3 & 1) << 1 | ((C >> 2 & 1) & ~(C >> 3) & 1); }
    else if ((C >> 2 & 3) == 3) { t[2] = 2; t[1] = 2; t[0] = C & 3; }
    else { t[2] = C >> 4 & 1; t[1] = C >> 2 & 3; t[0] = (C >> 1 & 1) << 1 | ((C & 1) & ~(C >> 1) & 1); }
}
static void quints_of(int Q, int q[3]) { /* zap_tex.h:1044-1053 */
    if ((Q >> 1 & 3) == 3 && (Q >> 5 & 3) == 0) {
        q[2] = (Q & 1) << 2 | ((Q >> 4 & 1) & ~Q & 1) << 1 | ((Q >> 3 & 1) & ~Q & 1); q[1] = 4; q[0] = 4;
        return;
    }
    int C;
    if ((Q >> 1 & 3) == 3) { q[2] = 4; C = (Q >> 3 & 3) << 3 | (~(Q >> 5) & 3) << 1 | (Q & 1); }
    else { q[2] = Q >> 5 & 3; C = Q & 31; }
    if ((C & 7) == 5) { q[1] = 4; q[0] = C >> 3 & 3; } else { q[1] = C >> 3 & 3; q[0] = C & 7; }
}
static const int TRIT_BITS[5] = { 2, 2, 1, 2, 1 }, QUINT_BITS[3] = { 3, 2, 2 };
typedef struct { uint8_t t[243], q[125]; } ise_tables;
static void ise_tables_init(ise_tables *k) { /* zap_tex.h:1060-1063: descending, so the smallest code wins */
    for (int T = 255; T >= 0; T--) { int t[5]; trits_of(T, t); k->t[t[0] + 3 * t[1] + 9 * t[2] + 27 * t[3] + 81 * t[4]] = (uint8_t)T; }
    for (int Q = 127; Q >= 0; Q--) { int q[3]; quints_of(Q, q); k->q[q[0] + 5 * q[1] + 25 * q[2]] = (uint8_t)Q; }
}
static void ise_read(const uint8_t *b, int *pos, const ise_range *r, int n, int *v) { /* zap_tex.h:1065-1078 */
    int gs = r->trit ? 5 : r->quint ? 3 : 1;
    const int *xb = r->trit ? TRIT_BITS : QUINT_BITS;
    for (int g = 0; g < n; g += gs) {
        int cnt = n - g < gs ? n - g : gs, m[5], X = 0, xp = 0, t[5];
        for (int k = 0; k < cnt; k++) {
            m[k] = (int)getbits(b, pos, r->bits);
            if (gs > 1) { X |= (int)getbits(b, pos, xb[k]) << xp; xp += xb[k]; }
        }
        if (r->trit) trits_of(X, t);
        else if (r->quint) quints_of(X, t);
        for (int k = 0; k < cnt; k++) v[g + k] = gs > 1 ? t[k] << r->bits | m[k] : m[k];
    }
}
static void ise_write(uint8_t *b, int *pos, const ise_range *r, int n, const int *v, const ise_tables *tab) { /* zap_tex.h:1080-1095 */
    int gs = r->trit ? 5 : r->quint ? 3 : 1, base = r->trit ? 3 : 5;
    const int *xb = r->trit ? TRIT_BITS : QUINT_BITS;
    for (int g = 0; g < n; g += gs) {
        int cnt = n - g < gs ? n - g : gs, X = 0, xp = 0;
        if (gs > 1) {
            int id = 0;
            for (int k = gs - 1; k >= 0; k--) id = id * base + (k < cnt ? v[g + k] >> r->bits : 0);
            X = r->trit ? tab->t[id] : tab->q[id];
        }
        for (int k = 0; k < cnt; k++) {
            putbits(b, pos, (unsigned)(v[g + k] & ((1 << r->bits) - 1)), r->bits);
            if (gs > 1) { putbits(b, pos, (unsigned)(X >> xp), xb[k]); xp += xb[k]; }
        }
    }
}
static uint32_t part_hash(uint32_t p) { /* zap_tex.h:1098-1101 (spec C.2.21) */
    p ^= p >> 15; p -= p << 17; p += p << 7; p += p << 4; p ^= p >> 5; p += p << 16; p ^= p >> 7; p ^= p >> 3; p ^= p << 6; p ^= p >> 17;
    return p;
}
static int astc_partition(int seed, int x, int y, int count) { /* zap_tex.h:1102-1120 */
    x <<= 1; y <<= 1;
    seed += (count - 1) * 1024;
    uint32_t rn = part_hash((uint32_t)seed);
    int s[12] = { (int)(rn & 15), (int)(rn >> 4 & 15), (int)(rn >> 8 & 15), (int)(rn >> 12 & 15), (int)(rn >> 16 & 15), (int)(rn >> 20 & 15),
                  (int)(rn >> 24 & 15), (int)(rn >> 28 & 15), (int)(rn >> 18 & 15), (int)(rn >> 22 & 15), (int)(rn >> 26 & 15), (int)((rn >> 30 | rn << 2) & 15) };
    for (int i = 0; i < 12; i++) s[i] *= s[i];
    int sh1, sh2;
    if (seed & 1) { sh1 = seed & 2 ? 4 : 5; sh2 = count == 3 ? 6 : 5; }
    else { sh1 = count == 3 ? 6 : 5; sh2 = seed & 2 ? 4 : 5; }
    int sh3 = seed & 0x10 ? sh1 : sh2;
    for (int i = 0; i < 8; i++) s[i] >>= i & 1 ? sh2 : sh1;
    for (int i = 8; i < 12; i++) s[i] >>= sh3;
    int a = (s[0] * x + s[1] * y + (int)(rn >> 14)) & 0x3F, b = (s[2] * x + s[3] * y + (int)(rn >> 10)) & 0x3F;
    int c = (s[4] * x + s[5] * y + (int)(rn >> 6)) & 0x3F, d = (s[6] * x + s[7] * y + (int)(rn >> 2)) & 0x3F;
    if (count < 4) d = 0;
    if (count < 3) c = 0;
    return a >= b && a >= c && a >= d ? 0 : b >= c && b >= d ? 1 : c >= d ? 2 : 3;
}
static void astc_endpoints(int cem, const int *v, int e0[4], int e1[4]) { /* zap_tex.h:1122-1129 */
    int a0 = cem == 12 ? v[6] : 255, a1 = cem == 12 ? v[7] : 255;
    if (v[1] + v[3] + v[5] >= v[0] + v[2] + v[4]) { e0[0] = v[0]; e0[1] = v[2]; e0[2] = v[4]; e0[3] = a0; e1[0] = v[1]; e1[1] = v[3]; e1[2] = v[5]; e1[3] = a1; }
    else {
        e0[0] = v[1]; e0[1] = v[3]; e0[2] = v[5]; e0[3] = a1; e1[0] = v[0]; e1[1] = v[2]; e1[2] = v[4]; e1[3] = a0;
        e0[0] = (e0[0] + e0[2]) >> 1; e0[1] = (e0[1] + e0[2]) >> 1; e1[0] = (e1[0] + e1[2]) >> 1; e1[1] = (e1[1] + e1[2]) >> 1;
    }
}
static int astc_lerp(int a, int b, int w) { return (((a * 257) * (64 - w) + (b * 257) * w + 32) >> 6) >> 8; }

static void astc_unpack(const uint8_t *b, px4 *px) { /* zap_tex.h:1135-1168 */
    int pos = 0, mode = (int)getbits(b, &pos, 11);
    for (int i = 0; i < 16; i++) { px[i][0] = 255; px[i][1] = 0; px[i][2] = 255; px[i][3] = 255; }
    if ((mode & 0x1FF) == 0x1FC) {
        int p2 = 64;
        for (int c = 0; c < 4; c++) { int x = (int)getbits(b, &p2, 16) >> 8; for (int i = 0; i < 16; i++) px[i][c] = (uint8_t)x; }
        return;
    }
    int R = (mode & 1) << 1 | (mode >> 1 & 1) << 2 | (mode >> 4 & 1), H = mode >> 9 & 1;
    if (!(mode & 3) || (mode >> 2 & 3) || (mode >> 5 & 3) != 2 || (mode >> 7 & 3) || (mode >> 10 & 1)) return;
    static const int WL[2][8] = { { 0, 0, 2, 3, 4, 5, 6, 8 }, { 0, 0, 10, 12, 16, 20, 24, 32 } };
    int levels = WL[H][R], wb = levels == 2 ? 1 : levels == 4 ? 2 : levels == 8 ? 3 : levels == 16 ? 4 : levels == 32 ? 5 : 0;
    int parts = (int)getbits(b, &pos, 2) + 1, seed = 0, cem;
    if (!wb || parts > 2) return;
    if (parts == 2) {
        seed = (int)getbits(b, &pos, 10);
        int cf = (int)getbits(b, &pos, 6);
        if (cf & 3) return;
        cem = cf >> 2;
    } else cem = (int)getbits(b, &pos, 4);
    int nv = cem == 8 ? 6 : cem == 12 ? 8 : 0, v[16], e[2][2][4];
    if (!nv) return;
    const ise_range *r = colour_range(nv * parts, 128 - pos - 16 * wb);
    if (!r) return;
    ise_read(b, &pos, r, nv * parts, v);
    for (int i = 0; i < nv * parts; i++) v[i] = colour_unquant(r, v[i]);
    for (int k = 0; k < parts; k++) astc_endpoints(cem, v + k * nv, e[k][0], e[k][1]);
    for (int i = 0; i < 16; i++) {
        int w = 0, k = parts == 2 ? astc_partition(seed, i & 3, i >> 2, 2) : 0;
        for (int j = 0; j < wb; j++) { int bit = 127 - (i * wb + j); w |= (b[bit >> 3] >> (bit & 7) & 1) << j; }
        w = weight_unquant(w, wb);
        for (int c = 0; c < 4; c++) px[i][c] = (uint8_t)astc_lerp(e[k][0][c], e[k][1][c], w);
    }
}

typedef struct { const ise_range *r; int wb, cem, nv, parts; uint8_t q[256]; } astc_config;
static void astc_config_init(astc_config *c, int cem, int wb, int parts) { /* zap_tex.h:1172-1180 */
    c->cem = cem; c->wb = wb; c->nv = cem == 8 ? 6 : 8; c->parts = parts;
    c->r = colour_range(c->nv * parts, 128 - (parts == 1 ? 17 : 29) - 16 * wb);
    for (int x = 0; x < 256; x++) {
        int best = 0, bd = 1 << 30;
        for (int v = 0; v < c->r->levels; v++) { int d = colour_unquant(c->r, v) - x; d = d < 0 ? -d : d; if (d < bd) { bd = d; best = v; } }
        c->q[x] = (uint8_t)best;
    }
}
static void astc_quantise(const astc_config *c, const float lo[4], const float hi[4], int *ise, int *uv) { /* zap_tex.h:1183-1194 */
    float s0 = lo[0] + lo[1] + lo[2], s1 = hi[0] + hi[1] + hi[2];
    const float *a = s1 >= s0 ? lo : hi, *b = s1 >= s0 ? hi : lo;
    for (int k = 0; k < c->nv / 2; k++) {
        int x0 = cvt_i32(a[k] + 0.5f, 2), x1 = cvt_i32(b[k] + 0.5f, 2);
        x0 = clampi(x0, 0, 255); x1 = clampi(x1, 0, 255);
        ise[2 * k] = c->q[x0]; ise[2 * k + 1] = c->q[x1];
        uv[2 * k] = colour_unquant(c->r, ise[2 * k]); uv[2 * k + 1] = colour_unquant(c->r, ise[2 * k + 1]);
    }
    OPS(5 + c->nv / 2 * 14 + 6);
    if (uv[1] + uv[3] + uv[5] < uv[0] + uv[2] + uv[4])
        for (int k = 0; k < c->nv / 2; k++) { int t = ise[2 * k]; ise[2 * k] = ise[2 * k + 1]; ise[2 * k + 1] = t; t = uv[2 * k]; uv[2 * k] = uv[2 * k + 1]; uv[2 * k + 1] = t; }
}
static int astc_weights(const px4 *px, const astc_config *c, const int *uv, const int *part, int nc, int w[16]) { /* zap_tex.h:1197-1210 */
    int L = 1 << c->wb, pal[2][32][4], err = 0;
    for (int k = 0; k < c->parts; k++) {
        int e0[4], e1[4];
        astc_endpoints(c->cem, uv + k * c->nv, e0, e1);
        for (int j = 0; j < L; j++) { int wt = weight_unquant(j, c->wb); for (int ch = 0; ch < 4; ch++) pal[k][j][ch] = astc_lerp(e0[ch], e1[ch], wt); }
    }
    for (int i = 0; i < 16; i++) {
        int bj = 0, be = 1 << 30, k = part[i];
        for (int j = 0; j < L; j++) { int d = 0; for (int ch = 0; ch < nc; ch++) { int x = px[i][ch] - pal[k][j][ch]; d += x * x; } if (d < be) { be = d; bj = j; } }
        w[i] = bj; err += be;
    }
    OPS(c->parts * (8 + L * 4 * 8) + 16 * (L * (3 * nc + 3) + 1));
    return err;
}
static int astc_trial(const px4 *px, const astc_config *c, const int *part, int ise[16], int w[16]) { /* zap_tex.h:1212-1244 */
    int nc = c->cem == 12 ? 4 : 3, uv[16], ti[16], tuv[16], tw[16];
    for (int k = 0; k < c->parts; k++) {
        int sel[16], cnt = 0;
        float lo[4], hi[4];
        for (int i = 0; i < 16; i++) if (part[i] == k) sel[cnt++] = i;
        if (!cnt) { sel[0] = 0; cnt = 1; }
        pca_endpoints(px, sel, cnt, nc, lo, hi);
        astc_quantise(c, lo, hi, ise + k * c->nv, uv + k * c->nv);
    }
    int err = astc_weights(px, c, uv, part, nc, w);
    for (int it = 0; it < 3 && err > 0; it++) {
        memcpy(ti, ise, sizeof ti); memcpy(tuv, uv, sizeof tuv);
        for (int k = 0; k < c->parts; k++) {
            float aa = 0, ab = 0, bb = 0, xa[4] = { 0, 0, 0, 0 }, xb[4] = { 0, 0, 0, 0 }, nlo[4], nhi[4];
            for (int i = 0; i < 16; i++) {
                if (part[i] != k) continue;
                float f = weight_unquant(w[i], c->wb) / 64.0f, g = 1 - f;
                aa += g * g; ab += g * f; bb += f * f;
                for (int ch = 0; ch < nc; ch++) { xa[ch] += g * px[i][ch]; xb[ch] += f * px[i][ch]; }
                OPS(3 + 6 + nc * 5);
            }
            float det = aa * bb - ab * ab;
            OPS(5);
            if (det < 1e-6f && det > -1e-6f) continue;
            LS_NEAR(1, det, 1e-4f);
            for (int ch = 0; ch < nc; ch++) { nlo[ch] = (xa[ch] * bb - xb[ch] * ab) / det; nhi[ch] = (xb[ch] * aa - xa[ch] * ab) / det; }
            OPS(nc * 8);
            for (int ch = nc; ch < 4; ch++) nlo[ch] = nhi[ch] = 255;
            astc_quantise(c, nlo, nhi, ti + k * c->nv, tuv + k * c->nv);
        }
        int e = astc_weights(px, c, tuv, part, nc, tw);
        if (e >= err) break;
        err = e; memcpy(ise, ti, sizeof ti); memcpy(uv, tuv, sizeof tuv); memcpy(w, tw, sizeof tw);
    }
    return err;
}
static void astc_write(const astc_config *c, int seed, const int *ise, const int w[16], const ise_tables *tab, uint8_t out[16]) { /* zap_tex.h:1246-1260 */
    static const int RH[6][2] = { { 0, 0 }, { 0, 0 }, { 4, 0 }, { 7, 0 }, { 4, 1 }, { 7, 1 } };
    int R = RH[c->wb][0], H = RH[c->wb][1], pos = 0;
    unsigned mode = (unsigned)((R >> 1 & 1) | (R >> 2 & 1) << 1 | (R & 1) << 4 | 2 << 5 | H << 9);
    uint8_t wbits[16] = { 0 };
    memset(out, 0, 16);
    putbits(out, &pos, mode, 11);
    putbits(out, &pos, (unsigned)(c->parts - 1), 2);
    if (c->parts == 2) { putbits(out, &pos, (unsigned)seed, 10); putbits(out, &pos, (unsigned)c->cem << 2, 6); }
    else putbits(out, &pos, (unsigned)c->cem, 4);
    ise_write(out, &pos, c->r, c->nv * c->parts, ise, tab);
    int wp = 0;
    for (int i = 0; i < 16; i++) putbits(wbits, &wp, (unsigned)w[i], c->wb);
    for (int i = 0; i < wp; i++) if (wbits[i >> 3] >> (i & 7) & 1) out[(127 - i) >> 3] |= (uint8_t)(1 << ((127 - i) & 7));
    OPS(80);
}
static unsigned two_means(const px4 *px) { /* zap_tex.h:1263-1283 */
    float c[2][3];
    int lo = 0, hi = 0;
    for (int i = 1; i < 16; i++) { int s = px[i][0] + px[i][1] + px[i][2]; if (s < px[lo][0] + px[lo][1] + px[lo][2]) lo = i; if (s > px[hi][0] + px[hi][1] + px[hi][2]) hi = i; }
    for (int k = 0; k < 3; k++) { c[0][k] = px[lo][k]; c[1][k] = px[hi][k]; }
    unsigned m = 0;
    for (int it = 0; it < 4; it++) {
        float sum[2][3] = { { 0 } };
        int n[2] = { 0, 0 };
        m = 0;
        for (int i = 0; i < 16; i++) {
            float d0 = 0, d1 = 0;
            for (int k = 0; k < 3; k++) { float a = px[i][k] - c[0][k], b = px[i][k] - c[1][k]; d0 += a * a; d1 += b * b; }
            int s = d1 < d0;
            m |= (unsigned)s << i; n[s]++;
            for (int k = 0; k < 3; k++) sum[s][k] += px[i][k];
        }
        for (int s = 0; s < 2; s++) if (n[s]) for (int k = 0; k < 3; k++) c[s][k] = sum[s][k] / n[s];
    }
    OPS(15 * 10 + 4 * (16 * (3 * 8 + 6 + 6) + 12));
    return m;
}
static int popcount16(unsigned x) {
    x = x - ((x >> 1) & 0x5555u); x = (x & 0x3333u) + ((x >> 2) & 0x3333u); x = (x + (x >> 4)) & 0x0F0Fu;
    return (int)((x + (x >> 8)) & 0x1Fu);
}

typedef struct { ise_tables tab; astc_config cfg[8]; uint16_t pmask[1024], pseed[1024]; int np; } astc_setup;
static void astc_setup_init(astc_setup *s) { /* zap_tex.h:1293-1305 */
    ise_tables_init(&s->tab);
    astc_config_init(&s->cfg[0], 8, 4, 1); astc_config_init(&s->cfg[1], 8, 3, 1); astc_config_init(&s->cfg[2], 8, 5, 1);
    astc_config_init(&s->cfg[3], 12, 4, 1); astc_config_init(&s->cfg[4], 12, 3, 1); astc_config_init(&s->cfg[5], 12, 2, 1);
    astc_config_init(&s->cfg[6], 8, 3, 2); astc_config_init(&s->cfg[7], 8, 2, 2);
    s->np = 0;
    for (int sd = 0; sd < 1024; sd++) {
        unsigned m = 0;
        for (int i = 0; i < 16; i++) m |= (unsigned)astc_partition(sd, i & 3, i >> 2, 2) << i;
        if (m && m != 0xFFFFu) { s->pmask[s->np] = (uint16_t)m; s->pseed[s->np++] = (uint16_t)sd; }
    }
}
static void astc_block(const astc_setup *S, const px4 *px, int opaque, uint8_t *out) { /* zap_tex.h:1318-1340 */
    int best = 1 << 30, bc = 0, bseed = 0, bi[16] = { 0 }, bwt[16] = { 0 }, ise[16], wt[16], part0[16] = { 0 };
    for (int k = opaque ? 0 : 3; k < (opaque ? 3 : 6); k++) {
        int e = astc_trial(px, &S->cfg[k], part0, ise, wt);
        if (e < best) { best = e; bc = k; memcpy(bi, ise, sizeof ise); memcpy(bwt, wt, sizeof wt); }
    }
    if (opaque && best > 16 * 3 * 4) {
        unsigned km = two_means(px);
        int cand[4] = { -1, -1, -1, -1 }, cd[4] = { 99, 99, 99, 99 };
        for (int pi = 0; pi < S->np && cd[3]; pi++) {
            int dd = popcount16(S->pmask[pi] ^ km), di = 16 - dd;
            int dist = dd < di ? dd : di;
            for (int j = 0; j < 4; j++) if (dist < cd[j]) { for (int t = 3; t > j; t--) { cd[t] = cd[t - 1]; cand[t] = cand[t - 1]; } cd[j] = dist; cand[j] = pi; break; }
            OPS(20);
        }
        for (int j = 0; j < 4 && cand[j] >= 0; j++) {
            int part[16];
            for (int i = 0; i < 16; i++) part[i] = S->pmask[cand[j]] >> i & 1;
            for (int k = 6; k < 8; k++) {
                int e = astc_trial(px, &S->cfg[k], part, ise, wt);
                if (e < best) { best = e; bc = k; bseed = S->pseed[cand[j]]; memcpy(bi, ise, sizeof ise); memcpy(bwt, wt, sizeof wt); }
            }
        }
    }
    astc_write(&S->cfg[bc], bseed, bi, bwt, &S->tab, out);
}

/* ======================================================================= mips (zap_tex.h:1358-1469) */

static float SRGB_LIN[256], SRGB_THR[255];
static double srgb_to_linear(double c) { return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4); }
/* The reference tables are float32(formula) except five entries one ulp away (zap_tex.h:1360-1427). */
static void srgb_init(void) {
    static int done;
    if (done) return;
    for (int v = 0; v < 256; v++) SRGB_LIN[v] = (float)srgb_to_linear(v / 255.0);
    for (int k = 0; k < 255; k++) SRGB_THR[k] = (float)srgb_to_linear((k + 0.5) / 255.0);
    static const int lin_fix[3] = { 98, 122, 164 };
    for (int i = 0; i < 3; i++) { uint32_t b; memcpy(&b, &SRGB_LIN[lin_fix[i]], 4); b -= 1; memcpy(&SRGB_LIN[lin_fix[i]], &b, 4); }
    { uint32_t b; memcpy(&b, &SRGB_THR[34], 4); b -= 1; memcpy(&SRGB_THR[34], &b, 4); }
    { uint32_t b; memcpy(&b, &SRGB_THR[133], 4); b += 1; memcpy(&SRGB_THR[133], &b, 4); }
    done = 1;
}
static uint8_t linear_to_srgb(float v) { /* zap_tex.h:1429-1433 */
    int lo = 0, hi = 255;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (v < SRGB_THR[mid]) hi = mid; else lo = mid + 1; }
    return (uint8_t)lo;
}

/* ======================================================================= public restatement */

static void gather(const uint8_t *rgba, int w, int h, size_t stride, int bx, int by, px4 *px) { /* zap_tex.h:1482-1488 */
    for (int i = 0; i < 16; i++) {
        int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
        if (x >= w) x = w - 1;
        if (y >= h) y = h - 1;
        memcpy(px[i], rgba + (size_t)y * stride + (size_t)x * 4, 4);
    }
}

const float *zto_srgb_lin(void) { srgb_init(); return SRGB_LIN; }
const float *zto_srgb_thr(void) { srgb_init(); return SRGB_THR; }
void zto_set_m1k(int k) { g_m1k = k < 1 ? 1 : k > 64 ? 64 : k; }
int zto_get_m1k(void) { return g_m1k; }

/* zap_bc_encode at rdo = 0 (zap_tex.h:1474-1506); rows [by0, by1) of blocks only (band support) */
void zto_bc_encode_rows(const uint8_t *rgba, int w, int h, size_t stride, int fmt, void *out_, int by0, int by1) {
    uint8_t *out = (uint8_t *)out_;
    int bw = (w + 3) / 4;
    size_t bs = fmt == 0 || fmt == 2 ? 8 : 16;
    p3_init();
    for (int by = by0; by < by1; by++)
        for (int bx = 0; bx < bw; bx++) {
            px4 px[16];
            uint8_t ch[2][16], *o = out + ((size_t)by * (size_t)bw + (size_t)bx) * bs;
            gather(rgba, w, h, stride, bx, by, px);
            if (fmt == 0) { uint16_t c0, c1; uint32_t ix; bc1_block((const px4 *)px, &c0, &c1, &ix); put_bc1(o, c0, c1, ix); }
            else if (fmt == 2 || fmt == 1 || fmt == 3) {
                uint8_t a0, a1; uint64_t ix;
                int first = fmt == 1 ? 3 : 0;
                for (int i = 0; i < 16; i++) { ch[0][i] = px[i][first]; ch[1][i] = px[i][1]; }
                bc4_block(ch[0], &a0, &a1, &ix); put_bc4(o, a0, a1, ix);
                if (fmt == 1) { uint16_t c0, c1; uint32_t cx; bc1_block((const px4 *)px, &c0, &c1, &cx); put_bc1(o + 8, c0, c1, cx); }
                if (fmt == 3) { bc4_block(ch[1], &a0, &a1, &ix); put_bc4(o + 8, a0, a1, ix); }
            } else bc7_block((const px4 *)px, o);
        }
}
void zto_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, int fmt, void *out) {
    zto_bc_encode_rows(rgba, w, h, stride, fmt, out, 0, (h + 3) / 4);
}

/* ======================================================================= RDO (rdo > 0)
 * zap_tex.h:223-306 (rate model, BC1 and BC4 sub-blocks), :520-579 (mode-6 repack, refit of indices or
 * endpoints), :786-815 (BC7), :1474-1506 (driver: raster order, one state per sub-block stream). */
typedef struct { int dist[2]; } rdo_state;

/* zap_tex.h:234-242: 0 when the field continues the current match, the cost of a new match when it equals a field
 * of the previous 16 blocks (the nearest one), else its size in literals */
static float field_rate(const uint8_t *f, int n, const uint8_t *base, size_t cur, size_t stride, size_t off, int dist, int *newdist) {
    *newdist = dist;
    if (dist > 0 && (size_t)dist <= cur && !memcmp(f, base + (cur - (size_t)dist) * stride + off, (size_t)n)) return 0.0f;
    for (size_t j = 1; j <= 16 && j <= cur; j++)
        if (!memcmp(f, base + (cur - j) * stride + off, (size_t)n)) { *newdist = (int)j; return n < 3 ? (float)n : 3.0f; }
    return (float)n;
}

/* zap_tex.h:245-276: candidates from each of the previous 16 blocks: its endpoints (indices refitted), this block's
 * endpoints with its indices, its whole block; 3-colour candidates that would use the transparent index are skipped */
static void bc1_rdo(const px4 *px, int force4, float lambda, uint8_t *base, size_t cur, size_t stride, size_t off, rdo_state *rs) {
    uint16_t c0, c1;
    uint32_t idx;
    int e0 = bc1_block(px, &c0, &c1, &idx);
    uint8_t best[8], cand[8];
    put_bc1(best, c0, c1, idx);
    int d0, d1, bd0 = 0, bd1 = 0;
    float bj = (float)e0 + lambda * (field_rate(best, 4, base, cur, stride, off, rs->dist[0], &bd0) +
                                     field_rate(best + 4, 4, base, cur, stride, off + 4, rs->dist[1], &bd1));
    for (size_t j = 1; j <= 16 && j <= cur; j++) {
        const uint8_t *pb = base + (cur - j) * stride + off;
        uint16_t p0 = (uint16_t)ld_le(pb, 2), p1 = (uint16_t)ld_le(pb + 2, 2);
        uint32_t pi = (uint32_t)ld_le(pb + 4, 4), fi;
        for (int kind = 0; kind < 3; kind++) {
            int e;
            if (kind == 0) { e = bc1_indices(px, p0, p1, force4, &fi); put_bc1(cand, p0, p1, fi); }
            else if (kind == 1) { e = bc1_score(px, c0, c1, force4, pi); put_bc1(cand, c0, c1, pi); }
            else { e = bc1_score(px, p0, p1, force4, pi); memcpy(cand, pb, 8); }
            if (!force4 && ld_le(cand, 2) <= ld_le(cand + 2, 2)) {
                uint32_t ci = (uint32_t)ld_le(cand + 4, 4);
                int bad = 0;
                for (int i = 0; i < 16; i++) bad |= ((ci >> (2 * i)) & 3) == 3;
                if (bad) continue;
            }
            float jj = (float)e;
            if (jj >= bj) continue;
            jj += lambda * (field_rate(cand, 4, base, cur, stride, off, rs->dist[0], &d0) +
                            field_rate(cand + 4, 4, base, cur, stride, off + 4, rs->dist[1], &d1));
            if (jj < bj) { bj = jj; memcpy(best, cand, 8); bd0 = d0; bd1 = d1; }
        }
    }
    rs->dist[0] = bd0; rs->dist[1] = bd1;
    memcpy(base + cur * stride + off, best, 8);
}

/* zap_tex.h:279-306 */
static void bc4_rdo(const uint8_t v[16], float lambda, uint8_t *base, size_t cur, size_t stride, size_t off, rdo_state *rs) {
    uint8_t a0, a1;
    uint64_t idx;
    int e0 = bc4_block(v, &a0, &a1, &idx);
    uint8_t best[8], cand[8];
    put_bc4(best, a0, a1, idx);
    int d0, d1, bd0 = 0, bd1 = 0;
    float bj = (float)e0 + lambda * (field_rate(best, 2, base, cur, stride, off, rs->dist[0], &bd0) +
                                     field_rate(best + 2, 6, base, cur, stride, off + 2, rs->dist[1], &bd1));
    for (size_t j = 1; j <= 16 && j <= cur; j++) {
        const uint8_t *pb = base + (cur - j) * stride + off;
        uint64_t pi = ld_le(pb + 2, 6), fi;
        for (int kind = 0; kind < 3; kind++) {
            int e;
            if (kind == 0) { e = bc4_indices(v, pb[0], pb[1], &fi); put_bc4(cand, pb[0], pb[1], fi); }
            else if (kind == 1) { e = bc4_score(v, a0, a1, pi); put_bc4(cand, a0, a1, pi); }
            else { e = bc4_score(v, pb[0], pb[1], pi); memcpy(cand, pb, 8); }
            float jj = (float)e;
            if (jj >= bj) continue;
            jj += lambda * (field_rate(cand, 2, base, cur, stride, off, rs->dist[0], &d0) +
                            field_rate(cand + 2, 6, base, cur, stride, off + 2, rs->dist[1], &d1));
            if (jj < bj) { bj = jj; memcpy(best, cand, 8); bd0 = d0; bd1 = d1; }
        }
    }
    rs->dist[0] = bd0; rs->dist[1] = bd1;
    memcpy(base + cur * stride + off, best, 8);
}

/* mode-6 block fields (zap_tex.h:520-543) */
static void mode6_pack(int q[2][4], const int pb[2], const int *idx, uint8_t out[16]) {
    int pos = 0;
    memset(out, 0, 16);
    putbits(out, &pos, 1 << 6, 7);
    for (int c = 0; c < 4; c++) for (int e = 0; e < 2; e++) putbits(out, &pos, (unsigned)q[e][c], 7);
    putbits(out, &pos, (unsigned)pb[0], 1); putbits(out, &pos, (unsigned)pb[1], 1);
    for (int i = 0; i < 16; i++) putbits(out, &pos, (unsigned)idx[i], i ? 4 : 3);
}
static int is_mode6(const uint8_t *b) { return (b[0] & 0x7F) == 0x40; }
static void mode6_unpack(const uint8_t *b, int q[2][4], int pb[2], int idx[16]) {
    int pos = 7;
    for (int c = 0; c < 4; c++) for (int e = 0; e < 2; e++) q[e][c] = (int)getbits(b, &pos, 7);
    pb[0] = (int)getbits(b, &pos, 1); pb[1] = (int)getbits(b, &pos, 1);
    for (int i = 0; i < 16; i++) idx[i] = (int)getbits(b, &pos, i ? 4 : 3);
}
static void mode6_palette(int q[2][4], const int pb[2], int pal[16][4]) { /* zap_tex.h:538-542 */
    int e[2][4];
    for (int k = 0; k < 2; k++) for (int c = 0; c < 4; c++) e[k][c] = expand_bits(q[k][c] << 1 | pb[k], 8);
    for (int j = 0; j < 16; j++) for (int c = 0; c < 4; c++) pal[j][c] = ((64 - W4[j]) * e[0][c] + W4[j] * e[1][c] + 32) >> 6;
}
/* zap_tex.h:545-555: best indices for fixed endpoints, pixel 0 limited to 0..7 */
static int mode6_refit_indices(const px4 *px, int q[2][4], const int pb[2], int idx[16]) {
    int pal[16][4], err = 0;
    mode6_palette(q, pb, pal);
    for (int i = 0; i < 16; i++) {
        int bi = 0, be = 1 << 30;
        for (int j = 0; j < (i ? 16 : 8); j++) { int d = sse_n(px[i], pal[j], 4); if (d < be) { be = d; bi = j; } }
        idx[i] = bi; err += be;
    }
    return err;
}
/* zap_tex.h:557-579: least-squares endpoints for fixed indices, each endpoint's p-bit chosen on its own */
static int mode6_refit_endpoints(const px4 *px, const int idx[16], int q[2][4], int pb[2]) {
    float aa = 0, ab = 0, bb = 0, ax[4] = { 0, 0, 0, 0 }, bx[4] = { 0, 0, 0, 0 }, lo[4], hi[4];
    for (int i = 0; i < 16; i++) {
        float w = W4[idx[i]] / 64.0f, v = 1 - w;
        aa += v * v; ab += v * w; bb += w * w;
        for (int c = 0; c < 4; c++) { ax[c] += v * px[i][c]; bx[c] += w * px[i][c]; }
    }
    float det = aa * bb - ab * ab;
    for (int c = 0; c < 4; c++) {
        if (det < 1e-6f && det > -1e-6f) lo[c] = hi[c] = (ax[c] + bx[c]) / 16.0f;
        else { lo[c] = (ax[c] * bb - bx[c] * ab) / det; hi[c] = (bx[c] * aa - ax[c] * ab) / det; }
        lo[c] = lo[c] < 0 ? 0 : lo[c] > 255 ? 255 : lo[c]; hi[c] = hi[c] < 0 ? 0 : hi[c] > 255 ? 255 : hi[c];
    }
    int e0[4], e1[4], t0[4], t1[4], b0 = 1 << 30, b1 = 1 << 30;
    for (int p = 0; p < 2; p++) {
        int x = quant_pbit(lo, 4, 8, p, t0, e0), y = quant_pbit(hi, 4, 8, p, t1, e1);
        if (x < b0) { b0 = x; memcpy(q[0], t0, sizeof t0); pb[0] = p; }
        if (y < b1) { b1 = y; memcpy(q[1], t1, sizeof t1); pb[1] = p; }
    }
    int pal[16][4], err = 0;
    mode6_palette(q, pb, pal);
    for (int i = 0; i < 16; i++) err += sse_n(px[i], pal[idx[i]], 4);
    return err;
}
/* zap_tex.h:788-790: bytes 0-7 and 9-15 are the two fields, byte 8 is always paid */
static float bc7_rate(const uint8_t *c, const uint8_t *base, size_t cur, const rdo_state *rs, int d[2]) {
    return field_rate(c, 8, base, cur, 16, 0, rs->dist[0], &d[0]) + field_rate(c + 9, 7, base, cur, 16, 9, rs->dist[1], &d[1]) + 1;
}
/* zap_tex.h:792-815: the whole previous block, or (mode 6) its endpoints with refitted indices, or its indices with
 * re-solved endpoints */
static void bc7_rdo(const px4 *px, float lambda, uint8_t *base, size_t cur, rdo_state *rs) {
    uint8_t best[16], cand[16];
    int e0 = bc7_block(px, best), bd[2] = { 0, 0 }, d[2];
    if (lambda > 0) {
        float bj = (float)e0 + lambda * bc7_rate(best, base, cur, rs, bd);
        for (size_t j = 1; j <= 16 && j <= cur; j++) {
            const uint8_t *pb = base + (cur - j) * 16;
            int pq[2][4], ppb[2], pidx[16], nq[2][4], npb[2], nidx[16];
            int six = is_mode6(pb);
            if (six) mode6_unpack(pb, pq, ppb, pidx);
            for (int kind = 0; kind < (six ? 3 : 1); kind++) {
                int e;
                if (kind == 0) { e = bc7_decoded_sse(px, pb); memcpy(cand, pb, 16); }
                else if (kind == 1) { e = mode6_refit_indices(px, pq, ppb, nidx); mode6_pack(pq, ppb, nidx, cand); }
                else { e = mode6_refit_endpoints(px, pidx, nq, npb); mode6_pack(nq, npb, pidx, cand); }
                if ((float)e >= bj) continue;
                float jj = (float)e + lambda * bc7_rate(cand, base, cur, rs, d);
                if (jj < bj) { bj = jj; memcpy(best, cand, 16); bd[0] = d[0]; bd[1] = d[1]; }
            }
        }
        rs->dist[0] = bd[0]; rs->dist[1] = bd[1];
    }
    memcpy(base + cur * 16, best, 16);
}

/* zap_bc_encode with rdo (zap_tex.h:1474-1506): raster order; BC3 = alpha BC4 stream (state 0) + 4-colour BC1
 * stream (state 1), BC5 = R and G BC4 streams */
void zto_bc_encode_rdo(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, void *out_) {
    uint8_t *out = (uint8_t *)out_;
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    size_t bs = fmt == 0 || fmt == 2 ? 8 : 16;
    rdo_state rs[2] = { { { 1, 1 } }, { { 1, 1 } } };
    p3_init();
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            px4 px[16];
            uint8_t ch[2][16];
            gather(rgba, w, h, stride, bx, by, px);
            size_t cur = (size_t)by * (size_t)bw + (size_t)bx;
            if (fmt == 0) bc1_rdo((const px4 *)px, 0, rdo, out, cur, bs, 0, &rs[0]);
            else if (fmt == 2) {
                for (int i = 0; i < 16; i++) ch[0][i] = px[i][0];
                bc4_rdo(ch[0], rdo, out, cur, bs, 0, &rs[0]);
            } else if (fmt == 1) {
                for (int i = 0; i < 16; i++) ch[0][i] = px[i][3];
                bc4_rdo(ch[0], rdo, out, cur, bs, 0, &rs[0]);
                bc1_rdo((const px4 *)px, 1, rdo, out, cur, bs, 8, &rs[1]);
            } else if (fmt == 3) {
                for (int i = 0; i < 16; i++) { ch[0][i] = px[i][0]; ch[1][i] = px[i][1]; }
                bc4_rdo(ch[0], rdo, out, cur, bs, 0, &rs[0]);
                bc4_rdo(ch[1], rdo, out, cur, bs, 8, &rs[1]);
            } else bc7_rdo((const px4 *)px, rdo, out, cur, &rs[0]);
        }
}

/* per-block BC7 stage probe: returns the chosen mode (6, 5 or 1) and the final SSE */
int zto_bc7_block(const uint8_t px64[64], uint8_t out[16], int *err) {
    p3_init();
    int e = bc7_block((const px4 *)px64, out);
    if (err) *err = e;
    return out[0] & 0x40 ? 6 : out[0] & 0x20 ? 5 : 1;
}
/* the SSE each BC7 mode encoder returns next to the SSE of its decoded block (identity check) */
void zto_bc7_mode_errors(const uint8_t px64[64], int part, int res[6]) {
    uint8_t b[16];
    p3_init();
    res[0] = bc7_mode6((const px4 *)px64, b); res[1] = bc7_decoded_sse((const px4 *)px64, b);
    res[2] = bc7_mode5((const px4 *)px64, b); res[3] = bc7_decoded_sse((const px4 *)px64, b);
    res[4] = bc7_mode1((const px4 *)px64, part, b); res[5] = bc7_decoded_sse((const px4 *)px64, b);
}
float zto_bc7_estimate(const uint8_t px64[64], int part) {
    const px4 *px = (const px4 *)px64;
    float mo[16][9], tot[9] = { 0 };
    for (int i = 0; i < 16; i++) {
        float r = px[i][0], g = px[i][1], b = px[i][2], v[9] = { r, g, b, r * r, g * g, b * b, r * g, r * b, g * b };
        for (int j = 0; j < 9; j++) { mo[i][j] = v[j]; tot[j] += v[j]; }
    }
    return mode1_estimate((const float(*)[9])mo, tot, part);
}

void zto_bc_decode(const void *bc_, int w, int h, int fmt, uint8_t *rgba, size_t stride) { /* zap_tex.h:1508-1547 */
    const uint8_t *bc = (const uint8_t *)bc_;
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    size_t bs = fmt == 0 || fmt == 2 ? 8 : 16;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            const uint8_t *b = bc + ((size_t)by * (size_t)bw + (size_t)bx) * bs;
            px4 px[16];
            uint8_t v0[16], v1[16];
            if (fmt == 0 || fmt == 1) {
                const uint8_t *cb = fmt == 1 ? b + 8 : b;
                int pal[4][3], n = bc1_palette((uint16_t)ld_le(cb, 2), (uint16_t)ld_le(cb + 2, 2), fmt == 1, pal);
                uint32_t ix = (uint32_t)ld_le(cb + 4, 4);
                for (int i = 0; i < 16; i++) {
                    int k = (ix >> (2 * i)) & 3;
                    px[i][0] = (uint8_t)pal[k][0]; px[i][1] = (uint8_t)pal[k][1]; px[i][2] = (uint8_t)pal[k][2];
                    px[i][3] = (uint8_t)(n == 3 && k == 3 ? 0 : 255);
                }
            }
            if (fmt == 4) bc7_unpack(b, px);
            else if (fmt != 0) {
                int pal[8];
                bc4_palette(b[0], b[1], pal);
                uint64_t ix = ld_le(b + 2, 6);
                for (int i = 0; i < 16; i++) v0[i] = (uint8_t)pal[(ix >> (3 * i)) & 7];
                if (fmt == 1) for (int i = 0; i < 16; i++) px[i][3] = v0[i];
                else {
                    if (fmt == 3) { bc4_palette(b[8], b[9], pal); ix = ld_le(b + 10, 6); for (int i = 0; i < 16; i++) v1[i] = (uint8_t)pal[(ix >> (3 * i)) & 7]; }
                    else memset(v1, 0, 16);
                    for (int i = 0; i < 16; i++) { px[i][0] = v0[i]; px[i][1] = v1[i]; px[i][2] = 0; px[i][3] = 255; }
                }
            }
            for (int i = 0; i < 16; i++) {
                int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
                if (x < w && y < h) memcpy(rgba + (size_t)y * stride + (size_t)x * 4, px[i], 4);
            }
        }
}

void zto_bc6h_encode_rows(const float *rgba, int w, int h, size_t stride, void *out_, int by0, int by1) { /* zap_tex.h:946-960 */
    uint8_t *out = (uint8_t *)out_;
    int bw = (w + 3) / 4;
    for (int by = by0; by < by1; by++)
        for (int bx = 0; bx < bw; bx++) {
            uint16_t t[16][3];
            for (int i = 0; i < 16; i++) {
                int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
                if (x >= w) x = w - 1;
                if (y >= h) y = h - 1;
                for (int c = 0; c < 3; c++) t[i][c] = half_of(rgba[(size_t)y * stride + (size_t)x * 4 + c]);
            }
            OPS(48 * 8);
            bc6_block((const uint16_t(*)[3])t, out + ((size_t)by * (size_t)bw + (size_t)bx) * 16);
        }
}
void zto_bc6h_encode(const float *rgba, int w, int h, size_t stride, void *out) { zto_bc6h_encode_rows(rgba, w, h, stride, out, 0, (h + 3) / 4); }
void zto_bc6h_decode(const void *bc_, int w, int h, uint16_t *rgba, size_t stride) { /* zap_tex.h:963-975 */
    const uint8_t *bc = (const uint8_t *)bc_;
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            uint16_t px[16][4];
            bc6_unpack(bc + ((size_t)by * (size_t)bw + (size_t)bx) * 16, px);
            for (int i = 0; i < 16; i++) {
                int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
                if (x < w && y < h) memcpy(rgba + (size_t)y * stride + (size_t)x * 4, px[i], 8);
            }
        }
}
uint16_t zto_f2h(float f) { return half_of(f); }
float zto_h2f(uint16_t h) { return float_of(h); }

static astc_setup *astc_shared(void) {
    static astc_setup S;
    static int done;
    if (!done) { astc_setup_init(&S); done = 1; }
    return &S;
}
void zto_astc_encode_rows(const uint8_t *rgba, int w, int h, size_t stride, void *out_, int by0, int by1) { /* zap_tex.h:1291-1342 */
    uint8_t *out = (uint8_t *)out_;
    const astc_setup *S = astc_shared();
    int bw = (w + 3) / 4;
    for (int by = by0; by < by1; by++)
        for (int bx = 0; bx < bw; bx++) {
            px4 px[16];
            int opaque = 1;
            gather(rgba, w, h, stride, bx, by, px);
            for (int i = 0; i < 16; i++) opaque &= px[i][3] == 255;
            astc_block(S, (const px4 *)px, opaque, out + ((size_t)by * (size_t)bw + (size_t)bx) * 16);
        }
}
void zto_astc_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out) { zto_astc_encode_rows(rgba, w, h, stride, out, 0, (h + 3) / 4); }
void zto_astc_decode(const void *bc_, int w, int h, uint8_t *rgba, size_t stride) { /* zap_tex.h:1344-1356 */
    const uint8_t *bc = (const uint8_t *)bc_;
    int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            px4 px[16];
            astc_unpack(bc + ((size_t)by * (size_t)bw + (size_t)bx) * 16, px);
            for (int i = 0; i < 16; i++) {
                int x = bx * 4 + (i & 3), y = by * 4 + (i >> 2);
                if (x < w && y < h) memcpy(rgba + (size_t)y * stride + (size_t)x * 4, px[i], 4);
            }
        }
}
int zto_astc_partition(int seed, int x, int y, int count) { return astc_partition(seed, x, y, count); }
int zto_astc_pattern_count(void) { return astc_shared()->np; }
/* ISE round trip probe: writes n values of range ri, reads them back; returns bits written or -1 on mismatch */
int zto_ise_roundtrip(int ri, int n, const int *v) {
    uint8_t b[32] = { 0 };
    int pos = 0, rpos = 0, got[16];
    ise_write(b, &pos, &ISE[ri], n, v, &astc_shared()->tab);
    if (pos != ise_bitcount(&ISE[ri], n)) return -1;
    ise_read(b, &rpos, &ISE[ri], n, got);
    if (rpos != pos || memcmp(v, got, sizeof(int) * (size_t)n)) return -1;
    return pos;
}

int zto_mip_levels(int w, int h) { /* zap_tex.h:1435-1439 */
    int n = 1;
    while (w > 1 || h > 1) { w = w > 1 ? w / 2 : 1; h = h > 1 ? h / 2 : 1; n++; }
    return n;
}
void zto_mip_next(const uint8_t *src, int w, int h, size_t stride, int srgb, uint8_t *dst, size_t dst_stride) { /* zap_tex.h:1442-1457 */
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    srgb_init();
    for (int y = 0; y < nh; y++)
        for (int x = 0; x < nw; x++) {
            int x0 = 2 * x, x1 = 2 * x + 1 < w ? 2 * x + 1 : w - 1, y0 = 2 * y, y1 = 2 * y + 1 < h ? 2 * y + 1 : h - 1;
            if (x0 >= w) x0 = w - 1;
            if (y0 >= h) y0 = h - 1;
            const uint8_t *p[4] = { src + (size_t)y0 * stride + (size_t)x0 * 4, src + (size_t)y0 * stride + (size_t)x1 * 4,
                                    src + (size_t)y1 * stride + (size_t)x0 * 4, src + (size_t)y1 * stride + (size_t)x1 * 4 };
            uint8_t *o = dst + (size_t)y * dst_stride + (size_t)x * 4;
            for (int c = 0; c < 4; c++) {
                if (srgb && c < 3) o[c] = linear_to_srgb((SRGB_LIN[p[0][c]] + SRGB_LIN[p[1][c]] + SRGB_LIN[p[2][c]] + SRGB_LIN[p[3][c]]) * 0.25f);
                else o[c] = (uint8_t)((p[0][c] + p[1][c] + p[2][c] + p[3][c] + 2) >> 2);
            }
        }
}
void zto_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t dst_stride) { /* zap_tex.h:1460-1469 */
    int nw = w > 1 ? w / 2 : 1, nh = h > 1 ? h / 2 : 1;
    for (int y = 0; y < nh; y++)
        for (int x = 0; x < nw; x++) {
            int x0 = 2 * x < w ? 2 * x : w - 1, x1 = 2 * x + 1 < w ? 2 * x + 1 : w - 1, y0 = 2 * y < h ? 2 * y : h - 1, y1 = 2 * y + 1 < h ? 2 * y + 1 : h - 1;
            for (int c = 0; c < 4; c++)
                dst[(size_t)y * dst_stride + (size_t)x * 4 + c] = 0.25f * (src[(size_t)y0 * stride + (size_t)x0 * 4 + c] + src[(size_t)y0 * stride + (size_t)x1 * 4 + c] +
                                                                           src[(size_t)y1 * stride + (size_t)x0 * 4 + c] + src[(size_t)y1 * stride + (size_t)x1 * 4 + c]);
        }
}

/* zap_tex.h:1551-1571: plane split / merge */
void zto_bc_shuffle(const uint8_t *in, uint8_t *out, size_t n, int fmt, int split) {
    static const int F[5][5] = { { 2, 4, 4 }, { 4, 2, 6, 4, 4 }, { 2, 2, 6 }, { 4, 2, 6, 2, 6 }, { 3, 8, 1, 7 } };
    int nf = F[fmt][0];
    size_t bs = fmt == 0 || fmt == 2 ? 8 : 16, nb = n / bs, plane = 0, off = 0;
    for (int k = 0; k < nf; k++) {
        size_t sz = (size_t)F[fmt][k + 1];
        for (size_t b = 0; b < nb; b++) {
            const uint8_t *s = split ? in + b * bs + off : in + plane + b * sz;
            uint8_t *d = split ? out + plane + b * sz : out + b * bs + off;
            memcpy(d, s, sz);
        }
        plane += nb * sz; off += sz;
    }
    memcpy(out + nb * bs, in + nb * bs, n - nb * bs);
}
