// zt_mlp.hip — the BC7 proposal network on gfx950: 64 -> 128 -> 128 -> 512 MLP (ReLU), BF16 weights and
// activations on MFMA (gfx950's v_mfma_f32_16x16x32_bf16) with FP32 accumulation, trained here with its own
// forward/backward/Adam kernels.  Input: the 4x4 RGBA8 block (64 values / 255); output: one logit per BC7
// candidate slot mode*64 + slot, of which the 285 used slots are read.
#include "zt_common.h"
#include "zt_kernels.h"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short s16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace {

ZT_DEV uint16_t to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
ZT_DEV float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One 16x16x32 MFMA step.  Its 32 K positions are used in one fixed order everywhere in this file: lane (r, q)
// supplies k = 4q + 0..3 in elements 0..3 and k = 16 + 4q + 0..3 in elements 4..7 (of its 32-wide K chunk), so a
// step's two halves are the operands two 16x16x16 steps (k = 4q + 0..3 of each 16) would take.  Every product
// then lands in the same position in the GEMM and the fused proposal kernel, which keeps their sums identical.
ZT_DEV bf16x8_t cat8(bf16x4_t lo, bf16x4_t hi) { return __builtin_bit_cast(bf16x8_t, (s16x8_t)__builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7)); }
ZT_DEV f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// candidate slot used by the exhaustive search (285 of 512)
ZT_DEV bool cand_valid(int c) {
    const int m = c >> 6, s = c & 63;
    return m == 0 ? s < 16 : m == 4 ? s < 8 : m == 5 ? s < 4 : m == 6 ? s == 0 : true;
}

ZT_DEV bf16x4_t load_a4(const zt_gemm_args &g, int m, int k) {
    bf16x4_t v = { 0, 0, 0, 0 };
    if (m >= g.M) return v;
    if (g.sak == 1 && k + 3 < g.K && (((uintptr_t)(g.A + (int64_t)m * g.sam + k)) & 7) == 0) {
        uint2 t = *(const uint2 *)(g.A + (int64_t)m * g.sam + k);
        v[0] = (short)(t.x & 0xFFFF); v[1] = (short)(t.x >> 16); v[2] = (short)(t.y & 0xFFFF); v[3] = (short)(t.y >> 16);
        return v;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) if (k + i < g.K) v[i] = (short)g.A[(int64_t)m * g.sam + (int64_t)(k + i) * g.sak];
    return v;
}
ZT_DEV bf16x4_t load_b4(const zt_gemm_args &g, int k, int n) {
    bf16x4_t v = { 0, 0, 0, 0 };
    if (n >= g.N) return v;
    if (g.sbk == 1 && k + 3 < g.K && (((uintptr_t)(g.B + (int64_t)n * g.sbn + k)) & 7) == 0) {
        uint2 t = *(const uint2 *)(g.B + (int64_t)n * g.sbn + k);
        v[0] = (short)(t.x & 0xFFFF); v[1] = (short)(t.x >> 16); v[2] = (short)(t.y & 0xFFFF); v[3] = (short)(t.y >> 16);
        return v;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) if (k + i < g.K) v[i] = (short)g.B[(int64_t)(k + i) * g.sbk + (int64_t)n * g.sbn];
    return v;
}

}  // namespace

// C(m, n) = sum_k A(m, k) B(k, n): 64x64 tile per workgroup, 32x32 per wave (2x2 MFMA 16x16x32 tiles).
// MFMA lane layout: A row = lane % 16, B col = lane % 16, k positions as in mfma32; D col = lane % 16,
// row = 4 (lane / 16) + v.  blockIdx.z splits K (accumulate mode adds into Cf atomically).
extern "C" __global__ void __launch_bounds__(256) zt_gemm_bf16(zt_gemm_args g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    const int kper = ((g.K + (int)gridDim.z - 1) / (int)gridDim.z + 31) / 32 * 32;
    const int kb = blockIdx.z * kper, ke = min(g.K, kb + kper);
    const int r = lane & 15, q = lane >> 4;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k = kb; k < ke; k += 32) {
        bf16x4_t a[2][2], b[2][2];   // [tile][half]: half h covers k + 16 h + 4q + 0..3
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int kk = k + 16 * h + 4 * q;
#pragma unroll
            for (int i = 0; i < 2; i++) a[i][h] = kk < ke ? load_a4(g, m0 + 16 * i + r, kk) : bf16x4_t{ 0, 0, 0, 0 };
#pragma unroll
            for (int j = 0; j < 2; j++) b[j][h] = kk < ke ? load_b4(g, kk, n0 + 16 * j + r) : bf16x4_t{ 0, 0, 0, 0 };
            // zero the tail of a partial last K step on both operands
            if (kk + 3 >= ke) {
#pragma unroll
                for (int v = 0; v < 4; v++) if (kk + v >= ke) { a[0][h][v] = a[1][h][v] = 0; b[0][h][v] = b[1][h][v] = 0; }
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma32(cat8(a[i][0], a[i][1]), cat8(b[j][0], b[j][1]), acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int n = n0 + 16 * j + r;
            if (n >= g.N) continue;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int m = m0 + 16 * i + 4 * q + v;
                if (m >= g.M) continue;
                float x = acc[i][j][v];
                if (g.accumulate) { atomicAdd(g.Cf + (int64_t)m * g.ldc + n, x); continue; }
                if (g.bias) x += g.bias[n];
                if (g.relu && x < 0.0f) x = 0.0f;
                if (g.mask && (g.mask[(int64_t)m * g.ldm + n] & 0x7FFF) == 0) x = 0.0f;  // ReLU backward: activation was 0
                if (g.Cf) g.Cf[(int64_t)m * g.ldc + n] = x;
                if (g.Cb) g.Cb[(int64_t)m * g.ldc + n] = to_bf16(x);
            }
        }
}

// block bytes (16 RGBA8 pixels) -> 64 bf16 features, one thread per block; cfg[0] picks the normalisation:
//   0: v / 255
//   1: (v - m_c) / 64, m_c the block's channel mean
//   2: (v - m_c) / max(1, R), R the block's largest channel range (max - min)
//   3: (v - m_c) / sqrt(64 max(1, R))   (contrast-normalised, keeps a monotone cue of the block's contrast)
//   4: colour as 3 with R over R, G, B only; alpha as (a - 255) / 255 (0 when opaque: modes 0-3 decode alpha 255)
namespace {
struct FeatScale { float mean[4], inv, ainv; };
// the block's normalisation from its channel sums, minima and maxima over the 16 pixels
ZT_DEV FeatScale feature_scale(const int sum[4], const int lo[4], const int hi[4], int mode) {
    FeatScale f;
    f.inv = 1.0f / 255.0f;
    int rng = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        f.mean[c] = mode ? (float)sum[c] / 16.0f : 0.0f;
        if (mode != 4 || c < 3) rng = max(rng, hi[c] - lo[c]);
    }
    const float r = (float)max(1, rng);
    if (mode == 1) f.inv = 1.0f / 64.0f;
    else if (mode == 2) f.inv = 1.0f / r;
    else if (mode >= 3) f.inv = 1.0f / sqrtf(64.0f * r);
    f.ainv = f.inv;
    if (mode == 4) { f.mean[3] = 255.0f; f.ainv = 1.0f / 255.0f; }
    return f;
}
ZT_DEV uint16_t feature(const FeatScale &f, uint32_t p, int c) { return to_bf16(((float)chan(p, c) - f.mean[c]) * (c == 3 ? f.ainv : f.inv)); }
ZT_DEV void block_features(const uint32_t px[16], int mode, uint16_t *x, int x_step) {
    int sum[4], lo[4], hi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        sum[c] = 0; lo[c] = 255; hi[c] = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) { const int v = chan(px[k], c); sum[c] += v; lo[c] = min(lo[c], v); hi[c] = max(hi[c], v); }
    }
    const FeatScale f = feature_scale(sum, lo, hi, mode);
#pragma unroll
    for (int k = 0; k < 16; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) x[(4 * k + c) * x_step] = feature(f, px[k], c);
}
}  // namespace

extern "C" __global__ void zt_mlp_features(const uint8_t *blocks, int n, const int *cfg, uint16_t *X) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    uint32_t px[16];
    const uint4 *src = (const uint4 *)(blocks + b * 64);
#pragma unroll
    for (int r = 0; r < 4; r++) { uint4 v = src[r]; px[4 * r] = v.x; px[4 * r + 1] = v.y; px[4 * r + 2] = v.z; px[4 * r + 3] = v.w; }
    block_features(px, cfg[0], X + b * 64, 1);
}

// softmax cross-entropy over the used candidate slots, one wave per row; dZ = (p - q) / B in bf16 for targets q;
// stats[0] += this row's loss / B, stats[1] += 1 for a top-1 hit.
// With `deltas` and risk_k > 0, the unused slot ZT_RISK_SLOT also learns log(1 + the smallest delta among the row's
// current top-risk_k slots) by squared error (weight risk_w): the SSE a top-k search would give away.
extern "C" __global__ void __launch_bounds__(256) zt_mlp_xent(const float *Z, const uint32_t *labels, const uint16_t *deltas, float inv_t,
                                                                int risk_k, float risk_w, int B, uint16_t *dZ, float *stats) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= B) return;
    const float *z = Z + (int64_t)row * 512;
    float v[8], mx = -3.0e38f;
    int am = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int c = lane + 64 * t;
        v[t] = cand_valid(c) ? z[c] : -3.0e38f;
        if (v[t] > mx) { mx = v[t]; am = c; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float om = __shfl_xor(mx, off);
        int oa = __shfl_xor(am, off);
        if (om > mx || (om == mx && oa < am)) { mx = om; am = oa; }
    }
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; t++) if (cand_valid(lane + 64 * t)) s += __expf(v[t] - mx);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    const uint16_t *d = deltas ? deltas + (int64_t)row * 512 : nullptr;
    // targets q: one-hot at the label, or softmax(-delta / T) over the used slots (delta 0 = an optimal candidate)
    float q[8], qs = 0.0f, qz = 0.0f;
    bool hit;
    if (d && inv_t > 0.0f) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = lane + 64 * t;
            q[t] = cand_valid(c) && d[c] != 65535 ? __expf(-(float)d[c] * inv_t) : 0.0f;
            qs += q[t];
            qz += q[t] * (cand_valid(c) ? v[t] : 0.0f);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) { qs += __shfl_xor(qs, off); qz += __shfl_xor(qz, off); }
        const float inv = 1.0f / qs;
#pragma unroll
        for (int t = 0; t < 8; t++) q[t] *= inv;
        qz *= inv;
        hit = d[am] == 0;
    } else {
        const int y = (int)labels[row];
#pragma unroll
        for (int t = 0; t < 8; t++) q[t] = lane + 64 * t == y ? 1.0f : 0.0f;
        qz = z[y];
        hit = am == y;
    }
    // risk target: the smallest delta among the top risk_k slots (largest logit first, lowest slot on ties)
    float rgrad = 0.0f, rloss = 0.0f;
    if (d && risk_k > 0) {
        float w[8];
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = v[t];
        int dmin = 65535;
        for (int k = 0; k < risk_k; k++) {
            float bm = -3.0e38f;
            int ba = 1 << 20;
#pragma unroll
            for (int t = 0; t < 8; t++) if (w[t] > bm) { bm = w[t]; ba = lane + 64 * t; }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                float om = __shfl_xor(bm, off);
                int oa = __shfl_xor(ba, off);
                if (om > bm || (om == bm && oa < ba)) { bm = om; ba = oa; }
            }
            dmin = min(dmin, (int)d[ba]);
#pragma unroll
            for (int t = 0; t < 8; t++) if (lane + 64 * t == ba) w[t] = -3.0e38f;
        }
        const float e = z[ZT_RISK_SLOT] - __logf(1.0f + (float)dmin);
        rgrad = 2.0f * risk_w * e;
        rloss = risk_w * e * e;
    }
    const float invB = 1.0f / (float)B;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int c = lane + 64 * t;
        float gr = cand_valid(c) ? (__expf(v[t] - mx) / s - q[t]) * invB : c == ZT_RISK_SLOT ? rgrad * invB : 0.0f;
        dZ[(int64_t)row * 512 + c] = to_bf16(gr);
    }
    if (lane == 0) {
        atomicAdd(stats, (mx + __logf(s) - qz + rloss) * invB);   // cross-entropy -sum q log p (targets sum to 1) + risk term
        if (hit) atomicAdd(stats + 1, 1.0f);
    }
}

// column sums of a bf16 [M x N] matrix into out[N] (bias gradients)
extern "C" __global__ void zt_mlp_colsum(const uint16_t *src, int M, int N, float *out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int rows = (M + gridDim.y - 1) / gridDim.y, r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
    float s = 0.0f;
    for (int m = r0; m < r1; m++) s += from_bf16(src[(int64_t)m * N + n]);
    atomicAdd(out + n, s);
}

// Adam on the fp32 master parameters; weights (not biases) are re-rounded into their bf16 copies
extern "C" __global__ void zt_mlp_adam(float *p, const float *gr, float *m1, float *m2, int n, float lr, float b1, float b2,
                                       float c1, float c2, float eps, uint16_t *wbf, const int *bf_index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g, b = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a; m2[i] = b;
    const float x = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = x;
    const int j = bf_index[i];
    if (j >= 0) wbf[j] = to_bf16(x);
}

// deterministic He-uniform initialisation (counter hash), biases zero; bf16 copies written
extern "C" __global__ void zt_mlp_init(float *p, int n, uint32_t seed, uint16_t *wbf, const int *bf_index, const float *scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = (uint32_t)i * 0x9E3779B9u ^ seed;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    const float u = (float)(h >> 8) * (1.0f / 16777216.0f);
    const float x = scale[i] * (2.0f * u - 1.0f);
    p[i] = x;
    const int j = bf_index[i];
    if (j >= 0) wbf[j] = to_bf16(x);
}

// re-round the bf16 weight copies from the fp32 master weights (after loading parameters)
extern "C" __global__ void zt_mlp_refresh(const float *p, int n, uint16_t *wbf, const int *bf_index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = bf_index[i];
    if (j >= 0) wbf[j] = to_bf16(p[i]);
}

// ---- inference: features, the three layers and the top-k proposals in one persistent kernel ----------------------
// Each layer runs transposed, H^T = W X^T: the weights are the MFMA's A operand (from LDS), the activations its B
// operand.  The accumulator of output tile j (lane (r, q) holds neurons 16 j + 4 q + i of block r) is then, after
// bias, ReLU and bf16 rounding, exactly one half of the B operand of the next layer's 32-wide K step j / 2 (mfma32's
// k order), so activations never leave registers.  The products, their positions in the 32-term MFMA steps and
// the K order are those of zt_gemm_bf16 (only the roles of A and B swap), and the epilogue is the same, so the
// logits are bit-identical to the three GEMMs.
// The weights stay in LDS for the kernel's life in fragment order (each wave reads 64 consecutive 16-byte operand
// pairs: no bank conflicts): W1, W2 and, of W3, only the rows the proposals read (the 285 candidate slots in
// ascending order, then the risk slot; 2 zero rows pad to 18 MFMA tiles).  A wave takes 32 blocks per pass, two
// 16-block column tiles sharing every weight fragment.  Layer-3 logits become distinct ranking keys (prop_key: the
// logit down to its 9 low bits, then the lower row) and stream into a sorted top-KT list of keys per lane, one
// median-of-three per entry; the four lanes of a block then merge their lists K times by the largest key.
#define ZT_PR_U 288   // W3 rows kept
#define ZT_PR_NC 2    // 16-block column tiles per wave pass
namespace {
// candidate slot (mode * 64 + slot) of kept W3 row u: the used slots in ascending order, then the risk slot
ZT_DEV int prop_slot(int u) {
    return u < 16 ? u : u < 80 ? 64 + (u - 16) : u < 144 ? 128 + (u - 80) : u < 208 ? 192 + (u - 144) : u < 216 ? 256 + (u - 208)
         : u < 220 ? 320 + (u - 216) : u == 220 ? 384 : u < 285 ? 448 + (u - 221) : u == 285 ? ZT_RISK_SLOT : -1;
}
ZT_DEV bf16x4_t bf4(uint32_t lo, uint32_t hi) { return __builtin_bit_cast(bf16x4_t, make_uint2(lo, hi)); }
ZT_DEV uint32_t bf16_pair(float a, float b) {
    const uint32_t ua = __float_as_uint(a), ub = __float_as_uint(b);
    return perm(ub + 0x7FFFu + ((ub >> 16) & 1u), ua + 0x7FFFu + ((ua >> 16) & 1u), 0x07060302u);
}
// weights [rows x IN] (bf16, row-major) -> LDS fragments: for row tile j, 32-wide K step p and lane l, 16 bytes =
// the A operand (row 16 j + l % 16, columns 32 p + 16 s + 4 (l / 16) + 0..3 for halves s = 0, 1: mfma32's k order)
template <int IN>
ZT_DEV void stage_fragments(uint4 *dst, const uint16_t *w, int rows_kept, int tid, int nthreads, bool w3) {
    constexpr int CP = IN / 32;
    for (int e = tid; e < rows_kept / 16 * CP * 64; e += nthreads) {
        const int l = e & 63, p = (e >> 6) % CP, j = (e >> 6) / CP;
        const int u = 16 * j + (l & 15), row = w3 ? prop_slot(u) : u, col = 32 * p + 4 * (l >> 4);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (row >= 0) {
            const uint2 a = *(const uint2 *)(w + (int64_t)row * IN + col), b = *(const uint2 *)(w + (int64_t)row * IN + col + 16);
            v = make_uint4(a.x, a.y, b.x, b.y);
        }
        dst[e] = v;
    }
}
// bias, optional ReLU and bf16 rounding of an accumulator tile (zt_gemm_bf16's epilogue) -> next layer's B operand
ZT_DEV bf16x4_t epilogue(f32x4_t d, const float *bias) {
    const float4 b = *(const float4 *)bias;
    float x0 = d[0] + b.x, x1 = d[1] + b.y, x2 = d[2] + b.z, x3 = d[3] + b.w;
    if (x0 < 0.0f) x0 = 0.0f;
    if (x1 < 0.0f) x1 = 0.0f;
    if (x2 < 0.0f) x2 = 0.0f;
    if (x3 < 0.0f) x3 = 0.0f;
    return bf4(bf16_pair(x0, x1), bf16_pair(x2, x3));
}
// one layer, NJ output row tiles over NK 16-wide K halves (NK / 2 MFMA steps), NC column tiles:
// out[ct][j] = epilogue(sum_k W[j][k] in[ct][k])
template <int NJ, int NK, int NC>
ZT_DEV void layer(const uint4 *Wf, const float *bias, const bf16x4_t (&in)[NC][NK], bf16x4_t (&out)[NC][NJ], int lane, int q) {
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        f32x4_t acc[NC];
#pragma unroll
        for (int ct = 0; ct < NC; ct++) acc[ct] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int p = 0; p < NK / 2; p++) {
            const bf16x8_t a = __builtin_bit_cast(bf16x8_t, Wf[(j * (NK / 2) + p) * 64 + lane]);
#pragma unroll
            for (int ct = 0; ct < NC; ct++) acc[ct] = mfma32(a, cat8(in[ct][2 * p], in[ct][2 * p + 1]), acc[ct]);
        }
#pragma unroll
        for (int ct = 0; ct < NC; ct++) out[ct][j] = epilogue(acc[ct], bias + 16 * j + 4 * q);
    }
}
// layer-3 logits of kept row tile t (bias added): lane (r, q) gets rows 16 t + 4 q + i of each column tile's block r
template <int NC>
ZT_DEV void logits_tile(const uint4 *W3f, const float *B3s, const bf16x4_t (&h2)[NC][8], int t, int lane, int q, float (&v)[NC][4]) {
    f32x4_t z[NC];
#pragma unroll
    for (int ct = 0; ct < NC; ct++) z[ct] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, W3f[(t * 4 + p) * 64 + lane]);
#pragma unroll
        for (int ct = 0; ct < NC; ct++) z[ct] = mfma32(a, cat8(h2[ct][2 * p], h2[ct][2 * p + 1]), z[ct]);
    }
    const float4 b = *(const float4 *)(B3s + 16 * t + 4 * q);
#pragma unroll
    for (int ct = 0; ct < NC; ct++) { v[ct][0] = z[ct][0] + b.x; v[ct][1] = z[ct][1] + b.y; v[ct][2] = z[ct][2] + b.z; v[ct][3] = z[ct][3] + b.w; }
}
// Ranking key of logit v on kernel row u (u < 512): v's bits with the low 9 replaced by 511 - u, or by u when v is
// negative.  Compared as floats, keys order like the logits down to those 9 bits (2^-14 of the logit) and then by
// ascending row; the keys of one block are distinct, so a top list needs no row payload.
ZT_DEV float prop_key(float v, int u) {
    const uint32_t b = __float_as_uint(v), s = (uint32_t)((int32_t)b >> 31);
    return __uint_as_float((b & ~511u) | ((s ^ (511u - (uint32_t)u)) & 511u));
}
ZT_DEV int prop_key_row(float k) {
    const uint32_t b = __float_as_uint(k), c = b & 511u;
    return (int)((b >> 31) ? c : 511u - c);
}
// sorted (largest first) top-KT of a stream of keys: one median-of-three per entry
template <int KT>
struct TopList {
    float t[KT];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < KT; i++) t[i] = -INFINITY;
    }
    __device__ __forceinline__ void insert(float x) {
#pragma unroll
        for (int i = KT - 1; i > 0; i--) t[i] = __builtin_amdgcn_fmed3f(t[i - 1], t[i], x);
        t[0] = fmaxf(t[0], x);
    }
    __device__ __forceinline__ void pop_if(bool pop) {
#pragma unroll
        for (int i = 0; i + 1 < KT; i++) t[i] = pop ? t[i + 1] : t[i];
        t[KT - 1] = pop ? -INFINITY : t[KT - 1];
    }
};

// the network in LDS for a kernel's life
struct PropLds {
    uint4 W1f[128 / 16 * 2 * 64];       // 16 KiB
    uint4 W2f[128 / 16 * 4 * 64];       // 32 KiB
    uint4 W3f[ZT_PR_U / 16 * 4 * 64];   // 72 KiB
    float B1[128], B2[128], B3[ZT_PR_U];
};
ZT_DEV void stage_network(PropLds &L, const uint16_t *w, const float *b1, const float *b2, const float *b3) {
    const int tid = threadIdx.x, nth = blockDim.x;
    stage_fragments<64>(L.W1f, w, 128, tid, nth, false);
    stage_fragments<128>(L.W2f, w + 64 * 128, 128, tid, nth, false);
    stage_fragments<128>(L.W3f, w + 64 * 128 + 128 * 128, ZT_PR_U, tid, nth, true);
    for (int i = tid; i < 128; i += nth) { L.B1[i] = b1[i]; L.B2[i] = b2[i]; }
    for (int u = tid; u < ZT_PR_U; u += nth) { const int id = prop_slot(u); L.B3[u] = id < 0 ? 0.0f : b3[id]; }
    __syncthreads();
}
// layer 1's B operands (16-wide K halves 0..3) for the 16 blocks of a column tile: lane (r, q) holds pixels q, 4 + q, 8 + q,
// 12 + q of block `row` (zeros when !valid); the block's channel sums and ranges combine its four lanes
ZT_DEV void tile_features(const uint4 *blocks, uint32_t row, bool valid, int q, int mode, bf16x4_t (&x)[4]) {
    uint32_t px[4];
    const uint32_t *src = (const uint32_t *)blocks + (uint64_t)row * 16 + q;
#pragma unroll
    for (int k = 0; k < 4; k++) px[k] = valid ? src[4 * k] : 0u;
    int sum[4], lo[4], hi[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        sum[c] = 0; lo[c] = 255; hi[c] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) { const int v = chan(px[k], c); sum[c] += v; lo[c] = min(lo[c], v); hi[c] = max(hi[c], v); }
#pragma unroll
        for (int off = 16; off <= 32; off <<= 1) {
            sum[c] += __shfl_xor(sum[c], off);
            lo[c] = min(lo[c], __shfl_xor(lo[c], off));
            hi[c] = max(hi[c], __shfl_xor(hi[c], off));
        }
    }
    const FeatScale f = feature_scale(sum, lo, hi, mode);
#pragma unroll
    for (int k = 0; k < 4; k++)
        x[k] = bf16x4_t{ (short)feature(f, px[k], 0), (short)feature(f, px[k], 1), (short)feature(f, px[k], 2), (short)feature(f, px[k], 3) };
}

template <int KT>
ZT_DEV void propose_body(const zt_propose_args &P) {
    __shared__ PropLds L;
    stage_network(L, P.w, P.b1, P.b2, P.b3);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int lane = tid & 63, r = lane & 15, q = lane >> 4, mode = P.cfg[0], K = P.topk;
    const uint32_t waves = gridDim.x * (nth >> 6), groups = (P.n + 16 * ZT_PR_NC - 1) / (16 * ZT_PR_NC);
    for (uint32_t g = blockIdx.x * (nth >> 6) + (tid >> 6); g < groups; g += waves) {
        bf16x4_t x[ZT_PR_NC][4];
#pragma unroll
        for (int ct = 0; ct < ZT_PR_NC; ct++) {
            const uint32_t row = (g * ZT_PR_NC + ct) * 16 + r;
            tile_features(P.blocks, row, row < P.n, q, mode, x[ct]);
        }
        bf16x4_t h1[ZT_PR_NC][8], h2[ZT_PR_NC][8];
        layer<8, 4, ZT_PR_NC>(L.W1f, L.B1, x, h1, lane, q);
        layer<8, 8, ZT_PR_NC>(L.W2f, L.B2, h1, h2, lane, q);
        // layer 3, one row tile at a time, straight into the top lists
        TopList<KT> T[ZT_PR_NC];
        float riskv[ZT_PR_NC];
#pragma unroll
        for (int ct = 0; ct < ZT_PR_NC; ct++) { T[ct].init(); riskv[ct] = 0.0f; }
#pragma unroll
        for (int t = 0; t < ZT_PR_U / 16; t++) {
            float v[ZT_PR_NC][4];
            logits_tile<ZT_PR_NC>(L.W3f, L.B3, h2, t, lane, q, v);
#pragma unroll
            for (int ct = 0; ct < ZT_PR_NC; ct++) {
                const uint32_t row = (g * ZT_PR_NC + ct) * 16 + r;
                if (P.logits && row < P.n)
                    *(float4 *)(P.logits + (uint64_t)row * ZT_PR_U + 16 * t + 4 * q) = make_float4(v[ct][0], v[ct][1], v[ct][2], v[ct][3]);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int u = 16 * t + 4 * q + i;
                    if (t == ZT_PR_U / 16 - 1) {   // the last tile holds rows 272..287: slots up to 284, the risk row 285
                        if (u == 285) riskv[ct] = v[ct][i];
                        T[ct].insert(u < 285 ? prop_key(v[ct][i], u) : -INFINITY);
                    } else {
                        T[ct].insert(prop_key(v[ct][i], u));
                    }
                }
            }
        }
#pragma unroll
        for (int ct = 0; ct < ZT_PR_NC; ct++) {
            const uint32_t row = (g * ZT_PR_NC + ct) * 16 + r;
            for (int k = 0; k < K; k++) {   // the block's largest key over its four lanes' lists, K times
                float hv = T[ct].t[0];
                hv = fmaxf(hv, __shfl_xor(hv, 16));
                hv = fmaxf(hv, __shfl_xor(hv, 32));
                if (P.cand && q == 0 && row < P.n) P.cand[(uint64_t)row * K + k] = (uint32_t)prop_slot(prop_key_row(hv));
                T[ct].pop_if(__float_as_uint(hv) == __float_as_uint(T[ct].t[0]));
            }
            if (P.risk && q == 3 && row < P.n) {
                const float x0 = riskv[ct];
                P.risk[row] = x0 <= 0.0f ? 0.0f : x0 >= 13.9f ? 1048576.0f : __expf(x0) - 1.0f;
            }
        }
    }
}

// order-preserving map of a float to uint32 (larger float -> larger key; -inf below every finite value)
ZT_DEV uint32_t ord_key(float v) { const uint32_t b = __float_as_uint(v); return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }

// The wide proposal of the listed blocks (the learned encode's fallback): each block's k candidate slots of largest
// logit among the 285, ties at the k-th logit to the lowest slots, written unordered to cand[i * k ..] for list
// entry i.  The k-th largest logit key T comes from a bisection over the 32-bit order-preserving keys (count of keys
// >= a trial value, per lane then over the block's four lanes); then every key > T is taken and, of the keys == T,
// the lowest slots until k are taken.
ZT_DEV void propose_wide_body(const zt_wide_args &P) {
    __shared__ PropLds L;
    stage_network(L, P.w, P.b1, P.b2, P.b3);
    const int tid = threadIdx.x, nth = blockDim.x;
    const int lane = tid & 63, r = lane & 15, q = lane >> 4, mode = P.cfg[0], K = P.k;
    const uint32_t n = P.count[0], waves = gridDim.x * (nth >> 6), groups = (n + 15) / 16;
    for (uint32_t g = blockIdx.x * (nth >> 6) + (tid >> 6); g < groups; g += waves) {
        const uint32_t i_list = g * 16 + r;
        const bool valid = i_list < n;
        const uint32_t row = valid ? P.list[i_list] : 0u;
        bf16x4_t x[1][4];
        tile_features(P.blocks, row, valid, q, mode, x[0]);
        bf16x4_t h1[1][8], h2[1][8];
        layer<8, 4, 1>(L.W1f, L.B1, x, h1, lane, q);
        layer<8, 8, 1>(L.W2f, L.B2, h1, h2, lane, q);
        uint32_t key[ZT_PR_U / 16][4];   // rows 285..287 (not candidates) get the key of -inf
#pragma unroll
        for (int t = 0; t < ZT_PR_U / 16; t++) {
            float v[1][4];
            logits_tile<1>(L.W3f, L.B3, h2, t, lane, q, v);
#pragma unroll
            for (int i = 0; i < 4; i++) key[t][i] = (t == ZT_PR_U / 16 - 1 && 4 * q + i >= 285 - 16 * t) ? ord_key(-INFINITY) : ord_key(v[0][i]);
            // one tile's LDS fragments at a time: hoisting all 18 tiles' loads ahead of their MFMAs spills
            __builtin_amdgcn_sched_barrier(0);
        }
        // bisection: T = the largest key value with at least K keys >= it (the K-th largest key of the block)
        uint32_t T = 0;
        for (int bit = 31; bit >= 0; bit--) {
            const uint32_t c = T | (1u << bit);
            int cnt = 0;
#pragma unroll
            for (int t = 0; t < ZT_PR_U / 16; t++)
#pragma unroll
                for (int i = 0; i < 4; i++) cnt += key[t][i] >= c ? 1 : 0;
            cnt += __shfl_xor(cnt, 16);
            cnt += __shfl_xor(cnt, 32);
            if (cnt >= K) T = c;
        }
        int gt = 0, eq = 0;
#pragma unroll
        for (int t = 0; t < ZT_PR_U / 16; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) { gt += key[t][i] > T ? 1 : 0; eq += key[t][i] == T ? 1 : 0; }
        int gt_b = gt + __shfl_xor(gt, 16), eq_b = eq + __shfl_xor(eq, 16);
        gt_b += __shfl_xor(gt_b, 32);
        eq_b += __shfl_xor(eq_b, 32);
        const int need = K - gt_b;   // keys == T to take, 1 <= need <= eq_b
        // taken ties as bits (t * 4 + i) of this lane; all of them unless the block has more ties than it needs
        uint32_t take[3] = { 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu };
        if (__any(eq_b > need)) {   // rare: walk the block's slots in ascending order (t, q, i) and take the first `need`
            take[0] = take[1] = take[2] = 0u;
            int rem = need;
#pragma unroll
            for (int t = 0; t < ZT_PR_U / 16; t++)
#pragma unroll
                for (int qq = 0; qq < 4; qq++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const bool took = q == qq && key[t][i] == T && rem > 0;
                        if (took) take[(t * 4 + i) >> 5] |= 1u << ((t * 4 + i) & 31);
                        rem -= __shfl(took ? 1 : 0, r + 16 * qq);
                    }
        }
        // this lane's selections, their offset among the block's four lanes, then the stores
        int mine = 0;
#pragma unroll
        for (int t = 0; t < ZT_PR_U / 16; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) mine += (key[t][i] > T || (key[t][i] == T && ((take[(t * 4 + i) >> 5] >> ((t * 4 + i) & 31)) & 1u))) ? 1 : 0;
        const int m16 = __shfl_xor(mine, 16), m32 = __shfl_xor(mine, 32), m48 = __shfl_xor(m16, 32);
        // exclusive prefix over q: lane q's partners are q ^ 1 (m16), q ^ 2 (m32), q ^ 3 (m48)
        const int off = q == 0 ? 0 : q == 1 ? m16 : q == 2 ? m32 + m48 : m16 + m32 + m48;
        if (valid) {
            uint32_t *dst = P.cand + (uint64_t)i_list * K + off;
#pragma unroll
            for (int t = 0; t < ZT_PR_U / 16; t++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (key[t][i] > T || (key[t][i] == T && ((take[(t * 4 + i) >> 5] >> ((t * 4 + i) & 31)) & 1u))) *dst++ = (uint32_t)prop_slot(16 * t + 4 * q + i);
        }
    }
}
}  // namespace

// K <= 4 and K <= 16
extern "C" __global__ void __launch_bounds__(512) zt_mlp_propose4(zt_propose_args P) { propose_body<4>(P); }
extern "C" __global__ void __launch_bounds__(512) zt_mlp_propose16(zt_propose_args P) { propose_body<16>(P); }
extern "C" __global__ void __launch_bounds__(256) zt_mlp_propose_wide(zt_wide_args P) { propose_wide_body(P); }
