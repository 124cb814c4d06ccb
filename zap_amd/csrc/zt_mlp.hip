// zt_mlp.hip — the BC7 proposal network on gfx950: 64 -> 128 -> 128 -> 512 MLP (ReLU), BF16 weights and
// activations on MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, trained here with its own
// forward/backward/Adam kernels.  Input: the 4x4 RGBA8 block (64 values / 255); output: one logit per BC7
// candidate slot mode*64 + slot, of which the 285 used slots are read.
#include "zt_common.h"
#include "zt_kernels.h"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace {

ZT_DEV uint16_t to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
ZT_DEV float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

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

// C(m, n) = sum_k A(m, k) B(k, n): 64x64 tile per workgroup, 32x32 per wave (2x2 MFMA 16x16x16 tiles).
// MFMA lane layout: A row = lane % 16, k = 4 (lane / 16) + v; B col = lane % 16, same k; D col = lane % 16,
// row = 4 (lane / 16) + v.  blockIdx.z splits K (accumulate mode adds into Cf atomically).
extern "C" __global__ void __launch_bounds__(256) zt_gemm_bf16(zt_gemm_args g) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    const int kper = ((g.K + (int)gridDim.z - 1) / (int)gridDim.z + 15) / 16 * 16;
    const int kb = blockIdx.z * kper, ke = min(g.K, kb + kper);
    const int r = lane & 15, q = lane >> 4;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k = kb; k < ke; k += 16) {
        bf16x4_t a[2], b[2];
        const int kk = k + 4 * q;
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = kk < ke ? load_a4(g, m0 + 16 * i + r, kk) : bf16x4_t{ 0, 0, 0, 0 };
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = kk < ke ? load_b4(g, kk, n0 + 16 * j + r) : bf16x4_t{ 0, 0, 0, 0 };
        // zero the tail of a partial last K step on both operands
        if (kk + 3 >= ke) {
#pragma unroll
            for (int v = 0; v < 4; v++) if (kk + v >= ke) { a[0][v] = a[1][v] = 0; b[0][v] = b[1][v] = 0; }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
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
ZT_DEV void block_features(const uint32_t px[16], int mode, uint16_t *x, int x_step) {
    float mean[4], inv = 1.0f / 255.0f;
    int rng = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int sum = 0, lo = 255, hi = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) { const int v = chan(px[k], c); sum += v; lo = min(lo, v); hi = max(hi, v); }
        mean[c] = mode ? (float)sum / 16.0f : 0.0f;
        if (mode != 4 || c < 3) rng = max(rng, hi - lo);
    }
    const float r = (float)max(1, rng);
    if (mode == 1) inv = 1.0f / 64.0f;
    else if (mode == 2) inv = 1.0f / r;
    else if (mode >= 3) inv = 1.0f / sqrtf(64.0f * r);
    float ainv = inv;
    if (mode == 4) { mean[3] = 255.0f; ainv = 1.0f / 255.0f; }
#pragma unroll
    for (int k = 0; k < 16; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) x[(4 * k + c) * x_step] = to_bf16(((float)chan(px[k], c) - mean[c]) * (c == 3 ? ainv : inv));
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

// top-K used candidate slots per row (ties to the lower slot) and the top-1 softmax probability
// top-K used slots per row (largest logit first, lowest slot on ties) and the row's predicted SSE cost of searching
// only them, expm1 of the risk slot (clamped to [0, 2^20])
extern "C" __global__ void __launch_bounds__(256) zt_mlp_topk(const float *Z, int B, int K, uint32_t *cand, float *risk) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= B) return;
    const float *z = Z + (int64_t)row * 512;
    float v[8];
#pragma unroll
    for (int t = 0; t < 8; t++) v[t] = cand_valid(lane + 64 * t) ? z[lane + 64 * t] : -3.0e38f;
    for (int k = 0; k < K; k++) {
        float mx = -3.0e38f;
        int am = 1 << 20;
#pragma unroll
        for (int t = 0; t < 8; t++) if (v[t] > mx) { mx = v[t]; am = lane + 64 * t; }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            float om = __shfl_xor(mx, off);
            int oa = __shfl_xor(am, off);
            if (om > mx || (om == mx && oa < am)) { mx = om; am = oa; }
        }
        if (lane == 0) cand[(int64_t)row * K + k] = (uint32_t)am;
#pragma unroll
        for (int t = 0; t < 8; t++) if (lane + 64 * t == am) v[t] = -3.0e38f;
    }
    if (lane == 0) {
        const float r = z[ZT_RISK_SLOT];
        risk[row] = r <= 0.0f ? 0.0f : r >= 13.9f ? 1048576.0f : __expf(r) - 1.0f;
    }
}

// re-round the bf16 weight copies from the fp32 master weights (after loading parameters)
extern "C" __global__ void zt_mlp_refresh(const float *p, int n, uint16_t *wbf, const int *bf_index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = bf_index[i];
    if (j >= 0) wbf[j] = to_bf16(p[i]);
}

// ---- inference: features, the three layers and the top-K proposals in one persistent kernel ----------------------
// The weights stay in LDS for the kernel's life: W1, W2 and, of W3, only the rows the proposals read (the 285
// candidate slots in ascending order, then the risk slot; 2 zero rows pad to 18 MFMA tiles).  Each wave takes
// 16 blocks at a time: features into its LDS rows, then each layer as 16x16x16 bf16 MFMAs over the same K order
// and with the same epilogue as zt_gemm_bf16, so logits, candidates and risk are bit-identical to the three GEMMs
// plus zt_mlp_topk; the logits never leave registers.
#define ZT_PR_S1 68   // LDS row strides in bf16 (padded against bank conflicts): W1
#define ZT_PR_S 132   // W2, W3 and the activation rows
#define ZT_PR_U 288   // W3 rows kept
namespace {
// candidate slot (mode * 64 + slot) of kept W3 row u: the used slots in ascending order, then the risk slot
ZT_DEV int prop_slot(int u) {
    return u < 16 ? u : u < 80 ? 64 + (u - 16) : u < 144 ? 128 + (u - 80) : u < 208 ? 192 + (u - 144) : u < 216 ? 256 + (u - 208)
         : u < 220 ? 320 + (u - 216) : u == 220 ? 384 : u < 285 ? 448 + (u - 221) : u == 285 ? ZT_RISK_SLOT : -1;
}
ZT_DEV bf16x4_t lds4(const uint16_t *p) { const uint2 t = *(const uint2 *)p; return bf16x4_t{ (short)(t.x & 0xFFFF), (short)(t.x >> 16), (short)(t.y & 0xFFFF), (short)(t.y >> 16) }; }
}  // namespace

extern "C" __global__ void __launch_bounds__(256, 1) zt_mlp_propose(zt_propose_args P) {
    __shared__ uint16_t W1s[128 * ZT_PR_S1];
    __shared__ uint16_t W2s[128 * ZT_PR_S];
    __shared__ uint16_t W3s[ZT_PR_U * ZT_PR_S];
    __shared__ float B1s[128], B2s[128], B3s[ZT_PR_U];
    __shared__ uint16_t Act[4][16 * ZT_PR_S];
    const int tid = threadIdx.x;
    for (int i = tid; i < 128 * 16; i += 256) *(uint2 *)(W1s + (i >> 4) * ZT_PR_S1 + 4 * (i & 15)) = *(const uint2 *)(P.w + (i >> 4) * 64 + 4 * (i & 15));
    for (int i = tid; i < 128 * 32; i += 256) *(uint2 *)(W2s + (i >> 5) * ZT_PR_S + 4 * (i & 31)) = *(const uint2 *)(P.w + 64 * 128 + (i >> 5) * 128 + 4 * (i & 31));
    for (int i = tid; i < ZT_PR_U * 32; i += 256) {
        const int u = i >> 5, id = prop_slot(u);
        *(uint2 *)(W3s + u * ZT_PR_S + 4 * (i & 31)) = id < 0 ? make_uint2(0, 0) : *(const uint2 *)(P.w + 64 * 128 + 128 * 128 + id * 128 + 4 * (i & 31));
    }
    for (int i = tid; i < 128; i += 256) { B1s[i] = P.b1[i]; B2s[i] = P.b2[i]; }
    for (int u = tid; u < ZT_PR_U; u += 256) { const int id = prop_slot(u); B3s[u] = id < 0 ? 0.0f : P.b3[id]; }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6, r = lane & 15, q = lane >> 4, mode = P.cfg[0], K = P.topk;
    uint16_t *A = Act[wv];
    const uint32_t ntile = (P.n + 15) / 16;
    for (uint32_t t0 = blockIdx.x * 4; t0 < ntile; t0 += gridDim.x * 4) {   // same trip count on every wave
        const uint32_t row0 = (t0 + wv) * 16;
        if (lane < 16) {
            uint32_t px[16];
            const uint32_t row = row0 + lane;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint4 v = row < P.n ? P.blocks[(uint64_t)row * 4 + k] : make_uint4(0, 0, 0, 0);
                px[4 * k] = v.x; px[4 * k + 1] = v.y; px[4 * k + 2] = v.z; px[4 * k + 3] = v.w;
            }
            block_features(px, mode, A + lane * ZT_PR_S, 1);
        }
        __syncthreads();
        // layer 1: 64 -> 128, ReLU
        f32x4_t h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            h[j] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int c = 0; c < 4; c++)
                h[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds4(A + r * ZT_PR_S + 16 * c + 4 * q), lds4(W1s + (16 * j + r) * ZT_PR_S1 + 16 * c + 4 * q), h[j], 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                float x = h[j][v] + B1s[16 * j + r];
                if (x < 0.0f) x = 0.0f;
                A[(4 * q + v) * ZT_PR_S + 16 * j + r] = to_bf16(x);
            }
        __syncthreads();
        // layer 2: 128 -> 128, ReLU
#pragma unroll
        for (int j = 0; j < 8; j++) {
            h[j] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int c = 0; c < 8; c++)
                h[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds4(A + r * ZT_PR_S + 16 * c + 4 * q), lds4(W2s + (16 * j + r) * ZT_PR_S + 16 * c + 4 * q), h[j], 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                float x = h[j][v] + B2s[16 * j + r];
                if (x < 0.0f) x = 0.0f;
                A[(4 * q + v) * ZT_PR_S + 16 * j + r] = to_bf16(x);
            }
        __syncthreads();
        // layer 3 on the kept rows: logits of rows 4q + v, columns 16 u + r
        f32x4_t z[ZT_PR_U / 16];
#pragma unroll
        for (int u = 0; u < ZT_PR_U / 16; u++) {
            z[u] = f32x4_t{ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int c = 0; c < 8; c++)
                z[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(lds4(A + r * ZT_PR_S + 16 * c + 4 * q), lds4(W3s + (16 * u + r) * ZT_PR_S + 16 * c + 4 * q), z[u], 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < ZT_PR_U / 16; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) z[u][v] += B3s[16 * u + r];
        // top-K of each row over the 285 candidate columns: largest logit first, lowest slot on ties (zt_mlp_topk)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const uint32_t row = row0 + 4 * q + v;
            for (int k = 0; k < K; k++) {
                float mx = -3.0e38f;
                int am = 1 << 20;
#pragma unroll
                for (int u = 0; u < ZT_PR_U / 16; u++) if (16 * u + r < 285 && z[u][v] > mx) { mx = z[u][v]; am = 16 * u + r; }
#pragma unroll
                for (int off = 1; off <= 8; off <<= 1) {
                    const float om = __shfl_xor(mx, off);
                    const int oa = __shfl_xor(am, off);
                    if (om > mx || (om == mx && oa < am)) { mx = om; am = oa; }
                }
                if (r == 0 && row < P.n) P.cand[(uint64_t)row * K + k] = am < 285 ? (uint32_t)prop_slot(am) : (uint32_t)am;
#pragma unroll
                for (int u = 0; u < ZT_PR_U / 16; u++) if (16 * u + r == am) z[u][v] = -3.0e38f;
            }
            if (r == 285 - 16 * 17 && row < P.n) {   // the risk column (row 285 of the kept W3) sits in tile 17
                const float x = z[17][v];
                P.risk[row] = x <= 0.0f ? 0.0f : x >= 13.9f ? 1048576.0f : __expf(x) - 1.0f;
            }
        }
        __syncthreads();
    }
}
