// zt_kernels.h — kernels and device tables shared between the HIP translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct zt_job;
struct zt_jobs;

// BC7 compaction lists (device memory, capacity = total blocks of the launch)
#define ZT_BC7_LIST1 (ZT_SHARDS * ZT_SHARD_STRIDE)
#define ZT_BC7_COUNTERS (2 * ZT_SHARDS * ZT_SHARD_STRIDE)
#define ZT_BC7_S6_CURSOR (ZT_BC7_LIST1 + 32)   // zt_bc7_stage6q's cursors, one per shard: a word of each of list 1's lines,
                                               // which stage 6 does not use
struct zt_bc7_lists {
    unsigned *counters;  // sharded lengths (ZT_SHARDS x ZT_SHARD_STRIDE): list 5 from 0, list 1 from ZT_BC7_LIST1
    uint32_t cap5, cap1; // slots per shard
    uint4 *meta5;        // by list-5 slot: (global block, mode-6 SSE, job, opaque)
    uint4 *px5;          // by list-5 slot: 4 x uint4 = the block's 16 RGBA8 pixels
    uint4 *meta1;        // by list-1 slot: (list-5 slot, current best SSE, 0, 0)
    uint8_t *cand1;      // 8 ranked partitions per list-1 entry (flat entry index, like sub_err and sub_state)
    int *sub_err;        // 16 subset fits per list-1 entry
    uint4 *sub_state;
};

// proposal network inference (zt_mlp.hip): dense blocks -> top-K candidate slots and the predicted risk per block
struct zt_propose_args {
    const uint4 *blocks;   // 64 B per block
    uint32_t n;
    int topk;
    const uint16_t *w;     // bf16 weight copies [out x in]: W1 at 0, W2 at 64 * 128, W3 at 64 * 128 + 128 * 128
    const float *b1, *b2, *b3;
    const int *cfg;        // [0] feature normalisation
    uint32_t *cand;        // [n * topk], or null
    float *risk;           // [n], or null
    float *logits;         // [n * 288] logits of the kept W3 rows (the candidate slots ascending, then the risk slot), or null
};
// the wide proposal of a block list (zt_mlp.hip): k candidate slots per listed block
struct zt_wide_args {
    const uint4 *blocks;      // 64 B per block
    const uint32_t *list;     // block indices
    const unsigned *count;    // list length (device)
    int k;                    // 1..285
    const uint16_t *w;
    const float *b1, *b2, *b3;
    const int *cfg;
    uint32_t *cand;           // [count * k], unordered per block
};
// bf16 GEMM (zt_mlp.hip): C(m, n) = sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]
struct zt_gemm_args {
    const uint16_t *A, *B;
    int64_t sam, sak, sbk, sbn;
    float *Cf;              // fp32 output (or accumulation target), may be null
    uint16_t *Cb;           // bf16 output, may be null
    int64_t ldc;
    const float *bias;      // per column, may be null
    const uint16_t *mask;   // bf16 activation; output zeroed where it is 0 (ReLU backward), may be null
    int64_t ldm;
    int M, N, K, relu, accumulate, pad;
};

// rate-distortion streams (zt_rdo.hip): one per image, or two for BC3 / BC5 (one per 8-byte half).  A stream is cut
// into segments of `seg` blocks walked speculatively in parallel (zt_rdo_spec), then checked and repaired in order
// (zt_rdo_fix).
enum { ZT_RDO_BC1 = 0, ZT_RDO_BC4 = 1, ZT_RDO_BC7 = 2 };
struct zt_rdo_state {           // the state a segment starts from: its 16 predecessors' bytes and the match distances
    uint4 ring[16];
    uint32_t dist, pad[3];
};
struct zt_rdo_stream {
    const uint8_t *src;         // RGBA8 rows
    uint8_t *out;               // the image's blocks (final)
    const uint8_t *plain;       // a copy of the image's plain encoding, laid out like out
    uint16_t *dists;            // per block: the distances (d0 | d1 << 8) its speculative decision left
    zt_rdo_state *states;       // per segment 1..: the state it assumed
    uint32_t *seg_end;          // per segment: the distances it left
    uint4 *cand;                // per block and lane 0..48: the candidate bytes of the speculative decision
    int *cand_err;              //   and its SSE (-1: not a valid candidate)
    uint64_t stride;
    int w, h, bw, bh;
    int bs, off;                // block stride (8 or 16) and this stream's byte offset in a block (0 or 8)
    int kind, chan, force4;     // chan: BC4's channel; force4: BC1 half of BC3 (4-colour palette always)
    uint32_t seg, warm, nseg;   // segment length (>= 16), warm-up blocks, segments
};
#define ZT_RDO_STREAMS 20
#define ZT_RDO_LANES 49         // candidates per block: 48 from the previous 16 blocks, and the plain block
struct zt_rdo_streams { zt_rdo_stream s[ZT_RDO_STREAMS]; };
extern "C" __global__ void zt_rdo_spec(zt_rdo_streams S, float lambda);
extern "C" __global__ void zt_rdo_fix(zt_rdo_streams S, float lambda);

// destination levels of a fused mip launch (packed rows, 4 bytes per pixel)
// one texture's level in a batched mip launch: source level (w x h, `stride` bytes per row) -> destination rows
// packed at (w/2) * 4 bytes; `first` = index of its first destination pixel among all jobs of the launch
struct zt_mipjob {
    const uint8_t *src;
    uint8_t *dst;
    uint64_t stride;
    uint64_t first;
    int32_t w, h, srgb, pad;
};
struct zt_mip_levels {
    uint8_t *p[16];
    int w[16], h[16];
};
// one texture's fused mip step in a batched launch: nlev levels from source level (src, w x h, stride bytes/row)
struct zt_mipstep {
    const uint8_t *src;
    uint64_t stride;
    int32_t w, h, srgb, nlev;
    zt_mip_levels L;
};

// ---- the persistent batch kernel (zt_batch.hip)
// One launch runs a whole batch (mip chains, BC1-family and BC7 encodes of many textures) as a queue of tasks, each
// `per` consecutive workgroups of one of the batch's kernels.  A segment is a run of tasks of one kind: workgroups
// vbase .. vbase + count - 1 of that kernel launched with vgrid workgroups.  Before a task starts, counter `wait`
// must reach `need`; a finished task adds 1 to counter `done`.  The host orders the queue so that every task waits
// only for tasks before it, which keeps the queue deadlock-free with any number of resident workgroups.
enum { ZT_T_MIP_TILE, ZT_T_MIP_LEVEL, ZT_T_MIP_TAIL, ZT_T_BC14, ZT_T_CLIMB, ZT_T_FINISH, ZT_T_S6, ZT_T_S5, ZT_T_S1, ZT_T_RANK,
       ZT_T_FIT, ZT_T_PICK };
#define ZT_NO_WAIT 0xFFFFFFFFu
#define ZT_TASK_SEG_SHIFT 24   // queue entry: segment << 24 | task within the segment
#define ZT_MAX_SEGS 256
struct zt_seg {
    uint32_t kind, vbase, count, per, vgrid, wait, need, done;
    uint32_t n, srgb;          // mip level jobs: job count, any sRGB
    const void *a, *b;         // mip: steps (tile, tail) or jobs (level); tile items
    uint64_t total;            // mip level jobs: destination pixels
};
// ctl words: [0] queue head, [1] finished workgroups, [2] dependency waits that timed out (kept across launches),
// [ZT_CTL_COUNTERS + i] counter i.  The last workgroup to finish clears the queue head, the finished count, the
// counters and the BC7 list counters for the next launch.
#define ZT_CTL_COUNTERS 4
struct zt_batch_args {
    const uint32_t *queue;
    uint32_t ntasks, nctl;     // tasks; ctl words in use
    const zt_seg *segs;
    unsigned *ctl;
    zt_jobs j14, j7;           // BC1/BC3/BC4/BC5 and BC7 job tables (device arrays)
    uint32_t tot14, tot7;
    uint4 *climb;
    int refill_min;
    uint32_t range;
    zt_bc7_lists L;
    int m1k;
};

extern "C" {
__global__ void zt_batch_persistent(zt_batch_args A);
__global__ void zt_bc7_stage6(zt_jobs jobs, uint32_t total, zt_bc7_lists L);
__global__ void zt_bc7_stage6q(zt_jobs jobs, uint32_t total, zt_bc7_lists L, uint32_t kc);   // 64 lanes per workgroup, any grid
__global__ void zt_bc7_stage5(zt_jobs jobs, zt_bc7_lists L);
__global__ void zt_bc7_stage1(zt_jobs jobs, zt_bc7_lists L, int m1k);
__global__ void zt_zero_u32(unsigned *p, int n);
__global__ void zt_bc7_s1_rank(zt_bc7_lists L, int m1k, uint8_t *cand);
__global__ void zt_bc7_s1_fit(zt_bc7_lists L, int m1k, const uint8_t *cand, int *sub_err, uint4 *sub_state);
__global__ void zt_bc7_s1_pick(zt_jobs jobs, zt_bc7_lists L, int m1k, const uint8_t *cand, const int *sub_err, const uint4 *sub_state);
__global__ void zt_bc7_s1_fitpick(zt_jobs jobs, zt_bc7_lists L, int m1k, const uint8_t *cand);   // M1K = 1, 2, 4 or 8
// exhaustive all-mode BC7 search: one launch per mode (mode 4: per index selector) over dense blocks, optionally a
// list of them, folding (SSE << 9 | candidate) into key[block]; zt_bc7_finish encodes the winners
#define ZT_EXH_DECL(NAME) __global__ void NAME(const uint4 *blocks, uint32_t total, const uint32_t *list, const unsigned *nlist, \
                                              uint32_t *key, uint32_t *errs);
ZT_EXH_DECL(zt_bc7_exh_m0) ZT_EXH_DECL(zt_bc7_exh_m1) ZT_EXH_DECL(zt_bc7_exh_m2) ZT_EXH_DECL(zt_bc7_exh_m3)
ZT_EXH_DECL(zt_bc7_exh_m4i0) ZT_EXH_DECL(zt_bc7_exh_m4i1) ZT_EXH_DECL(zt_bc7_exh_m5) ZT_EXH_DECL(zt_bc7_exh_m6) ZT_EXH_DECL(zt_bc7_exh_m7)
#undef ZT_EXH_DECL
__global__ void zt_bc7_finish(zt_jobs jobs, const uint4 *blocks, uint32_t total, const uint32_t *key, uint32_t *best_cand, const uint2 *items,
                              const unsigned *count);
__global__ void zt_bc7_deltas(const uint32_t *errs, const uint32_t *key, uint32_t total, uint16_t *deltas);

__global__ void zt_gemm_bf16(zt_gemm_args g);
__global__ void zt_mlp_features(const uint8_t *blocks, int n, const int *cfg, uint16_t *X);
// the proposal network's unused output slot (mode 6, slot 1) that predicts log(1 + SSE given away by a top-k search)
#define ZT_RISK_SLOT 385
__global__ void zt_mlp_xent(const float *Z, const uint32_t *labels, const uint16_t *deltas, float inv_t, int risk_k, float risk_w, int B,
                            uint16_t *dZ, float *stats);
__global__ void zt_mlp_colsum(const uint16_t *src, int M, int N, float *out);
__global__ void zt_mlp_adam(float *p, const float *gr, float *m1, float *m2, int n, float lr, float b1, float b2, float c1, float c2,
                            float eps, uint16_t *wbf, const int *bf_index);
__global__ void zt_mlp_init(float *p, int n, uint32_t seed, uint16_t *wbf, const int *bf_index, const float *scale);
__global__ void zt_mlp_propose4(zt_propose_args P);    // topk <= 4
__global__ void zt_mlp_propose16(zt_propose_args P);   // topk <= 16
__global__ void zt_mlp_propose_wide(zt_wide_args P);
__global__ void zt_mlp_refresh(const float *p, int n, uint16_t *wbf, const int *bf_index);
// learned-proposal BC7 (zt_bc7_full.hip).  A bucketing counter array: [0..7] items per mode, [8] fallback blocks,
// [24..31] each mode's run start, then one counter per (class 0..8, shard) on its own line from ZT_BUCKET_SHARDS_AT.
#define ZT_BUCKET_SHARDS_AT 64
#define ZT_BUCKET_WORDS (ZT_BUCKET_SHARDS_AT + 9 * ZT_SHARDS * ZT_SHARD_STRIDE)
__global__ void zt_bc7_gather(zt_jobs jobs, uint32_t total, uint4 *blocks);
__global__ void zt_bc7_bucket(const uint32_t *cand, const float *risk, uint32_t total, int K, float threshold, unsigned *counters,
                              uint2 *items, uint32_t pass, uint32_t *fallback, uint32_t *slot);
__global__ void zt_bc7_bucket_scan(unsigned *counters);
__global__ void zt_bc7_bucket_wide(const uint32_t *cand, const uint32_t *key, const uint32_t *list, const unsigned *count, int K,
                                   unsigned *counters, uint2 *items, uint32_t pass);
__global__ void zt_fill_u32(uint32_t *p, uint64_t n, uint32_t v);
#define ZT_PROP_DECL(NAME) __global__ void NAME(const uint4 *blocks, const uint2 *items, const unsigned *counters, uint32_t *key, uint4 *enc);
ZT_PROP_DECL(zt_bc7_prop_m0) ZT_PROP_DECL(zt_bc7_prop_m1) ZT_PROP_DECL(zt_bc7_prop_m2) ZT_PROP_DECL(zt_bc7_prop_m3)
ZT_PROP_DECL(zt_bc7_prop_m4) ZT_PROP_DECL(zt_bc7_prop_m5) ZT_PROP_DECL(zt_bc7_prop_m6) ZT_PROP_DECL(zt_bc7_prop_m7)
#undef ZT_PROP_DECL
__global__ void zt_bc7_finish_learned(zt_jobs jobs, uint32_t total, int K, const uint32_t *cand, const uint32_t *slot, const uint4 *enc,
                                      const uint32_t *key);

// largest block range per wave of zt_bc1_climb (a wave's lanes refill from its range)
#define ZT_CLIMB_RANGE 1024u
__global__ void zt_bc14_encode(zt_jobs jobs, uint32_t total, uint4 *climb);
__global__ void zt_bc14_encode2(zt_jobs jobs, uint32_t total, uint4 *climb);   // two lanes per block
__global__ void zt_bc1_climb(zt_jobs jobs, uint32_t total, uint4 *climb, int refill_min, uint32_t range);
__global__ void zt_bc1_climb4(zt_jobs jobs, uint32_t total, uint4 *climb, uint32_t range);   // four lanes per block
__global__ void zt_bc1_climb8(zt_jobs jobs, uint32_t total, uint4 *climb, uint32_t range);   // eight lanes per block
__global__ void zt_bc1_finish(zt_jobs jobs, uint32_t total, const uint4 *climb);
__global__ void zt_bc6h_encode(zt_jobs jobs, uint32_t total);
// ASTC 4x4 encode in stages (workspace regions, each sized for `total` blocks):
//   classify   gather each block once (px), list opaque and translucent blocks
//   single     every 1-partition trial in one launch, configuration-major items (K 0..2 over the opaque list,
//              3..5 over the translucent one): error and packed block per (K % 3, list entry), so a wave's
//              lanes write consecutive results
//   pick1      over the list entries: the reference's first smallest error; opaque blocks that go on to the
//              2-partition search are listed, the others are stored
//   search     2-means split and the 4 stored patterns nearest to it, 16 lanes per listed block
//   pair       every 2-partition trial in one launch: configurations 6 then 7 over (listed block, pattern) items
//   pick2      the first smallest error over the 1-partition winner and the 8 pattern trials, stored
struct zt_astc_ws {
    unsigned *counters;     // sharded list lengths (ZT_SHARDS x ZT_SHARD_STRIDE each): opaque, translucent, listed
    uint32_t cap;           // slots per shard of each list
    uint32_t n;             // blocks
    uint32_t *lo, *lt;      // block lists: opaque, translucent
    uint32_t *lp;           // opaque-list entries listed for the 2-partition search
    uint4 *px;              // [block * 4]: 16 pixels
    int *err1;              // [(configuration % 3) * n + u]: u = opaque-list entry, or n_opaque + translucent-list entry
    uint4 *blk1;            // the same
    uint32_t *cand;         // [listed * 4]: pattern mask | seed << 16 (all ones: none)
    int *err2;              // [listed * 8]: pattern j, configuration 6 + k at 2j + k
    uint4 *blk2;            // [listed * 8]
};
__global__ void zt_astc_classify(zt_jobs jobs, uint32_t total, zt_astc_ws W);
__global__ void zt_astc_single(zt_astc_ws W);
__global__ void zt_astc_pick1(zt_jobs jobs, uint32_t total, zt_astc_ws W);
__global__ void zt_astc_search(zt_astc_ws W);
__global__ void zt_astc_pair(zt_astc_ws W);
__global__ void zt_astc_pick2(zt_jobs jobs, zt_astc_ws W);

__global__ void zt_bc_decode(const uint8_t *bc, int w, int h, int fmt, uint8_t *rgba, uint64_t stride);
__global__ void zt_bc6h_decode(const uint8_t *bc, int w, int h, uint16_t *rgba, uint64_t stride);
__global__ void zt_astc_decode(const uint8_t *bc, int w, int h, uint8_t *rgba, uint64_t stride);
__global__ void zt_bc_shuffle(const uint8_t *in, uint8_t *out, uint64_t nb, int fmt, int split);

// mips: one level (any size); tile chain (levels 1..k inside a workgroup tile); float level
__global__ void zt_mip_level(const uint8_t *src, int w, int h, uint64_t stride, int srgb, uint8_t *dst, uint64_t dst_stride);
__global__ void zt_mip_tiles(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev, zt_mip_levels L);
__global__ void zt_mip_tiles32(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev, zt_mip_levels L);
__global__ void zt_mip_tail(const uint8_t *src, int w, int h, uint64_t stride, int srgb, int nlev, zt_mip_levels L);
__global__ void zt_mip_batch(const zt_mipjob *jobs, int njobs, uint64_t total, int any_srgb);
__global__ void zt_mip_tiles_batch(const zt_mipstep *steps, const uint2 *items);
__global__ void zt_mip_tiles32_batch(const zt_mipstep *steps, const uint2 *items);
__global__ void zt_mip_tail_batch(const zt_mipstep *steps);
__global__ void zt_mip_level_f(const float *src, int w, int h, uint64_t stride, float *dst, uint64_t dst_stride);
}

// host-side table uploads (called once per device by the C-ABI layer); return hipError_t
hipError_t zt_mip_upload_tables(const float lin[256], const float thr[255]);
hipError_t zt_astc_upload_tables(const void *setup, size_t bytes);
