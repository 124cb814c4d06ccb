/* zaptex.h — C ABI of libzaptex, the gfx950 (MI355X) implementation of zap_tex.h's encoders and mips.
 *
 * Plain C types only; every function returns a status (ZAPTEX_OK = 0) and never falls back to the CPU:
 * without a usable HIP device the call fails with ZAPTEX_E_DEVICE and zaptex_last_error() says why.
 * Output bytes equal the reference encoder's (WatchDogStudios/zap 1.3.0, zap_tex.h), rdo > 0 included.
 *
 * Two families of entry points:
 *   host  — pointers are ordinary (pageable) host memory, exactly like zap_tex.h; the call is synchronous.
 *           Each replaces the zap_tex.h function cited beside it.  Thread-safe: concurrent calls run side by
 *           side on the device (up to 8 per device, each with its own stream and staging buffers; more wait).
 *           Any size the kernels can index: sides up to 2^30, fewer than 2^31 blocks.
 *   _dev  — pointers are device memory on the current HIP device; work is enqueued on `stream`
 *           (a hipStream_t, NULL = the default stream) and the call returns without waiting.
 */
#ifndef ZAPTEX_H
#define ZAPTEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ZAPTEX_OK = 0,
    ZAPTEX_E_ARG = 1,         /* invalid argument (size, format, alignment, workspace too small) */
    ZAPTEX_E_DEVICE = 2,      /* no HIP device, or a HIP runtime / kernel launch error */
    ZAPTEX_E_NOMEM = 3,       /* device or pinned allocation failed */
    ZAPTEX_E_UNSUPPORTED = 4  /* a reference feature outside this library (none of zap_tex.h's entry points today) */
};

/* formats: zap_bc_format values (zap_tex.h:32), plus the two non-BCn encoders */
enum { ZAPTEX_BC1 = 0, ZAPTEX_BC3 = 1, ZAPTEX_BC4 = 2, ZAPTEX_BC5 = 3, ZAPTEX_BC7 = 4, ZAPTEX_BC6H = 5, ZAPTEX_ASTC = 6 };

/* options */
enum {
    ZAPTEX_OPT_BC7_M1K = 1,  /* mode-1 partitions fully fitted after ranking, 1..64 (zap_tex.h:750-752, default 4) */
    ZAPTEX_OPT_PROFILE = 2,  /* 1: time every kernel with HIP events on its launch stream (zaptex_profile_*) */
    ZAPTEX_OPT_CLIMB_REFILL = 3, /* BC1 hill climb: idle lanes (1..64) a wave gathers before it refills them */
    ZAPTEX_OPT_RDO_SEGMENT = 4,  /* rdo > 0: blocks per speculative segment, 16..2^24, or 0 = automatic (default) */
    ZAPTEX_OPT_RDO_WARMUP = 5,   /* rdo > 0: blocks a segment's speculative walk starts early, 0..2^24 (default 512);
                                    segment and warm-up change speed only, never the bytes */
    /* batches (zaptex_batch_create reads these; none changes the bytes): */
    ZAPTEX_OPT_BATCH_PERSISTENT = 6, /* 1: one persistent kernel over a task queue; 0 (default): a graph of the
                                        per-stage kernels (also used when ZAPTEX_OPT_BC7_M1K > 8) */
    ZAPTEX_OPT_BATCH_MIX = 7,        /* level-0 encode tasks queued per round-0 mip task, 0..64 (default 2) */
    ZAPTEX_OPT_BATCH_CHUNKS = 8,     /* 256-block chunks per BC1-family / BC7 stage-6 task, 1..64 (default 1) */
    ZAPTEX_OPT_BATCH_GROUPS = 9,     /* persistent workgroups per CU, 0 (default: as many as fit) or 1..32 */
    ZAPTEX_OPT_MIP_TILE = 10,        /* level-1 tile side of the fused mip passes: 32 (default, 6 levels a pass) or 64
                                        (7 levels a pass); the bytes are the same either way */
    ZAPTEX_OPT_CLIMB4_MAX = 11,      /* a launch with at most this many BC1/BC3/BC4/BC5 blocks runs the BC1 starts with two
                                        lanes per block and the hill climb with four (default 2^20; 0: always one lane
                                        per block) */
    ZAPTEX_OPT_CLIMB4_RANGE = 12,    /* blocks per wave of the grouped climb, a multiple of 8 (of 64 / lanes per block),
                                        or 0 = automatic */
    ZAPTEX_OPT_CLIMB_GROUP = 13,     /* lanes per block of the grouped climb: 4 (default) or 8 */
    ZAPTEX_OPT_S6_CHUNKS = 14        /* BC7 mode-6 stage: 0 = one pass per block, k = 1..64: the waves take k chunks of 64
                                        blocks at a time until none are left, and run the least-squares refits that can
                                        change a block over 64 queued blocks at a time (default 2) */
};

/* ---------------- host entry points (drop-in bodies of zap_tex.h) */

/* zap_tex.h:1474 zap_bc_encode, rdo > 0 included (zap_tex.h:223-306, :786-815). */
int zaptex_bc_encode(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, void *out);
/* zap_tex.h:1526 zap_bc_decode */
int zaptex_bc_decode(const void *bc, int w, int h, int fmt, uint8_t *rgba, size_t stride);
/* zap_tex.h:946 zap_bc6h_encode (stride in floats per row) */
int zaptex_bc6h_encode(const float *rgba, int w, int h, size_t stride, void *out);
/* zap_tex.h:963 zap_bc6h_decode (stride in uint16 per row) */
int zaptex_bc6h_decode(const void *bc, int w, int h, uint16_t *rgba, size_t stride);
/* zap_tex.h:1291 zap_astc_encode */
int zaptex_astc_encode(const uint8_t *rgba, int w, int h, size_t stride, void *out);
/* zap_tex.h:1344 zap_astc_decode */
int zaptex_astc_decode(const void *bc, int w, int h, uint8_t *rgba, size_t stride);
/* zap_tex.h:1442 zap_mip_next */
int zaptex_mip_next(const uint8_t *src, int w, int h, size_t stride, int srgb, uint8_t *dst, size_t dst_stride);
/* zap_tex.h:1460 zap_mip_next_f (strides in floats per row) */
int zaptex_mip_next_f(const float *src, int w, int h, size_t stride, float *dst, size_t dst_stride);
/* zap_tex.h:1570-1571 zap_bc_split / zap_bc_merge */
int zaptex_bc_split(const void *bc, size_t n, int fmt, void *out);
int zaptex_bc_merge(const void *planes, size_t n, int fmt, void *out);

/* ---------------- texture files (zap.c:309-355, the `zap tex` command; zap.c is the reference's CLI, not zap_tex.h)
 * The bytes `zap tex` writes for an RGBA8 image: for BC1/BC3/BC4/BC5/BC7 a DDS file (the DX10 extension for BC7 and
 * sRGB formats, DXGI 71/77/80/83/98, +1 for sRGB) holding level 0 and, with `mips`, every level of the box chain
 * down to 1x1 (filtered like zap_mip_next with `srgb`; BC4/BC5 are never sRGB); for ZAPTEX_ASTC a .astc file
 * holding one level (`mips` ignored, as there).  The chain is generated and encoded on the device. */
size_t zaptex_tex_file_size(int w, int h, int fmt, int mips, int srgb);   /* 0 for invalid arguments */
int zaptex_tex_file(const uint8_t *rgba, int w, int h, size_t stride, int fmt, float rdo, int mips, int srgb, void *out, size_t out_size);

/* ---------------- device-resident entry points (no reference counterpart: the reference is CPU-only;
 * these are what the host entry points run, exposed for callers whose data already lives in HBM) */

/* One image to encode: RGBA8 pixels (RGBA float for ZAPTEX_BC6H), `stride` bytes per row; blocks written
 * row-major to dst.  src rows must be 4-byte aligned (16 for BC6H), dst 16-byte aligned. */
typedef struct {
    const void *src;
    void *dst;
    size_t stride;
    int w, h, fmt;
} zaptex_image;

/* workspace bytes zaptex_encode_dev needs for these images */
size_t zaptex_workspace_size(const zaptex_image *imgs, int n);
/* Encode n (1..16) images, any mix of formats, on `stream`.  The workspace is device memory of at least
 * zaptex_workspace_size bytes (256-byte aligned), owned by the caller; calls sharing a workspace must not
 * overlap. */
int zaptex_encode_dev(const zaptex_image *imgs, int n, void *workspace, size_t ws_bytes, void *stream);
/* zap_tex.h:1474 zap_bc_encode with rdo (zap_tex.h:223-306, :786-815) on device memory: zaptex_encode_dev, then the
 * rate-distortion pass of every image (BCn formats only).  Each image is its own zap_bc_encode call: candidates come
 * from its own previous blocks, in raster order.  The pass cuts every 8-byte block stream into segments walked in
 * parallel from guessed starting states, then checks the guesses in order and re-walks a segment from the true state
 * until it rejoins the guess, so the bytes are the reference's whatever the guesses.  The workspace is device memory
 * of zaptex_rdo_workspace_size bytes (zaptex_workspace_size when rdo <= 0). */
size_t zaptex_rdo_workspace_size(const zaptex_image *imgs, int n);
int zaptex_encode_rdo_dev(const zaptex_image *imgs, int n, float rdo, void *workspace, size_t ws_bytes, void *stream);
/* zaptex_encode_dev without a workspace argument and without any allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The workspace is a scratch array inside the library, ZAPTEX_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends; images whose zaptex_workspace_size exceeds it are rejected with
 * ZAPTEX_E_ARG.  Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (another thread,
 * another stream) race and must be ordered by the caller.  Kernel timing (ZAPTEX_OPT_PROFILE) is skipped. */
#define ZAPTEX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int zaptex_encode_dev_noalloc(const zaptex_image *imgs, int n, void *stream);
/* Exhaustive all-mode BC7 search (no reference counterpart: the reference tries modes 6, 5 and 1 only).
 * 285 candidates per block (mode*64 + slot); never worse than zaptex_encode_dev's reference search.
 * best_cand (device, one uint32 per block of all images in order, may be NULL) receives the chosen candidate;
 * deltas (device, 512 uint16 per block in candidate-slot order, may be NULL) receives every candidate's
 * min(65535, SSE - best SSE), 65535 in the 227 unused slots.  The workspace (device memory of at least
 * zaptex_bc7_full_workspace_size bytes, with_deltas = whether deltas will be asked for) is the caller's. */
size_t zaptex_bc7_full_workspace_size(const zaptex_image *imgs, int n, int with_deltas);
int zaptex_bc7_full_dev(const zaptex_image *imgs, int n, uint32_t *best_cand, uint16_t *deltas, void *ws, size_t ws_bytes, void *stream);
/* ---------------- learned BC7 proposal (no reference counterpart)
 * A 64 -> 128 -> 128 -> 512 ReLU MLP (BF16 on MFMA, FP32 accumulate) scores the exhaustive search's candidate
 * slots from the block's pixels; the top-k are encoded exactly and the best kept.  One unused output slot (385)
 * predicts log(1 + the SSE the top-k give away against the exhaustive optimum); blocks whose predicted SSE
 * exceeds `threshold` are searched exhaustively instead (threshold < 0: every block; huge: none).  The network
 * state is one device buffer of zaptex_mlp_state_size() bytes: fp32 parameters
 * [W1 64x128 | b1 | W2 128x128 | b2 | W3 128x512 | b3], gradients, Adam moments, tables and the bf16 weight
 * copies. */
size_t zaptex_mlp_state_size(void);
size_t zaptex_mlp_workspace_size(int batch);
int zaptex_mlp_init(void *state, uint32_t seed, void *stream);
/* load fp32 parameters (90,880 floats, host memory) into a state buffer */
int zaptex_mlp_set_params(void *state, const float *host_params, void *stream);
/* input normalisation of the network (kept in the state; zaptex_mlp_init sets 0), per 4x4 block, m_c = channel
 * mean, R = largest channel range: 0 v/255; 1 (v - m_c)/64; 2 (v - m_c)/max(1, R); 3 (v - m_c)/sqrt(64 max(1, R));
 * 4 colour as 3 with R over R, G, B only, alpha (a - 255)/255 */
int zaptex_mlp_set_features(void *state, int mode, void *stream);
/* logits [batch x 512] for batch blocks of 16 RGBA8 pixels (64 bytes each) */
int zaptex_mlp_forward(const void *state, const void *blocks, int batch, float *logits, void *ws, size_t ws_bytes, void *stream);
/* The proposal itself, fused (features, three layers and top-k in one kernel, no workspace): cand [n x topk] gets
 * each block's topk candidate slots of largest logit among the 285 the search uses, largest first (logits are
 * ranked on their float32 bits above the 9 lowest, so logits equal down to those bits rank by ascending slot), risk [n] = exp(logit 385) - 1 (0 below 0, 2^20 from 13.9 up).  The logits are bit-identical to
 * zaptex_mlp_forward's.  topk 1..16. */
int zaptex_mlp_propose(const void *state, const void *blocks, int n, int topk, uint32_t *cand, float *risk, void *stream);
/* test hook: the fused kernel's logits of the kept output rows (the 285 candidate slots ascending, the risk slot,
 * two zero rows) [n x 288] */
int zaptex_mlp_propose_logits(const void *state, const void *blocks, int n, float *logits, void *stream);
/* Training targets (device arrays, one row per block):
 *   labels       uint32 slot per block: one-hot cross-entropy targets (used when temperature <= 0);
 *   deltas       zaptex_bc7_full_dev's 512 uint16 per block (may be NULL without temperature or risk_k);
 *   temperature  > 0: soft targets softmax(-delta / temperature) over the used slots instead of the labels;
 *   risk_k       1..16: the risk slot also learns log(1 + min delta over the current top-risk_k slots), squared
 *                error weighted by risk_weight; 0: off. */
typedef struct {
    const uint32_t *labels;
    const uint16_t *deltas;
    float temperature;
    int risk_k;
    float risk_weight;
} zaptex_mlp_targets;
/* Loss gradients (mean over the batch) into the state's gradient area; stats (device, 2 floats, may be NULL)
 * receive this batch's (mean loss, top-1 hits), a hit being the label, or with soft targets a slot of delta 0. */
int zaptex_mlp_gradients(void *state, const void *blocks, const zaptex_mlp_targets *targets, int batch, void *ws, size_t ws_bytes,
                         float *stats, void *stream);
/* gradients then one Adam step (step counts from 1) */
int zaptex_mlp_train_step(void *state, const void *blocks, const zaptex_mlp_targets *targets, int batch, float lr, int step, void *ws,
                          size_t ws_bytes, float *stats, void *stream);
/* every block of the images as 64 contiguous bytes (training inputs) */
int zaptex_bc7_gather_dev(const zaptex_image *imgs, int n, void *blocks, void *stream);
/* The learned encode.  Blocks whose predicted risk exceeds `threshold` fall back: with fallback_k = 0 to the
 * exhaustive search, with fallback_k = 1..285 to the network's top fallback_k slots (largest logits, ties to the
 * lowest slot) encoded exactly; fallback_k = 285 evaluates every candidate and equals the exhaustive search. */
size_t zaptex_bc7_learned_workspace_size(const zaptex_image *imgs, int n, int topk, int fallback_k);
int zaptex_bc7_learned_dev(const zaptex_image *imgs, int n, const void *state, int topk, float threshold, int fallback_k, void *ws,
                           size_t ws_bytes, void *stream);

/* ---------------- batch: many textures with their mip chains in one graph launch (no reference counterpart:
 * the reference encodes one image per call and its callers loop)
 * A texture: level 0 in device memory; `levels` levels are encoded (1..zap_mip_levels(w, h)), levels 1.. generated
 * with zap_mip_next's box filter (`srgb` as there) into the `mips` arena (zaptex_batch_mip_bytes; level l at the
 * sum of the 256-byte-aligned tight sizes of levels 1..l-1), and every level's blocks go to the `out` arena
 * (zaptex_batch_out_bytes; level l at the sum of the 256-byte-aligned block sizes of levels 0..l-1).  Arenas
 * 256-byte aligned; formats BC1, BC3, BC4, BC5, BC7.  Bytes equal zap_mip_next + zap_bc_encode per level. */
typedef struct {
    const void *base;
    size_t stride;
    int w, h, fmt, srgb, levels;
    void *mips;
    void *out;
} zaptex_texture;
typedef struct zaptex_batch zaptex_batch;
size_t zaptex_batch_mip_bytes(int w, int h, int levels);
size_t zaptex_batch_out_bytes(int w, int h, int fmt, int levels);
/* the workspace a plan of these textures needs under the launch shapes now in force (zaptex_set_launch): a plan
 * made after the shapes change can need more */
size_t zaptex_batch_workspace_size(const zaptex_texture *t, int n);
/* Builds the plan: device job tables (uploaded once into the caller's workspace) and one captured graph:
 * the mip chains of every texture (each chain cut into fused tile passes, single levels and a <= 128x128 tail;
 * step s of all textures shares the launches of round s), then the BC1/3/4/5 encoder, the BC1 hill climb and the
 * BC7 stages over all blocks.  The launch shapes in force (zaptex_set_launch) are baked in.  The texture buffers
 * and the workspace must outlive the plan. */
int zaptex_batch_create(const zaptex_texture *t, int n, void *ws, size_t ws_bytes, zaptex_batch **out);
/* one launch of the graph on `stream`; runs of one plan must not overlap (they share the workspace) */
int zaptex_batch_run(zaptex_batch *b, void *stream);
void zaptex_batch_destroy(zaptex_batch *b);
/* launch shape of a grid-stride kernel ("zt_bc14_encode", "zt_bc1_climb", "zt_bc7_stage6", "zt_bc7_stage5",
 * "zt_bc7_s1", "zt_mip_batch"): workgroup size 64/128/256 and the cap on workgroups (the autotuner sets these) */
int zaptex_set_launch(const char *kernel, int wg, int max_groups);
int zaptex_get_launch(const char *kernel, int *wg, int *max_groups);

/* zap_mip_next on device memory */
int zaptex_mip_next_dev(const void *src, int w, int h, size_t stride, int srgb, void *dst, size_t dst_stride, void *stream);
/* zap_mip_next_f on device memory (strides in floats) */
int zaptex_mip_next_f_dev(const void *src, int w, int h, size_t stride, void *dst, size_t dst_stride, void *stream);
/* Whole chain: levels[i] receives level i+1 (tightly packed rows), i < zap_mip_levels(w, h) - 1. */
int zaptex_mip_chain_dev(const void *base, int w, int h, size_t stride, int srgb, void *const *levels, int nlevels, void *stream);
/* zap_bc_decode on device memory (out rows 4-byte aligned) */
int zaptex_bc_decode_dev(const void *bc, int w, int h, int fmt, void *rgba, size_t stride, void *stream);

/* ---------------- library */
int zaptex_set_option(int key, int value);
int zaptex_get_option(int key);
const char *zaptex_last_error(void);  /* message of this thread's last failed call */
int zaptex_device_count(void);
const char *zaptex_version(void);

/* ---------------- kernel timing (no reference counterpart; for benchmarks) */
int zaptex_profile_collect(void);  /* wait for the recorded events and fold them into per-kernel totals */
int zaptex_profile_get(const char *kernel, double *ms_total, int *launches);
void zaptex_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif
