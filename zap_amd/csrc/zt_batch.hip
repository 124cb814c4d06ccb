// zt_batch.hip — the persistent batch kernel: a whole batch (the mip chains, BC1-family and BC7 encodes of many
// textures) in one launch over a task queue built by the host (zt_kernels.h: zt_seg, zt_batch_args).
//
// This file is the one translation unit of the mip, BC1-family and BC7 kernels: it includes their sources, so the
// batch kernel runs their bodies directly and every kernel shares one copy of the device tables.
//
// Each workgroup takes the next task (one atomic), waits for the task's dependency counter if it has one (an
// agent-scope acquire: the producers may sit on other XCDs, whose L2 caches are separate), runs the body, and
// publishes it with one agent-scope release (an L2 write-back) on its segment's counter.  Tasks wait only for
// tasks earlier in the queue, and a task is taken only by a running workgroup, so the earliest unfinished task
// can always run: the queue drains with any grid size, and every workgroup leaves when the queue is empty.  A wait
// is bounded all the same (kWaitSpins); one that runs out is counted in ctl[2], the task goes ahead, and every
// later wait gives up at once.
#include "zt_mip.hip"
#include "zt_bc14.hip"
#include "zt_bc7.hip"

namespace {
constexpr uint32_t kWaitSpins = 1u << 22;   // about a second of s_sleep 8: far beyond any task's duration

ZT_DEV void run_task(const zt_batch_args &A, const zt_seg &S, uint32_t vb) {
    switch (S.kind) {
    case ZT_T_MIP_TILE: tiles_item_body<64>((const zt_mipstep *)S.a, (const uint2 *)S.b, vb); break;
    case ZT_T_MIP_LEVEL: mip_batch_body((const zt_mipjob *)S.a, (int)S.n, S.total, (int)S.srgb, vb, S.vgrid); break;
    case ZT_T_MIP_TAIL: tail_step_body((const zt_mipstep *)S.a, vb); break;
    case ZT_T_BC14: bc14_encode_body(A.j14, A.tot14, A.climb, vb, S.vgrid); break;
    case ZT_T_CLIMB: bc1_climb_body(A.j14, A.tot14, A.climb, A.refill_min, A.range, vb, S.vgrid); break;
    case ZT_T_FINISH: bc1_finish_body(A.j14, A.tot14, A.climb, vb, S.vgrid); break;
    case ZT_T_S6: stage6_body(A.j7, A.tot7, A.L, vb, S.vgrid); break;
    case ZT_T_S5: stage5_body(A.j7, A.L, vb, S.vgrid); break;
    case ZT_T_RANK: s1_rank_body(A.L, A.m1k, A.L.cand1, vb, S.vgrid); break;
    case ZT_T_FIT: s1_fit_body(A.L, A.m1k, A.L.cand1, A.L.sub_err, A.L.sub_state, vb, S.vgrid); break;
    default: s1_pick_body(A.j7, A.L, A.m1k, A.L.cand1, A.L.sub_err, A.L.sub_state, vb, S.vgrid); break;
    }
}
}  // namespace

extern "C" __global__ void __launch_bounds__(256) zt_batch_persistent(zt_batch_args A) {
    __shared__ uint32_t task, last;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t t = __hip_atomic_fetch_add(A.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t < A.ntasks) {
                const zt_seg &S = A.segs[A.queue[t] >> ZT_TASK_SEG_SHIFT];
                if (S.wait != ZT_NO_WAIT) {
                    const unsigned *c = A.ctl + ZT_CTL_COUNTERS + S.wait;
                    for (uint32_t spins = 0; __hip_atomic_load(c, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < S.need;) {
                        if (++spins == kWaitSpins) {
                            atomicAdd(A.ctl + 2, 1u);
                            break;
                        }
                        // after one time-out every wait gives up at once: a broken plan costs a second, not a hang
                        if (__hip_atomic_load(A.ctl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                        __builtin_amdgcn_s_sleep(8);
                    }
                }
            }
            task = t;
        }
        __syncthreads();
        // uniform across the workgroup: the task's descriptors then come through scalar loads and the switch below
        // is a scalar branch (a task kind the compiler took for divergent would have every case's registers live)
        const uint32_t t = __builtin_amdgcn_readfirstlane(task);
        if (t >= A.ntasks) break;
        const uint32_t q = A.queue[t];
        const zt_seg &S = A.segs[q >> ZT_TASK_SEG_SHIFT];
        const uint32_t v0 = S.vbase + (q & ((1u << ZT_TASK_SEG_SHIFT) - 1u)) * S.per, v1 = min(S.vbase + S.count, v0 + S.per);
        for (uint32_t vb = v0; vb < v1; vb++) {
            if (vb != v0) __syncthreads();
            run_task(A, S, vb);
        }
        // every wave's stores written back to memory before the barrier, then the task is published
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_fetch_add(A.ctl + ZT_CTL_COUNTERS + S.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the last workgroup out (every task has finished) clears the queue state for the next launch
    if (threadIdx.x == 0) last = __hip_atomic_fetch_add(A.ctl + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    __syncthreads();
    if (last) {
        for (uint32_t i = threadIdx.x; i < A.nctl; i += blockDim.x)
            if (i != 2) A.ctl[i] = 0;
        if (A.L.counters)
            for (uint32_t i = threadIdx.x; i < ZT_BC7_COUNTERS; i += blockDim.x) A.L.counters[i] = 0;
    }
}
