"""The persistent batch kernel's task queue, planned on the host.  zaptex_batch_workspace_size runs the planner and
its deadlock check without touching a device: every task that waits must find all the tasks counting towards its
counter earlier in the queue (so the earliest unfinished task can always run), and every task must be queued once.
A plan that fails the check is refused (size 0) and never reaches the GPU."""
import ctypes as C

import pytest

from zap_amd import batch, binding

OPTS = {"persistent": 6, "mix": 7, "chunks": 8, "groups": 9}


def textures(specs):
    """zaptex_texture descriptors with placeholder addresses (the planner reads sizes and formats only)."""
    arr = (binding.ZapTexture * len(specs))()
    for k, (w, h, fmt, levels) in enumerate(specs):
        arr[k] = binding.ZapTexture(0x10000000, w * 4, w, h, binding.fmt_id(fmt), 0, levels, 0x20000000, 0x30000000)
    return arr


def plan_ok(zt, specs):
    n = zt.lib.zaptex_batch_workspace_size(textures(specs), len(specs))
    assert n > 0, zt.lib.zaptex_last_error().decode()
    return n


@pytest.mark.parametrize("mix,chunks", [(2, 1), (0, 1), (5, 4), (64, 64)])
def test_c4_plan_passes_the_deadlock_check(zt, mix, chunks):
    specs = [(w, h, f, batch.mip_levels(w, h)) for (w, h, f, _s, _k) in batch.c4_specs(512, 4)]
    saved = {k: zt.get_option(v) for k, v in OPTS.items()}
    try:
        zt.set_option(OPTS["mix"], mix)
        zt.set_option(OPTS["chunks"], chunks)
        plan_ok(zt, specs)
    finally:
        for k, v in saved.items():
            zt.set_option(OPTS[k], v)


@pytest.mark.parametrize("specs", [
    [(4096, 4096, "bc7", 13)],                                   # one chain, every step kind
    [(64, 64, "bc1", 1), (5, 3, "bc7", 1)],                      # no mips at all
    [(300, 200, "bc1", 9), (1000, 750, "bc5", 10)],              # BC1 family only, odd sizes
    [(1, 4096, "bc7", 13), (4096, 1, "bc3", 13), (129, 2, "bc4", 8)],   # single-level steps (a side of 1)
    [(257, 4096, "bc7", 5), (33, 17, "bc1", 6), (2048, 64, "bc7", 12)],
])
def test_edge_plans_pass_the_deadlock_check(zt, specs):
    plan_ok(zt, specs)


def test_options_are_checked(zt):
    assert zt.lib.zaptex_set_option(OPTS["mix"], 65) == 1
    assert zt.lib.zaptex_set_option(OPTS["chunks"], 0) == 1
    assert zt.lib.zaptex_set_option(OPTS["groups"], 33) == 1
    assert zt.get_option(OPTS["persistent"]) == 0 and zt.get_option(OPTS["mix"]) == 2 and zt.get_option(OPTS["chunks"]) == 1


def test_workspace_size_follows_the_launch_shapes(zt):
    """The task queue is planned from the launch shapes, so the workspace a plan needs changes with
    zaptex_set_launch: Batch.replan sizes it again (the GPU test of a re-plan relies on this growth)."""
    specs = [(64, 64, "bc1", 7), (100, 36, "bc7", 7), (257, 129, "bc3", 9), (1024, 512, "bc7", 11), (512, 512, "bc1", 10)]
    names = ["zt_bc14_encode", "zt_bc1_climb", "zt_bc7_stage6", "zt_bc7_stage5", "zt_bc7_s1", "zt_mip_batch"]
    saved = {k: zt.get_launch(k) for k in names}
    try:
        for k in names:
            zt.set_launch(k, 256, 1)
        small = plan_ok(zt, specs)
        for k in names:
            zt.set_launch(k, 64, 1 << 20)
        assert plan_ok(zt, specs) > small
    finally:
        for k, (wg, cap) in saved.items():
            zt.set_launch(k, wg, cap)
