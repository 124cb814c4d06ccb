"""Writes tests/golden/ops.json: algorithmic op counts of the workloads, from the op-counting oracle build.

One op = one scalar arithmetic, compare/select or convert operation of the reference algorithm (loads, stores
and branches count zero), summed while the oracle encodes the exact workload input.  These are the numerators
of the VALU roofline (ops / kernel time / 39.3e12) that bench.py reports.
Usage: python tests/golden/make_ops.py
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import loader  # noqa: E402
from zap_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ops.json")


def main():
    lib = loader.oracle(counting=True)
    L = lib.lib
    L.zto_ops_reset.restype = None
    L.zto_stage_ops.argtypes = [C.c_void_p]
    res = {"definition": "scalar arithmetic + compare/select + convert ops of the reference algorithm (loads/stores/branches = 0)"}
    t0 = time.time()
    c2 = synth.c2_bc7(4096, 2)
    chain = loader.oracle().mip_chain(c2, 0)
    L.zto_ops_reset()
    st = np.zeros(4, np.uint64)
    per_level = []
    blocks = 0
    for lvl in chain:
        before = int(L.zto_ops_get())
        lib.bc_encode(lvl, "bc7")
        per_level.append(int(L.zto_ops_get()) - before)
        blocks += ((lvl.shape[1] + 3) // 4) * ((lvl.shape[0] + 3) // 4)
    L.zto_stage_ops(st.ctypes.data_as(C.c_void_p))
    res["c2_bc7_chain"] = {"blocks": blocks, "pixels": int(sum(l.shape[0] * l.shape[1] for l in chain)),
                           "ops_total": int(sum(per_level)), "ops_per_level": per_level,
                           "ops_stage6": int(st[0]), "ops_stage5": int(st[1]), "ops_stage1": int(st[2]), "mode1_blocks": int(st[3])}
    print(f"c2 {time.time() - t0:.1f}s", flush=True)
    c1 = synth.c1_bc1(1024, 1)
    L.zto_ops_reset()
    lib.bc_encode(c1, "bc1")
    res["c1_bc1_1024"] = {"blocks": 65536, "ops_total": int(L.zto_ops_get())}
    with open(OUT, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
