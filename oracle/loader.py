"""ctypes bindings for the CPU oracle (oracle/_lib) and the compiled reference (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
Both libraries expose the zap_tex.h encode/decode/mip entry points with the reference's argument meaning;
this module wraps them with numpy in / numpy out.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FMT = {"bc1": 0, "bc3": 1, "bc4": 2, "bc5": 3, "bc7": 4}
BLOCK_BYTES = {0: 8, 1: 16, 2: 8, 3: 16, 4: 16}


def build(reference: str | None = "/root/reference") -> None:
    """Compile the oracle (and oracle/_ref when the reference checkout exists)."""
    env = dict(os.environ)
    args = ["make", "-s", "-C", HERE]
    if reference:
        args.append(f"REF={reference}")
    subprocess.run(args, check=True, env=env)


def bc_size(w: int, h: int, fmt: int) -> int:
    return ((w + 3) // 4) * ((h + 3) // 4) * BLOCK_BYTES[fmt]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class TexLib:
    """One loaded implementation (prefix 'zto_' = oracle restatement, 'zref_' = reference build)."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path, self.prefix = path, prefix
        self.lib = C.CDLL(path)
        L = self.lib
        vp, i, sz, f = C.c_void_p, C.c_int, C.c_size_t, C.c_float
        p = prefix
        self._enc = getattr(L, p + "bc_encode")
        if prefix == "zref_":
            self._enc.argtypes = [vp, i, i, sz, i, f, vp]
        else:
            self._enc.argtypes = [vp, i, i, sz, i, vp]
        self._dec = getattr(L, p + "bc_decode"); self._dec.argtypes = [vp, i, i, i, vp, sz]
        self._b6e = getattr(L, p + "bc6h_encode"); self._b6e.argtypes = [vp, i, i, sz, vp]
        self._b6d = getattr(L, p + "bc6h_decode"); self._b6d.argtypes = [vp, i, i, vp, sz]
        self._ase = getattr(L, p + "astc_encode"); self._ase.argtypes = [vp, i, i, sz, vp]
        self._asd = getattr(L, p + "astc_decode"); self._asd.argtypes = [vp, i, i, vp, sz]
        self._mip = getattr(L, p + "mip_next"); self._mip.argtypes = [vp, i, i, sz, i, vp, sz]
        self._mipf = getattr(L, p + "mip_next_f"); self._mipf.argtypes = [vp, i, i, sz, vp, sz]
        self._lev = getattr(L, p + "mip_levels"); self._lev.argtypes = [i, i]; self._lev.restype = i
        self._f2h = getattr(L, p + "f2h"); self._f2h.argtypes = [f]; self._f2h.restype = C.c_uint16
        self._h2f = getattr(L, p + "h2f"); self._h2f.argtypes = [C.c_uint16]; self._h2f.restype = f
        for name in ("srgb_lin", "srgb_thr"):
            fn = getattr(L, p + name); fn.restype = C.POINTER(C.c_float)
        if prefix == "zto_":
            L.zto_bc_encode_rows.argtypes = [vp, i, i, sz, i, vp, i, i]
            L.zto_bc_encode_rdo.argtypes = [vp, i, i, sz, i, f, vp]
            L.zto_bc6h_encode_rows.argtypes = [vp, i, i, sz, vp, i, i]
            L.zto_astc_encode_rows.argtypes = [vp, i, i, sz, vp, i, i]
            L.zto_bc7_block.argtypes = [vp, vp, C.POINTER(i)]; L.zto_bc7_block.restype = i
            L.zto_bc7_mode_errors.argtypes = [vp, i, vp]
            L.zto_bc7_estimate.argtypes = [vp, i]; L.zto_bc7_estimate.restype = f
            L.zto_set_m1k.argtypes = [i]
            L.zto_bc7_candidate.argtypes = [vp, i, vp]; L.zto_bc7_candidate.restype = i
            L.zto_bc7_full_block.argtypes = [vp, vp, C.POINTER(i)]; L.zto_bc7_full_block.restype = i
            L.zto_bc7_full_encode.argtypes = [vp, i, i, sz, vp]
            L.zto_astc_partition.argtypes = [i, i, i, i]; L.zto_astc_partition.restype = i
            L.zto_ise_roundtrip.argtypes = [i, i, vp]; L.zto_ise_roundtrip.restype = i
            L.zto_bc_shuffle.argtypes = [vp, vp, sz, i, i]
            if hasattr(L, "zto_ops_get"):
                L.zto_ops_get.restype = C.c_ulonglong
        else:
            L.zref_bc7_block.argtypes = [vp, vp]; L.zref_bc7_block.restype = i
            L.zref_bc7_est1.argtypes = [vp, i]; L.zref_bc7_est1.restype = f
            L.zref_astc_part.argtypes = [i, i, i, i]; L.zref_astc_part.restype = i
            L.zref_bc_split.argtypes = [vp, sz, i, vp]
            L.zref_bc_merge.argtypes = [vp, sz, i, vp]

    # ---- encoders (img: HxWx4 uint8 / float32, C-contiguous rows; stride may exceed W*4)
    def bc_encode(self, img: np.ndarray, fmt, w: int | None = None, h: int | None = None, rdo: float = 0.0) -> np.ndarray:
        fmt = FMT.get(fmt, fmt)
        h = img.shape[0] if h is None else h
        w = img.shape[1] if w is None else w
        img = np.ascontiguousarray(img)
        out = np.zeros(bc_size(w, h, fmt), dtype=np.uint8)
        stride = img.strides[0]
        if self.prefix == "zref_":
            self._enc(_ptr(img), w, h, stride, fmt, float(rdo), _ptr(out))
        elif rdo > 0:
            self.lib.zto_bc_encode_rdo(_ptr(img), w, h, stride, fmt, float(rdo), _ptr(out))
        else:
            self._enc(_ptr(img), w, h, stride, fmt, _ptr(out))
        return out

    def bc7_full_encode(self, img: np.ndarray) -> np.ndarray:
        """Exhaustive all-mode BC7 search (project-defined; oracle only)."""
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        out = np.zeros(bc_size(w, h, 4), dtype=np.uint8)
        self.lib.zto_bc7_full_encode(_ptr(img), w, h, img.strides[0], _ptr(out))
        return out

    def bc7_candidate(self, block16x4: np.ndarray, cand: int):
        px = np.ascontiguousarray(block16x4, dtype=np.uint8)
        out = np.zeros(16, np.uint8)
        e = self.lib.zto_bc7_candidate(_ptr(px), cand, _ptr(out))
        return e, out

    def bc7_full_block(self, block16x4: np.ndarray):
        px = np.ascontiguousarray(block16x4, dtype=np.uint8)
        out = np.zeros(16, np.uint8)
        c = C.c_int()
        e = self.lib.zto_bc7_full_block(_ptr(px), _ptr(out), C.byref(c))
        return e, c.value, out

    def bc_encode_rows(self, img, fmt, out, by0, by1):
        fmt = FMT.get(fmt, fmt)
        self.lib.zto_bc_encode_rows(_ptr(img), img.shape[1], img.shape[0], img.strides[0], fmt, _ptr(out), by0, by1)

    def bc_decode(self, bc: np.ndarray, w: int, h: int, fmt) -> np.ndarray:
        fmt = FMT.get(fmt, fmt)
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._dec(_ptr(np.ascontiguousarray(bc)), w, h, fmt, _ptr(out), w * 4)
        return out

    def bc6h_encode(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img, dtype=np.float32)
        h, w = img.shape[:2]
        out = np.zeros(bc_size(w, h, 4), dtype=np.uint8)
        self._b6e(_ptr(img), w, h, img.strides[0] // 4, _ptr(out))
        return out

    def bc6h_decode(self, bc: np.ndarray, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.uint16)
        self._b6d(_ptr(np.ascontiguousarray(bc)), w, h, _ptr(out), w * 4)
        return out

    def astc_encode(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        out = np.zeros(bc_size(w, h, 4), dtype=np.uint8)
        self._ase(_ptr(img), w, h, img.strides[0], _ptr(out))
        return out

    def astc_decode(self, bc: np.ndarray, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._asd(_ptr(np.ascontiguousarray(bc)), w, h, _ptr(out), w * 4)
        return out

    def mip_levels(self, w: int, h: int) -> int:
        return self._lev(w, h)

    def mip_next(self, img: np.ndarray, srgb: int = 0) -> np.ndarray:
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        nw, nh = max(1, w // 2), max(1, h // 2)
        out = np.zeros((nh, nw, 4), dtype=np.uint8)
        self._mip(_ptr(img), w, h, img.strides[0], srgb, _ptr(out), nw * 4)
        return out

    def mip_next_f(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img, dtype=np.float32)
        h, w = img.shape[:2]
        nw, nh = max(1, w // 2), max(1, h // 2)
        out = np.zeros((nh, nw, 4), dtype=np.float32)
        self._mipf(_ptr(img), w, h, w * 4, _ptr(out), nw * 4)
        return out

    def mip_chain(self, img: np.ndarray, srgb: int = 0) -> list[np.ndarray]:
        levels = [np.ascontiguousarray(img)]
        for _ in range(self.mip_levels(img.shape[1], img.shape[0]) - 1):
            levels.append(self.mip_next(levels[-1], srgb))
        return levels

    def f2h(self, x: float) -> int:
        return int(self._f2h(x))

    def h2f(self, h: int) -> float:
        return float(self._h2f(h))

    def srgb_tables(self):
        lin = np.ctypeslib.as_array(getattr(self.lib, self.prefix + "srgb_lin")(), shape=(256,)).copy()
        thr = np.ctypeslib.as_array(getattr(self.lib, self.prefix + "srgb_thr")(), shape=(255,)).copy()
        return lin, thr


def oracle(counting: bool = False) -> TexLib:
    return TexLib(os.path.join(HERE, "_lib", "libzto_count.so" if counting else "libzto.so"), "zto_")


def reference(m1k64: bool = False) -> TexLib:
    return TexLib(os.path.join(HERE, "_ref", "libzapref_m1k64.so" if m1k64 else "libzapref.so"), "zref_")


def reference_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libzapref.so"))
