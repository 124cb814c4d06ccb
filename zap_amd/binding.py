"""ctypes binding of libzaptex (include/zaptex.h), mirroring zap_tex.h's API with numpy and torch buffers.

Host functions take and return numpy arrays and keep the reference's argument meaning (zap_tex.h:4-23):
RGBA8 images are HxWx4 uint8, BC6H input is HxWx4 float32, decoded BC6H is HxWx4 uint16 half bits.
Device functions take torch tensors that live on the GPU and enqueue on the current torch stream.
There is no CPU fallback: a missing library or device raises ZapTexError.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libzaptex.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "zaptex.h")

BC1, BC3, BC4, BC5, BC7, BC6H, ASTC = range(7)
FORMATS = {"bc1": BC1, "bc3": BC3, "bc4": BC4, "bc5": BC5, "bc7": BC7, "bc6h": BC6H, "astc": ASTC}
OPT_BC7_M1K = 1
OPT_PROFILE = 2
STATUS = {0: "ok", 1: "bad argument", 2: "device error", 3: "out of memory", 4: "unsupported"}


class ZapTexError(RuntimeError):
    pass


def fmt_id(fmt) -> int:
    return FORMATS[fmt.lower()] if isinstance(fmt, str) else int(fmt)


def bc_block_bytes(fmt) -> int:  # zap_tex.h:34
    return 8 if fmt_id(fmt) in (BC1, BC4) else 16


def bc_size(w: int, h: int, fmt) -> int:  # zap_tex.h:35-37 (BC6H / ASTC use 16-byte blocks)
    return ((w + 3) // 4) * ((h + 3) // 4) * bc_block_bytes(fmt)


def mip_levels(w: int, h: int) -> int:  # zap_tex.h:1435-1439
    n = 1
    while w > 1 or h > 1:
        w, h = max(1, w // 2), max(1, h // 2)
        n += 1
    return n


def mip_sizes(w: int, h: int) -> list[tuple[int, int]]:
    out = [(w, h)]
    while w > 1 or h > 1:
        w, h = max(1, w // 2), max(1, h // 2)
        out.append((w, h))
    return out


class ZapImage(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("stride", C.c_size_t),
                ("w", C.c_int), ("h", C.c_int), ("fmt", C.c_int)]


class ZapTexture(C.Structure):
    _fields_ = [("base", C.c_void_p), ("stride", C.c_size_t), ("w", C.c_int), ("h", C.c_int), ("fmt", C.c_int), ("srgb", C.c_int),
                ("levels", C.c_int), ("mips", C.c_void_p), ("out", C.c_void_p)]


class MlpTargets(C.Structure):
    _fields_ = [("labels", C.c_void_p), ("deltas", C.c_void_p), ("temperature", C.c_float), ("risk_k", C.c_int),
                ("risk_weight", C.c_float)]


def declared_functions(header: str = HEADER) -> list[str]:
    """Names of the functions include/zaptex.h declares."""
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s+\*?\s*(zaptex_[a-z_0-9]+)\s*\(", text, re.M)))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Library:
    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise ZapTexError(f"{path} is not built (run __graft_entry__.build() / python zap_amd/build.py)")
        self.path = path
        # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 and
        # its libraries name them without the version suffix.  Loaded first, torch's copies carry the sonames
        # libzaptex.so asks for, so both share them; loaded second, torch would start a second HSA runtime
        # that cannot open the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        self.lib = L = C.CDLL(path)
        vp, i, sz, f = C.c_void_p, C.c_int, C.c_size_t, C.c_float
        sig = {
            "zaptex_bc_encode": [vp, i, i, sz, i, f, vp], "zaptex_bc_decode": [vp, i, i, i, vp, sz],
            "zaptex_bc6h_encode": [vp, i, i, sz, vp], "zaptex_bc6h_decode": [vp, i, i, vp, sz],
            "zaptex_astc_encode": [vp, i, i, sz, vp], "zaptex_astc_decode": [vp, i, i, vp, sz],
            "zaptex_mip_next": [vp, i, i, sz, i, vp, sz], "zaptex_mip_next_f": [vp, i, i, sz, vp, sz],
            "zaptex_bc_split": [vp, sz, i, vp], "zaptex_bc_merge": [vp, sz, i, vp],
            "zaptex_encode_dev": [C.POINTER(ZapImage), i, vp, sz, vp], "zaptex_encode_dev_noalloc": [C.POINTER(ZapImage), i, vp],
            "zaptex_encode_rdo_dev": [C.POINTER(ZapImage), i, f, vp, sz, vp],
            "zaptex_mip_next_dev": [vp, i, i, sz, i, vp, sz, vp], "zaptex_mip_next_f_dev": [vp, i, i, sz, vp, sz, vp],
            "zaptex_mip_chain_dev": [vp, i, i, sz, i, C.POINTER(vp), i, vp],
            "zaptex_bc_decode_dev": [vp, i, i, i, vp, sz, vp],
            "zaptex_bc7_full_dev": [C.POINTER(ZapImage), i, vp, vp, vp, sz, vp],
            "zaptex_mlp_init": [vp, C.c_uint32, vp], "zaptex_mlp_set_params": [vp, vp, vp],
            "zaptex_mlp_set_features": [vp, i, vp],
            "zaptex_mlp_forward": [vp, vp, i, vp, vp, sz, vp], "zaptex_mlp_propose": [vp, vp, i, i, vp, vp, vp],
            "zaptex_mlp_propose_logits": [vp, vp, i, vp, vp],
            "zaptex_mlp_gradients": [vp, vp, C.POINTER(MlpTargets), i, vp, sz, vp, vp],
            "zaptex_mlp_train_step": [vp, vp, C.POINTER(MlpTargets), i, f, i, vp, sz, vp, vp],
            "zaptex_bc7_gather_dev": [C.POINTER(ZapImage), i, vp, vp],
            "zaptex_bc7_learned_dev": [C.POINTER(ZapImage), i, vp, i, f, i, vp, sz, vp],
            "zaptex_set_option": [i, i], "zaptex_get_option": [i],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.zaptex_profile_collect.restype = C.c_int
        L.zaptex_profile_get.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.zaptex_profile_get.restype = C.c_int
        L.zaptex_profile_reset.restype = None
        L.zaptex_mlp_state_size.restype = C.c_size_t
        L.zaptex_mlp_workspace_size.argtypes = [i]; L.zaptex_mlp_workspace_size.restype = C.c_size_t
        L.zaptex_bc7_learned_workspace_size.argtypes = [C.POINTER(ZapImage), i, i, i]
        L.zaptex_bc7_learned_workspace_size.restype = C.c_size_t
        for name, args in {"zaptex_batch_mip_bytes": [i, i, i], "zaptex_batch_out_bytes": [i, i, i, i],
                           "zaptex_batch_workspace_size": [C.POINTER(ZapTexture), i]}.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_size_t
        L.zaptex_batch_create.argtypes = [C.POINTER(ZapTexture), i, vp, sz, C.POINTER(vp)]
        L.zaptex_batch_create.restype = C.c_int
        L.zaptex_batch_run.argtypes = [vp, vp]
        L.zaptex_batch_run.restype = C.c_int
        L.zaptex_batch_destroy.argtypes = [vp]
        L.zaptex_batch_destroy.restype = None
        L.zaptex_set_launch.argtypes = [C.c_char_p, i, i]
        L.zaptex_set_launch.restype = C.c_int
        L.zaptex_get_launch.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.zaptex_get_launch.restype = C.c_int
        L.zaptex_bc7_full_workspace_size.argtypes = [C.POINTER(ZapImage), i, i]
        L.zaptex_bc7_full_workspace_size.restype = C.c_size_t
        L.zaptex_tex_file_size.argtypes = [i, i, i, i, i]
        L.zaptex_tex_file_size.restype = C.c_size_t
        L.zaptex_tex_file.argtypes = [vp, i, i, sz, i, f, i, i, vp, sz]
        L.zaptex_tex_file.restype = C.c_int
        L.zaptex_rdo_workspace_size.argtypes = [C.POINTER(ZapImage), i]
        L.zaptex_rdo_workspace_size.restype = C.c_size_t
        L.zaptex_workspace_size.argtypes = [C.POINTER(ZapImage), i]
        L.zaptex_workspace_size.restype = C.c_size_t
        L.zaptex_last_error.restype = C.c_char_p
        L.zaptex_version.restype = C.c_char_p
        L.zaptex_device_count.restype = C.c_int

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.zaptex_last_error().decode(errors="replace")
            raise ZapTexError(f"{what}: {STATUS.get(rc, rc)}: {msg}")

    # ------------------------------------------------------------------ host API (zap_tex.h)
    def bc_encode(self, img: np.ndarray, fmt, rdo: float = 0.0, w: int | None = None, h: int | None = None) -> np.ndarray:
        f = fmt_id(fmt)
        img = np.ascontiguousarray(img) if img.strides[-1] != 1 or img.strides[-2] != 4 else img
        h = img.shape[0] if h is None else h
        w = img.shape[1] if w is None else w
        out = np.zeros(bc_size(w, h, f), dtype=np.uint8)
        self._check(self.lib.zaptex_bc_encode(_ptr(img), w, h, img.strides[0], f, rdo, _ptr(out)), "zap_bc_encode")
        return out

    def bc_decode(self, bc: np.ndarray, w: int, h: int, fmt) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(self.lib.zaptex_bc_decode(_ptr(np.ascontiguousarray(bc)), w, h, fmt_id(fmt), _ptr(out), w * 4), "zap_bc_decode")
        return out

    def bc6h_encode(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img, dtype=np.float32)
        h, w = img.shape[:2]
        out = np.zeros(bc_size(w, h, BC7), dtype=np.uint8)
        self._check(self.lib.zaptex_bc6h_encode(_ptr(img), w, h, img.strides[0] // 4, _ptr(out)), "zap_bc6h_encode")
        return out

    def bc6h_decode(self, bc: np.ndarray, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.uint16)
        self._check(self.lib.zaptex_bc6h_decode(_ptr(np.ascontiguousarray(bc)), w, h, _ptr(out), w * 4), "zap_bc6h_decode")
        return out

    def astc_encode(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        out = np.zeros(bc_size(w, h, BC7), dtype=np.uint8)
        self._check(self.lib.zaptex_astc_encode(_ptr(img), w, h, img.strides[0], _ptr(out)), "zap_astc_encode")
        return out

    def astc_decode(self, bc: np.ndarray, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(self.lib.zaptex_astc_decode(_ptr(np.ascontiguousarray(bc)), w, h, _ptr(out), w * 4), "zap_astc_decode")
        return out

    def mip_next(self, img: np.ndarray, srgb: int = 0) -> np.ndarray:
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        nw, nh = max(1, w // 2), max(1, h // 2)
        out = np.zeros((nh, nw, 4), dtype=np.uint8)
        self._check(self.lib.zaptex_mip_next(_ptr(img), w, h, img.strides[0], srgb, _ptr(out), nw * 4), "zap_mip_next")
        return out

    def mip_next_f(self, img: np.ndarray) -> np.ndarray:
        img = np.ascontiguousarray(img, dtype=np.float32)
        h, w = img.shape[:2]
        nw, nh = max(1, w // 2), max(1, h // 2)
        out = np.zeros((nh, nw, 4), dtype=np.float32)
        self._check(self.lib.zaptex_mip_next_f(_ptr(img), w, h, w * 4, _ptr(out), nw * 4), "zap_mip_next_f")
        return out

    def bc_split(self, bc: np.ndarray, fmt) -> np.ndarray:
        bc = np.ascontiguousarray(bc)
        out = np.zeros_like(bc)
        self._check(self.lib.zaptex_bc_split(_ptr(bc), bc.size, fmt_id(fmt), _ptr(out)), "zap_bc_split")
        return out

    def bc_merge(self, planes: np.ndarray, fmt) -> np.ndarray:
        planes = np.ascontiguousarray(planes)
        out = np.zeros_like(planes)
        self._check(self.lib.zaptex_bc_merge(_ptr(planes), planes.size, fmt_id(fmt), _ptr(out)), "zap_bc_merge")
        return out

    def tex_file(self, img: np.ndarray, fmt, mips: bool = False, srgb: bool = False, rdo: float = 0.0) -> bytes:
        """The file `zap tex` writes (zap.c:309-355): DDS for BC1/3/4/5/7 (with mips: the whole chain), .astc for ASTC."""
        img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        f = fmt_id(fmt)
        n = int(self.lib.zaptex_tex_file_size(w, h, f, int(mips), int(srgb)))
        out = np.zeros(max(1, n), dtype=np.uint8)
        self._check(self.lib.zaptex_tex_file(_ptr(img), w, h, img.strides[0], f, rdo, int(mips), int(srgb), _ptr(out), n), "zaptex_tex_file")
        return out[:n].tobytes()

    def set_option(self, key: int, value: int) -> None:
        self._check(self.lib.zaptex_set_option(key, value), "zaptex_set_option")

    def get_option(self, key: int) -> int:
        return self.lib.zaptex_get_option(key)

    def set_m1k(self, k: int) -> None:
        self.set_option(OPT_BC7_M1K, k)

    # ------------------------------------------------------------------ device API (torch tensors on the GPU)
    @staticmethod
    def _stream(stream=None) -> int:
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return s.cuda_stream

    @staticmethod
    def images(specs) -> C.Array:
        """specs: iterable of (src_tensor, dst_tensor, w, h, row_stride_bytes, fmt)."""
        specs = list(specs)
        arr = (ZapImage * len(specs))()
        for k, (src, dst, w, h, stride, fmt) in enumerate(specs):
            arr[k] = ZapImage(src.data_ptr(), dst.data_ptr(), stride, w, h, fmt_id(fmt))
        return arr

    def workspace_size(self, imgs) -> int:
        return int(self.lib.zaptex_workspace_size(imgs, len(imgs)))

    def encode_dev(self, imgs, workspace=None, stream=None) -> None:
        ws_ptr = workspace.data_ptr() if workspace is not None else None
        ws_len = workspace.numel() * workspace.element_size() if workspace is not None else 0
        self._check(self.lib.zaptex_encode_dev(imgs, len(imgs), ws_ptr, ws_len, self._stream(stream)), "zaptex_encode_dev")

    def rdo_workspace_size(self, imgs) -> int:
        return int(self.lib.zaptex_rdo_workspace_size(imgs, len(imgs)))

    def encode_rdo_dev(self, imgs, rdo: float, workspace=None, stream=None) -> None:
        """zaptex_encode_dev, then the rate-distortion pass of every image (zap_bc_encode with rdo, per image)."""
        ws_ptr = workspace.data_ptr() if workspace is not None else None
        ws_len = workspace.numel() * workspace.element_size() if workspace is not None else 0
        self._check(self.lib.zaptex_encode_rdo_dev(imgs, len(imgs), float(rdo), ws_ptr, ws_len, self._stream(stream)), "zaptex_encode_rdo_dev")

    NOALLOC_SCRATCH_BYTES = 64 << 20   # ZAPTEX_NOALLOC_SCRATCH_BYTES

    def encode_dev_noalloc(self, imgs, stream=None) -> None:
        """zaptex_encode_dev on the library's own scratch array: no workspace, no allocation."""
        self._check(self.lib.zaptex_encode_dev_noalloc(imgs, len(imgs), self._stream(stream)), "zaptex_encode_dev_noalloc")

    def bc7_full_workspace(self, imgs, with_deltas: bool = False):
        import torch
        n = int(self.lib.zaptex_bc7_full_workspace_size(imgs, len(imgs), int(with_deltas)))
        return torch.empty(n, dtype=torch.uint8, device="cuda")

    def bc7_full_dev(self, imgs, best_cand=None, deltas=None, ws=None, stream=None) -> None:
        """Exhaustive all-mode search; best_cand: int32 [blocks], deltas: int16/uint16 [blocks, 512] (optional)."""
        if ws is None:
            ws = self.bc7_full_workspace(imgs, deltas is not None)
        self._check(self.lib.zaptex_bc7_full_dev(imgs, len(imgs), best_cand.data_ptr() if best_cand is not None else None,
                                                 deltas.data_ptr() if deltas is not None else None, ws.data_ptr(), ws.numel(),
                                                 self._stream(stream)), "zaptex_bc7_full_dev")

    def mip_chain_dev(self, base, w: int, h: int, stride: int, levels: list, srgb: int = 0, stream=None) -> None:
        ptrs = (C.c_void_p * len(levels))(*[t.data_ptr() for t in levels])
        self._check(self.lib.zaptex_mip_chain_dev(base.data_ptr(), w, h, stride, srgb, ptrs, len(levels), self._stream(stream)),
                    "zaptex_mip_chain_dev")

    def bc_decode_dev(self, bc, w: int, h: int, fmt, out, stream=None) -> None:
        self._check(self.lib.zaptex_bc_decode_dev(bc.data_ptr(), w, h, fmt_id(fmt), out.data_ptr(), w * 4, self._stream(stream)),
                    "zaptex_bc_decode_dev")

    # ------------------------------------------------------------------ learned proposal (torch tensors on the GPU)
    MLP_PARAMS = 90880

    def mlp_state(self, seed: int | None = 1, params: np.ndarray | None = None, features: int = 0):
        import torch
        st = torch.zeros(int(self.lib.zaptex_mlp_state_size()), dtype=torch.uint8, device="cuda")
        if params is not None:
            p = np.ascontiguousarray(params, dtype=np.float32)
            assert p.size == self.MLP_PARAMS
            self._check(self.lib.zaptex_mlp_init(st.data_ptr(), 1, self._stream()), "zaptex_mlp_init")
            self._check(self.lib.zaptex_mlp_set_params(st.data_ptr(), _ptr(p), self._stream()), "zaptex_mlp_set_params")
        else:
            self._check(self.lib.zaptex_mlp_init(st.data_ptr(), seed, self._stream()), "zaptex_mlp_init")
        self._check(self.lib.zaptex_mlp_set_features(st.data_ptr(), features, self._stream()), "zaptex_mlp_set_features")
        return st

    @staticmethod
    def mlp_params(state) -> np.ndarray:
        import torch
        return state[: Library.MLP_PARAMS * 4].view(torch.float32).cpu().numpy().copy()

    @staticmethod
    def mlp_grads(state) -> np.ndarray:
        import torch
        n = Library.MLP_PARAMS * 4
        return state[n: 2 * n].view(torch.float32).cpu().numpy().copy()

    def mlp_workspace(self, batch: int):
        import torch
        return torch.empty(int(self.lib.zaptex_mlp_workspace_size(batch)), dtype=torch.uint8, device="cuda")

    def mlp_forward(self, state, blocks, batch: int, logits, ws) -> None:
        self._check(self.lib.zaptex_mlp_forward(state.data_ptr(), blocks.data_ptr(), batch, logits.data_ptr(), ws.data_ptr(),
                                                ws.numel(), self._stream()), "zaptex_mlp_forward")

    def mlp_propose(self, state, blocks, n: int, topk: int, cand, risk) -> None:
        """top-k candidate slots [n x topk] (uint32 as int32) and predicted risk [n] of n 64-byte blocks, fused"""
        self._check(self.lib.zaptex_mlp_propose(state.data_ptr(), blocks.data_ptr(), n, topk, cand.data_ptr(), risk.data_ptr(),
                                                self._stream()), "zaptex_mlp_propose")

    def mlp_propose_logits(self, state, blocks, n: int, logits) -> None:
        """the fused kernel's logits of the kept rows [n x 288] (test hook)"""
        self._check(self.lib.zaptex_mlp_propose_logits(state.data_ptr(), blocks.data_ptr(), n, logits.data_ptr(), self._stream()),
                    "zaptex_mlp_propose_logits")

    @staticmethod
    def _targets(labels, deltas, temperature: float, risk_k: int, risk_weight: float) -> MlpTargets:
        return MlpTargets(labels.data_ptr() if labels is not None else None, deltas.data_ptr() if deltas is not None else None,
                          temperature, risk_k, risk_weight)

    def mlp_gradients(self, state, blocks, labels, batch: int, ws, stats, deltas=None, temperature: float = 0.0, risk_k: int = 0,
                      risk_weight: float = 0.0) -> None:
        """labels: int32 [batch]; deltas: int16 [batch, 512] (soft targets at temperature > 0, and/or the risk slot)."""
        t = self._targets(labels, deltas, temperature, risk_k, risk_weight)
        self._check(self.lib.zaptex_mlp_gradients(state.data_ptr(), blocks.data_ptr(), C.byref(t), batch, ws.data_ptr(), ws.numel(),
                                                  stats.data_ptr(), self._stream()), "zaptex_mlp_gradients")

    def mlp_train_step(self, state, blocks, labels, batch: int, lr: float, step: int, ws, stats, deltas=None,
                       temperature: float = 0.0, risk_k: int = 0, risk_weight: float = 0.0) -> None:
        t = self._targets(labels, deltas, temperature, risk_k, risk_weight)
        self._check(self.lib.zaptex_mlp_train_step(state.data_ptr(), blocks.data_ptr(), C.byref(t), batch, lr, step, ws.data_ptr(),
                                                   ws.numel(), stats.data_ptr(), self._stream()), "zaptex_mlp_train_step")

    def bc7_gather_dev(self, imgs, blocks) -> None:
        self._check(self.lib.zaptex_bc7_gather_dev(imgs, len(imgs), blocks.data_ptr(), self._stream()), "zaptex_bc7_gather_dev")

    def bc7_learned_workspace(self, imgs, topk: int, fallback_k: int = 0):
        import torch
        n = int(self.lib.zaptex_bc7_learned_workspace_size(imgs, len(imgs), topk, fallback_k))
        return torch.empty(n, dtype=torch.uint8, device="cuda")

    def bc7_learned_dev(self, imgs, state, topk: int, threshold: float, ws, fallback_k: int = 0, stream=None) -> None:
        """fallback_k = 0: blocks over the risk threshold get the exhaustive search; 1..285: the network's top fallback_k"""
        self._check(self.lib.zaptex_bc7_learned_dev(imgs, len(imgs), state.data_ptr(), topk, threshold, fallback_k, ws.data_ptr(),
                                                    ws.numel(), self._stream(stream)), "zaptex_bc7_learned_dev")

    # ------------------------------------------------------------------ batch (device buffers as torch tensors)
    def batch_mip_bytes(self, w: int, h: int, levels: int) -> int:
        return int(self.lib.zaptex_batch_mip_bytes(w, h, levels))

    def batch_out_bytes(self, w: int, h: int, fmt, levels: int) -> int:
        return int(self.lib.zaptex_batch_out_bytes(w, h, fmt_id(fmt), levels))

    @staticmethod
    def textures(specs) -> C.Array:
        """specs: (base tensor, stride, w, h, fmt, srgb, levels, mips tensor or None, out tensor)"""
        specs = list(specs)
        arr = (ZapTexture * len(specs))()
        for k, (base, stride, w, h, fmt, srgb, levels, mips, out) in enumerate(specs):
            arr[k] = ZapTexture(base.data_ptr(), stride, w, h, fmt_id(fmt), int(srgb), levels, mips.data_ptr() if mips is not None else None,
                                out.data_ptr())
        return arr

    def batch_workspace_size(self, texs) -> int:
        return int(self.lib.zaptex_batch_workspace_size(texs, len(texs)))

    def batch_create(self, texs, ws) -> int:
        h = C.c_void_p()
        self._check(self.lib.zaptex_batch_create(texs, len(texs), ws.data_ptr(), ws.numel(), C.byref(h)), "zaptex_batch_create")
        return h.value

    def batch_run(self, handle: int, stream=None) -> None:
        self._check(self.lib.zaptex_batch_run(handle, self._stream(stream)), "zaptex_batch_run")

    def batch_destroy(self, handle: int) -> None:
        self.lib.zaptex_batch_destroy(handle)

    def set_launch(self, kernel: str, wg: int, max_groups: int) -> None:
        self._check(self.lib.zaptex_set_launch(kernel.encode(), wg, max_groups), "zaptex_set_launch")

    def get_launch(self, kernel: str) -> tuple[int, int]:
        wg, mg = C.c_int(), C.c_int()
        self._check(self.lib.zaptex_get_launch(kernel.encode(), C.byref(wg), C.byref(mg)), "zaptex_get_launch")
        return wg.value, mg.value

    # ------------------------------------------------------------------ kernel timing
    def profile(self, on: bool) -> None:
        self.set_option(OPT_PROFILE, 1 if on else 0)

    def profile_reset(self) -> None:
        self.lib.zaptex_profile_reset()

    def profile_get(self, kernel: str) -> tuple[float, int]:
        """(total milliseconds, launches) of `kernel` since the last reset."""
        self._check(self.lib.zaptex_profile_collect(), "zaptex_profile_collect")
        ms, n = C.c_double(), C.c_int()
        self._check(self.lib.zaptex_profile_get(kernel.encode(), C.byref(ms), C.byref(n)), "zaptex_profile_get")
        return ms.value, n.value

    def version(self) -> str:
        return self.lib.zaptex_version().decode()

    def device_count(self) -> int:
        return int(self.lib.zaptex_device_count())


_LIB: Library | None = None


def load() -> Library:
    global _LIB
    if _LIB is None:
        _LIB = Library()
    return _LIB
