"""CPU oracle of the BC7 proposal network (numpy float64).  TEST INFRASTRUCTURE ONLY.

Mirrors zap_amd/csrc/zt_mlp.hip + the C ABI's training step: inputs v/255 rounded to bf16; weights used in
bf16; hidden activations rounded to bf16 after bias + ReLU; logits in full precision; dZ = (softmax - q)/B
over the 285 used candidate slots, rounded to bf16, for targets q = one-hot labels or, given per-slot SSE
deltas, softmax(-delta / T) over the used slots whose delta is below 65535; with deltas and risk_k > 0, the
unused slot RISK regresses log(1 + min delta over the row's top-risk_k used slots) with squared error weighted by
risk_w; back-propagated activation gradients rounded to bf16.
`bf16_points=False` gives the exact-arithmetic network, which tests/test_mlp_oracle.py checks against
central finite differences.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 64, 128, 512
OFF = {"W1": (0, (IN, H)), "b1": (IN * H, (H,)), "W2": (IN * H + H, (H, H)), "b2": (IN * H + H + H * H, (H,)),
       "W3": (IN * H + 2 * H + H * H, (H, OUT)), "b3": (IN * H + 2 * H + H * H + H * OUT, (OUT,))}
NPARAM = IN * H + H + H * H + H + H * OUT + OUT
RISK = 385   # unused slot (mode 6, slot 1) holding the risk output
VALID = np.array([(c >> 6) not in (0, 4, 5, 6) or (c & 63) < {0: 16, 4: 8, 5: 4, 6: 1}[c >> 6] for c in range(512)])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray) -> dict:
    return {k: p[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in OFF.items()}


def features(blocks: np.ndarray, bf16_points: bool = True, mode: int = 0) -> np.ndarray:
    """Per-block input normalisation, float32 in the kernel's order: 0 v/255; 1 (v - m_c)/64; 2 (v - m_c)/max(1, R);
    3 (v - m_c)/sqrt(64 max(1, R)); 4 colour as 3 with R over R, G, B only, alpha (a - 255)/255; m_c the channel mean
    over the 16 pixels, R the largest channel range."""
    v = blocks.reshape(-1, 16, 4).astype(np.float32)
    if mode == 0:
        x = v * (np.float32(1.0) / np.float32(255.0))
    else:
        b = blocks.reshape(-1, 16, 4).astype(np.int32)
        mean = b.sum(1).astype(np.float32) / np.float32(16.0)
        ch = 3 if mode == 4 else 4
        r = np.maximum(1, (b.max(1) - b.min(1))[:, :ch].max(1)).astype(np.float32)[:, None, None]
        inv = {1: np.full_like(r, np.float32(1.0) / np.float32(64.0)), 2: np.float32(1.0) / r,
               3: np.float32(1.0) / np.sqrt(np.float32(64.0) * r), 4: np.float32(1.0) / np.sqrt(np.float32(64.0) * r)}[mode]
        inv = np.broadcast_to(inv, v.shape).copy()
        if mode == 4:
            mean[:, 3] = np.float32(255.0)
            inv[:, :, 3] = np.float32(1.0) / np.float32(255.0)
        x = (v - mean[:, None, :]) * inv
    x = x.reshape(-1, 64).astype(np.float32)
    return (bf16(x) if bf16_points else x).astype(np.float64)


def forward(p: np.ndarray, blocks: np.ndarray, bf16_points: bool = True, feat: int = 0):
    P = unpack(p)
    q = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if bf16_points else (lambda a: a)
    X = features(blocks, bf16_points, feat)
    W1, W2, W3 = (q(P[k]) for k in ("W1", "W2", "W3"))
    H1 = q(np.maximum(X @ W1 + P["b1"], 0.0))
    H2 = q(np.maximum(H1 @ W2 + P["b2"], 0.0))
    Z = H2 @ W3 + P["b3"]
    return X, H1, H2, Z, (W1, W2, W3)


def targets(B: int, labels: np.ndarray | None = None, deltas: np.ndarray | None = None, T: float | None = None) -> np.ndarray:
    """[B, 512] target distribution: one-hot labels, or softmax(-delta / T) over used slots with delta < 65535."""
    if deltas is None:
        t = np.zeros((B, OUT))
        t[np.arange(B), labels] = 1.0
        return t
    d = deltas.reshape(B, OUT).astype(np.float64)
    use = VALID[None, :] & (d < 65535)
    e = np.where(use, np.exp(-d / T), 0.0)
    return e / e.sum(1, keepdims=True)


def propose_topk(Z: np.ndarray, k: int) -> np.ndarray:
    """[B, k] used slots as the fused proposal kernel ranks them (zaptex_mlp_propose): each used slot's float32
    logit with its low 9 bits replaced by 511 - u, or by u for a negative logit (u = the slot's rank among the
    used slots), compared as floats.  Larger logits first; logits equal above their low 9 bits by ascending slot."""
    slots = np.nonzero(VALID)[0]
    v = np.ascontiguousarray(Z[:, slots], dtype=np.float32).view(np.uint32)
    u = np.arange(slots.size, dtype=np.uint32)[None, :]
    code = np.where(v >> np.uint32(31), u, np.uint32(511) - u).astype(np.uint32)
    key = ((v & ~np.uint32(511)) | code).view(np.float32).astype(np.float64)
    return slots[np.argsort(-key, axis=1, kind="stable")[:, :k]]


def topk(Z: np.ndarray, k: int) -> np.ndarray:
    """[B, k] used slots by descending logit, lowest slot first on ties."""
    Zm = np.where(VALID[None, :], Z, -np.inf)
    return np.argsort(-Zm, axis=1, kind="stable")[:, :k]


def loss_and_grads(p: np.ndarray, blocks: np.ndarray, labels: np.ndarray | None, bf16_points: bool = True,
                   deltas: np.ndarray | None = None, T: float | None = None, risk_k: int = 0, risk_w: float = 0.0, feat: int = 0):
    q = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if bf16_points else (lambda a: a)
    X, H1, H2, Z, (W1, W2, W3) = forward(p, blocks, bf16_points, feat)
    B = Z.shape[0]
    Zm = np.where(VALID[None, :], Z, -np.inf)
    mx = Zm.max(1, keepdims=True)
    e = np.where(VALID[None, :], np.exp(Zm - mx), 0.0)
    s = e.sum(1, keepdims=True)
    prob = e / s
    tq = targets(B, labels, deltas if T else None, T)
    per_row = np.log(s[:, 0]) + mx[:, 0] - (tq * np.where(VALID[None, :], Z, 0.0)).sum(1)
    g_z = np.where(VALID[None, :], (prob - tq) / B, 0.0)
    if risk_k:
        d = deltas.reshape(B, OUT).astype(np.float64)
        dmin = np.take_along_axis(d, topk(Z, risk_k), 1).min(1)
        e = Z[:, RISK] - np.log1p(dmin)
        per_row = per_row + risk_w * e * e
        g_z[:, RISK] = 2.0 * risk_w * e / B
    loss = float(np.mean(per_row))
    dZ = q(g_z)
    g = {"W3": H2.T @ dZ, "b3": dZ.sum(0)}
    dH2 = q((dZ @ W3.T) * (H2 > 0))
    g["W2"], g["b2"] = H1.T @ dH2, dH2.sum(0)
    dH1 = q((dH2 @ W2.T) * (H1 > 0))
    g["W1"], g["b1"] = X.T @ dH1, dH1.sum(0)
    flat = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        flat[o:o + int(np.prod(sh))] = g[k].reshape(-1)
    return loss, flat
