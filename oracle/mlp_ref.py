"""CPU oracle of the BC7 proposal network (numpy float64).  TEST INFRASTRUCTURE ONLY.

Mirrors zap_amd/csrc/zt_mlp.hip + the C ABI's training step: inputs v/255 rounded to bf16; weights used in
bf16; hidden activations rounded to bf16 after bias + ReLU; logits in full precision; dZ = (softmax - onehot)/B
over the 285 used candidate slots, rounded to bf16; back-propagated activation gradients rounded to bf16.
`bf16_points=False` gives the exact-arithmetic network, which tests/test_mlp_oracle.py checks against
central finite differences.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 64, 128, 512
OFF = {"W1": (0, (IN, H)), "b1": (IN * H, (H,)), "W2": (IN * H + H, (H, H)), "b2": (IN * H + H + H * H, (H,)),
       "W3": (IN * H + 2 * H + H * H, (H, OUT)), "b3": (IN * H + 2 * H + H * H + H * OUT, (OUT,))}
NPARAM = IN * H + H + H * H + H + H * OUT + OUT
VALID = np.array([(c >> 6) not in (0, 4, 5, 6) or (c & 63) < {0: 16, 4: 8, 5: 4, 6: 1}[c >> 6] for c in range(512)])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray) -> dict:
    return {k: p[o:o + int(np.prod(s))].reshape(s).astype(np.float64) for k, (o, s) in OFF.items()}


def features(blocks: np.ndarray, bf16_points: bool = True) -> np.ndarray:
    x = blocks.reshape(-1, 64).astype(np.float32) / np.float32(255.0)
    return (bf16(x) if bf16_points else x).astype(np.float64)


def forward(p: np.ndarray, blocks: np.ndarray, bf16_points: bool = True):
    P = unpack(p)
    q = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if bf16_points else (lambda a: a)
    X = features(blocks, bf16_points)
    W1, W2, W3 = (q(P[k]) for k in ("W1", "W2", "W3"))
    H1 = q(np.maximum(X @ W1 + P["b1"], 0.0))
    H2 = q(np.maximum(H1 @ W2 + P["b2"], 0.0))
    Z = H2 @ W3 + P["b3"]
    return X, H1, H2, Z, (W1, W2, W3)


def loss_and_grads(p: np.ndarray, blocks: np.ndarray, labels: np.ndarray, bf16_points: bool = True):
    q = (lambda a: bf16(a.astype(np.float32)).astype(np.float64)) if bf16_points else (lambda a: a)
    X, H1, H2, Z, (W1, W2, W3) = forward(p, blocks, bf16_points)
    B = Z.shape[0]
    Zm = np.where(VALID[None, :], Z, -np.inf)
    mx = Zm.max(1, keepdims=True)
    e = np.where(VALID[None, :], np.exp(Zm - mx), 0.0)
    s = e.sum(1, keepdims=True)
    prob = e / s
    loss = float(np.mean(np.log(s[:, 0]) + mx[:, 0] - Z[np.arange(B), labels]))
    onehot = np.zeros_like(Z)
    onehot[np.arange(B), labels] = 1.0
    dZ = q(np.where(VALID[None, :], (prob - onehot) / B, 0.0))
    g = {"W3": H2.T @ dZ, "b3": dZ.sum(0)}
    dH2 = q((dZ @ W3.T) * (H2 > 0))
    g["W2"], g["b2"] = H1.T @ dH2, dH2.sum(0)
    dH1 = q((dH2 @ W2.T) * (H1 > 0))
    g["W1"], g["b1"] = X.T @ dH1, dH1.sum(0)
    flat = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        flat[o:o + int(np.prod(sh))] = g[k].reshape(-1)
    return loss, flat
