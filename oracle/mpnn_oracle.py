"""CPU oracle of the learned warm-start network (test infrastructure only).

A numpy restatement of the message-passing network that
rockfit-dem_amd/rfd/warmstart.py runs on the device with the kernels of
csrc/rfd_mpnn.hip.  It has no counterpart in the reference: the network is
new.  The dense products round their operands to BF16 (nearest even) exactly
where the device's matrix-core kernels do, and accumulate in float64; set
``bf16=False`` for the exact float64 network that the finite-difference pins
check the backward pass against.

Network (hidden width 64, two message rounds):
  H0 = relu(X W_enc + b_enc)
  round r:  Z = H Wa_r;  q_h = Z[j(h)] + E_h We_r + bm_r  (h in row i)
            A_i = sum over row i of relu(q_h);  C = [H | A] Wu_r + bu_r
            H <- H + relu(C)
  Y = H W_dec + b_dec          (first 6 columns: corrections of du, dtheta)
Loss: mean over N x 6 of (Y - T)^2.
"""
from __future__ import annotations

import numpy as np

HID = 64
ROUNDS = 2
NIN = 32
NOUT = 32
NUSED = 6


def param_shapes() -> dict[str, tuple[int, ...]]:
    s = {"W_enc": (NIN, HID), "b_enc": (HID,)}
    for r in range(ROUNDS):
        s[f"Wa{r}"] = (HID, HID)
        s[f"We{r}"] = (4, HID)
        s[f"bm{r}"] = (HID,)
        s[f"Wu{r}"] = (2 * HID, HID)
        s[f"bu{r}"] = (HID,)
    s["W_dec"] = (HID, NOUT)
    s["b_dec"] = (NOUT,)
    return s


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class Graph:
    """Half-bond rows: hoff[N+1], partner hj[H], mirror half hm[H]."""

    def __init__(self, hoff, hj, hm):
        self.hoff = np.asarray(hoff, np.int64)
        self.hj = np.asarray(hj, np.int64)
        self.hm = np.asarray(hm, np.int64)
        self.N = len(self.hoff) - 1
        self.row = np.repeat(np.arange(self.N), np.diff(self.hoff))


def _mm(a, b, bf):
    return (bf16(a) @ bf16(b)) if bf else (a.astype(np.float64) @ b.astype(np.float64))


def forward(p: dict, g: Graph, X, E, bf16_ops: bool = True):
    """Returns Y [N x NOUT] and the cache for backward."""
    X = np.asarray(X, np.float64)
    E = np.asarray(E, np.float64)
    cache = {"X": X, "E": E}
    P0 = _mm(X, p["W_enc"], bf16_ops) + p["b_enc"]
    H = np.maximum(P0, 0.0)
    cache["P0"] = P0
    for r in range(ROUNDS):
        Z = _mm(H, p[f"Wa{r}"], bf16_ops)
        q = Z[g.hj] + E @ np.asarray(p[f"We{r}"], np.float64) + p[f"bm{r}"]
        A = np.zeros_like(H)
        np.add.at(A, g.row, np.maximum(q, 0.0))
        HA = np.concatenate([H, A], axis=1)
        C = _mm(HA, p[f"Wu{r}"], bf16_ops) + p[f"bu{r}"]
        cache[r] = (H, q, HA, C)
        H = H + np.maximum(C, 0.0)
    cache["H"] = H
    Y = _mm(H, p["W_dec"], bf16_ops) + p["b_dec"]
    return Y, cache


def loss_and_grad_y(Y, T):
    d = Y[:, :NUSED] - T[:, :NUSED]
    n = d.size
    dY = np.zeros_like(Y)
    dY[:, :NUSED] = 2.0 * d / n
    return float((d * d).sum() / n), dY


def backward(p: dict, g: Graph, cache, dY, bf16_ops: bool = True) -> dict:
    """Gradients of the loss (through dY) with respect to every parameter."""
    gr = {}
    H = cache["H"]
    tn = (lambda a, b: bf16(a).T @ bf16(b)) if bf16_ops else (lambda a, b: a.T @ b)
    gr["W_dec"] = tn(H, dY)
    gr["b_dec"] = dY.sum(0)
    dH = _mm(dY, np.asarray(p["W_dec"]).T, bf16_ops)
    for r in reversed(range(ROUNDS)):
        Hin, q, HA, C = cache[r]
        dC = dH * (C > 0)
        gr[f"Wu{r}"] = tn(HA, dC)
        gr[f"bu{r}"] = dC.sum(0)
        dHA = _mm(dC, np.asarray(p[f"Wu{r}"]).T, bf16_ops)
        dHin = dH + dHA[:, :HID]
        dA = dHA[:, HID:]
        dq = dA[g.row] * (q > 0)
        gr[f"We{r}"] = cache["E"].T @ dq
        gr[f"bm{r}"] = dq.sum(0)
        dZ = np.zeros_like(Hin)
        np.add.at(dZ, g.hj, dq)
        gr[f"Wa{r}"] = tn(Hin, dZ)
        dHin = dHin + _mm(dZ, np.asarray(p[f"Wa{r}"]).T, bf16_ops)
        dH = dHin
    dP0 = dH * (cache["P0"] > 0)
    gr["W_enc"] = tn(cache["X"], dP0)
    gr["b_enc"] = dP0.sum(0)
    return gr
