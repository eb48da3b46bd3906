"""Learned FIRE warm start: a message-passing network over the bond graph that
predicts each particle's displacement and rotation increment for the next
quasi-static load step, trained on the device on relaxations that plain FIRE
produced.  There is no counterpart in the reference.

Every arithmetic operation runs in the HIP kernels of csrc/rfd_mpnn.hip
(BF16 matrix-core products with FP32 accumulation, message step, Adam) and
csrc/rfd_native.hip (inputs); torch only allocates device memory.

Inputs per particle: the last two step displacements and rotations, scaled by
the platen displacement of one step (su) and su / r_mean (sth); the largest
and the mean utilisation of its intact bonds (E strain / tensile strength,
how close the stretch alone brings each to breaking) and the share of its
bonds already broken.  Inputs per
half-bond: bond strain x 1e3 (0 once broken) and the unit direction.  The
network predicts corrections: du = du_last + su * y[0:3], dth = dth_last +
sth * y[3:6], so an untrained (near zero) output repeats the last increment.
The FIRE residual judges every prediction: NativeSystem.fire_load_step keeps
the plain start when the predicted one has a larger residual.

Network (hidden 64, two rounds):
  H0 = relu(X W_enc + b_enc)
  round r:  Z = H Wa_r;  q_h = Z[j] + E_h We_r + bm_r;  A_i = sum_row relu(q_h)
            H <- H + relu([H | A] Wu_r + bu_r)
  Y = H W_dec + b_dec
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import SIGNATURES, lib

HID = 64
ROUNDS = 2
NIN = 32
NOUT = 32
NUSED = 6

_u32, _f32, _int, _vp = C.c_uint32, C.c_float, C.c_int, C.c_void_p
MPNN_SIGNATURES = {
    "rfd_mm_bf16": [_vp, _vp, _u32, _vp, _u32, _u32, _vp, _vp, _u32, _u32, _u32, _u32, _int],
    "rfd_mm_tn_bf16": [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _u32, _u32, _int],
    "rfd_colsum": [_vp, _vp, _u32, _vp, _u32, _u32, _int],
    "rfd_mpnn_ewise": [_vp, _int, _u32, _u32, _f32, _vp, _u32, _vp, _u32, _vp, _u32],
    "rfd_mpnn_edge_fwd": [_vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32],
    "rfd_mpnn_edge_bwd": [_vp, _u32, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp,
                          _int],
    "rfd_mse_loss": [_vp, _vp, _u32, _vp, _u32, _u32, _u32, _vp, _u32, C.POINTER(C.c_double)],
    "rfd_adam": [_vp, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _f32, _u32],
}


def _bind():
    L = lib()
    if not getattr(L, "_mpnn_bound", False):
        for name, args in MPNN_SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L._mpnn_bound = True
    return L


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


def _p(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def ewise(ctx, op, M, nc, a, b, out, alpha=1.0):
    """rfd_mpnn_ewise on row-strided 2-D device views."""
    _bind()
    ctx.call("rfd_mpnn_ewise", op, M, nc, alpha, _p(a), a.stride(0),
             _p(b) if b is not None else None, b.stride(0) if b is not None else 0,
             _p(out), out.stride(0))


@dataclass
class GraphView:
    """Half-bond rows on the device: hoff [N+1], partner at hj[h * stride],
    mirror half hm [H].  The arrays belong to ``owner`` (a NativeSystem),
    which the view keeps alive."""
    N: int
    H: int
    hoff: torch.Tensor
    hj: torch.Tensor
    stride: int
    hm: torch.Tensor
    owner: object = None


class WarmStartNet:
    """Parameters (FP32 master copy), gradients and Adam moments live in one
    flat device buffer each; ``views`` maps names to slices."""

    def __init__(self, ctx, seed: int = 1, dec_scale: float = 0.0):
        self.ctx = ctx
        self.lib = _bind()
        self.shapes = param_shapes()
        n = sum(int(np.prod(s)) for s in self.shapes.values())
        self.n = n
        dev = "cuda"
        self.flat = torch.zeros(n, dtype=torch.float32, device=dev)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.t = 0
        self.views, self.gviews = {}, {}
        off = 0
        for k, s in self.shapes.items():
            sz = int(np.prod(s))
            self.views[k] = self.flat[off:off + sz].view(*s)
            self.gviews[k] = self.grad[off:off + sz].view(*s)
            off += sz
        rng = np.random.default_rng(seed)
        init = {}
        for k, s in self.shapes.items():
            if k.startswith("b"):
                init[k] = np.zeros(s, np.float32)
            elif k == "W_dec":
                init[k] = (rng.standard_normal(s) * dec_scale).astype(np.float32)
            else:
                init[k] = (rng.standard_normal(s) * math.sqrt(2.0 / s[0])).astype(np.float32)
        self.set_params(init)
        self._bufs = {}

    # -- parameters --------------------------------------------------------
    def set_params(self, p: dict):
        for k, v in p.items():
            self.views[k].copy_(torch.as_tensor(np.asarray(v, np.float32)))

    def params(self) -> dict:
        return {k: v.detach().cpu().numpy().copy() for k, v in self.views.items()}

    def grads(self) -> dict:
        return {k: v.detach().cpu().numpy().copy() for k, v in self.gviews.items()}

    # -- buffers -----------------------------------------------------------
    def _buf(self, name, rows, cols):
        t = self._bufs.get(name)
        if t is None or t.shape[0] < rows:
            t = torch.zeros((rows, cols), dtype=torch.float32, device="cuda")
            self._bufs[name] = t
        return t[:rows]

    def _call(self, name, *args):
        self.ctx.call(name, *args)

    def _mm(self, X, W, bias, Y, M, K, Nc, acc=0, trans=False):
        ldx, ldy = X.stride(0), Y.stride(0)
        if trans:   # B = W^T with W [Nc x K] row-major
            wsk, wsc = 1, W.stride(0)
        else:
            wsk, wsc = W.stride(0), 1
        self._call("rfd_mm_bf16", _p(X), ldx, _p(W), wsk, wsc,
                   _p(bias) if bias is not None else None, _p(Y), ldy, M, K, Nc, acc)

    def _tn(self, X, D, G, M, K, Nc, acc=0):
        self._call("rfd_mm_tn_bf16", _p(X), X.stride(0), _p(D), D.stride(0), _p(G), M, K, Nc, acc)

    def _ew(self, op, M, nc, a, b, out, alpha=1.0):
        ewise(self.ctx, op, M, nc, a, b, out, alpha)

    # -- forward / backward ------------------------------------------------
    def forward(self, g: GraphView, X: torch.Tensor, E: torch.Tensor, train: bool = False):
        """Y [N x 32] (device); keeps what backward needs when train."""
        N, v = g.N, self.views
        P0 = self._buf("P0", N, HID)
        self._mm(X, v["W_enc"], v["b_enc"], P0, N, NIN, HID)
        Hs = [self._buf("H0", N, HID)]
        self._ew(0, N, HID, P0, None, Hs[0])
        for r in range(ROUNDS):
            H = Hs[-1]
            Z = self._buf(f"Z{r}", N, HID)
            self._mm(H, v[f"Wa{r}"], None, Z, N, HID, HID)
            HA = self._buf(f"HA{r}", N, 2 * HID)
            self._ew(4, N, HID, H, None, HA[:, :HID])
            Q = self._buf(f"Q{r}", g.H, HID) if train else None
            self._call("rfd_mpnn_edge_fwd", N, _p(g.hoff), _p(g.hj), g.stride, _p(E),
                       _p(v[f"We{r}"]), _p(v[f"bm{r}"]), _p(Z), _p(Q) if Q is not None else None,
                       C.c_void_p(HA.data_ptr() + 4 * HID), HA.stride(0))
            Cb = self._buf(f"C{r}", N, HID)
            self._mm(HA, v[f"Wu{r}"], v[f"bu{r}"], Cb, N, 2 * HID, HID)
            Hn = self._buf(f"H{r + 1}", N, HID)
            self._ew(1, N, HID, Cb, H, Hn)
            Hs.append(Hn)
        Y = self._buf("Y", N, NOUT)
        self._mm(Hs[-1], v["W_dec"], v["b_dec"], Y, N, HID, NOUT)
        self._fw = (g, X, E, N)
        return Y

    def backward(self, dY: torch.Tensor, accumulate: bool = False):
        """Gradients into self.grad (summed into it when accumulate)."""
        g, X, E, N = self._fw
        v, gv = self.views, self.gviews
        acc = 1 if accumulate else 0
        H2 = self._bufs[f"H{ROUNDS}"][:N]
        self._tn(H2, dY, gv["W_dec"], N, HID, NOUT, acc)
        self._call("rfd_colsum", _p(dY), dY.stride(0), _p(gv["b_dec"]), N, NOUT, acc)
        dH = self._buf("dH", N, HID)
        self._mm(dY, v["W_dec"], None, dH, N, NOUT, HID, trans=True)
        for r in reversed(range(ROUNDS)):
            Hin = self._bufs[f"H{r}"][:N]
            HA = self._bufs[f"HA{r}"][:N]
            Cb = self._bufs[f"C{r}"][:N]
            Q = self._bufs[f"Q{r}"][:g.H]
            dC = self._buf("dC", N, HID)
            self._ew(2, N, HID, dH, Cb, dC)
            self._tn(HA, dC, gv[f"Wu{r}"], N, 2 * HID, HID, acc)
            self._call("rfd_colsum", _p(dC), dC.stride(0), _p(gv[f"bu{r}"]), N, HID, acc)
            dHA = self._buf("dHA", N, 2 * HID)
            self._mm(dC, v[f"Wu{r}"], None, dHA, N, HID, 2 * HID, trans=True)
            dZ = self._buf("dZ", N, HID)
            self._call("rfd_mpnn_edge_bwd", N, _p(g.hoff), _p(g.hj), g.stride, _p(g.hm), _p(E),
                       _p(Q), C.c_void_p(dHA.data_ptr() + 4 * HID), dHA.stride(0), _p(dZ),
                       _p(gv[f"We{r}"]), _p(gv[f"bm{r}"]), acc)
            self._tn(Hin, dZ, gv[f"Wa{r}"], N, HID, HID, acc)
            # dH_in = dH + dHA[:, :64] + dZ Wa^T
            dHn = self._buf("dHn", N, HID)
            self._ew(3, N, HID, dH, dHA[:, :HID], dHn)
            self._mm(dZ, v[f"Wa{r}"], None, dHn, N, HID, HID, acc=1, trans=True)
            self._ew(4, N, HID, dHn, None, dH)
        dP0 = self._buf("dP0", N, HID)
        self._ew(2, N, HID, dH, self._bufs["P0"][:N], dP0)
        self._tn(X, dP0, gv["W_enc"], N, NIN, HID, acc)
        self._call("rfd_colsum", _p(dP0), dP0.stride(0), _p(gv["b_enc"]), N, HID, acc)

    def loss(self, Y: torch.Tensor, T: torch.Tensor, want_grad: bool = True):
        N = Y.shape[0]
        dY = self._buf("dY", N, NOUT)
        if want_grad:
            dY.zero_()
        out = C.c_double()
        self._call("rfd_mse_loss", _p(Y), Y.stride(0), _p(T), T.stride(0), N, NUSED,
                   _p(dY) if want_grad else None, dY.stride(0), C.byref(out))
        return out.value, dY

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        self._call("rfd_adam", _p(self.flat), _p(self.grad), _p(self.m), _p(self.v), self.n,
                   lr, b1, b2, eps, self.t)


# ---------------------------------------------------------------- samples
def graph_of(sys) -> GraphView:
    """The bond graph of a NativeSystem as a GraphView."""
    hoff, half, _, _ = sys.bond_graph()
    hm = sys.half_mirror()
    return GraphView(sys.N, half.shape[0], hoff, half, 2, hm, owner=sys)


def scales(sys, dstrain: float) -> tuple[float, float]:
    L = sys.loading()
    su = abs(dstrain) * L.size0
    r_mean = float(sys.mean_radius())
    return su, su / r_mean


@dataclass
class Sample:
    """One load step of one specimen: inputs at the step start, targets."""
    g: GraphView
    X: torch.Tensor
    E: torch.Tensor
    T: torch.Tensor


def capture_inputs(sys, dstrain: float):
    """(X, E) of the system's current state for a step of dstrain."""
    su, sth = scales(sys, dstrain)
    X = torch.zeros((sys.N, NIN), dtype=torch.float32, device="cuda")
    H = sys.bonds.n * 2
    E = torch.zeros((H, 4), dtype=torch.float32, device="cuda")
    sys.mpnn_inputs(su, sth, X, E)
    return X, E


def targets_after_step(sys, dstrain: float, X: torch.Tensor) -> torch.Tensor:
    """Corrections the network should output: (du_k - du_{k-1}) / su and
    (dth_k - dth_{k-1}) / sth, from the buffers after step k (X holds the
    scaled du_{k-1}, dth_{k-1} of the step start)."""
    su, sth = scales(sys, dstrain)
    du, _, dth, _ = sys.fire_buffers()
    T = torch.zeros((sys.N, NOUT), dtype=torch.float32, device="cuda")
    ewise(sys.ctx, 6, sys.N, 3, du, X[:, 0:3], T[:, 0:3], 1.0 / su)
    ewise(sys.ctx, 6, sys.N, 3, dth, X[:, 6:9], T[:, 3:6], 1.0 / sth)
    return T


def predict(net: WarmStartNet, sys, dstrain: float):
    """Predicted (dU, dTh) device tensors [N x 3] for the next load step."""
    su, sth = scales(sys, dstrain)
    X, E = capture_inputs(sys, dstrain)
    Y = net.forward(graph_of(sys), X, E, train=False)
    dU = torch.empty((sys.N, 3), dtype=torch.float32, device="cuda")
    dTh = torch.empty((sys.N, 3), dtype=torch.float32, device="cuda")
    ewise(sys.ctx, 5, sys.N, 3, X[:, 0:3], Y[:, 0:3], dU, su)
    ewise(sys.ctx, 5, sys.N, 3, X[:, 6:9], Y[:, 3:6], dTh, sth)
    return dU, dTh


def collect_samples(sys, dstrain: float, fp, nsteps: int, warm=None,
                    log=None) -> tuple[list, list]:
    """Run nsteps load steps of dstrain (plain FIRE, or warm-started by
    ``warm(sys) -> (dU, dTh)`` when given) and keep, for every step that has
    two previous increments, the network inputs at its start and the
    corrections it needed.  ``log(k, result)`` is called after each step.
    Returns (samples, results)."""
    samples, results = [], []
    g = graph_of(sys)
    for k in range(nsteps):
        X = E = None
        if k >= 2:
            X, E = capture_inputs(sys, dstrain)
        dU = dTh = None
        if warm is not None and k >= 2:
            dU, dTh = warm(sys)
        r = sys.fire_load_step(dstrain, fp, dU, dTh)
        results.append(r)
        if log is not None:
            log(k, r)
        if X is not None:
            samples.append(Sample(g, X, E, targets_after_step(sys, dstrain, X)))
    return samples, results


def scale_targets(ctx, samples: list, ts: float):
    """Divide every sample's targets by ts (in place, on the device)."""
    for s in samples:
        ewise(ctx, 7, s.T.shape[0], NUSED, s.T, None, s.T, 1.0 / ts)


def mean_square(net: WarmStartNet, samples: list, zero: bool = False) -> float:
    """Mean loss of the network (or of a zero output) over samples."""
    tot = 0.0
    for s in samples:
        Y = net._buf("Yzero", s.T.shape[0], NOUT) if zero else net.forward(s.g, s.X, s.E)
        if zero:
            Y.zero_()
        tot += net.loss(Y, s.T, want_grad=False)[0]
    return tot / max(len(samples), 1)


def train(net: WarmStartNet, samples: list, epochs: int, lr: float = 1e-3,
          val: list | None = None) -> dict:
    """Adam on normalised corrections (one update per sample per epoch).

    The targets are divided by their RMS over the training set; the decoder
    is rescaled by it at the end, so the trained network outputs corrections
    in the units of ``targets_after_step``.  With validation samples the
    parameters of the best validation epoch are kept, and the correction head
    is switched off (zero decoder: the start repeats the last increment) when
    it does not beat a zero correction on them.  Returns the loss history."""
    ts = math.sqrt(max(mean_square(net, samples, zero=True), 1e-30))
    scale_targets(net.ctx, samples, ts)
    if val:
        scale_targets(net.ctx, val, ts)
    hist = {"train": [], "val": [], "val_zero": None, "scale": ts, "head": True}
    best, best_val = net.flat.clone(), float("inf")
    if val:
        hist["val_zero"] = mean_square(net, val, zero=True)
    for _ in range(epochs):
        tot = 0.0
        for s in samples:
            Y = net.forward(s.g, s.X, s.E, train=True)
            loss, dY = net.loss(Y, s.T)
            net.backward(dY)
            net.adam(lr)
            tot += loss
        hist["train"].append(tot / max(len(samples), 1))
        if val:
            v = mean_square(net, val)
            hist["val"].append(v)
            if v < best_val:
                best_val = v
                best.copy_(net.flat)
    if val:
        net.flat.copy_(best)
        if not best_val < hist["val_zero"]:
            hist["head"] = False
            net.views["W_dec"].zero_()
            net.views["b_dec"].zero_()
    scale_targets(net.ctx, samples, 1.0 / ts)
    if val:
        scale_targets(net.ctx, val, 1.0 / ts)
    for k in ("W_dec", "b_dec"):
        ewise(net.ctx, 7, 1, net.views[k].numel(), net.views[k].view(1, -1), None,
              net.views[k].view(1, -1), ts)
    return hist
