"""GPU checks of the learned warm start (csrc/rfd_mpnn.hip through the C ABI):

* BF16 matrix-core products against the CPU oracle's BF16-rounded products
  (FP32 vs FP64 accumulation: 1e-4 of the largest entry);
* the network's forward output and every parameter gradient against the CPU
  oracle (oracle/mpnn_oracle.py, itself pinned by finite differences) on a
  real bond graph, 5e-3 relative (Frobenius) per tensor;
* training on plain-FIRE relaxations lowers the loss, and the trained network
  reaches the same residual as plain FIRE in fewer iterations.
"""
import numpy as np
import pytest
import torch

from oracle import mpnn_oracle as mo
from rfd import native, native_cases, warmstart as ws

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.as_tensor(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("M,K,Nc,trans", [(200, 64, 64, False), (131, 32, 32, False),
                                          (77, 128, 64, False), (150, 64, 128, True),
                                          (64, 32, 64, True)])
def test_gemm_matches_bf16_oracle(ctx, M, K, Nc, trans):
    net = ws.WarmStartNet(ctx)
    rng = np.random.default_rng(M + K)
    X = rng.standard_normal((M, K)).astype(np.float32)
    W = rng.standard_normal((Nc, K) if trans else (K, Nc)).astype(np.float32)
    b = rng.standard_normal(Nc).astype(np.float32)
    Y = torch.zeros((M, Nc), device="cuda")
    Xd, Wd, bd = _t(X), _t(W), _t(b)
    net._mm(Xd, Wd, bd, Y, M, K, Nc, trans=trans)
    ref = mo.bf16(X) @ mo.bf16(W.T if trans else W) + b
    got = Y.cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-4 * np.abs(ref).max())


@pytest.mark.parametrize("M,K,Nc", [(5000, 128, 64), (300, 32, 64), (4097, 64, 32)])
def test_transposed_product_and_colsum(ctx, M, K, Nc):
    net = ws.WarmStartNet(ctx)
    rng = np.random.default_rng(M)
    X = rng.standard_normal((M, K)).astype(np.float32)
    D = rng.standard_normal((M, Nc)).astype(np.float32)
    G = torch.zeros((K, Nc), device="cuda")
    Xd, Dd = _t(X), _t(D)
    net._tn(Xd, Dd, G, M, K, Nc)
    ref = mo.bf16(X).T @ mo.bf16(D)
    np.testing.assert_allclose(G.cpu().numpy(), ref, rtol=0, atol=1e-4 * np.abs(ref).max())
    g = torch.zeros(Nc, device="cuda")
    ctx.call("rfd_colsum", ws._p(Dd), Nc, ws._p(g), M, Nc, 0)
    np.testing.assert_allclose(g.cpu().numpy(), D.astype(np.float64).sum(0), rtol=0,
                               atol=1e-4 * np.abs(D).sum(0).max())


def _small_system(ctx, seed=3):
    nc = native_cases.uniaxial_cylinder(ctx, 5.0, 7.0, seed=seed, strength=(1e12, 1e12))
    return nc


def test_forward_backward_match_oracle(ctx):
    nc = _small_system(ctx)
    s = nc.sys
    g = ws.graph_of(s)
    rng = np.random.default_rng(11)
    N, H = g.N, g.H
    X = np.zeros((N, 32), np.float32)
    X[:, :12] = rng.standard_normal((N, 12))
    E = rng.standard_normal((H, 4)).astype(np.float32)
    T = (rng.standard_normal((N, 32)) * 0.3).astype(np.float32)
    net = ws.WarmStartNet(ctx, seed=5, dec_scale=0.3)
    p = net.params()
    Y = net.forward(g, _t(X), _t(E), train=True)
    loss, dY = net.loss(Y, _t(T))
    net.backward(dY)
    og = mo.Graph(g.hoff.cpu().numpy(), g.hj[:, 0].cpu().numpy(), g.hm.cpu().numpy())
    Yo, cache = mo.forward(p, og, X, E, bf16_ops=True)
    lo, dYo = mo.loss_and_grad_y(Yo, T)
    gro = mo.backward(p, og, cache, dYo, bf16_ops=True)
    Yg = Y.cpu().numpy()
    assert np.linalg.norm(Yg - Yo) <= 5e-3 * np.linalg.norm(Yo)
    assert loss == pytest.approx(lo, rel=5e-3)
    grg = net.grads()
    for k in gro:
        a, b = grg[k].astype(np.float64), gro[k]
        assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b) + 1e-12, k


def test_training_lowers_loss_and_saves_iterations(ctx):
    """Train on the plain-FIRE relaxations of two specimens (validation on a
    third), then run load steps of a fourth specimen warm-started by the
    network: every step reaches the same residual as plain FIRE on its twin,
    in at most half the FIRE iterations over the warm-started steps."""
    dstrain = -1e-5

    def spec(seed, D=5.0, H=7.0):
        nc = native_cases.uniaxial_cylinder(ctx, D, H, seed=seed, strength=(1e12, 1e12))
        nc.sys.seat_platens(preload=0.05)
        return nc

    a = spec(21)
    fp = native.fire_params(a.mat.E, 0.5, a.mat.dt, rel_tol=1e-6, max_iter=50000,
                            check_every=100)
    samples, _ = ws.collect_samples(a.sys, dstrain, fp, 12)
    more, _ = ws.collect_samples(spec(24, 6.0, 8.0).sys, dstrain, fp, 12)
    val, _ = ws.collect_samples(spec(23).sys, dstrain, fp, 8)
    net = ws.WarmStartNet(ctx, seed=2)
    hist = ws.train(net, samples + more, epochs=40, lr=1e-3, val=val)
    assert hist["train"][-1] < hist["train"][0], hist["train"][::8]
    b_plain, b_warm = spec(22), spec(22)
    _, rp = ws.collect_samples(b_plain.sys, dstrain, fp, 8)
    _, rw = ws.collect_samples(b_warm.sys, dstrain, fp, 8,
                               warm=lambda s: ws.predict(net, s, dstrain))
    for r in rp + rw:
        assert r.converged and r.fmax < fp.ftol
    it_p = sum(r.iterations for r in rp[2:])
    it_w = sum(r.iterations for r in rw[2:])
    assert 2 * it_w <= it_p, (it_w, it_p, hist)
    assert sum(r.warm_used for r in rw[2:]) >= 4
