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
    third), then run load steps of a fourth specimen three ways: plain FIRE,
    the skip path alone (the untrained network, whose zero correction head
    repeats the last step's increment) and the trained network.  Every step
    of every run reaches the same residual; the skip path takes at most half
    the iterations of plain FIRE; the trained network is kept to the skip
    path's gain: with its correction head on (it beat a zero correction on
    the validation steps) it takes no more than 1.25x the skip path's
    iterations, with the head off it is the skip path, step for step.  The
    iterations of the three runs are printed: on this elastic specimen the
    head is switched off, so what the network adds over the skip path is
    nothing."""
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
    skip = ws.WarmStartNet(ctx, seed=2)
    runs = {}
    for name, nt in (("plain", None), ("skip", skip), ("learned", net)):
        b = spec(22)
        warm = (lambda s, nt=nt: ws.predict(nt, s, dstrain)) if nt is not None else None
        runs[name] = ws.collect_samples(b.sys, dstrain, fp, 8, warm=warm)[1]
    for name, rr in runs.items():
        for r in rr:
            assert r.converged and r.fmax < fp.ftol, name
    it = {k: sum(r.iterations for r in rr[2:]) for k, rr in runs.items()}
    print(f"[warm start] head {'on' if hist['head'] else 'off'}; FIRE iterations over the "
          f"warm-started steps: {it}", flush=True)
    assert 2 * it["skip"] <= it["plain"], it
    assert sum(r.warm_used for r in runs["skip"][2:]) >= 4
    if hist["head"]:
        assert it["learned"] <= 1.25 * it["skip"], it
    else:
        assert [r.iterations for r in runs["learned"]] == [r.iterations for r in runs["skip"]]


def test_mpnn_inputs_match_host(ctx):
    """rfd_system_mpnn_inputs against the same quantities formed on the host
    after three load steps with weak bonds (some broken): the scaled last two
    increments, the largest and mean utilisation E strain / tensile strength
    of each particle's intact bonds, the broken share of its records, and per
    half-bond the scaled strain and the unit direction."""
    nc = native_cases.uniaxial_cylinder(ctx, 5.0, 7.0, seed=25, strength=(8e6, 16e6))
    s = nc.sys
    s.seat_platens(preload=0.02)
    fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-5, max_iter=20000,
                            check_every=100)
    for _ in range(3):
        s.fire_load_step(-1e-3, fp)
    su, sth = ws.scales(s, -1e-3)
    X, E = ws.capture_inputs(s, -1e-3)
    X, E = X.cpu().numpy(), E.cpu().numpy()
    N = s.N
    du, dup, dth, dthp = (t.cpu().numpy().reshape(N, 3) for t in s.fire_buffers())
    np.testing.assert_allclose(X[:, 0:3], du / np.float32(su), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(X[:, 3:6], dup / np.float32(su), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(X[:, 6:9], dth / np.float32(sth), rtol=1e-6, atol=1e-12)
    assert np.all(X[:, 15:] == 0)
    R = s.state64()["R"]
    _, fl = s.bond_state()
    hoff, half, bond, _ = s.bond_graph()
    hoff = hoff.cpu().numpy().astype(np.int64)
    half = half.cpu().numpy()
    bond = bond.cpu().numpy()
    il = bond[:, 0].astype(np.float64)
    sc = bond[:, 2].astype(np.float64)
    broken = 0
    for i in range(N):
        us, nb = [], 0
        for h in range(hoff[i], hoff[i + 1]):
            j, q = int(half[h, 0]), int(half[h, 1])
            r = R[j] - R[i]
            L = np.sqrt(r @ r)
            strain = L * il[q] - 1.0
            if fl[q] == 1:
                us.append(s.params.E * strain / sc[q])
                assert E[h, 0] == pytest.approx(strain * 1e3, rel=1e-5, abs=1e-7)
            else:
                nb += 1
                assert E[h, 0] == 0
            np.testing.assert_allclose(E[h, 1:], r / L, rtol=1e-6, atol=1e-7)
        n = len(us) + nb
        broken += nb
        if us:
            assert X[i, 12] == pytest.approx(max(0.0, max(us)), rel=1e-5, abs=1e-7)
            assert X[i, 13] == pytest.approx(np.mean(us), rel=1e-4, abs=1e-6)
        assert X[i, 14] == pytest.approx(nb / n if n else 0.0, rel=1e-6)
    assert broken > 0
