"""Pins for the CPU oracle of the warm-start network (oracle/mpnn_oracle.py):
BF16 rounding against hand-computed values, and the backward pass against
central finite differences of the exact float64 network on a random graph, so
that a wrong transpose, a dropped residual path or a wrong mirror index in the
restated gradients fails."""
import numpy as np
import pytest

from oracle import mpnn_oracle as mo


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # bf16 spacing at 1.0
    x = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + ulp / 2 + 2 ** -20, -3.0,
                  3.14159265, 0.0], np.float32)
    got = mo.bf16(x)
    want = [1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -3.0, 3.140625, 0.0]
    np.testing.assert_array_equal(got, np.array(want))
    assert mo.bf16(np.array([one]))[0] == 1.0


def _random_graph(rng, N=12, extra=20):
    pairs = set()
    for i in range(N - 1):
        pairs.add((i, i + 1))
    while len(pairs) < N - 1 + extra:
        a, b = sorted(rng.choice(N, 2, replace=False))
        pairs.add((int(a), int(b)))
    pairs = sorted(pairs)
    rows = [[] for _ in range(N)]
    for b, (i, j) in enumerate(pairs):
        rows[i].append((j, b))
        rows[j].append((i, b))
    hoff, hj, hb = [0], [], []
    for r in rows:
        r.sort()
        for j, b in r:
            hj.append(j)
            hb.append(b)
        hoff.append(len(hj))
    hm = np.zeros(len(hj), np.int64)
    where = {}
    for h, b in enumerate(hb):
        where.setdefault(b, []).append(h)
    for b, (h1, h2) in where.items():
        hm[h1], hm[h2] = h2, h1
    return mo.Graph(hoff, hj, hm)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    g = _random_graph(rng)
    N, H = g.N, len(g.hj)
    X = np.zeros((N, mo.NIN))
    X[:, :12] = rng.standard_normal((N, 12))
    E = rng.standard_normal((H, 4))
    p = {k: rng.standard_normal(s) * (0.5 if len(s) > 1 else 0.1)
         for k, s in mo.param_shapes().items()}
    T = rng.standard_normal((N, mo.NOUT)) * 0.1

    def loss(pp):
        Y, _ = mo.forward(pp, g, X, E, bf16_ops=False)
        return mo.loss_and_grad_y(Y, T)[0]

    Y, cache = mo.forward(p, g, X, E, bf16_ops=False)
    L0, dY = mo.loss_and_grad_y(Y, T)
    gr = mo.backward(p, g, cache, dY, bf16_ops=False)
    eps = 1e-6
    for k, s in mo.param_shapes().items():
        flat = p[k].reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for q in idx:
            pp = {kk: v.copy() for kk, v in p.items()}
            pp[k].reshape(-1)[q] += eps
            lp = loss(pp)
            pp[k].reshape(-1)[q] -= 2 * eps
            lm = loss(pp)
            fd = (lp - lm) / (2 * eps)
            an = gr[k].reshape(-1)[q]
            assert an == pytest.approx(fd, rel=1e-5, abs=1e-9), (k, q)
