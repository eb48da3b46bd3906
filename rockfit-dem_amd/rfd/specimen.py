"""Seeded synthetic specimens.

The reference packs its samples with a wall-clock-seeded cuRAND fill and a
soft-sphere relaxation (md_createcylindersample.cu:136-343) and loads them
from text files that it does not ship.  Specimens here are reproducible:
particle centres on a jittered face-centred-cubic lattice clipped to the
specimen shape (a cylinder along z, or a disc whose axis is z and whose load
axis is y), radii drawn from U[r_lo, r_hi], all from a Philox counter-based
generator seeded with ``seed``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BASE_SEED = 20261203


@dataclass
class Specimen:
    R: np.ndarray            # (N, 3) float64 centres, mm
    radius: np.ndarray       # (N,) float64, mm
    seed: int
    shape: str
    D: float
    H: float

    @property
    def N(self) -> int:
        return int(self.R.shape[0])

    def soa_f32(self) -> np.ndarray:
        """Positions as the reference's SoA-by-component float array [3N]."""
        return np.ascontiguousarray(self.R.T.astype(np.float32)).reshape(-1)


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def fcc_points(extent: tuple[float, float, float], d_nn: float) -> np.ndarray:
    """FCC lattice points with nearest-neighbour distance d_nn filling the box
    [-e/2, e/2]^3 (plus one cell of margin)."""
    a = d_nn * np.sqrt(2.0)
    n = [int(np.ceil(e / a)) + 2 for e in extent]
    ii, jj, kk = np.meshgrid(*(np.arange(-m // 2 - 1, m // 2 + 2) for m in n), indexing="ij")
    base = np.stack([ii, jj, kk], -1).reshape(-1, 3).astype(np.float64) * a
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    pts = (base[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    return pts


def cylinder(D: float, H: float, center=(0.0, 0.0, 0.0), d_nn: float = 1.0,
             r_lo: float = 0.4, r_hi: float = 0.6, jitter: float = 0.02,
             seed: int = BASE_SEED, monodisperse: float | None = None,
             flat_ends: bool = False) -> Specimen:
    """Cylinder of diameter D and height H (axis z).  Particles are kept when
    the whole sphere of the mean radius lies inside.  ``monodisperse`` sets
    every radius to that value instead of drawing U[r_lo, r_hi].
    ``flat_ends`` keeps the lowest and highest lattice layers unjittered in z
    (capped ends), so that a platen placed at the end layer holds the whole
    layer from the first step instead of a few high particles."""
    g = rng(seed)
    pts = fcc_points((D, D, H), d_nn)
    rm = monodisperse if monodisperse is not None else 0.5 * (r_lo + r_hi)
    keep = (pts[:, 0] ** 2 + pts[:, 1] ** 2 <= (0.5 * D - rm) ** 2) & \
           (np.abs(pts[:, 2]) <= 0.5 * H - rm)
    pts = pts[keep]
    # deterministic order: sort by (z, y, x) lattice position
    order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))
    pts = pts[order]
    jit = g.uniform(-jitter, jitter, size=pts.shape)
    if flat_ends:
        z = pts[:, 2]
        jit[(z == z.max()) | (z == z.min()), 2] = 0.0
    pts = pts + jit
    if monodisperse is not None:
        rad = np.full(len(pts), monodisperse)
    else:
        rad = g.uniform(r_lo, r_hi, size=len(pts))
    return Specimen(pts + np.asarray(center, dtype=np.float64), rad, seed, "cylinder", D, H)


def disc(D: float, T: float, center=(0.0, 0.0, 0.0), **kw) -> Specimen:
    """Brazilian disc of diameter D and thickness T: axis along z, the platens
    act along y (md_braziltest_simple2.cu:30-34)."""
    s = cylinder(D, T, center=center, **kw)
    s.shape = "disc"
    return s


def bonds_by_gap(R: np.ndarray, radius: np.ndarray, gap: float) -> np.ndarray:
    """Undirected pairs (i<j) whose surface gap |rj-ri| - ri - rj is below
    ``gap`` (brute force over a cell grid, host side)."""
    N = len(R)
    rmax = float(radius.max())
    cut = 2 * rmax + gap
    lo = R.min(0)
    cell = np.floor((R - lo) / cut).astype(np.int64)
    dims = cell.max(0) + 1
    key = (cell[:, 2] * dims[1] + cell[:, 1]) * dims[0] + cell[:, 0]
    order = np.argsort(key, kind="stable")
    skey = key[order]
    starts = np.searchsorted(skey, np.arange(dims.prod() + 1))
    pairs = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb = cell + np.array([dx, dy, dz])
                ok = np.all((nb >= 0) & (nb < dims), axis=1)
                nkey = (nb[:, 2] * dims[1] + nb[:, 1]) * dims[0] + nb[:, 0]
                idx = np.nonzero(ok)[0]
                s0 = starts[nkey[idx]]
                s1 = starts[nkey[idx] + 1]
                cnt = s1 - s0
                rep_i = np.repeat(idx, cnt)
                offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
                rep_j = order[np.repeat(s0, cnt) + offs]
                m = rep_i < rep_j
                pairs.append(np.stack([rep_i[m], rep_j[m]], 1))
    P = np.concatenate(pairs, 0)
    d = np.linalg.norm(R[P[:, 1]] - R[P[:, 0]], axis=1)
    m = d - radius[P[:, 0]] - radius[P[:, 1]] < gap
    P = P[m]
    P = P[np.lexsort((P[:, 1], P[:, 0]))]
    return P.astype(np.uint32)


def truncated_normal(g: np.random.Generator, mean, sd, n: int) -> np.ndarray:
    """N(mean, sd) truncated to > 0 by resampling."""
    mean = np.broadcast_to(np.asarray(mean, dtype=np.float64), (n,)).copy()
    sd = np.broadcast_to(np.asarray(sd, dtype=np.float64), (n,)).copy()
    out = g.normal(mean, sd)
    bad = out <= 0
    while bad.any():
        out[bad] = g.normal(mean[bad], sd[bad])
        bad = out <= 0
    return out
