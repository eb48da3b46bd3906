"""Native-tier specimens (polydisperse, per-bond strengths) for tests and the
benchmark."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import specimen
from .native import (Bonds, NativeSystem, PackParams, loading_from_compression, morton_order,
                     params_from_material, strengths)
from .params import Compression, Material, SampleBox, STRESS, VELOCITY


def material_c(r_min: float = 0.4) -> Material:
    """BASELINE material: E = 30 GPa, nu = 0.25, rho = 2650 kg/m^3; the time
    step follows the reference rule (md_data_types.h:166-176) at r_min."""
    return Material(E=30e9 / STRESS, nu=0.25, rho=2650 / 1e3, p_r=r_min, b_r=r_min,
                    b_d=2 * r_min, sc=20e6 / STRESS, tc=40e6 / STRESS)


@dataclass
class NativeCase:
    sys: NativeSystem
    mat: Material
    box: SampleBox
    cmp: Compression
    perm: np.ndarray


def uniaxial_cylinder(ctx, D: float, H: float, seed: int = specimen.BASE_SEED,
                      bond_gap: float | None = None, refmode: int = 0,
                      strength=(20e6, 40e6), morton: bool = True, rebuild_every: int = 20,
                      v_load: float = -10.0 * 0.02 / VELOCITY,
                      flat_ends: bool = False, fp64: bool = False,
                      skin: float = 0.05) -> NativeCase:
    """Polydisperse cylinder (r ~ U[0.4, 0.6] mm) on a jittered FCC lattice of
    spacing 1 mm.  bond_gap=None joins every nearest lattice neighbour (Nb/N
    close to 6); otherwise pairs with surface gap < bond_gap are joined."""
    mat = material_c(0.4)
    g = specimen.rng(seed + 1)
    a = 2.0 * 2.9 * 0.6
    box = SampleBox(D, D, H, space=(D + 16.0, D + 16.0, H + 12.0), cell_a=a, p_r=0.5)
    sp = specimen.cylinder(D, H, center=box.center, d_nn=1.0, seed=seed, flat_ends=flat_ends)
    R, rad = sp.R, sp.radius
    perm = morton_order(R, 2.0) if morton else np.arange(len(R))
    R, rad = R[perm], rad[perm]
    p = params_from_material(mat, refmode=refmode, rebuild_every=rebuild_every, fp64=fp64,
                             skin=skin)
    p.farcut = 1.45
    sysn = NativeSystem(ctx, p, R, rad)
    gap = bond_gap if bond_gap is not None else 1.0 - 0.8 + 0.01
    pairs = sysn.find_pairs(gap)
    n = len(pairs)
    sc, tc = strengths(g, n, strength[0], strength[1])
    br = np.minimum(rad[pairs[:, 0]], rad[pairs[:, 1]]).astype(np.float32)
    sysn.set_bonds(Bonds(pairs[:, 0].copy(), pairs[:, 1].copy(), br, sc, tc))
    cmp = Compression("uniaxial", box.L, box.center, 0.5, mat.dt, v_load=v_load,
                      min_stress=15e6 / STRESS, max_strain=0.1)
    sysn.set_loading(loading_from_compression(cmp, box.center[2] + box.spacesized2[2], box))
    return NativeCase(sysn, mat, box, cmp, perm)


@dataclass
class BatchCase:
    sys: NativeSystem
    mat: Material
    first: np.ndarray          # specimen k owns particles [first[k], first[k+1])
    n_target: np.ndarray       # drawn particle counts
    D: np.ndarray              # disc diameters (mm)
    T: np.ndarray              # disc thicknesses (mm)
    mu_t: np.ndarray           # mean bond tensile strength per specimen (MPa)
    seeds: np.ndarray
    loadings: list
    bond_first: np.ndarray     # specimen k owns bond records [bond_first[k], bond_first[k+1])


def disc_diameter_for(n: float, d_nn: float = 1.0) -> float:
    """Diameter of a disc of thickness D/4 holding about n particles of an FCC
    lattice of spacing d_nn (number density sqrt(2) / d_nn^3)."""
    return float((16.0 * n * d_nn ** 3 / (np.sqrt(2.0) * np.pi)) ** (1.0 / 3.0))


def _disc_count(D: float, T: float) -> int:
    """Lattice sites of a disc of diameter D and thickness T (the keep rule of
    specimen.cylinder at the mean radius 0.5 mm)."""
    pts = specimen.fcc_points((D, D, T), 1.0)
    return int(np.count_nonzero((pts[:, 0] ** 2 + pts[:, 1] ** 2 <= (0.5 * D - 0.5) ** 2) &
                                (np.abs(pts[:, 2]) <= 0.5 * T - 0.5)))


def _disc_holding(n: float) -> tuple[float, float]:
    """(D, T): thickness a quarter of the nominal diameter for n particles,
    diameter bisected until the disc holds n (the count grows by rings in D)."""
    D0 = disc_diameter_for(n)
    T = D0 / 4.0
    lo, hi = 0.7 * D0, 2.0 * D0
    for _ in range(16):
        mid = 0.5 * (lo + hi)
        if _disc_count(mid, T) < n:
            lo = mid
        else:
            hi = mid
    return hi, T


def brazil_batch(ctx, S: int = 256, seed: int = specimen.BASE_SEED, n_lo: float = 1000.0,
                 n_hi: float = 40000.0, mu_range=(10.0, 30.0), cv: float = 0.2,
                 spacing: float = 3.0, v_load: float = -0.5 / VELOCITY,
                 rebuild_every: int = 20, min_stress_mpa: float = 1.0,
                 hist_cap: int = 1 << 13, build: bool = True):
    """Batched Brazilian discs: S specimens with particle counts log-uniform in
    [n_lo, n_hi], each from its own seed (packing jitter and radii
    r ~ U[0.4, 0.6] mm) and with its own bond strengths N(mu, cv mu) MPa
    (tensile, mu ~ U(mu_range); shear twice that).  Disc k has diameter D_k
    holding its drawn count at thickness about D_k / 4 (axis z); the discs are
    stacked along z with ``spacing`` mm between them, so that no contact forms
    between specimens; platens act along y on each disc separately.
    build=False returns the geometry only (no device work)."""
    g = specimen.rng(seed + 17)
    n_target = np.exp(g.uniform(np.log(n_lo), np.log(n_hi), S))
    mu_t = g.uniform(mu_range[0], mu_range[1], S)
    seeds = seed + 1000 + np.arange(S)
    # the lattice inset at the boundary loses particles: each diameter is
    # bisected until the disc holds its drawn count
    DT = np.array([_disc_holding(n) for n in n_target])
    D, T = DT[:, 0], DT[:, 1]
    mat = material_c(0.4)
    Rs, rads, first, loadings = [], [], [0], []
    z = 0.0
    cx = cy = 0.5 * float(D.max()) + 10.0
    for k in range(S):
        cz = z + 0.5 * T[k]
        z += T[k] + spacing
        sp = specimen.disc(float(D[k]), float(T[k]), center=(cx, cy, cz), d_nn=1.0,
                           seed=int(seeds[k]))
        perm = morton_order(sp.R, 2.0)
        Rs.append(sp.R[perm])
        rads.append(sp.radius[perm])
        first.append(first[-1] + len(perm))
        size = (float(D[k]) - 1.0, float(D[k]) - 1.0, float(T[k]) - 1.0)
        cmp = Compression("brazil", size, (cx, cy, cz), 0.5, mat.dt, v_load=v_load,
                          min_stress=min_stress_mpa * 1e6 / STRESS, max_strain=0.1)
        L = loading_from_compression(cmp, 1e30)
        # platens start on the outermost particle centres (they hold a particle
        # once its centre reaches them), so that the load builds from the first
        # step instead of arriving with an impact after the fast approach
        y = sp.R[:, 1]
        L.top, L.bottom = float(y.max()), float(y.min())
        L.top_r, L.bottom_r = L.top - 0.5, L.bottom + 0.5
        L.size = L.size0 = L.top - L.bottom
        loadings.append(L)
    first = np.asarray(first, np.int64)
    if not build:
        return BatchCase(None, mat, first, n_target, D, T, mu_t, seeds, loadings, None)
    R = np.concatenate(Rs)
    rad = np.concatenate(rads)
    p = params_from_material(mat, refmode=0, rebuild_every=rebuild_every)
    p.farcut = 1.45
    sysn = NativeSystem(ctx, p, R, rad)
    pairs = sysn.find_pairs(1.0 - 0.8 + 0.01)
    spec_of = np.searchsorted(first, pairs[:, 0], side="right") - 1
    if np.any(spec_of != np.searchsorted(first, pairs[:, 1], side="right") - 1):
        raise ValueError("a bond joins two specimens: increase the spacing")
    bond_first = np.searchsorted(spec_of, np.arange(S + 1))
    sc = np.empty(len(pairs), np.float32)
    tc = np.empty(len(pairs), np.float32)
    for k in range(S):
        a, b = bond_first[k], bond_first[k + 1]
        sc[a:b], tc[a:b] = strengths(specimen.rng(int(seeds[k]) + 1), int(b - a),
                                     mu_t[k] * 1e6, 2.0 * mu_t[k] * 1e6, cv)
    br = np.minimum(rad[pairs[:, 0]], rad[pairs[:, 1]]).astype(np.float32)
    sysn.set_bonds(Bonds(pairs[:, 0].copy(), pairs[:, 1].copy(), br, sc, tc))
    sysn.set_specimens(first, loadings, hist_cap)
    return BatchCase(sysn, mat, first, n_target, D, T, mu_t, seeds, loadings, bond_first)


def batch_member(ctx, case: BatchCase, k: int, fp64: bool = False) -> NativeSystem:
    """Specimen k of a batch as a system of its own, at the same place, with
    the batch's cell grid, bonds and loading control: stepped alone it follows
    the batch's specimen k bit for bit.  Call it before the batch steps (the
    bond rest lengths are taken from the current positions).  fp64: on the
    FP64 path instead."""
    bs = case.sys
    a, b = int(case.first[k]), int(case.first[k + 1])
    st = bs.state()
    N = bs.N
    R = st["R"].reshape(3, N).T[a:b]
    params = type(bs.params).from_buffer_copy(bs.params)
    params.fp64 = int(fp64)
    sysn = NativeSystem(ctx, params, R, bs.radius[a:b])
    origin, cell, dims = bs.grid()
    sysn.set_grid(origin, cell, dims)
    bb = bs.bonds
    m = (bb.bi >= a) & (bb.bi < b)
    sysn.set_bonds(Bonds((bb.bi[m] - a).astype(np.uint32), (bb.bj[m] - a).astype(np.uint32),
                         bb.br[m], bb.sc[m], bb.tc[m], None if bb.L0 is None else bb.L0[m]))
    sysn.set_loading(case.loadings[k])
    return sysn


@dataclass
class Packing:
    R: np.ndarray        # (N, 3) positions, Morton order
    radius: np.ndarray   # (N,)
    phi: float           # solid fraction of the container
    result: dict         # the packer's report


def mean_sphere_volume(r_lo: float, r_hi: float) -> float:
    """E[4/3 pi r^3] for r ~ U[r_lo, r_hi]."""
    if r_hi == r_lo:
        return 4.0 / 3.0 * np.pi * r_lo ** 3
    return 4.0 / 3.0 * np.pi * (r_hi ** 4 - r_lo ** 4) / (4.0 * (r_hi - r_lo))


def packed_cylinder(ctx, D: float, H: float, phi: float = 0.6, seed: int = specimen.BASE_SEED,
                    r_lo: float = 0.4, r_hi: float = 0.6, center=(0.0, 0.0, 0.0),
                    s0: float = 0.7, growth: float = 1.005, beta: float = 0.5,
                    momentum: float = 0.9, tol: float = 0.05, stage_iters: int = 64,
                    final_iters: int = 3000) -> Packing:
    """Seeded random packing of a cylinder of diameter D and height H (axis z)
    at solid fraction phi: N = phi V / E[sphere volume] spheres with radii
    U[r_lo, r_hi], uniform random centres (Philox, `seed`), packed on the
    device by radius growth with overlap removal (rfd_system_pack).  At
    phi = 0.6 in small containers the wall layers bring the interior close to
    jamming, so a packing ends with small overlaps: at most `tol` of the
    smaller radius (the mean is about 1 %), the scale of the bonding gap
    0.05 r_mean."""
    g = specimen.rng(seed + 7)
    V = np.pi * (0.5 * D) ** 2 * H
    N = int(round(phi * V / mean_sphere_volume(r_lo, r_hi)))
    # radii as the device keeps them (FP32 values)
    rad = g.uniform(r_lo, r_hi, N).astype(np.float32).astype(np.float64)
    # uniform points in the cylinder, clear of the wall at the initial scale
    rr = 0.5 * D - s0 * rad
    rho = rr * np.sqrt(g.uniform(0.0, 1.0, N))
    th = g.uniform(0.0, 2 * np.pi, N)
    z = g.uniform(-0.5 * H + s0 * rad, 0.5 * H - s0 * rad)
    c = np.asarray(center, np.float64)
    R = np.stack([c[0] + rho * np.cos(th), c[1] + rho * np.sin(th), c[2] + z], 1)
    perm = morton_order(R, 2.0)
    R, rad = R[perm], rad[perm]
    mat = material_c(r_lo)
    p = params_from_material(mat, refmode=0, contact_capacity=32)
    sysn = NativeSystem(ctx, p, R, rad)
    sysn.set_bonds(Bonds(np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                         np.zeros(0, np.float32), np.zeros(0, np.float32),
                         np.zeros(0, np.float32)))
    pp = PackParams(cx=c[0], cy=c[1], radius=0.5 * D, z0=c[2] - 0.5 * H, z1=c[2] + 0.5 * H,
                    s0=s0, growth=growth, beta=beta, momentum=momentum, tol=tol,
                    stage_iters=stage_iters,
                    final_iters=final_iters)
    res = sysn.pack(pp).as_dict()
    Rp = sysn.state()["R"].reshape(3, N).T.copy()
    radius = np.ascontiguousarray(rad, np.float64)
    sysn.close()
    perm = morton_order(Rp, 2.0)
    phi_out = float((4.0 / 3.0 * np.pi * radius ** 3).sum() / V)
    return Packing(Rp[perm], radius[perm], phi_out, res)


def packed_uniaxial_cylinder(ctx, D: float, H: float, phi: float = 0.6,
                             seed: int = specimen.BASE_SEED, gap_frac: float = 0.05,
                             strength=(20e6, 40e6), rebuild_every: int = 20,
                             v_load: float = -10.0 * 0.02 / VELOCITY, fp64: bool = False,
                             packing: Packing | None = None) -> NativeCase:
    """Uniaxial specimen on a seeded random packing (packed_cylinder): bonds
    join pairs with surface gap < gap_frac * r_mean, bond radius min(ri, rj),
    strengths N(mu, 0.2 mu).  Platens start on the outermost particle centres
    along z."""
    pk = packing or packed_cylinder(ctx, D, H, phi, seed)
    assert np.allclose(0.5 * (pk.R.max(0) + pk.R.min(0))[:2], 0.0, atol=0.5 * D)
    R, rad = pk.R, pk.radius
    mat = material_c(0.4)
    a = 2.0 * 2.9 * 0.6
    box = SampleBox(D, D, H, space=(D + 16.0, D + 16.0, H + 12.0), cell_a=a, p_r=0.5)
    R = R + np.asarray(box.center)  # the packing is centred on the origin
    p = params_from_material(mat, refmode=0, rebuild_every=rebuild_every, fp64=fp64)
    p.farcut = 1.45
    sysn = NativeSystem(ctx, p, R, rad.astype(np.float32))
    pairs = sysn.find_pairs(gap_frac * float(rad.mean()))
    g = specimen.rng(seed + 1)
    sc, tc = strengths(g, len(pairs), strength[0], strength[1])
    br = np.minimum(rad[pairs[:, 0]], rad[pairs[:, 1]]).astype(np.float32)
    sysn.set_bonds(Bonds(pairs[:, 0].copy(), pairs[:, 1].copy(), br, sc, tc))
    cmp = Compression("uniaxial", box.L, box.center, 0.5, mat.dt, v_load=v_load,
                      min_stress=15e6 / STRESS, max_strain=0.1)
    L = loading_from_compression(cmp, box.center[2] + box.spacesized2[2], box)
    z = R[:, 2]
    L.top, L.bottom = float(z.max()), float(z.min())
    L.top_r, L.bottom_r = L.top - 0.5, L.bottom + 0.5
    L.size = L.size0 = L.top - L.bottom
    sysn.set_loading(L)
    return NativeCase(sysn, mat, box, cmp, np.arange(len(R)))


def packed_brazil_disc(ctx, D: float = 20.0, T: float = 5.0, phi: float = 0.6,
                       seed: int = specimen.BASE_SEED, gap_frac: float = 0.05,
                       strength=(20e6, 40e6), v_load: float = -0.5 / VELOCITY,
                       fp64: bool = False, rebuild_every: int = 20) -> NativeCase:
    """Brazilian disc on a seeded random packing (packed_cylinder, axis z):
    bonds join pairs with surface gap < gap_frac * r_mean, bond radius
    min(ri, rj), strengths N(mu, 0.2 mu) truncated > 0; diametral platens along
    y close at v_load from the outermost particle centres."""
    pk = packed_cylinder(ctx, D, T, phi, seed)
    mat = material_c(0.4)
    c = np.array([0.5 * D + 10.0, 0.5 * D + 10.0, 0.5 * T + 10.0])
    R, rad = pk.R + c, pk.radius
    p = params_from_material(mat, refmode=0, rebuild_every=rebuild_every, fp64=fp64)
    p.farcut = 1.45
    sysn = NativeSystem(ctx, p, R, rad.astype(np.float32))
    pairs = sysn.find_pairs(gap_frac * float(rad.mean()))
    sc, tc = strengths(specimen.rng(seed + 1), len(pairs), strength[0], strength[1])
    br = np.minimum(rad[pairs[:, 0]], rad[pairs[:, 1]]).astype(np.float32)
    sysn.set_bonds(Bonds(pairs[:, 0].copy(), pairs[:, 1].copy(), br, sc, tc))
    cmp = Compression("brazil", (D - 1.0, D - 1.0, T - 1.0), tuple(c), 0.5, mat.dt,
                      v_load=v_load, min_stress=1e6 / STRESS, max_strain=0.1)
    L = loading_from_compression(cmp, 1e30)
    y = R[:, 1]
    L.top, L.bottom = float(y.max()), float(y.min())
    L.top_r, L.bottom_r = L.top - 0.5, L.bottom + 0.5
    L.size = L.size0 = L.top - L.bottom
    sysn.set_loading(L)
    return NativeCase(sysn, mat, None, cmp, np.arange(len(R)))
