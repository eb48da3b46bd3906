"""Native-tier specimens (polydisperse, per-bond strengths) for tests and the
benchmark."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import specimen
from .native import (Bonds, NativeSystem, loading_from_compression, morton_order,
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
                      flat_ends: bool = False, fp64: bool = False) -> NativeCase:
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
    p = params_from_material(mat, refmode=refmode, rebuild_every=rebuild_every, fp64=fp64)
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
