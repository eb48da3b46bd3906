"""Ready-made reference-layout test cases (monodisperse, the reference's own
material model), shared by the drivers, tests and the benchmark."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import specimen
from .params import Compression, Material, SampleBox, STRESS, VELOCITY


@dataclass
class RefCase:
    R: np.ndarray          # SoA float32 [3N]
    mat: Material
    box: SampleBox
    cmp: Compression
    K: int

    @property
    def N(self) -> int:
        return self.R.size // 3

    def grid(self):
        from .reflayout import Grid

        g = self.box.grid
        return Grid(float(np.float32(self.box.cell_a)), g.inv_a, g.Nx, g.Ny, g.Nz)


def reference_uniaxial(D: float = 12.5, H: float = 25.0, seed: int = specimen.BASE_SEED,
                       K: int = 32, mat: Material | None = None) -> RefCase:
    """A D x H cylinder loaded along z as in CalculateStepCG
    (mp_determineconjgradient.cu:96-101): VLoad = -0.2 m/s, MinStress 15 MPa,
    MaxStrain 0.1.  Particles: radius p_r on a jittered FCC lattice of spacing
    2 p_r, so every lattice neighbour is within the bond distance b_d."""
    mat = mat or Material.reference_granite()
    a = 2.0 * mat.a_farcut
    box = SampleBox(D, D, H, space=(D + 16.0, D + 16.0, H + 12.0), cell_a=a, p_r=mat.p_r)
    sp = specimen.cylinder(D, H, center=box.center, d_nn=2.0 * mat.p_r, jitter=0.02,
                           seed=seed, monodisperse=mat.p_r)
    cmp = Compression("uniaxial", box.L, box.center, mat.p_r, mat.dt,
                      v_load=-10.0 * 0.02 / VELOCITY, min_stress=15e6 / STRESS, max_strain=0.1)
    return RefCase(sp.soa_f32(), mat, box, cmp, K)


def reference_brazil(D: float = 25.0, T: float = 10.0, seed: int = specimen.BASE_SEED,
                     K: int = 32, mat: Material | None = None) -> RefCase:
    """A Brazilian disc (axis z, platens along y) as in CalculateStepCG
    (mp_determineconjgradient.cu:116-121): VLoad = -2 m/s, MinStress 2.1 MPa."""
    mat = mat or Material.reference_granite()
    a = 2.0 * mat.a_farcut
    box = SampleBox(D, D, T, space=(D + 16.0, D + 16.0, T + 12.0), cell_a=a, p_r=mat.p_r)
    sp = specimen.disc(D, T, center=box.center, d_nn=2.0 * mat.p_r, jitter=0.02, seed=seed,
                       monodisperse=mat.p_r)
    cmp = Compression("brazil", box.L, box.center, mat.p_r, mat.dt,
                      v_load=-100.0 * 0.02 / VELOCITY, min_stress=2.1e6 / STRESS, max_strain=0.1)
    return RefCase(sp.soa_f32(), mat, box, cmp, K)
