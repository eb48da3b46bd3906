"""Host-side parameter sets of the BPM step, mirroring md_data_types.h.

Units are the reference's (md_phys_constants.h:5-13): length mm, mass mg,
time us; so force is 1e3 N, stress 1e9 Pa, velocity 1e3 m/s and density
1e3 kg/m^3 (2650 kg/m^3 -> 2.65).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

LENGTH = 1e-3
MASS = 1e-6
TIME = 1e-6
DENSITY = MASS / LENGTH**3          # 1e3
FORCE = MASS * LENGTH / TIME**2     # 1e3 N
STRESS = FORCE / LENGTH**2          # 1e9 Pa
VELOCITY = LENGTH / TIME            # 1e3 m/s


@dataclass
class Material:
    """potential_data (md_data_types.h:146-186); CalculateParameters :153-185."""

    E: float                 # Young's modulus (stress units)
    nu: float                # Poisson ratio
    rho: float               # density
    p_r: float               # particle radius
    b_r: float               # bond radius
    b_d: float               # bond creation distance
    sc: float                # bond tensile strength  (m_Ec)
    tc: float                # bond shear strength    (m_Gc)
    p_e: float = 0.6         # restitution
    p_mu: float = 0.45       # sliding friction
    p_mur: float = 0.05      # rolling friction
    a_farcut: float = 0.0    # far-link cut (default 2.9 p_r)
    # derived
    G: float = field(init=False)
    b_S: float = field(init=False)
    p_V: float = field(init=False)
    p_m: float = field(init=False)
    p_I: float = field(init=False)
    b_dd: float = field(init=False)
    hm_E: float = field(init=False)
    hm_G: float = field(init=False)
    p_A: float = field(init=False)
    aa_farcut: float = field(init=False)
    T_rayleigh: float = field(init=False)
    T_sim: float = field(init=False)
    dt: float = field(init=False)
    dt_d_m: float = field(init=False)
    dt_d_I: float = field(init=False)

    def __post_init__(self):
        if self.a_farcut <= 0:
            self.a_farcut = 2.9 * self.p_r
        self.G = 0.5 * self.E / (1.0 + self.nu)
        self.b_S = math.pi * self.b_r * self.b_r
        self.p_V = 4.0 * (1.0 / 3.0) * math.pi * self.p_r**3
        self.p_m = self.rho * self.p_V
        self.b_dd = self.b_d * self.b_d
        self.p_I = 0.4 * self.p_m * self.p_r * self.p_r
        self.hm_E = 0.5 * self.E / (1.0 - self.nu * self.nu)
        self.hm_G = 0.5 * self.G / (2.0 - self.nu)
        le = math.log(self.p_e)
        self.p_A = le / math.sqrt(math.pi * math.pi + le * le)
        self.aa_farcut = self.a_farcut * self.a_farcut
        # time step rule, md_data_types.h:166-176
        self.T_rayleigh = (math.pi * self.p_r * math.sqrt(2.0 * self.rho * (1.0 + self.nu))
                           / (math.sqrt(self.E) * (0.163 * self.nu + 0.8766)))
        k = max(self.E * self.b_S / (2.0 * self.p_r), self.G * self.b_S / (2.0 * self.p_r))
        self.T_sim = 2.0 * math.sqrt(self.p_m / k)
        self.dt = 0.1 * min(self.T_rayleigh, self.T_sim)
        self.dt_d_m = self.dt / self.p_m
        self.dt_d_I = self.dt / self.p_I

    @classmethod
    def reference_granite(cls) -> "Material":
        """SetInitialParameters (mp_determineconjgradient.cu:15-33), start point
        of the calibration (mp_determineconjgradient.cu:150-154)."""
        p_r = 0.5 * 0.001 / LENGTH
        return cls(E=1.95255e11 / STRESS, nu=0.247, rho=2610 / DENSITY, p_r=p_r,
                   b_r=0.5 * 0.000330015 / LENGTH, b_d=2.0 * p_r + 5.93169e-05 / LENGTH,
                   sc=5.1447e8 / STRESS, tc=3.7131e8 / STRESS)


@dataclass
class CellGrid:
    """cell_data::CalculateParameters (md_data_types.h:42-55)."""

    a: float
    Bx: float
    By: float
    Bz: float

    def __post_init__(self):
        self.inv_a = float(_f32(1.0 / _f32(self.a)))
        ia = self.inv_a
        self.Nx = int(math.floor(_f32(self.Bx * ia))) + 1
        self.Ny = int(math.floor(_f32(self.By * ia))) + 1
        self.Nz = int(math.floor(_f32(self.Bz * ia))) + 3
        self.C = self.Nx * self.Ny * self.Nz


def _f32(x: float) -> float:
    import numpy as np

    return float(np.float32(x))


@dataclass
class SampleBox:
    """sample_data (md_data_types.h:78-122) for a specimen of outer size L
    (diameter x diameter x height, mm) inside a padded space box."""

    Lx: float
    Ly: float
    Lz: float
    space: tuple[float, float, float]
    cell_a: float
    p_r: float

    def __post_init__(self):
        a = self.cell_a
        self.A = (1.01 * a, 1.01 * a, 1.01 * a)
        self.B = tuple(self.A[d] + self.space[d] for d in range(3))
        self.center = tuple(0.5 * (self.A[d] + self.B[d]) for d in range(3))
        self.L = (self.Lx - 2 * self.p_r, self.Ly - 2 * self.p_r, self.Lz - 2 * self.p_r)
        self.spacesized2 = tuple(0.5 * s for s in self.space)
        grid = CellGrid(a, *self.B)
        # CellDistributionInit (md_cell_distribution.cu:21-23)
        self.hidden_point = (self.A[0] + 0.5 * self.space[0], self.A[1] + 0.5 * self.space[1],
                             (grid.Nz - 0.5) * a)
        self.grid = grid


@dataclass
class Compression:
    """compression_data (md_data_types.h:188-294): platen kinematics and the
    fracture detector of SumForcesULoading (md_results.cu:34-74)."""

    mode: str                # "uniaxial" (load along z) or "brazil" (along y)
    size: tuple[float, float, float]
    center: tuple[float, float, float]
    p_r: float
    dt: float
    v_load: float
    min_stress: float
    max_strain: float

    def __post_init__(self):
        sx, sy, sz = self.size
        self.size_d = list(self.size)
        self.center_d = list(self.center)
        ax = 2 if self.mode == "uniaxial" else 1
        self.axis = ax
        self.size0 = self.size_d[ax]
        self.bottom = self.center_d[ax] - 0.5 * self.size_d[ax]
        self.bottom_r = self.bottom + self.p_r
        self.top = self.center_d[ax] + 0.5 * self.size_d[ax]
        self.top_r = self.top - self.p_r
        hx = 0.5 * sx
        if self.mode == "uniaxial":
            self.R = 1.5 * hx
            self.area = math.pi * (hx + self.p_r) ** 2
            self.zero_stress = 0.33 * self.min_stress
        else:
            hz = 0.5 * sz
            self.R = hx if hx > hz else 1.05 * hz
            self.area = 0.5 * 2 * math.pi * (hx + self.p_r) * (sz + 2.0 * self.p_r)
            self.zero_stress = 0.5 * self.min_stress
        self.RR = self.R * self.R
        self.inv_area = 1.0 / self.area
        self.V = 0.0
        self.stress = 0.0
        self.strain = 0.0
        self.loading = False
        self.fracture = False
        self.fracture_stress = 0.0
        self.fracture_strain = 0.0
        self.history: list[tuple[float, float, float]] = []

    def sum_forces(self, sums: tuple[float, float]) -> None:
        """SumForcesULoading (md_results.cu:34-74) after the platen sums."""
        s0, s1 = sums
        sz0 = self.size0
        self.history.append(((self.size_d[self.axis] - sz0) / sz0, s0 * self.inv_area,
                             s1 * self.inv_area))
        self.stress = 0.5 * (abs(s0) + abs(s1)) * self.inv_area
        if not self.loading and self.stress > self.min_stress:
            self.loading = True
        if self.fracture_stress < self.stress:
            self.fracture_stress = self.stress
            self.fracture_strain = (sz0 - self.size_d[self.axis]) / sz0
        if self.loading and self.stress < self.zero_stress:
            self.fracture = True

    def compress(self) -> None:
        """CompressU / CompressB (md_data_types.h:223-244, 272-293)."""
        self.V = 10.0 * self.v_load if self.stress < 1e-9 else self.v_load
        ax = self.axis
        self.size_d[ax] += self.V * self.dt
        self.top += self.V * self.dt
        self.top_r = self.top - 2.0 * self.p_r
        self.center_d[ax] += self.V * self.dt
        self.strain = (self.size_d[ax] - self.size0) / self.size0
