"""Calibration outer loop: a nonlinear conjugate-gradient fit of the BPM
parameters (E, tensile and shear bond strength, bond radius, bond gap) to
target test results (uniaxial peak strain and stress, Brazilian peak stress),
with finite-difference gradients, as DetermineConjGradient does
(mp_determineconjgradient.cu:139-196, mp_task.h:114-294).

The reference runs the 6 parameter sets of a step (the current one and one
variation per parameter) through 6 uniaxial and 6 Brazilian simulations one
after the other (CalculateStepCG, mp_determineconjgradient.cu:84-137).  Here
the sets that share E run as one batch per test type on one system
(rfd_system_set_specimens); the set with the varied E has its own material and
runs as its own system.  Every specimen of a step uses the smallest time step
among the sets (the reference recomputes it per set).

Two places of the reference's update read as mistakes and are kept only with
`reference_quirks=True`:
* the finite-difference derivative of result j by parameter i divides by the
  variation of parameter j (mp_task.h:222-230, index `NParam * (j + 1) + j`);
  the default divides by the variation of parameter i;
* SetMaterialContractions (mp_task.h:93-113) sets the tensile strength to
  (1 - 0.4) E whenever (E - Ec) / E > 0.4, which holds for any rock; the
  default keeps the lower bounds only.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import native_cases, specimen
from .native import Bonds, NativeSystem, loading_from_compression, params_from_material
from .params import LENGTH, STRESS, VELOCITY, Compression, Material

NPARAM = 5   # E, Ec (tensile), Gc (shear), bond radius, bond gap
NRESULT = 3  # uniaxial peak strain, uniaxial peak stress, Brazilian peak stress

# DetermineConjGradient (mp_determineconjgradient.cu:143-152): granite
GRANITE_GOALS = np.array([0.00480002, 104.0497482e6 / STRESS, 10.5e6 / STRESS])
GRANITE_START = np.array([1.95255e11 / STRESS, 5.1447e8 / STRESS, 3.7131e8 / STRESS,
                          0.5 * 0.000330015 / LENGTH, 5.93169e-05 / LENGTH])


def contractions(p: np.ndarray, reference_quirks: bool = False) -> np.ndarray:
    """Bounds on a parameter set (SetMaterialContractions, mp_task.h:93-113)."""
    p = np.array(p, np.float64)
    p[0] = max(p[0], 1e5 / STRESS)
    p[1] = max(p[1], 1e5 / STRESS)
    if reference_quirks:
        d = 0.4
        a = (p[0] - p[1]) / p[0]
        if a < d:
            p[1] = (1.0 + d) * p[0]
        elif a > d:
            p[1] = (1.0 - d) * p[0]
    p[2] = max(p[2], 1e8 / STRESS)
    p[3] = max(p[3], 1e-6 / LENGTH)
    p[4] = max(p[4], 1e-5 / LENGTH)
    return p


@dataclass
class CGState:
    """mp_task_data's arrays for one fit (mp_task.h:114-128, 200-294)."""

    params: np.ndarray
    goals: np.ndarray
    delta: float = 0.05                      # MaterialDeltaFactor
    reference_quirks: bool = False
    grad: np.ndarray = field(default_factory=lambda: np.zeros(NPARAM))   # dFunctional
    direction: np.ndarray = field(default_factory=lambda: np.zeros(NPARAM))  # ddFunctional
    history: list = field(default_factory=list)

    @property
    def weights(self) -> np.ndarray:
        return 1.0 / (self.goals * self.goals)

    def variations(self) -> np.ndarray:
        """Parameter set k+1 varies parameter k by delta times its value
        (CalculateStartParameters, mp_task.h:114-128, all variation flags on)."""
        return self.delta * self.params

    def parameter_sets(self) -> np.ndarray:
        """(NPARAM + 1) x NPARAM: the current set, then one variation each."""
        v = self.variations()
        sets = np.repeat(self.params[None, :], NPARAM + 1, axis=0)
        for k in range(NPARAM):
            sets[k + 1, k] += v[k]
        return sets

    def functional(self, R0: np.ndarray) -> float:
        D = self.goals - R0
        return float((self.weights * D * D).sum())

    def update(self, R: np.ndarray) -> np.ndarray:
        """One CG step from the results R[(NPARAM + 1) x NRESULT] of the
        parameter sets (CalculateNewParameters2, mp_task.h:200-287):
        finite-difference Jacobian, gradient of the weighted squared misfit,
        Fletcher-Reeves direction, step length from the linearised misfit
        along it, new parameters with the contractions."""
        R = np.asarray(R, np.float64)
        v = self.variations()
        J = np.zeros((NPARAM, NRESULT))  # J[i, j] = dR_j / dp_i
        for i in range(NPARAM):
            for j in range(NRESULT):
                den = v[j] if self.reference_quirks else v[i]
                J[i, j] = (R[i + 1, j] - R[0, j]) / den if den > 1e-12 else 0.0
        w = self.weights
        D = self.goals - R[0]
        g_old = self.grad.copy()
        g = -2.0 * (J * (w * D)[None, :]).sum(1)
        n0 = float((g_old * g_old).sum())
        beta = float((g * g).sum()) / n0 if n0 > 1e-12 else 0.0
        d = g + beta * self.direction
        a1 = float((w[None, :] * D[None, :] * J * d[:, None]).sum())
        Jd = (J * d[:, None]).sum(0)  # directional derivative of each result
        a0 = float((w * Jd * Jd).sum())
        t = -a1 / a0 if a0 != 0.0 else 0.0
        self.history.append({"params": self.params.copy(), "results": R[0].copy(),
                             "functional": self.functional(R[0]), "t": t, "beta": beta})
        self.grad, self.direction = g, d
        self.params = contractions(self.params - t * d, self.reference_quirks)
        return self.params


# ------------------------------------------------------------ simulations
@dataclass
class TestGeometry:
    """The two specimens of the fit (mp_determineconjgradient.cu:157-167):
    a uniaxial cylinder and a Brazilian disc, packed at solid fraction 0.61
    (PN = (1 - 0.39) V / p_V) with monodisperse spheres of radius p_r."""

    uni_D: float = 25.0
    uni_H: float = 50.0
    bra_D: float = 50.0
    bra_T: float = 20.0
    phi: float = 0.61
    seed: int = specimen.BASE_SEED
    uni_v: float = -10.0 * 0.02 / VELOCITY     # VLoad of the uniaxial runs
    bra_v: float = -100.0 * 0.02 / VELOCITY    # VLoad of the Brazilian runs
    uni_min_stress: float = 15e6 / STRESS
    bra_min_stress: float = 2.1e6 / STRESS


def _material(p: np.ndarray) -> Material:
    base = Material.reference_granite()
    return Material(E=float(p[0]), nu=base.nu, rho=base.rho, p_r=base.p_r, b_r=float(p[3]),
                    b_d=2.0 * base.p_r + float(p[4]), sc=float(p[1]), tc=float(p[2]))


class Evaluator:
    """Runs parameter sets through the uniaxial and Brazilian tests, batched."""

    def __init__(self, ctx, geom: TestGeometry | None = None, max_steps: int = 200000,
                 chunk: int = 2000):
        self.ctx = ctx
        self.geom = geom or TestGeometry()
        self.max_steps, self.chunk = max_steps, chunk
        g = self.geom
        r = Material.reference_granite().p_r
        self.uni = native_cases.packed_cylinder(ctx, g.uni_D, g.uni_H, g.phi, g.seed, r, r)
        self.bra = native_cases.packed_cylinder(ctx, g.bra_D, g.bra_T, g.phi, g.seed + 1, r, r)

    def _specimen(self, pk, mode: str, p: np.ndarray, offset: np.ndarray, tmp):
        """Positions, radii, bonds (local ids) and loading of one specimen at
        `offset` (the packing is centred on the origin)."""
        g = self.geom
        R = pk.R + offset
        rad = pk.radius
        pairs = tmp[id(pk)]
        d = np.linalg.norm(R[pairs[:, 1]] - R[pairs[:, 0]], axis=1)
        keep = d - rad[pairs[:, 0]] - rad[pairs[:, 1]] < p[4]
        pr = pairs[keep]
        n = len(pr)
        bonds = (pr, np.full(n, p[3], np.float32), np.full(n, p[1], np.float32),
                 np.full(n, p[2], np.float32))
        D, H = (g.uni_D, g.uni_H) if mode == "uniaxial" else (g.bra_D, g.bra_T)
        size = (D - 1.0, D - 1.0, H - 1.0)
        v = g.uni_v if mode == "uniaxial" else g.bra_v
        ms = g.uni_min_stress if mode == "uniaxial" else g.bra_min_stress
        cmp = Compression(mode, size, tuple(offset), 0.5, 1.0, v_load=v, min_stress=ms,
                          max_strain=0.1)
        L = loading_from_compression(cmp, 1e30 if mode == "brazil" else offset[2] + H)
        ax = 2 if mode == "uniaxial" else 1
        L.top, L.bottom = float(R[:, ax].max()), float(R[:, ax].min())
        L.top_r, L.bottom_r = L.top - 0.5, L.bottom + 0.5
        L.size = L.size0 = L.top - L.bottom
        return R, rad, bonds, L

    def _candidates(self, pk):
        """Pairs within the largest bond gap any set may use (host, brute force
        over a cell grid)."""
        return specimen.bonds_by_gap(pk.R, pk.radius, 0.5)

    def run(self, sets: np.ndarray, mode: str, dt: float) -> list[tuple[float, float]]:
        """(peak strain, peak stress) of each parameter set in one batch of
        `mode` specimens sharing E = sets[0, 0] (their other parameters may
        differ), time step dt."""
        assert np.all(sets[:, 0] == sets[0, 0])
        pk = self.uni if mode == "uniaxial" else self.bra
        tmp = {id(pk): self._candidates(pk)}
        ext = pk.R.max(0) - pk.R.min(0) + 4.0
        Rs, rads, bis, bjs, brs, scs, tcs, Ls, first = [], [], [], [], [], [], [], [], [0]
        for k, p in enumerate(sets):
            off = np.array([0.0, 0.0, k * ext[2]]) + np.array([100.0, 100.0, 100.0])
            R, rad, (pr, br, sc, tc), L = self._specimen(pk, mode, p, off, tmp)
            base = first[-1]
            Rs.append(R)
            rads.append(rad)
            bis.append(pr[:, 0] + base)
            bjs.append(pr[:, 1] + base)
            brs.append(br)
            scs.append(sc)
            tcs.append(tc)
            Ls.append(L)
            first.append(base + len(R))
        mat = _material(sets[0])
        prm = params_from_material(mat, refmode=0)
        prm.dt = dt
        prm.farcut = 1.45
        s = NativeSystem(self.ctx, prm, np.concatenate(Rs), np.concatenate(rads))
        s.set_bonds(Bonds(np.concatenate(bis).astype(np.uint32),
                          np.concatenate(bjs).astype(np.uint32), np.concatenate(brs),
                          np.concatenate(scs), np.concatenate(tcs)))
        for L in Ls:
            L.v = 0.0
        s.set_specimens(np.array(first), Ls, hist_cap=1)
        done = 0
        while done < self.max_steps:
            s.step(self.chunk)
            done += self.chunk
            out = s.specimen_loadings()
            if all(L.fracture for L in out):
                break
        res = [(float(L.fracture_strain), float(L.fracture_stress)) for L in out]
        s.close()
        return res

    def evaluate(self, sets: np.ndarray) -> np.ndarray:
        """Results [(NPARAM + 1) x NRESULT] of the parameter sets: one batch
        per test type for the sets sharing the first set's E, one more system
        per test type for each other E."""
        dt = min(_material(p).dt for p in sets)
        R = np.zeros((len(sets), NRESULT))
        for E in np.unique(sets[:, 0]):
            idx = np.flatnonzero(sets[:, 0] == E)
            uni = self.run(sets[idx], "uniaxial", dt)
            bra = self.run(sets[idx], "brazil", dt)
            for k, i in enumerate(idx):
                R[i, 0], R[i, 1] = uni[k]
                R[i, 2] = bra[k][1]
        return R


def calibrate(ctx, steps: int = 1, geom: TestGeometry | None = None,
              reference_quirks: bool = False, evaluator: Evaluator | None = None,
              log=None) -> CGState:
    """`steps` CG steps from the granite start towards the granite goals."""
    ev = evaluator or Evaluator(ctx, geom)
    st = CGState(params=GRANITE_START.copy(), goals=GRANITE_GOALS.copy(),
                 reference_quirks=reference_quirks)
    for k in range(steps):
        R = ev.evaluate(st.parameter_sets())
        st.update(R)
        if log:
            log(f"CG step {k}: functional {st.history[-1]['functional']:.4g}, "
                f"results {R[0].tolist()}")
    return st
