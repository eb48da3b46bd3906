"""CPU restatement, in float64, of one quasi-static load step of the native
tier (rfd_system_fire_load_step): platen move, start pass, standard FIRE over
translations and rotations, bonds judged at equilibrium, relax-and-break
rounds.

TEST INFRASTRUCTURE ONLY: the GPU tests compare the FP64 native path against
it; the product never imports it.

The reference has no quasi-static load step (its FIRE, md_firerelaxation.cu:
18-84, 263-304, relaxes unbonded packings); this step is the build's own, and
what it restates is the build's definition of it:

* the top platen moves by dstrain * size0 along the load axis; particles held
  by a platen follow it, particles past a platen are put on it (the platen
  clamp of md_uniaxialcompression3_simple.cu:19-76 / md_braziltest_simple2.cu
  :19-92 with the platens at rest and no friction);
* one force pass with the time step dt0 and the rate (R - R0) / dt0 carries
  the move into the incremental springs (bond shear, twist and bending,
  contact springs); velocities are then zero;
* FIRE (Bitzek et al. 2006, half-step back-off on uphill steps): per iteration
  a force pass with the time step and the rate of the last move, bonds not
  judged, no viscous damping and no rolling resistance; the platen clamp; the
  power P = F.V + M.W and norms; the control; mixing, kick and move;
* converged when max |F_i| and max |M_i| / r_i are below ftol;
* a pass with no increments (dt = 0) then judges the bonds
  (md_calculateforcesdem3.cu:20-36); while any break, FIRE relaxes again from
  rest (at most 200 rounds).
The bond and contact laws are the oracle's orc_native_bonds_f64 and
orc_native_contacts_f64, through OracleNativeStep.forces.
"""
from __future__ import annotations

import math

import numpy as np

MAX_BREAK_ROUNDS = 200


class FireState:
    """The FIRE control state of a relaxation (one per relaxation round)."""

    def __init__(self, fp):
        self.dt = self.dt_last = fp.dt0
        self.alpha = self.alpha0 = fp.alpha0
        self.dt_max, self.dt_min = fp.dt_max, 1e-3 * fp.dt0
        self.f_inc, self.f_dec, self.f_alpha = fp.f_inc, fp.f_dec, fp.f_alpha
        self.ftol, self.n_delay, self.max_iter = fp.ftol, fp.n_delay, fp.max_iter
        self.n_pos = self.iter = 0
        self.done = self.converged = self.backoff = False
        self.first = True
        self.cV = self.cW = 0.0
        self.fmax = self.mmax = self.fmax0 = 0.0
        self.sum_bottom = self.sum_top = 0.0
        self.dt_move = fp.dt0  # time step of the last move, taken by the next force pass

    def copy(self):
        c = FireState.__new__(FireState)
        c.__dict__.update(self.__dict__)
        return c


class OracleFireLoadStep:
    """Quasi-static load steps on an OracleNativeStep o (positions, bonds,
    contact springs, loading control) with FIRE parameters fp (the fields of
    rfd_fire_params)."""

    def __init__(self, o, fp):
        self.o = o
        self.fp = fp
        N = o.N
        self.Vf = np.zeros((N, 3))   # FIRE's own velocity
        self.Th = np.zeros((N, 3))   # accumulated rotation of the step

    # ------------------------------------------------------------ pieces
    def _platen_rest(self, R, V, F, friction):
        L = dict(self.o.L)
        L["v"] = 0.0
        return self.o._platen(R, V, F, friction=friction, L=L)

    def _measure(self, f):
        o = self.o
        sb, st = self._platen_rest(o.R, self.Vf, o.F, friction=False)
        F, V, M, W = o.F, self.Vf, o.M, o.W
        P = float((F * V).sum() + (M * W).sum())
        f2 = (F * F).sum(1)
        m2 = (M * M).sum(1)
        v = [P, float(f2.sum()), float((V * V).sum()), float(m2.sum()), float((W * W).sum()),
             float(f2.max()), float((m2 / (o.rad * o.rad)).max()), sb, st]
        return v

    def _control(self, f, v):
        f.fmax = math.sqrt(v[5])
        f.mmax = math.sqrt(v[6])
        f.sum_bottom, f.sum_top = v[7], v[8]
        if f.first:
            f.fmax0 = f.fmax
            f.first = False
        f.converged = f.fmax < f.ftol and f.mmax < f.ftol
        if f.converged or f.iter >= f.max_iter:
            f.done = True
            self.o.L["sum_bottom"], self.o.L["sum_top"] = v[7], v[8]
            return
        f.iter += 1
        f.dt_last = f.dt
        if v[2] == 0.0 and v[4] == 0.0:
            f.backoff = False
            f.cV = f.cW = 0.0
        elif v[0] > 0.0:
            f.backoff = False
            f.cV = f.alpha * math.sqrt(v[2] / v[1]) if v[1] > 0 else 0.0
            f.cW = f.alpha * math.sqrt(v[4] / v[3]) if v[3] > 0 else 0.0
            f.n_pos += 1
            if f.n_pos > f.n_delay:
                f.dt = min(f.dt * f.f_inc, f.dt_max)
                f.alpha *= f.f_alpha
        else:
            f.backoff = True
            f.n_pos = 0
            f.dt = max(f.dt * f.f_dec, f.dt_min)
            f.alpha = f.alpha0
        f.dt_move = f.dt

    def _update(self, f):
        o = self.o
        R0 = o.R.copy()
        R = o.R
        if f.backoff:
            R -= self.Vf * (0.5 * f.dt_last)
            self.Vf[:] = 0.0
            o.W = np.zeros_like(o.W)
        else:
            a1 = 1.0 - f.alpha
            self.Vf = self.Vf * a1 + o.F * f.cV
            o.W = o.W * a1 + o.M * f.cW
        self.Vf = self.Vf + o.F * (f.dt / o.m)[:, None]
        o.W = o.W + o.M * (f.dt / o.I)[:, None]
        R += self.Vf * f.dt
        o.V = (R - R0) * (1.0 / f.dt)
        self.Th = self.Th + o.W * f.dt

    def _relax(self, f, forces_current):
        o = self.o
        while True:
            if forces_current:
                forces_current = False
            else:
                o.forces(dt=f.dt_move, quasi=True, nobreak=True)
            v = self._measure(f)
            self._control(f, v)
            if f.done:
                return f
            self._update(f)

    # ------------------------------------------------------------- step
    def load_step(self, dstrain):
        """One load step; returns a dict of the rfd_fire_result fields."""
        o, fp = self.o, self.fp
        L = o.L
        dtop = 0.0
        if L["mode"]:
            d = dstrain * L["size0"]
            dtop = d
            L["size"] += d
            L["top"] += d
            L["top_r"] = L["top"] - L["band"]
            L["strain"] = (L["size"] - L["size0"]) / L["size0"]
            L["v"] = 0.0
        P0 = o.R.copy()
        f = FireState(fp)
        self._begin(P0, dtop, f.dt)
        o.forces(dt=f.dt_move, quasi=True, nobreak=True)
        o.V = np.zeros_like(o.V)
        o.W = np.zeros_like(o.W)
        f_start = f.copy()
        self._relax(f, forces_current=True)
        f_first = f
        iters, breaks, rounds = f.iter, 0, 0
        while f.converged:
            o.forces(dt=0.0, quasi=True, nobreak=False)
            nb = len(o.broke_last)
            if nb == 0 or rounds >= MAX_BREAK_ROUNDS:
                break
            breaks += nb
            rounds += 1
            f = f_start.copy()
            o.V = np.zeros_like(o.V)
            o.W = np.zeros_like(o.W)
            self.Vf[:] = 0.0
            self._relax(f, forces_current=True)
            iters += f.iter
        self.du = o.R - P0
        o.V = np.zeros_like(o.V)
        o.W = np.zeros_like(o.W)
        return {"iterations": iters, "converged": f.converged, "breaks": breaks,
                "break_rounds": rounds, "fmax0": f_first.fmax0, "fmax": f.fmax,
                "mmax": f.mmax, "sum_bottom": f.sum_bottom, "sum_top": f.sum_top,
                "strain": L.get("strain", 0.0)}

    def _begin(self, P0, dtop, dt0):
        o = self.o
        L = dict(o.L)
        L["v"] = 0.0
        R = P0.copy()
        mode = L["mode"]
        if mode:
            a = 2 if mode == 1 else 1
            for i in range(o.N):
                if mode == 1:
                    ra, rb, h0, zc = P0[i, 0] - L["cx"], P0[i, 1] - L["cy"], P0[i, 2], P0[i, 2]
                else:
                    ra, rb, h0 = P0[i, 0] - L["cx"], P0[i, 2] - L["cz"], P0[i, 1]
                    zc = P0[i, 2] - L["cz"]
                if ra * ra + rb * rb < L["RR"] and zc < L["zcut"]:
                    if h0 - (L["top"] - dtop) > -1e-9:
                        R[i, a] = L["top"]
                    elif h0 - L["bottom"] < 1e-9:
                        R[i, a] = L["bottom"]
            Vd = np.zeros_like(R)
            F = np.zeros_like(R)
            o._platen(R, Vd, F, friction=True, L=L)
        o.R = R
        o.V = (R - P0) * (1.0 / dt0)
        o.W = np.zeros_like(o.W)
        self.Th = np.zeros_like(self.Th)
        self.Vf = np.zeros_like(self.Vf)
