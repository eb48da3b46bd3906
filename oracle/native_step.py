"""CPU restatement, in float64, of the native tier's explicit step.

TEST INFRASTRUCTURE ONLY: the GPU tests compare the native system against it;
the product never imports it.

One step is what rfd_system_step does for one specimen: a force evaluation
(bond law with the break check, the contact law of broken bonds and of
never-bonded pairs), the particle pass (platens, symplectic Euler, cut box) and
the loading control.  It is built from the oracle's two FP64 laws:

* orc_native_bonds_f64 for every intact bond (md_calculateforcesdem3.cu:20-197
  per undirected bond, previous direction at R - V dt);
* orc_native_contacts_f64 for every directed pair in contact
  (md_calculateforcesdem3.cu:199-336 with per-particle radii): a bond broken
  before this evaluation, one breaking in it (from a fresh spring, in the same
  evaluation), and never-bonded pairs within r_i + r_j (found by brute force
  over a k-d tree, springs kept per pair while in contact and cleared when
  apart).  A broken bond's contact reaches min(r_i + r_j, L0) in the physical
  mode and r_i + r_j in the reference mode; the two directions of a pair carry
  opposite springs.

The particle pass restates d_braziltest_simple2 (md_braziltest_simple2.cu:19-92)
and d_UniaxialCompression3_simple (md_uniaxialcompression3_simple.cu:19-76) on
FP64 positions, d_CalculateIncrementsDEM (md_calculateincrementsdem.cu:19-49)
with per-particle m = rho 4/3 pi r^3 and I = 2/5 m r^2, and the cut box of
md_parallelepipedrestriction.cu:73-158; the control restates
SumForcesULoading (md_results.cu:34-74) with CompressU/B
(md_data_types.h:223-244, 272-293).
"""
from __future__ import annotations

import math

import numpy as np

from oracle.oracle import lib, native_contacts, ptr

FRICTION = float(np.float32(0.45))  # the platens' 0.45f, widened to double


class OracleNativeStep:
    """State of one specimen: R (N, 3), radii, bonds (bi < bj) with their
    float32 rest-length reciprocal, radius and strengths, and a loading
    control given as a dict of rfd_loading's fields (mode 0: none)."""

    def __init__(self, R, rad, bi, bj, iL, br, sc, tc, *, E, nu, rho, mu, mur, e, dt,
                 refmode=0, loading=None, V=None, W=None):
        self.N = N = len(R)
        self.R = np.array(R, np.float64)
        self.V = np.zeros((N, 3)) if V is None else np.array(V, np.float64)
        self.W = np.zeros((N, 3)) if W is None else np.array(W, np.float64)
        self.rad = np.asarray(rad, np.float64)
        self.bi = np.ascontiguousarray(bi, np.uint32)
        self.bj = np.ascontiguousarray(bj, np.uint32)
        self.iL = np.ascontiguousarray(iL, np.float64)
        self.br = np.ascontiguousarray(br, np.float64)
        self.sc = np.ascontiguousarray(sc, np.float64)
        self.tc = np.ascontiguousarray(tc, np.float64)
        self.Nb = len(self.bi)
        self.state = np.zeros(8 * self.Nb)
        self.act = np.ones(self.Nb, np.uint8)
        self.Ob = np.zeros((self.Nb, 3))  # spring of a broken bond, seen from bi
        self.removed = np.zeros(self.Nb, bool)  # far links dropped by the device
        self.springs: dict[tuple[int, int], np.ndarray] = {}
        G = 0.5 * E / (1.0 + nu)
        le = math.log(e)
        self.E, self.G, self.dt, self.refmode = E, G, dt, refmode
        self.Estar = 0.5 * E / (1.0 - nu * nu)
        self.Gstar = 0.5 * G / (2.0 - nu)
        self.pA = le / math.sqrt(math.pi * math.pi + le * le)
        self.mu, self.mur = mu, mur
        self.rho43pi = rho * 4.0 / 3.0 * math.pi
        self.m = self.rho43pi * self.rad ** 3
        self.I = 0.4 * self.m * self.rad ** 2
        self.L = dict(loading) if loading else {"mode": 0}
        self.L.setdefault("steps", 0)
        self.L.setdefault("hist_len", 0)
        self.history: list[tuple[float, float, float]] = []
        self._bond_keys = np.sort(self.bi.astype(np.int64) * N + self.bj.astype(np.int64))
        self.F = np.zeros((N, 3))
        self.M = np.zeros((N, 3))
        self.broke_last = np.zeros(0, np.int64)
        # smallest |Coulomb margin| of the last evaluation's contacts: below
        # about 1e-12 the capped / damped branch is decided by rounding, and a
        # second implementation of the same law may take the other branch
        self.cap_tie = np.inf

    def load_bond_state(self, state12, flag):
        """Take the bond state of a device system (rfd_system_get_bond_state:
        per bond n, O, Mn, Mt as seen from bi, and the flag 1 bond, 0 broken,
        2 removed): O, the twist Mn . n and Mt of intact bonds, the contact
        spring O of broken ones."""
        st = np.asarray(state12, np.float64).reshape(-1, 12)
        flag = np.asarray(flag)
        n = st[:, 0:3]
        nn = (n * n).sum(1)
        mn = (st[:, 6:9] * n).sum(1) / np.where(nn > 0, nn, 1.0)
        S = self.state.reshape(-1, 8)
        S[:] = 0.0
        S[:, 0:3] = st[:, 3:6]
        S[:, 3] = mn
        S[:, 4:7] = st[:, 9:12]
        S[flag != 1] = 0.0
        self.act = (flag == 1).astype(np.uint8)
        self.removed = flag == 2
        self.Ob = np.where((flag == 0)[:, None], st[:, 3:6], 0.0)

    # --------------------------------------------------------------- forces
    def _soa(self, a):
        return np.ascontiguousarray(a.T).reshape(-1)

    def forces(self, dt=None, quasi=False, nobreak=False):
        """One force evaluation.  dt: the time step the increments take
        (default the system's); quasi: a quasi-static pass of a FIRE load step
        (no viscous damping, no rolling resistance); nobreak: bonds are not
        judged."""
        N = self.N
        dt = self.dt if dt is None else dt
        R, V, W = self._soa(self.R), self._soa(self.V), self._soa(self.W)
        Fb, Mb = np.zeros(3 * N), np.zeros(3 * N)
        before = self.act.copy()
        sc, tc = self.sc, self.tc
        if nobreak:
            sc = tc = np.full(self.Nb, np.inf)
        lib().orc_native_bonds_f64(ptr(R), ptr(V), ptr(W), N, ptr(self.bi), ptr(self.bj),
                                   ptr(self.iL), ptr(self.br), ptr(sc), ptr(tc),
                                   ptr(self.state), ptr(self.act), self.Nb, dt, self.E,
                                   self.G, self.refmode, ptr(Fb), ptr(Mb))
        newly = np.nonzero((before == 1) & (self.act == 0))[0]
        self.broke_last = newly
        self.Ob[newly] = 0.0  # a breaking bond meets the contact law with a fresh spring
        F = Fb.reshape(3, N).T.copy()
        M = Mb.reshape(3, N).T.copy()
        # broken bonds in contact
        brk = np.nonzero((self.act == 0) & ~self.removed)[0]
        bi, bj = self.bi[brk].astype(np.int64), self.bj[brk].astype(np.int64)
        sumr = self.rad[bi] + self.rad[bj]
        reach_b = sumr if self.refmode else np.minimum(sumr, 1.0 / self.iL[brk])
        Lb = np.linalg.norm(self.R[bj] - self.R[bi], axis=1)
        touch_b = (Lb < reach_b) & (Lb > 1e-5)
        self.Ob[brk[~touch_b]] = 0.0
        # never-bonded pairs in contact
        from scipy.spatial import cKDTree

        cand = cKDTree(self.R).query_pairs(2.0 * float(self.rad.max()), output_type="ndarray")
        keep = []
        if len(cand):
            ci, cj = cand[:, 0].astype(np.int64), cand[:, 1].astype(np.int64)
            d = self.R[cj] - self.R[ci]
            Lc = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
            hit = (Lc < self.rad[ci] + self.rad[cj]) & (Lc > 1e-5)
            hit &= ~np.isin(ci * N + cj, self._bond_keys)
            ci, cj = ci[hit], cj[hit]
            order = np.argsort(ci * N + cj)
            keep = list(zip(ci[order].tolist(), cj[order].tolist()))
        self.springs = {k: self.springs.get(k, np.zeros(3)) for k in keep}
        pi = np.concatenate([bi[touch_b], np.array([k[0] for k in keep], np.int64)])
        pj = np.concatenate([bj[touch_b], np.array([k[1] for k in keep], np.int64)])
        reach = np.concatenate([reach_b[touch_b], self.rad[pi[len(brk[touch_b]):]] +
                                self.rad[pj[len(brk[touch_b]):]]])
        O = np.concatenate([self.Ob[brk[touch_b]],
                            np.array([self.springs[k] for k in keep]).reshape(-1, 3)])
        P = len(pi)
        anyc = np.zeros(N, bool)
        self.cap_tie = np.inf
        if P:
            pairs = np.concatenate([np.stack([pi, pj], 1), np.stack([pj, pi], 1)])
            O2 = np.ascontiguousarray(np.concatenate([O, -O]))
            cap = np.zeros(2 * P)
            Fc, Mc = native_contacts(self.R, self.V, self.W, self.rad, pairs.astype(np.uint32),
                                     np.concatenate([reach, reach]), O2, dt, self.Estar,
                                     self.Gstar, 0.0 if quasi else self.pA, self.mu,
                                     0.0 if quasi else self.mur, self.rho43pi, cap=cap)
            self.cap_tie = float(np.abs(cap).min())
            nb = int(touch_b.sum())
            self.Ob[brk[touch_b]] = O2[:nb]
            for q, k in enumerate(keep):
                self.springs[k] = O2[nb + q].copy()
            anyc[pi] = True
            anyc[pj] = True
            if self.refmode:  # rolling-moment wipe (md_calculateforcesdem3.cu:330-335)
                still = (self.W * self.W).sum(1) <= 1e-18
                Mc[anyc & still] = 0.0
            F += Fc
            M += Mc
        self.F, self.M = F, M
        return F, M

    # ------------------------------------------------------- particle pass
    def _platen(self, R, V, F, friction=True, L=None):
        L = self.L if L is None else L
        mode = L["mode"]
        sb = st = 0.0
        a = 2 if mode == 1 else 1
        for i in range(self.N):
            if mode == 1:
                ra, rb, h, zc = R[i, 0] - L["cx"], R[i, 1] - L["cy"], R[i, 2], R[i, 2]
            else:
                ra, rb, h, zc = R[i, 0] - L["cx"], R[i, 2] - L["cz"], R[i, 1], R[i, 2] - L["cz"]
            if not (ra * ra + rb * rb < L["RR"] and zc < L["zcut"]):
                continue
            if h - L["top_r"] > -1e-9:
                f = F[i, a]
                if f > 0:
                    st += f
                    if mode == 2 and friction:
                        vv = V[i, 0] * V[i, 0] + V[i, 2] * V[i, 2]
                        if vv > 1e-9:
                            vv = 1.0 / math.sqrt(vv)
                            F[i, 0] = F[i, 0] - FRICTION * f * V[i, 0] * vv
                            F[i, 2] = F[i, 2] - FRICTION * f * V[i, 2] * vv
                if h - L["top"] > -1e-9:
                    R[i, a], V[i, a], F[i, a] = L["top"], L["v"], 0.0
            if h - L["bottom_r"] < 1e-9:
                f = F[i, a]
                if f < 0:
                    sb += f
                    if mode == 2 and friction:
                        vv = V[i, 0] * V[i, 0] + V[i, 2] * V[i, 2]
                        if vv > 1e-9:
                            vv = 1.0 / math.sqrt(vv)
                            F[i, 0] = F[i, 0] + FRICTION * f * V[i, 0] * vv
                            F[i, 2] = F[i, 2] + FRICTION * f * V[i, 2] * vv
                if h - L["bottom"] < 1e-9:
                    R[i, a], V[i, a], F[i, a] = L["bottom"], 0.0, 0.0
        return sb, st

    def particle_pass(self):
        R, V, F = self.R, self.V, self.F
        sums = self._platen(R, V, F) if self.L["mode"] else (0.0, 0.0)
        dt = self.dt
        V += F * (dt / self.m)[:, None]
        R += V * dt
        self.W = self.W + self.M * (dt / self.I)[:, None]
        if self.L.get("cut"):
            c, h = np.asarray(self.L["box_c"]), np.asarray(self.L["box_h"])
            d = R - c
            out = np.any((d - h > 0) | (d + h < 0), axis=1)
            R[out] = np.asarray(self.L["hidden"])
            V[out] = 0.0
            F[out] = 0.0
        return sums

    def control(self, s0, s1):
        L = self.L
        L["steps"] += 1
        if not L["mode"]:
            return
        L["sum_bottom"], L["sum_top"] = s0, s1
        self.history.append(((L["size"] - L["size0"]) / L["size0"], s0 * L["inv_area"],
                             s1 * L["inv_area"]))
        L["stress"] = 0.5 * (abs(s0) + abs(s1)) * L["inv_area"]
        if not L["loading"] and L["stress"] > L["min_stress"]:
            L["loading"] = 1
        if L["fracture_stress"] < L["stress"]:
            L["fracture_stress"] = L["stress"]
            L["fracture_strain"] = (L["size0"] - L["size"]) / L["size0"]
        if L["loading"] and L["stress"] < L["zero_stress"]:
            L["fracture"] = 1
        L["v"] = 10.0 * L["v_load"] if L["stress"] < 1e-9 else L["v_load"]
        L["size"] += L["v"] * self.dt
        L["top"] += L["v"] * self.dt
        L["top_r"] = L["top"] - L["band"]
        L["strain"] = (L["size"] - L["size0"]) / L["size0"]

    def step(self):
        self.forces()
        s0, s1 = self.particle_pass()
        self.control(s0, s1)
        return s0, s1
