"""numpy/ctypes wrapper of the CPU oracle (oracle/rfd_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to compare against.  The product never imports it.

``OracleLayout`` holds the reference slot-layout arrays in host memory and
has the same stage methods as rfd.reflayout.RefLayout, so that a parity test
runs both through the same sequence and compares the arrays.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "librfd_oracle.so"
LIB_OMP_PATH = HERE / "_build" / "librfd_oracle_omp.so"  # per-particle loops on OpenMP

_u32, _f32, _int = C.c_uint32, C.c_float, C.c_int
_P = C.c_void_p
_D = C.POINTER(C.c_double)

_SIG = {
    "orc_cell_distribution": ([_P, _u32, _f32, _u32, _u32, _u32, _P, _P, _P], _int),
    "orc_il_construct": ([_P, _u32, _P, _P, _P, _u32, _f32, _f32, _f32, _u32, _u32, _u32], _int),
    "orc_il_reconstruct_bpm": ([_P, _u32, _P, _P, _P, _P, _u32, _f32, _f32, _f32, _u32, _u32,
                                _u32, _int], _int),
    "orc_bonds_construct": ([_P] * 11 + [_u32, _u32, _f32], None),
    "orc_delete_far_links": ([_P, _u32] + [_P] * 9 + [_u32, _f32], None),
    "orc_forces_bpm": ([_P] * 12 + [_u32, _u32] + [_f32] * 6, None),
    "orc_forces_contact": ([_P] * 8 + [_u32, _u32] + [_f32] * 8, None),
    "orc_forces_slots": ([C.c_int] + [_P] * 13 + [_u32, _u32] + [_f32] * 9, None),
    "orc_platen_uniaxial_penalty": ([_P, _P, _P, _u32] + [_f32] * 8 + [_D], None),
    "orc_sumup_forces": ([_P] * 7 + [_u32, _u32], None),
    "orc_check_break": ([_P] * 7 + [_u32, _f32, _f32, _f32], None),
    "orc_increments": ([_P] * 5 + [_u32, _f32, _f32, _f32], None),
    "orc_platen_uniaxial": ([_P, _P, _P, _u32] + [_f32] * 9 + [_D], None),
    "orc_platen_brazil": ([_P, _P, _P, _u32] + [_f32] * 9 + [_D], None),
    "orc_cut_restriction": ([_P, _P, _P, _u32, _P, _P, _P], None),
    "orc_fire_increments": ([_P, _P, _P, _u32, _f32, _f32, _f32], None),
    "orc_fire_halfstep": ([_P, _P, _u32, _f32], None),
    "orc_fire_power": ([_P, _P, _P, _P, _u32, _D], None),
    "orc_norm3": ([_f32, _f32, _f32], _f32),
    "orc_rsqrt": ([_f32], _f32),
    "orc_threads": ([], _int),
    "orc_native_bonds_f64": ([_P] * 3 + [_u32] + [_P] * 8 + [_u32, C.c_double, C.c_double,
                                                             C.c_double, _int, _P, _P], None),
    "orc_native_contacts_f64": ([_P] * 4 + [_u32] + [_P] * 4 + [_u32] + [C.c_double] * 7
                                + [_P, _P, _P], None),
}

_libs: dict = {}


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib(omp: bool = False):
    """The oracle library; omp=True: the build whose per-particle loops run on
    OpenMP threads (the same results)."""
    path = LIB_OMP_PATH if omp else LIB_PATH
    if path not in _libs:
        if not path.exists():
            build()
        h = C.CDLL(str(path))
        for name, (args, res) in _SIG.items():
            f = getattr(h, name)
            f.argtypes = args
            f.restype = res
        _libs[path] = h
    return _libs[path]


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


ERR = {0: "ok", -10: "cell range", -11: "row full"}


class OracleLayout:
    """Host twin of rfd.reflayout.RefLayout evaluated by the C oracle."""

    def __init__(self, R: np.ndarray, K: int, grid, omp: bool = False):
        self._lib = lib(omp)
        R = np.asarray(R, dtype=np.float32).reshape(-1).copy()
        self.N = N = R.size // 3
        self.K = K
        self.grid = grid
        self.R = R
        self.V = np.zeros(3 * N, np.float32)
        self.W = np.zeros(3 * N, np.float32)
        self.F = np.zeros(3 * N, np.float32)
        self.M = np.zeros(3 * N, np.float32)
        self.CI = np.zeros(N, np.uint32)
        self.CIs = np.zeros(2 * N, np.uint32)
        self.pnC = np.zeros(2 * grid.C, np.uint32)
        S = N * K
        self.IL = np.full(S, 0xFFFFFFFF, np.uint32)
        self.ILtype = np.zeros(S, np.uint8)
        self.iL = np.zeros(S, np.float32)
        self.slot = {n: np.zeros(3 * S, np.float32)
                     for n in ("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd")}
        self.last_status = 0

    def _s(self, *names):
        return [ptr(self.slot[n]) for n in names]

    def cell_distribution(self):
        g = self.grid
        self.last_status = self._lib.orc_cell_distribution(ptr(self.R), self.N, g.inv_a, g.Nx, g.Ny,
                                                       g.Nz, ptr(self.CI), ptr(self.CIs),
                                                       ptr(self.pnC))

    def il_construct(self, aacut):
        g = self.grid
        self.last_status = self._lib.orc_il_construct(ptr(self.R), self.N, ptr(self.CIs),
                                                  ptr(self.pnC), ptr(self.IL), self.K, g.a,
                                                  g.inv_a, aacut, g.Nx, g.Ny, g.Nz)

    def il_reconstruct_bpm(self, aacut, quirk_p0=True):
        g = self.grid
        self.last_status = self._lib.orc_il_reconstruct_bpm(
            ptr(self.R), self.N, ptr(self.CIs), ptr(self.pnC), ptr(self.IL), ptr(self.ILtype),
            self.K, g.a, g.inv_a, aacut, g.Nx, g.Ny, g.Nz, int(quirk_p0))

    def bonds_construct(self, bdd):
        self._lib.orc_bonds_construct(ptr(self.R), ptr(self.IL), ptr(self.ILtype), ptr(self.iL),
                                  *self._s("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt",
                                           "Mijadd"), self.N, self.K, bdd)

    def delete_far_links(self, aa_farcut):
        self._lib.orc_delete_far_links(ptr(self.R), self.N, ptr(self.IL), ptr(self.ILtype),
                                   *self._s("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt",
                                            "Mijadd"), self.K, aa_farcut)

    def zero_fm(self):
        self.F[:] = 0
        self.M[:] = 0

    def forces_bpm(self, m):
        self._lib.orc_forces_bpm(ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                             ptr(self.ILtype), ptr(self.iL), *self._s("rij", "Oijt", "Mijn",
                                                                       "Mijt"),
                             ptr(self.F), ptr(self.M), self.N, self.K, m.dt, m.E, m.G, m.b_r,
                             m.sc, m.tc)

    def forces_contact(self, m):
        self._lib.orc_forces_contact(ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                                 ptr(self.ILtype), ptr(self.slot["Oijt"]), ptr(self.F),
                                 ptr(self.M), self.N, self.K, m.dt, m.hm_E, m.hm_G, m.p_A,
                                 m.p_mu, m.p_mur, m.p_m, m.p_r)

    def forces_slots(self, mode, m):
        E, G = (m.hm_E, m.hm_G) if mode == 1 else (m.E, m.G)
        self._lib.orc_forces_slots(int(mode), ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                                   ptr(self.ILtype), ptr(self.iL),
                                   *self._s("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt",
                                            "Mijadd"),
                                   self.N, self.K, m.dt, E, G, m.b_r, m.p_A, m.p_mu, m.p_mur,
                                   m.p_m, m.p_r)

    def platen_uniaxial_penalty(self, center, RR, Hd2, stiffness, mu, zcut):
        out = (C.c_double * 2)()
        self._lib.orc_platen_uniaxial_penalty(ptr(self.R), ptr(self.V), ptr(self.F), self.N,
                                              center[0], center[1], center[2], RR, Hd2,
                                              stiffness, mu, zcut, out)
        return out[0], out[1]

    def sumup_forces(self):
        self._lib.orc_sumup_forces(*self._s("Fijn", "Fijt", "Mijn", "Mijt", "Mijadd"), ptr(self.F),
                               ptr(self.M), self.N, self.K)

    def check_break(self, sc, tc, b_r):
        self._lib.orc_check_break(*self._s("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt"),
                              ptr(self.ILtype), self.N * self.K, sc, tc, b_r)

    def increments(self, m):
        self._lib.orc_increments(ptr(self.F), ptr(self.V), ptr(self.R), ptr(self.M), ptr(self.W),
                             self.N, m.dt_d_m, m.dt, m.dt_d_I)

    def platen_uniaxial(self, cmp, zcut):
        out = (C.c_double * 2)()
        self._lib.orc_platen_uniaxial(ptr(self.R), ptr(self.V), ptr(self.F), self.N, cmp.center_d[0],
                                  cmp.center_d[1], cmp.RR, cmp.top, cmp.bottom, cmp.top_r,
                                  cmp.bottom_r, cmp.V, zcut, out)
        return out[0], out[1]

    def platen_brazil(self, cmp, zcut):
        out = (C.c_double * 2)()
        self._lib.orc_platen_brazil(ptr(self.R), ptr(self.V), ptr(self.F), self.N, cmp.center_d[0],
                                cmp.center_d[2], cmp.RR, cmp.top, cmp.bottom, cmp.top_r,
                                cmp.bottom_r, cmp.V, zcut, out)
        return out[0], out[1]

    def cut_restriction(self, center, sized2, hidden_point):
        a = [np.asarray(x, np.float32) for x in (center, sized2, hidden_point)]
        self._lib.orc_cut_restriction(ptr(self.R), ptr(self.V), ptr(self.F), self.N, *map(ptr, a))

    def fire_power(self, rotations=False):
        out = (C.c_double * 4)()
        self._lib.orc_fire_power(ptr(self.F), ptr(self.V), ptr(self.M) if rotations else None,
                             ptr(self.W) if rotations else None, self.N, out)
        return tuple(out)

    def host(self):
        d = {k: getattr(self, k) for k in ("R", "V", "W", "F", "M", "CI", "CIs", "pnC", "IL",
                                           "ILtype", "iL")}
        d.update(self.slot)
        return d


class OracleUniaxialDriver:
    """The oracle's restatement of the driver loop
    (md_calculationuniaxialcompression.cu:131-189), with the same host logic
    objects (rfd.params.Compression) as the product driver."""

    def __init__(self, R, mat, box, cmp, grid, K=32, omp: bool = False):
        self.mat, self.box, self.cmp = mat, box, cmp
        self.s = OracleLayout(R, K, grid, omp)
        self.aacut = float(np.float32(mat.a_farcut) ** 2)
        self.zcut = float(box.center[2] + box.spacesized2[2])

    def setup(self):
        s, m = self.s, self.mat
        s.cell_distribution()
        s.il_construct(self.aacut)
        s.bonds_construct(m.b_dd)
        s.delete_far_links(m.aa_farcut)

    def step(self):
        s, m, cmp, box = self.s, self.mat, self.cmp, self.box
        s.cell_distribution()
        s.il_reconstruct_bpm(self.aacut)
        s.zero_fm()
        s.forces_bpm(m)
        s.forces_contact(m)
        sums = (s.platen_uniaxial(cmp, self.zcut) if cmp.mode == "uniaxial"
                else s.platen_brazil(cmp, self.zcut))
        cmp.sum_forces(sums)
        s.increments(m)
        s.cut_restriction(box.center, box.spacesized2, box.hidden_point)
        s.delete_far_links(m.aa_farcut)
        cmp.compress()
        return sums


def native_contacts(R, V, W, rad, pairs, reach, O, dt, Estar, Gstar, pA, mu, mur, rho43pi,
                    cap=None):
    """Native-tier contact law per directed pair (orc_native_contacts_f64):
    R, V, W are (N, 3), pairs (P, 2) with i = pairs[:, 0] seeing j =
    pairs[:, 1], reach (P,), O (P, 3) updated in place.  Returns F, M (N, 3)
    of i's load in float64; cap, a float64 (P,) array if given, receives each
    pair's Coulomb-test margin."""
    R, V, W = (np.ascontiguousarray(np.asarray(a, np.float64).T) for a in (R, V, W))
    N = R.shape[1]
    rad = np.ascontiguousarray(rad, np.float64)
    pi = np.ascontiguousarray(pairs[:, 0], np.uint32)
    pj = np.ascontiguousarray(pairs[:, 1], np.uint32)
    reach = np.ascontiguousarray(reach, np.float64)
    assert O.dtype == np.float64 and O.flags["C_CONTIGUOUS"] and O.shape == (len(pi), 3)
    F = np.zeros((3, N))
    M = np.zeros((3, N))
    if cap is not None:
        assert cap.dtype == np.float64 and cap.flags["C_CONTIGUOUS"] and cap.shape == (len(pi),)
    lib().orc_native_contacts_f64(ptr(R), ptr(V), ptr(W), ptr(rad), N, ptr(pi), ptr(pj),
                                  ptr(reach), ptr(O), len(pi), dt, Estar, Gstar, pA, mu, mur,
                                  rho43pi, ptr(F), ptr(M), ptr(cap))
    return F.T.copy(), M.T.copy()


def pack_iteration(R, rad, beta, cx, cy, rc, z0, z1, reach_pairs=None):
    """One iteration of the packer (k_pack_disp + k_pack_apply), restated in
    float64: every particle moves by -beta/2 * overlap along the pair
    direction for each overlapping partner (Jacobi, from the positions at the
    start of the iteration; the displacement is rounded to float32 as the
    device keeps it), then its sphere is clamped inside the cylinder
    (md_cylinderrestriction.cu:74-118 applied to sphere surfaces).  Pairs are
    found by brute force (all pairs, or the given candidate pairs).  Returns
    the new positions and the largest overlap relative to the smaller radius."""
    R = np.asarray(R, np.float64)
    rad = np.asarray(rad, np.float64)
    N = len(R)
    if reach_pairs is None:
        ii, jj = np.triu_indices(N, 1)
    else:
        ii, jj = reach_pairs[:, 0], reach_pairs[:, 1]
    r = R[jj] - R[ii]
    L = np.linalg.norm(r, axis=1)
    ov = rad[ii] + rad[jj] - L
    m = (ov > 0) & (L > 1e-12)
    ii, jj, r, L, ov = ii[m], jj[m], r[m], L[m], ov[m]
    c = (-0.5 * beta * ov / L)[:, None] * r
    D = np.zeros_like(R)
    np.add.at(D, ii, c)
    np.add.at(D, jj, -c)
    mo = float((ov / np.minimum(rad[ii], rad[jj])).max()) if len(ov) else 0.0
    out = R + D.astype(np.float32).astype(np.float64)
    ex, ey = out[:, 0] - cx, out[:, 1] - cy
    lim = rc - rad
    rho = np.sqrt(ex * ex + ey * ey)
    f = np.where(rho > lim, lim / np.where(rho > 0, rho, 1.0), 1.0)
    out[:, 0] = cx + ex * f
    out[:, 1] = cy + ey * f
    out[:, 2] = np.minimum(out[:, 2], z1 - rad)
    out[:, 2] = np.maximum(out[:, 2], z0 + rad)
    return out, mo


class FireRefState:
    """The reference FIRE loop's settings and state (firerelax_data,
    md_data_types.h:296-301, float fields), as SetFIREData leaves them
    (md_firerelaxation.cu:72-84) with the given dt0 and initial alpha."""

    def __init__(self, dt0, alpha0):
        f32 = np.float32
        self.dt0 = f32(dt0)
        self.dtmax = f32(100.0 * float(self.dt0))
        self.dtmin = f32(0.001 * float(self.dt0))
        self.dtgrow, self.dtshrink = f32(1.1), f32(0.5)
        self.alpha0, self.alphashrink = f32(alpha0), f32(0.99)
        self.n_negative_max, self.n_delay = 30000, 20
        self.dt, self.alpha = self.dt0, self.alpha0
        self.n_positive = self.n_negative = self.steps = 0
        self.stopped = False
        self.power = 0.0


def fire_relax_reference(F, V, R, N, dt_m, f, iterations, forces=None, restriction=None):
    """The reference's FIRE loop (md_firerelaxation.cu:201-304) on host arrays
    (float32, SoA by component): forces() fills F, then the power F.V (summed
    in double), the P > 0 / P <= 0 rules on the float dt and alpha, the
    half-step back-off and velocity reset, d_CalculateIncrementsFIRE and
    restriction()."""
    f32 = np.float32
    out = (C.c_double * 4)()
    for _ in range(iterations):
        if f.stopped:
            break
        if forces is not None:
            forces()
        lib().orc_fire_power(ptr(F), ptr(V), None, None, N, out)
        f.power = out[0]
        if out[0] > 0:
            f.n_positive += 1
            f.n_negative = 0
            if f.n_positive > f.n_delay:
                g = f32(f.dt * f.dtgrow)
                f.dt = g if g < f.dtmax else f.dtmax
                f.alpha = f32(f.alpha * f.alphashrink)
        else:
            f.n_positive = 0
            f.n_negative += 1
            if f.n_negative > f.n_negative_max:
                f.stopped = True
                break
            if f.steps > f.n_delay:
                g = f32(f.dt * f.dtshrink)
                f.dt = g if g > f.dtmin else f.dtmin
                f.alpha = f.alpha0
            lib().orc_fire_halfstep(ptr(V), ptr(R), N, f32(0.5 * float(f.dt)))
            V[:] = 0
        lib().orc_fire_increments(ptr(F), ptr(V), ptr(R), N, f32(dt_m), f.dt, f.alpha)
        if restriction is not None:
            restriction()
        f.steps += 1
    return f

