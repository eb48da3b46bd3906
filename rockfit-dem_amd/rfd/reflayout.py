"""Reference-layout tier: the reference's device buffers and host wrappers,
driven through the C ABI.

``RefLayout`` owns the buffers a reference host driver owns
(particle_data / cell_data / interaction_list_data, md_data_types.h:12-56,
124-144) as torch device tensors, and exposes one method per reference host
wrapper or launch site.  ``UniaxialDriver`` / ``BrazilDriver`` re-express the
test-driver loops h_CalculationUniaxialCompression
(md_calculationuniaxialcompression.cu:113-197) and h_CalculationBrazilTest
(md_calculationbraziltest.cu:113-198) on top of it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import RfdError, lib
from .params import Compression, Material, SampleBox

SENTINEL = 0xFFFFFFFF
FLAG_CELL_RANGE = 0x1
FLAG_ROW_FULL = 0x2


def _p(t: torch.Tensor) -> int:
    return t.data_ptr()


class Context:
    """An rfd_ctx bound to one device and (by default) the null stream."""

    def __init__(self, device: int = 0, stream: int | None = None):
        self.lib = lib()
        h = C.c_void_p()
        st = self.lib.rfd_ctx_create(C.byref(h), device, stream)
        if st != 0:
            raise RfdError("rfd_ctx_create", st, "cannot create context")
        self.h = h
        self.device = device

    def call(self, name: str, *args) -> None:
        st = getattr(self.lib, name)(self.h, *args)
        if st != 0:
            msg = self.lib.rfd_last_error(self.h)
            raise RfdError(name, st, msg.decode() if msg else "")

    def flags(self) -> int:
        f = C.c_uint32()
        self.call("rfd_read_flags", C.byref(f))
        return f.value

    def sync(self) -> None:
        self.call("rfd_sync")

    def close(self) -> None:
        if self.h:
            self.lib.rfd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Grid:
    a: float
    inv_a: float
    Nx: int
    Ny: int
    Nz: int

    @property
    def C(self) -> int:
        return self.Nx * self.Ny * self.Nz


class RefLayout:
    """Reference slot-layout buffers on the device (all torch tensors)."""

    def __init__(self, ctx: Context, R: np.ndarray, K: int, grid: Grid, device="cuda"):
        self.ctx = ctx
        R = np.asarray(R, dtype=np.float32).reshape(-1)
        self.N = N = R.size // 3
        self.K = K
        self.grid = grid
        dev = torch.device(device)
        f32, i32, u8 = torch.float32, torch.int32, torch.uint8
        self.R = torch.from_numpy(R.copy()).to(dev)
        self.V = torch.zeros(3 * N, dtype=f32, device=dev)
        self.W = torch.zeros(3 * N, dtype=f32, device=dev)
        self.F = torch.zeros(3 * N, dtype=f32, device=dev)
        self.M = torch.zeros(3 * N, dtype=f32, device=dev)
        self.CI = torch.zeros(N, dtype=i32, device=dev)
        self.CIs = torch.zeros(2 * N, dtype=i32, device=dev)
        self.pnC = torch.zeros(2 * grid.C, dtype=i32, device=dev)
        S = N * K
        self.IL = torch.full((S,), -1, dtype=i32, device=dev)
        self.ILtype = torch.zeros(S, dtype=u8, device=dev)
        self.iL = torch.zeros(S, dtype=f32, device=dev)
        self.slot = {name: torch.zeros(3 * S, dtype=f32, device=dev)
                     for name in ("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd")}

    # -- host wrappers / launch sites -------------------------------------
    def cell_distribution(self):
        g = self.grid
        self.ctx.call("rfd_cell_distribution", _p(self.R), self.N, g.inv_a, g.Nx, g.Ny, g.Nz,
                      _p(self.CI), _p(self.CIs), _p(self.pnC))

    def il_construct(self, aacut: float):
        g = self.grid
        self.ctx.call("rfd_il_construct", _p(self.R), self.N, _p(self.CIs), _p(self.pnC),
                      _p(self.IL), self.K, g.a, g.inv_a, aacut, g.Nx, g.Ny, g.Nz)

    def il_reconstruct_bpm(self, aacut: float, quirk_p0: bool = True):
        g = self.grid
        self.ctx.call("rfd_il_reconstruct_bpm", _p(self.R), self.N, _p(self.CIs), _p(self.pnC),
                      _p(self.IL), _p(self.ILtype), self.K, g.a, g.inv_a, aacut, g.Nx, g.Ny,
                      g.Nz, int(quirk_p0))

    def _slots(self, *names):
        return [_p(self.slot[n]) for n in names]

    def bonds_construct(self, bdd: float):
        self.ctx.call("rfd_bonds_construct", _p(self.R), _p(self.IL), _p(self.ILtype),
                      _p(self.iL), *self._slots("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt",
                                                "Mijadd"), self.N, self.K, bdd)

    def delete_far_links(self, aa_farcut: float):
        self.ctx.call("rfd_delete_far_links", _p(self.R), self.N, _p(self.IL), _p(self.ILtype),
                      *self._slots("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd"),
                      self.K, aa_farcut)

    def zero_fm(self):
        self.ctx.call("rfd_zero_fm", _p(self.F), _p(self.M), self.N)

    def forces_bpm(self, m: Material):
        self.ctx.call("rfd_forces_bpm", _p(self.R), _p(self.V), _p(self.W), _p(self.IL),
                      _p(self.ILtype), _p(self.iL), *self._slots("rij", "Oijt", "Mijn", "Mijt"),
                      _p(self.F), _p(self.M), self.N, self.K, m.dt, m.E, m.G, m.b_r, m.sc, m.tc)

    def forces_contact(self, m: Material):
        self.ctx.call("rfd_forces_contact", _p(self.R), _p(self.V), _p(self.W), _p(self.IL),
                      _p(self.ILtype), _p(self.slot["Oijt"]), _p(self.F), _p(self.M), self.N,
                      self.K, m.dt, m.hm_E, m.hm_G, m.p_A, m.p_mu, m.p_mur, m.p_m, m.p_r)

    def forces_slots(self, mode: int, m: Material):
        """The superseded per-slot force path (md_calculateforcesdem.cu): every
        slot's law to the slot arrays; mode 0 d_CalculateForcesDEM (bonds, empty
        slots zeroed), 1 d_CalculateForcesDEM_1 (contacts, Hertz-Mindlin moduli),
        2 d_CalculateForcesDEM_2 (bonds only).  F and M follow from
        sumup_forces."""
        E, G = (m.hm_E, m.hm_G) if mode == 1 else (m.E, m.G)
        self.ctx.call("rfd_forces_slots", int(mode), _p(self.R), _p(self.V), _p(self.W),
                      _p(self.IL), _p(self.ILtype), _p(self.iL),
                      *self._slots("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd"),
                      self.N, self.K, m.dt, E, G, m.b_r, m.p_A, m.p_mu, m.p_mur, m.p_m, m.p_r)

    def sumup_forces(self):
        """Stand-alone sum-up (d_SumUpForcesDEM): F, M from the slot loads."""
        self.ctx.call("rfd_sumup_forces", *self._slots("Fijn", "Fijt", "Mijn", "Mijt", "Mijadd"),
                      _p(self.F), _p(self.M), self.N, self.K)

    def check_break(self, sc: float, tc: float, b_r: float):
        """Stand-alone break check (d_CheckBreakConditionDEM) over all slots."""
        self.ctx.call("rfd_check_break", *self._slots("rij", "Oijt", "Fijn", "Fijt", "Mijn",
                                                      "Mijt"), _p(self.ILtype),
                      self.N * self.K, sc, tc, b_r)

    def increments(self, m: Material):
        self.ctx.call("rfd_increments", _p(self.F), _p(self.V), _p(self.R), _p(self.M),
                      _p(self.W), self.N, m.dt_d_m, m.dt, m.dt_d_I)

    def platen_uniaxial(self, cmp: Compression, zcut: float) -> tuple[float, float]:
        out = (C.c_double * 2)()
        self.ctx.call("rfd_platen_uniaxial", _p(self.R), _p(self.V), _p(self.F), self.N,
                      cmp.center_d[0], cmp.center_d[1], cmp.RR, cmp.top, cmp.bottom, cmp.top_r,
                      cmp.bottom_r, cmp.V, zcut, out)
        return out[0], out[1]

    def platen_uniaxial_penalty(self, center, RR: float, Hd2: float, stiffness: float,
                                mu: float, zcut: float) -> tuple[float, float]:
        """The superseded penalty platens (d_UniaxialCompression_simple): returns
        the (bottom, top) reaction sums."""
        out = (C.c_double * 2)()
        self.ctx.call("rfd_platen_uniaxial_penalty", _p(self.R), _p(self.V), _p(self.F), self.N,
                      center[0], center[1], center[2], RR, Hd2, stiffness, mu, zcut, out)
        return out[0], out[1]

    def platen_brazil(self, cmp: Compression, zcut: float) -> tuple[float, float]:
        out = (C.c_double * 2)()
        self.ctx.call("rfd_platen_brazil", _p(self.R), _p(self.V), _p(self.F), self.N,
                      cmp.center_d[0], cmp.center_d[2], cmp.RR, cmp.top, cmp.bottom, cmp.top_r,
                      cmp.bottom_r, cmp.V, zcut, out)
        return out[0], out[1]

    def cut_restriction(self, center, sized2, hidden_point):
        f3 = C.c_float * 3
        self.ctx.call("rfd_cut_restriction", _p(self.R), _p(self.V), _p(self.F), self.N,
                      f3(*center), f3(*sized2), f3(*hidden_point))

    def fire_power(self, rotations: bool = False) -> tuple[float, float, float, float]:
        out = (C.c_double * 4)()
        M = _p(self.M) if rotations else None
        W = _p(self.W) if rotations else None
        self.ctx.call("rfd_fire_power", _p(self.F), _p(self.V), M, W, self.N, out)
        return tuple(out)

    # -- host views --------------------------------------------------------
    def host(self) -> dict[str, np.ndarray]:
        torch.cuda.synchronize()
        d = {k: getattr(self, k).cpu().numpy() for k in ("R", "V", "W", "F", "M")}
        d["CI"] = self.CI.cpu().numpy().view(np.uint32)
        d["CIs"] = self.CIs.cpu().numpy().view(np.uint32)
        d["pnC"] = self.pnC.cpu().numpy().view(np.uint32)
        d["IL"] = self.IL.cpu().numpy().view(np.uint32)
        d["ILtype"] = self.ILtype.cpu().numpy()
        d["iL"] = self.iL.cpu().numpy()
        for k, v in self.slot.items():
            d[k] = v.cpu().numpy()
        return d

    def bond_slots(self) -> int:
        return int((self.ILtype == 1).sum().item())


class ReferenceFire:
    """The reference's FIRE relaxation loop (rfd_fire_relax_reference,
    md_firerelaxation.cu:201-304) on a RefLayout's F, V and R: per iteration
    forces() fills lay.F, then the power, the dt / alpha rules, the back-off
    and d_CalculateIncrementsFIRE run in the library, then restriction().
    The settings and the loop state (float dt and alpha, positive and
    negative counters, iteration count) persist between run() calls, as the
    reference keeps them in firerelax_data."""

    def __init__(self, dt0: float, alpha0: float = 0.25):
        from ._lib import FireRef

        self.state = FireRef()
        lib().rfd_fire_ref_init(C.byref(self.state), dt0, alpha0)

    def run(self, lay: "RefLayout", dt_m: float, iterations: int, forces=None,
            restriction=None) -> None:
        from ._lib import CALLBACK

        errors = []

        def wrap(fn):
            if fn is None:
                return None

            def cb(_user):
                try:
                    fn()
                    return 0
                except Exception as e:  # reported after the call returns
                    errors.append(e)
                    return 1
            return CALLBACK(cb)

        cf, cr = wrap(forces), wrap(restriction)
        try:
            lay.ctx.call("rfd_fire_relax_reference", _p(lay.F), _p(lay.V), _p(lay.R), lay.N,
                         dt_m, C.byref(self.state), iterations, cf, cr, None)
        except RfdError:
            if errors:
                raise errors[0]
            raise


class UniaxialDriver:
    """h_CalculationUniaxialCompression (md_calculationuniaxialcompression.cu:113-197)
    on the reference layout.  Setup (:131-143): CellDistribution,
    InteractionListConstruct, bonds, far-link prune.  Step (:169-189)."""

    def __init__(self, ctx: Context, R: np.ndarray, mat: Material, box: SampleBox,
                 cmp: Compression, K: int = 32):
        g = box.grid
        self.grid = Grid(float(np.float32(box.cell_a)), g.inv_a, g.Nx, g.Ny, g.Nz)
        self.mat, self.box, self.cmp = mat, box, cmp
        self.s = RefLayout(ctx, R, K, self.grid)
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

    def run(self, max_steps: int) -> int:
        n = 0
        while n < max_steps and self.cmp.strain < self.cmp.max_strain and not self.cmp.fracture:
            self.step()
            n += 1
        return n
