"""The reference's on-disk formats, for interchange with its outputs.

* Sample_%u.txt: particle count, then one line "x y z" per particle in metres
  (SaveParticleDATAtxt / LoadParticleDATAtxt, md_createcylindersample.cu:104-133);
* Uniaxial.txt / Brazil.txt: the stress-strain record "EpsZ StressZ_bottom
  StressZ_top" with stresses in Pa (SaveSumForcesLoading, md_results.cu:119-131);
* LAMMPS data with bonds (SaveLammpsDATASimple, md_savelammpsdata.cpp:8-74);
* particle position files: N whitespace-separated "x y z" triples in metres,
  placed about the sample centre (ReadParticlesCSV, md_readparticlesfile.cu:9-72);
* DATA.dat: the binary checkpoint of the slot layout (SaveAllData /
  LoadAllData and the write_/read_ helpers, md_savealldata.cpp:12-277).  The
  reference is built with MSVC, where uint_fast32_t is 4 bytes.

Numbers in the text formats are written as C++ streams write them by default
(6 significant digits, "%g").  Positions and stresses are converted with the
unit constants of rfd.params (1 length unit = 1 mm, 1 stress unit = 1e9 Pa).
"""
from __future__ import annotations

import pathlib
from dataclasses import dataclass

import numpy as np

from .params import LENGTH, STRESS

ENERGY = 1.0  # energy unit in J (force unit 1e3 N x length unit 1e-3 m)


def _g(x: float) -> str:
    return "%g" % float(x)


# ------------------------------------------------------------ Sample_%u.txt
def write_sample(path, R: np.ndarray) -> None:
    """R: (N, 3) positions in length units."""
    R = np.asarray(R, np.float64)
    with open(path, "w") as f:
        f.write(f"{len(R)}\n")
        for x, y, z in R * LENGTH:
            f.write(f"{_g(x)} {_g(y)} {_g(z)}\n")


def read_sample(path) -> np.ndarray:
    """(N, 3) positions in length units (the reference reads them into FP32)."""
    tok = pathlib.Path(path).read_text().split()
    n = int(tok[0])
    v = np.array(tok[1:1 + 3 * n], np.float64).reshape(n, 3)
    return (v * (1.0 / LENGTH)).astype(np.float32).astype(np.float64)


def read_particles_csv(path, n: int, center=(0.0, 0.0, 0.0)):
    """Positions of the first n triples of a particle file, shifted to the
    sample centre, and the bounding box the reference derives from them
    (A_real, B_real, size_real, center_real; md_readparticlesfile.cu:42-72).
    Returns (R (n, 3), dict of the four float3)."""
    tok = pathlib.Path(path).read_text().replace(",", " ").split()
    if len(tok) < 3 * n:
        raise ValueError(f"{path}: {len(tok) // 3} triples, {n} asked for")
    v = np.array(tok[:3 * n], np.float64).reshape(n, 3)
    R = (v / LENGTH + np.asarray(center, np.float64)).astype(np.float32)
    lo, hi = R.min(0).astype(np.float64), R.max(0).astype(np.float64)
    box = {"A_real": lo, "B_real": hi, "size_real": hi - lo, "center_real": 0.5 * (hi + lo)}
    return R.astype(np.float64), box


# ------------------------------------------------- Uniaxial.txt / Brazil.txt
def write_stress_strain(path, hist: np.ndarray) -> None:
    """hist: (n, 3) rows (strain, bottom stress, top stress) in stress units,
    as rfd_system_get_history returns them."""
    with open(path, "w") as f:
        f.write("EpsZ StressZ_bottom StressZ_top\n")
        for e, sb, st in np.asarray(hist, np.float64):
            f.write(f"{_g(e)} {_g(sb * STRESS)} {_g(st * STRESS)}\n")


def read_stress_strain(path) -> np.ndarray:
    lines = pathlib.Path(path).read_text().splitlines()
    assert lines[0].split() == ["EpsZ", "StressZ_bottom", "StressZ_top"], lines[0]
    v = np.array([ln.split() for ln in lines[1:] if ln.strip()], np.float64).reshape(-1, 3)
    v[:, 1:] /= STRESS
    return v


# --------------------------------------------------------------- LAMMPS data
def write_lammps(path, R, bonds, box_lo, box_hi, dt, V=None, W=None, mass=None,
                 inertia=None) -> None:
    """Atoms with positions in metres; per-atom translational, rotational and
    total kinetic energy in the Velocities section when V and W are given
    (zeros otherwise); bonds as type 2 (the reference writes ILtype + 1 for
    its bond slots).  bonds: (Nb, 2) zero-based pairs."""
    R = np.asarray(R, np.float64)
    N = len(R)
    bonds = np.asarray(bonds, np.int64).reshape(-1, 2)
    with open(path, "w") as f:
        f.write(f"LAMMPS Description T0={_g(dt)}  (1st line of file)\n\n")
        f.write(f"{N} atoms\n{len(bonds)} bonds\n0 angles\n0 dihedrals\n0 impropers\n")
        f.write("1 atom types\n2 bond types\n")
        for d, name in enumerate("xyz"):
            f.write(f"{_g(box_lo[d] * LENGTH)} {_g(box_hi[d] * LENGTH)} {name}lo {name}hi\n")
        f.write("Masses\n\n1 1\n")
        f.write("Bond Coeffs\n\n2 1 2\n\n")
        f.write("Atoms\n\n")
        for i, (x, y, z) in enumerate(R * LENGTH):
            f.write(f"{i + 1} 1 {_g(x)} {_g(y)} {_g(z)}\n")
        f.write("Velocities\n\n")
        for i in range(N):
            if V is not None and W is not None:
                m = mass[i] if np.ndim(mass) else mass
                I = inertia[i] if np.ndim(inertia) else inertia
                et = 0.5 * m * float(np.dot(V[i], V[i])) * ENERGY
                er = 0.5 * I * float(np.dot(W[i], W[i])) * ENERGY
            else:
                et = er = 0.0
            f.write(f"{i + 1} {_g(et)} {_g(er)} {_g(et + er)}\n")
        f.write("\nBonds\n\n")
        for k, (a, b) in enumerate(bonds):
            f.write(f"{k + 1} 2 {a + 1} {b + 1}\n")


def write_lammps_system(path, sys, box_lo=None, box_hi=None) -> None:
    """LAMMPS data of a native-tier system (rfd.native.NativeSystem): its
    current positions, kinetic energies from its density and radii, and the
    bonds still intact (flag 1).  The box defaults to the particles' bounds."""
    st = sys.state()
    N = sys.N
    R = st["R"].reshape(3, N).T
    V = st["V"].reshape(3, N).T.astype(np.float64)
    W = st["W"].reshape(3, N).T.astype(np.float64)
    r = sys.radius.astype(np.float64)
    mass = sys.params.rho * (4.0 / 3.0) * np.pi * r ** 3
    inertia = 0.4 * mass * r * r
    if sys.bonds is not None and sys.bonds.n:
        _, flag = sys.bond_state()
        keep = flag == 1
        bonds = np.stack([sys.bonds.bi[keep], sys.bonds.bj[keep]], 1)
    else:
        bonds = np.zeros((0, 2), np.int64)
    lo = R.min(0) if box_lo is None else box_lo
    hi = R.max(0) if box_hi is None else box_hi
    write_lammps(path, R, bonds, lo, hi, sys.params.dt, V, W, mass, inertia)


@dataclass
class LammpsData:
    R: np.ndarray       # (N, 3) in length units
    bonds: np.ndarray   # (Nb, 2) zero-based
    box_lo: np.ndarray
    box_hi: np.ndarray
    energy: np.ndarray  # (N, 3): translational, rotational, total


def read_lammps(path) -> LammpsData:
    lines = pathlib.Path(path).read_text().splitlines()
    n = int(lines[2].split()[0])
    nb = int(lines[3].split()[0])
    lo, hi = np.zeros(3), np.zeros(3)
    for d in range(3):
        a, b = lines[9 + d].split()[:2]
        lo[d], hi[d] = float(a) / LENGTH, float(b) / LENGTH

    def section(name, count, cols):
        k = lines.index(name) + 2
        return np.array([ln.split()[:cols] for ln in lines[k:k + count]], np.float64)

    at = section("Atoms", n, 5)
    ve = section("Velocities", n, 4)
    bo = section("Bonds", nb, 4) if nb else np.zeros((0, 4))
    return LammpsData(at[:, 2:5] / LENGTH, bo[:, 2:4].astype(np.int64) - 1, lo, hi, ve[:, 1:4])


# ------------------------------------------------------------------ DATA.dat
SAMPLE_FIELDS = ("L", "A", "B", "axis", "center", "size", "R0", "H0")
SLOT_VEC = ("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd")


@dataclass
class Checkpoint:
    """The slot-layout state of DATA.dat: sample data (8 float3), particle
    arrays (SoA float32, 3N each) and the quaternion array (N float4), the
    interaction list (IL, ILtype, 1/L per slot and 7 float3 per slot)."""

    sample: dict
    F: np.ndarray
    V: np.ndarray
    R: np.ndarray
    M: np.ndarray
    W: np.ndarray
    Q: np.ndarray
    IonP: int
    acut: float
    IL: np.ndarray
    ILtype: np.ndarray
    iL: np.ndarray
    slot: dict  # name -> (S, 3) float32


def write_checkpoint(path, c: Checkpoint) -> None:
    f32, u32 = np.float32, np.uint32
    with open(path, "wb") as f:
        for k in SAMPLE_FIELDS:
            f.write(np.asarray(c.sample[k], f32).reshape(3).tobytes())
        N = len(c.R) // 3
        f.write(np.array([N], u32).tobytes())
        for a in (c.F, c.V, c.R, c.M, c.W):
            f.write(np.asarray(a, f32).reshape(3 * N).tobytes())
        f.write(np.asarray(c.Q, f32).reshape(4 * N).tobytes())
        S = len(c.IL)
        f.write(np.array([S, c.IonP], u32).tobytes())
        f.write(np.array([c.acut], f32).tobytes())
        f.write(np.asarray(c.IL, u32).tobytes())
        f.write(np.asarray(c.ILtype, np.uint8).tobytes())
        f.write(np.asarray(c.iL, f32).tobytes())
        for k in SLOT_VEC:
            f.write(np.asarray(c.slot[k], f32).reshape(3 * S).tobytes())


def read_checkpoint(path) -> Checkpoint:
    b = pathlib.Path(path).read_bytes()
    o = 0

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(b, dtype, n, o).copy()
        o += a.nbytes
        return a

    sample = {k: take(np.float32, 3) for k in SAMPLE_FIELDS}
    N = int(take(np.uint32, 1)[0])
    F, V, R, M, W = (take(np.float32, 3 * N) for _ in range(5))
    Q = take(np.float32, 4 * N)
    S, IonP = (int(x) for x in take(np.uint32, 2))
    acut = float(take(np.float32, 1)[0])
    IL = take(np.uint32, S)
    ILtype = take(np.uint8, S)
    iL = take(np.float32, S)
    slot = {k: take(np.float32, 3 * S).reshape(S, 3) for k in SLOT_VEC}
    assert o == len(b), f"{len(b) - o} trailing bytes"
    return Checkpoint(sample, F, V, R, M, W, Q, IonP, acut, IL, ILtype, iL, slot)


def _host(a) -> np.ndarray:
    """A numpy copy of a layout array (a torch tensor on any device, or numpy)."""
    if isinstance(a, np.ndarray):
        return a.copy()
    return a.detach().cpu().numpy()


def checkpoint_from_layout(lay, sample: dict | None = None, acut: float = 0.0) -> Checkpoint:
    """The DATA.dat state of a system in the reference's slot layout (the
    particle arrays and the interaction list of rfd.reflayout.RefLayout, or
    of any object with the same arrays).  Quaternions are not evolved by the
    per-step loop and are written as zeros; `sample` holds the 8 float3 of
    sample_data (zeros when not given)."""
    if hasattr(lay, "ctx"):
        import torch
        torch.cuda.synchronize()
    N, S = lay.N, lay.N * lay.K
    smp = sample or {k: np.zeros(3) for k in SAMPLE_FIELDS}
    return Checkpoint(smp, *(_host(getattr(lay, k)) for k in ("F", "V", "R", "M", "W")),
                      np.zeros(4 * N, np.float32), lay.K, acut,
                      _host(lay.IL).view(np.uint32), _host(lay.ILtype), _host(lay.iL),
                      {k: _host(lay.slot[k]).reshape(S, 3) for k in SLOT_VEC})


def restore_layout(lay, c: Checkpoint) -> None:
    """Load a checkpoint into a slot-layout system of the same size (the
    arrays LoadAllData fills; the cell arrays are rebuilt by the next cell
    distribution, which opens every step)."""
    N, S = lay.N, lay.N * lay.K
    if len(c.R) != 3 * N or len(c.IL) != S or c.IonP != lay.K:
        raise ValueError(f"checkpoint holds N={len(c.R) // 3}, K={c.IonP}; "
                         f"the layout has N={N}, K={lay.K}")

    def put(dst, src):
        src = np.ascontiguousarray(src)
        if isinstance(dst, np.ndarray):
            dst[...] = src.view(dst.dtype)
        else:
            import torch
            dst.copy_(torch.from_numpy(src.view(np.uint8)).to(dst.device).view(dst.dtype))

    for k in ("R", "V", "W", "F", "M"):
        put(getattr(lay, k), np.asarray(getattr(c, k), np.float32).reshape(-1))
    put(lay.IL, np.asarray(c.IL, np.uint32))
    put(lay.ILtype, np.asarray(c.ILtype, np.uint8))
    put(lay.iL, np.asarray(c.iL, np.float32))
    for k in SLOT_VEC:
        put(lay.slot[k], np.asarray(c.slot[k], np.float32).reshape(-1))
