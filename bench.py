#!/usr/bin/env python3
"""Benchmark of the BPM explicit step on one MI355X.

Prints ONE JSON line (rank 0):
  metric  bond-updates/s over whole explicit steps of a uniaxial-compression
          run (cell sort, list reconstruct, bond + contact forces with the
          break check, platens, integration, cut, prune, loading control)
  roofline  the bond-force kernel: algorithmic bytes per launch / its average
          duration measured with HIP events on the launch stream
  cpu_baseline  the oracle (single-threaded C restatement) on a bounded
          sample of the same kind of step, on this host's cores

One step = one pass of the hot path over the whole specimen.  Inputs are
synthetic (seeded FCC-lattice cylinder) and resident in HBM before timing.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "rockfit-dem_amd"))
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12  # B/s, MI355X


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--tier", choices=["ref", "native"], default="native")
    p.add_argument("--D", type=float, default=100.0)
    p.add_argument("--H", type=float, default=200.0)
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-warmstart", action="store_true",
                   help="skip the learned-warm-start load-step measurement")
    p.add_argument("--warmstart-only", action="store_true",
                   help="only the learned-warm-start load-step measurement")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the batched Brazilian-disc measurement")
    p.add_argument("--batch-only", action="store_true",
                   help="only the batched Brazilian-disc measurement")
    p.add_argument("--calibrate", action="store_true",
                   help="only the calibration outer loop (CG steps on batched tests)")
    p.add_argument("--disc-only", action="store_true",
                   help="only the 2k-sphere Brazilian disc (FP64 and FP32 GPU, CPU oracle)")
    p.add_argument("--fire-only", action="store_true",
                   help="only the plain-FIRE load steps on the large cylinder and the "
                        "FP32 / FP64 FIRE comparison")
    p.add_argument("--c2-peak-only", action="store_true",
                   help="only the FP32 / FP64 quasi-static run of the D25xH50 cylinder to past "
                        "its peak")
    p.add_argument("--packed-only", action="store_true",
                   help="only the random-packing cylinder measurement")
    p.add_argument("--copy-calib", action="store_true",
                   help="only run a plain device copy of 1 GiB (counter calibration and "
                        "stream-copy ceiling)")
    return p.parse_args()


COPY_BYTES = 1 << 30


def stream_copy(reps: int = 5) -> dict:
    """HBM ceiling: the library's streaming copy (rfd_stream_copy, 16-byte
    non-temporal loads and stores) of 1 GiB, best of a sweep of grid sizes,
    and torch's device copy beside it; rates count bytes read + written."""
    import torch

    from rfd.reflayout import Context

    ctx = Context(0)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def rate(fn):
        fn()
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return 2 * COPY_BYTES * reps / (a.elapsed_time(b) * 1e-3)

    sweep = {}
    for blocks in (1024, 2048, 4096, 8192, 16384, 32768):
        sweep[blocks] = rate(lambda: ctx.call("rfd_stream_copy", dst.data_ptr(), src.data_ptr(),
                                              COPY_BYTES, blocks))
    assert torch.equal(dst, src)
    best = max(sweep, key=sweep.get)
    out = {"stream_copy_GBps": sweep[best] / 1e9, "stream_copy_blocks": best,
           "torch_copy_GBps": rate(lambda: dst.copy_(src)) / 1e9}
    ctx.close()
    return out


def force_alg_bytes(kernel: str, bonds: int, particles: int, contacts: int = 0) -> int:
    """Algorithmic bytes of one native force launch: 128 B per bond and 72 B
    per particle; the fused particle update adds 1/m, 1/I (8 B) read and the
    next R, V, W (48 B) written per particle; each contact-row entry (a
    never-bonded pair within reach) its partner id (4 B) and spring read and
    written (32 B)."""
    return (128 * bonds + (72 + (56 if kernel.endswith("<fused>") else 0)) * particles +
            36 * contacts)


def load_traffic(kernel: str, particles: int, bonds: int):
    """HBM bytes per launch of `kernel` measured by the FETCH_SIZE / WRITE_SIZE
    counter runs recorded under profiles/ for this workload, or None."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    for rec in json.loads(f.read_text()):
        if rec["kernel"] == kernel and rec["particles"] == particles and rec["bonds"] == bonds:
            return rec["bytes_per_launch"]
    return None


def cpu_model() -> str:
    """The host CPU's model name, from /proc/cpuinfo ("unknown" elsewhere)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_ref(seconds: float, omp: bool = False) -> dict:
    """Oracle driver on a 25 x 50 mm cylinder, the size of the reference's own
    28,593-particle sample (kernel.cu:166): single thread as the reference's
    CPU path runs, or (omp) its per-particle loops on every OpenMP thread."""
    from oracle.oracle import OracleUniaxialDriver, lib
    from rfd import cases

    c = cases.reference_uniaxial(25.0, 50.0)
    d = OracleUniaxialDriver(c.R, c.mat, c.box, c.cmp, c.grid(), c.K, omp=omp)
    threads = lib(omp).orc_threads()
    d.setup()
    d.step()
    nb = int((d.s.ILtype == 1).sum()) // 2
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        d.step()
        n += 1
    dt = time.perf_counter() - t0
    return {"value": nb * n / dt, "unit": "bond-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle step loop, {n} explicit steps of a {c.N}-particle "
                      f"D25xH50 mm cylinder ({nb} bonds), "
                      + (f"{threads} OpenMP threads" if omp else "single thread")}


def bench_ref(args) -> dict:
    import numpy as np
    import torch

    from rfd import cases
    from rfd.reflayout import Context, UniaxialDriver

    ctx = Context(0)
    case = cases.reference_uniaxial(args.D, args.H)
    drv = UniaxialDriver(ctx, case.R, case.mat, case.box, case.cmp, case.K)
    drv.setup()
    print(f"[bench] ref tier: N={case.N} K={case.K}", file=sys.stderr, flush=True)
    s = drv.s
    # HIP events around each bond-force launch (same null stream as the ctx)
    ev = []
    orig = s.forces_bpm

    def timed_forces(m):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        orig(m)
        b.record()
        ev.append((a, b))

    s.forces_bpm = timed_forces
    for _ in range(args.warmup):
        drv.step()
    torch.cuda.synchronize()
    ev.clear()
    nb = s.bond_slots() // 2
    t0 = time.perf_counter()
    for _ in range(args.steps):
        drv.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    nb_end = s.bond_slots() // 2
    ms = (t1 - t0) * 1e3 / args.steps
    N = case.N
    # k_forces_bpm in the slot layout: IL + ILtype of every slot (5 B); per
    # particle R, V, W read and F, M read and written (84 B); per bonded slot
    # (two per bond) iL, rij, Oijt, Mijn, Mijt read (52 B) and rij, Oijt,
    # Mijn, Mijt written (48 B)
    alg_bytes = 5 * N * case.K + 84 * N + 100 * 2 * nb
    achieved = alg_bytes / (kern_ms * 1e-3)
    return {
        "value": nb * args.steps / (t1 - t0),
        "ms_per_step": ms,
        "dtype": "f32",
        "config": {"workload": f"uniaxial cylinder D{args.D:g}xH{args.H:g} mm, reference slot "
                               f"layout K={case.K}, monodisperse r=0.5 mm",
                   "particles": N, "bonds": nb, "bonds_end": nb_end, "tier": "ref"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None,
                     "kernel": "k_forces_bpm", "kernel_ms": kern_ms,
                     "alg_bytes_per_launch": alg_bytes},
    }


def bench_native(args) -> dict:
    """Native tier: polydisperse cylinder (r ~ U[0.4,0.6] mm, FCC lattice of
    spacing 1 mm, every lattice neighbour pair bonded so Nb/N ~ 6), bond radius
    min(ri,rj), strengths N(20,4)/N(40,8) MPa; uniaxial loading control on the
    device; contact rows rebuilt every 20 steps."""
    from rfd import native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    nc = native_cases.uniaxial_cylinder(ctx, args.D, args.H)
    s = nc.sys
    N, Nb = s.N, s.bonds.n
    print(f"[bench] native tier: N={N} Nb={Nb} setup {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    s.step(args.warmup)
    ctx.sync()
    s.timing(True)
    t0 = time.perf_counter()
    s.step(args.steps)
    ctx.sync()
    t1 = time.perf_counter()
    kern_ms, launches = s.timing_read()
    kern_ms /= max(launches, 1)
    active, broken, removed = s.bond_counts()
    assert ctx.flags() == 0, "device flags raised during the benchmark"
    kernel = s.force_kernel()
    alg_bytes = force_alg_bytes(kernel, Nb, N)
    achieved = alg_bytes / (kern_ms * 1e-3)
    return {
        "value": active * args.steps / (t1 - t0),
        "ms_per_step": (t1 - t0) * 1e3 / args.steps,
        "dtype": "f32",
        "config": {"workload": f"uniaxial cylinder D{args.D:g}xH{args.H:g} mm, polydisperse "
                               "r~U[0.4,0.6] mm, native edge-list tier",
                   "particles": N, "bonds": Nb, "bonds_active_end": active,
                   "bonds_broken_end": broken, "tier": "native",
                   "precision": "f32 bond/contact law; f64 positions, stretch, break check",
                   "contact_rebuild_every": 20,
                   "contact_entries_per_particle": s.contact_entries() / N},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None,
                     "kernel": kernel, "kernel_ms": kern_ms,
                     "alg_bytes_per_launch": alg_bytes},
    }


def bench_c2(steps: int = 500) -> dict:
    """Explicit steps on the D25xH50 cylinder, the size of the CPU baseline's
    sample: the reference-layout tier on the same monodisperse specimen the
    CPU loop runs, and the native tier on the polydisperse one."""
    import argparse

    out = {}
    for name, fn in (("ref_tier", bench_ref), ("native_tier", bench_native)):
        r = fn(argparse.Namespace(D=25.0, H=50.0, steps=steps, warmup=20))
        out[name] = {"bond_updates_per_s": r["value"], "ms_per_step": r["ms_per_step"],
                     "particles": r["config"]["particles"], "bonds": r["config"]["bonds"],
                     "workload": r["config"]["workload"]}
    out["ref_tier"]["driver"] = "Python (ctypes) calls, one per reference launch site"
    out["ref_tier_cpp"] = bench_ref_cpp(25.0, 50.0, steps, out["ref_tier"]["bonds"])
    return out


def bench_ref_cpp(D: float, H: float, steps: int, bonds: int) -> dict:
    """The reference-layout tier driven by the C++ host driver (host/rfd_run),
    as the reference's own C++ test loop would call the library: one ABI call
    per reference launch site, the blocking platen-sum read each step
    (md_results.cu:36).  Same monodisperse specimen as the Python-driven line."""
    import re
    import subprocess
    import tempfile

    import numpy as np

    from rfd import cases, formats
    from rfd.build import RUN

    c = cases.reference_uniaxial(D, H)
    N, K = c.N, c.K
    S = N * K
    z = np.zeros(3 * N, np.float32)
    slots = {k: np.zeros((S, 3), np.float32) for k in formats.SLOT_VEC}
    ck = formats.Checkpoint({k: np.zeros(3) for k in formats.SAMPLE_FIELDS}, z, z, c.R, z, z,
                            np.zeros(4 * N, np.float32), K, 0.0,
                            np.full(S, 0xFFFFFFFF, np.uint32), np.zeros(S, np.uint8),
                            np.zeros(S, np.float32), slots)
    with tempfile.TemporaryDirectory() as td:
        formats.write_checkpoint(f"{td}/in.dat", ck)
        r = subprocess.run([str(RUN), "--mode", "uniaxial", "--D", str(D), "--H", str(H),
                            "--in", f"{td}/in.dat", "--out", td, "--steps", str(steps)],
                           capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"rfd_run failed: {r.stderr[-2000:]}")
    m = re.search(r"us_per_step=([0-9.]+)", r.stdout)
    us = float(m.group(1))
    return {"driver": "C++ host driver (host/rfd_run), one ABI call per reference launch site",
            "ms_per_step": us / 1e3, "particles": N, "bonds": bonds, "steps": steps,
            "device_flags_clear": "flags=0" in r.stdout,
            "bond_updates_per_s": bonds * 1e6 / us,
            "workload": f"uniaxial cylinder D{D:g}xH{H:g} mm, reference slot layout K={K}, "
                        "monodisperse r=0.5 mm"}


def bench_warmstart(D: float = 61.0, H: float = 122.0, steps: int = 12) -> dict:
    """Quasi-static load steps (strain increment -1e-5, FIRE to |F|max <
    1e-6 F_ref) on a ~500k-particle cylinder from three starts: plain GPU
    FIRE (the platen move alone), the skip path (the last step's increment
    repeated: the network with a zero correction head) and the trained
    network (skip path + learned correction).  The network is trained on the
    device on plain-FIRE relaxations of four cylinders of 19k to 207k
    particles and validated on a fifth; its correction head is kept only if
    it beats a zero correction on the validation steps, so with the head off
    the "learned" start is the skip path, and the line says so.  All runs
    start from the same seeded specimen; the first two steps (seating, first
    increment) are plain and not timed."""
    from rfd import native, native_cases, warmstart as ws
    from rfd.reflayout import Context

    ctx = Context(0)
    dstrain = -1e-5
    t0 = time.perf_counter()

    def spec(d, h, seed):
        nc = native_cases.uniaxial_cylinder(ctx, d, h, seed=seed, flat_ends=True)
        nc.sys.seat_platens(preload=0.0)
        return nc

    def log(tag):
        def f(k, r):
            print(f"[bench] {tag} step {k}: {r.iterations} FIRE iterations, {r.seconds:.2f}s, "
                  f"converged {r.converged}, warm {r.warm_used}", file=sys.stderr, flush=True)
        return f

    train_specs = ((21.0, 101), (28.0, 102), (36.0, 105), (46.0, 106))
    samples, sizes = [], []
    fp = None
    for d, seed in train_specs:
        c = spec(d, 2 * d, seed)
        if fp is None:
            fp = native.fire_params(c.mat.E, 0.5, c.mat.dt, rel_tol=1e-6, max_iter=100000,
                                    check_every=200)
        s_, _ = ws.collect_samples(c.sys, dstrain, fp, 8, log=log(f"train N={c.sys.N}"))
        samples += s_
        sizes.append(c.sys.N)
        del c
    vs = spec(28.0, 56.0, 103)
    val, _ = ws.collect_samples(vs.sys, dstrain, fp, 8, log=log("val"))
    net = ws.WarmStartNet(ctx, seed=7)
    t1 = time.perf_counter()
    hist = ws.train(net, samples, epochs=30, lr=1e-3, val=val)
    ctx.sync()
    t_train = time.perf_counter() - t1
    print(f"[bench] warm start: {len(samples)} samples, trained in {t_train:.1f}s, "
          f"head {'on' if hist['head'] else 'off'}", file=sys.stderr, flush=True)
    skip = ws.WarmStartNet(ctx, seed=7)  # zero decoder: repeats the last increment
    out = {}
    for name, nt in (("plain", None), ("skip", skip), ("learned", net)):
        c = spec(D, H, 104)
        warm = (lambda s_, nt=nt: ws.predict(nt, s_, dstrain)) if nt is not None else None
        _, res = ws.collect_samples(c.sys, dstrain, fp, 2 + steps, warm=warm,
                                    log=log(f"{name} N={c.sys.N}"))
        timed = res[2:]
        sec = sum(r.seconds for r in timed)
        out[name] = {"particles": c.sys.N, "bonds": c.sys.bonds.n,
                     "load_steps_per_s": len(timed) / sec, "s_per_step": sec / len(timed),
                     "fire_iterations_per_step": sum(r.iterations for r in timed) / len(timed),
                     "fire_iterations": [int(r.iterations) for r in timed],
                     "all_converged": all(bool(r.converged) for r in timed),
                     "max_final_residual": max(r.fmax for r in timed),
                     "warm_starts_taken": sum(r.warm_used for r in timed)}
        print(f"[bench] warm start {name}: {out[name]}", file=sys.stderr, flush=True)
        del c
    head = bool(hist["head"])
    if not head:
        out["learned"]["note"] = ("correction head off (it did not beat a zero correction on "
                                  "the validation steps): this start is the skip path")
    return {"metric": "converged load-steps/s", "unit": "load-steps/s",
            "workload": f"uniaxial cylinder D{D:g}xH{H:g} mm, strain step {dstrain:g}, "
                        "FIRE to |F|max < 1e-6 E pi r_mean^2",
            "ftol": fp.ftol, **out,
            "speedup_wall": out["plain"]["s_per_step"] / out["learned"]["s_per_step"],
            "speedup_skip_vs_plain": out["plain"]["s_per_step"] / out["skip"]["s_per_step"],
            "speedup_learned_vs_skip": out["skip"]["s_per_step"] / out["learned"]["s_per_step"],
            "training": {"specimens": "D21xH42 (101), D28xH56 (102), D36xH72 (105), "
                                      "D46xH92 (106); validation D28xH56 (103)",
                         "particles": sizes, "samples": len(samples), "epochs": 30,
                         "seconds": t_train, "correction_head": head,
                         "val_loss": min(hist["val"]) if hist["val"] else None,
                         "val_loss_zero_correction": hist["val_zero"]},
            "total_seconds": time.perf_counter() - t0}


def bench_warmstart_damage(pre: int = 18, steps: int = 26) -> dict:
    """The warm start where bonds break: load steps of strain -5e-5 from the
    onset of breaking (after `pre` plain steps) towards the peak.  The
    network is trained on two D21xH42 cylinders and validated on a third over
    the same range, then three copies of a D28xH56 cylinder run the range
    with plain FIRE, with the skip path alone (an untrained network, whose
    zero decoder repeats the last increment), and with the trained network.
    The first two steps of the range are plain in all three and not timed."""
    from rfd import native, native_cases, warmstart as ws
    from rfd.reflayout import Context

    ctx = Context(0)
    ds = -5e-5
    t0 = time.perf_counter()

    def spec(d, h, seed):
        nc = native_cases.uniaxial_cylinder(ctx, d, h, seed=seed, flat_ends=True)
        nc.sys.seat_platens(preload=0.0)
        return nc

    first = spec(21.0, 42.0, 101)
    fp = native.fire_params(first.mat.E, 0.5, first.mat.dt, rel_tol=1e-6, max_iter=60000,
                            check_every=200)

    def loaded(c):
        for _ in range(pre):
            c.sys.fire_load_step(ds, fp)
        return c

    samples = []
    for c in (first, spec(21.0, 42.0, 102)):
        samples += ws.collect_samples(loaded(c).sys, ds, fp, steps)[0]
    val = ws.collect_samples(loaded(spec(21.0, 42.0, 103)).sys, ds, fp, steps)[0]
    net = ws.WarmStartNet(ctx, seed=7)
    hist = ws.train(net, samples, epochs=40, lr=1e-3, val=val)
    skip = ws.WarmStartNet(ctx, seed=7)
    out = {}
    for name, nt in (("plain", None), ("skip", skip), ("learned", net)):
        c = loaded(spec(28.0, 56.0, 104))
        warm = (lambda s_, nt=nt: ws.predict(nt, s_, ds)) if nt is not None else None
        res = ws.collect_samples(c.sys, ds, fp, steps, warm=warm)[1][2:]
        sec = sum(r.seconds for r in res)
        out[name] = {"s_per_step": sec / len(res),
                     "fire_iterations_per_step": sum(r.iterations for r in res) / len(res),
                     "fire_iterations": [int(r.iterations) for r in res],
                     "bonds_broken": sum(r.breaks for r in res),
                     "all_converged": all(bool(r.converged) for r in res),
                     "warm_starts_taken": sum(r.warm_used for r in res)}
        print(f"[bench] damage-regime warm start {name}: {out[name]}", file=sys.stderr,
              flush=True)
    return {"workload": f"uniaxial cylinder D28xH56 mm (seed 104), load steps {pre + 2}-"
                        f"{pre + steps - 1} of strain {ds:g} (onset of breaking to near the "
                        "peak), FIRE to |F|max < 1e-6 E pi r_mean^2",
            "particles": c.sys.N, **out,
            "speedup_learned_vs_plain": out["plain"]["s_per_step"] / out["learned"]["s_per_step"],
            "speedup_learned_vs_skip": out["skip"]["s_per_step"] / out["learned"]["s_per_step"],
            "training": {"specimens": "D21xH42 seeds 101, 102; validation seed 103",
                         "samples": len(samples), "epochs": 40,
                         "correction_head": hist["head"],
                         "val_loss": min(hist["val"]), "val_loss_zero_correction":
                             hist["val_zero"]},
            "total_seconds": time.perf_counter() - t0}


def bench_fire_large(D: float = 100.0, H: float = 200.0, steps: int = 2) -> dict:
    """Plain GPU FIRE load steps on the large cylinder: strain increment
    -1e-5, FIRE to |F|max < 1e-6 F_ref, contact rows rebuilt every 20 FIRE
    iterations.  The first load step (seating) is not timed.  Bond-updates/s
    counts active bonds x FIRE iterations (one force pass each)."""
    from rfd import native, native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    nc = native_cases.uniaxial_cylinder(ctx, D, H, flat_ends=True)
    s = nc.sys
    s.seat_platens(preload=0.0)
    fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-6, max_iter=100000,
                            check_every=200)
    res = []
    for k in range(1 + steps):
        if k == 1:
            s.timing(True)
        r = s.fire_load_step(-1e-5, fp)
        print(f"[bench] fire D{D:g}xH{H:g} step {k}: {r.iterations} iterations, "
              f"{r.seconds:.2f}s, converged {r.converged}", file=sys.stderr, flush=True)
        res.append(r)
    kern_ms, launches = s.timing_read()
    kern_ms /= max(launches, 1)
    timed = res[1:]
    sec = sum(r.seconds for r in timed)
    it = sum(r.iterations for r in timed)
    active = s.bond_counts()[0]
    kernel = s.force_kernel().replace("<fused>", "")  # FIRE passes run the force pass alone
    alg = force_alg_bytes(kernel, s.bonds.n, s.N)
    assert ctx.flags() == 0, "device flags raised in the FIRE run"
    out = {"workload": f"uniaxial cylinder D{D:g}xH{H:g} mm (as the main line), plain FIRE "
                       "load steps of strain -1e-5 to |F|max < 1e-6 E pi r_mean^2",
           "particles": s.N, "bonds": s.bonds.n, "load_steps_per_s": len(timed) / sec,
           "s_per_step": sec / len(timed), "fire_iterations_per_step": it / len(timed),
           "all_converged": all(bool(r.converged) for r in timed),
           "max_final_residual": max(r.fmax for r in timed), "ftol": fp.ftol,
           "bond_updates_per_s": active * it / sec, "ms_per_iteration": sec * 1e3 / it,
           "force_kernel": kernel, "force_ms": kern_ms,
           "force_roofline_frac": alg / (kern_ms * 1e-3) / PEAK_HBM}
    s.close()
    return out


def bench_fire_fp64(D: float = 25.0, H: float = 50.0, steps: int = 4) -> dict:
    """Quasi-static load steps (strain -1e-5, FIRE to |F|max < 1e-6 F_ref) on
    the same seeded cylinder on the FP32 path and on the FP64 GPU path: the
    relaxed platen reactions of each step and their largest relative
    difference."""
    from rfd import native, native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    out = {}
    for name, fp64 in (("fp32", False), ("fp64", True)):
        nc = native_cases.uniaxial_cylinder(ctx, D, H, flat_ends=True, fp64=fp64)
        s = nc.sys
        s.seat_platens(preload=0.0)
        fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-6, max_iter=100000,
                                check_every=200)
        rr = [s.fire_load_step(-1e-5, fp) for _ in range(1 + steps)][1:]
        out[name] = {"reaction_top": [r.sum_top for r in rr],
                     "reaction_bottom": [r.sum_bottom for r in rr],
                     "fire_iterations": [r.iterations for r in rr],
                     "s_per_step": sum(r.seconds for r in rr) / len(rr),
                     "all_converged": all(bool(r.converged) for r in rr),
                     "particles": s.N, "bonds": s.bonds.n}
        print(f"[bench] FIRE {name} D{D:g}xH{H:g}: {out[name]['fire_iterations']} iterations, "
              f"{out[name]['s_per_step']:.3f}s per step", file=sys.stderr, flush=True)
        s.close()
    a = [out["fp32"]["reaction_top"], out["fp32"]["reaction_bottom"]]
    b = [out["fp64"]["reaction_top"], out["fp64"]["reaction_bottom"]]
    scale = max(abs(x) for row in b for x in row)
    diff = max(abs(x - y) for ra, rb in zip(a, b) for x, y in zip(ra, rb)) / scale
    return {"workload": f"uniaxial cylinder D{D:g}xH{H:g} mm, load steps 2-{steps + 1} of "
                        "strain -1e-5, FIRE to |F|max < 1e-6 E pi r_mean^2",
            **out, "max_relative_reaction_difference": diff}


def bench_c2_peak(D: float = 25.0, H: float = 50.0, dstrain: float = -1e-5,
                  max_steps: int = 600, budget_s: float = 150.0) -> dict:
    """Quasi-static uniaxial compression of the D25 x H50 mm cylinder (lattice
    of spacing 1 mm, r ~ U[0.4, 0.6] mm, strengths N(20,4)/N(40,8) MPa) to past
    its peak: load steps of `dstrain`, each FIRE-relaxed to |F|max < 1e-6 F_ref
    with bonds judged at equilibrium, on the FP32 path and on the FP64 GPU path
    from the same seeded specimen.  A run stops once the stress has fallen
    below 90 % of its peak, after max_steps steps or after budget_s seconds.
    Reports both stress-strain curves, peak stress and strain, and the
    relative difference of the peaks."""
    from rfd import native, native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    out = {}
    for name, fp64 in (("fp32", False), ("fp64", True)):
        nc = native_cases.uniaxial_cylinder(ctx, D, H, flat_ends=True, fp64=fp64)
        s = nc.sys
        s.seat_platens(preload=0.0)
        L0 = s.loading()
        area_inv = L0.inv_area
        fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-6, max_iter=100000,
                                check_every=200)
        curve, iters, breaks, conv, unconv = [], 0, 0, True, []
        peak, peak_strain, t0 = 0.0, 0.0, time.perf_counter()
        for k in range(max_steps):
            r = s.fire_load_step(dstrain, fp)
            stress = 0.5 * (abs(r.sum_bottom) + abs(r.sum_top)) * area_inv
            strain = -r.strain
            curve.append((strain, stress * 1e3))
            iters += r.iterations
            breaks += r.breaks
            conv = conv and bool(r.converged)
            if not r.converged:
                unconv.append(strain)
            if stress > peak:
                peak, peak_strain = stress, strain
            if k % 10 == 0:
                print(f"[bench] C2 {name} step {k}: strain {strain:.2e} stress "
                      f"{stress * 1e3:.2f} MPa, {r.iterations} iterations, {r.breaks} breaks",
                      file=sys.stderr, flush=True)
            if stress < 0.9 * peak or time.perf_counter() - t0 > budget_s:
                break
        sec = time.perf_counter() - t0
        out[name] = {"peak_stress_MPa": peak * 1e3, "peak_strain": peak_strain,
                     "steps": len(curve), "past_peak": curve[-1][1] < 0.9 * peak * 1e3,
                     "fire_iterations": iters, "bonds_broken": breaks,
                     "all_converged": conv, "unconverged_steps": len(unconv),
                     "unconverged_before_peak": sum(1 for e in unconv if e <= peak_strain),
                     "first_unconverged_strain": unconv[0] if unconv else None,
                     "seconds": sec,
                     "load_steps_per_s": len(curve) / sec, "particles": s.N,
                     "bonds": s.bonds.n, "curve_strain_stressMPa": curve}
        s.close()
    out["peak_relative_difference"] = abs(out["fp32"]["peak_stress_MPa"] /
                                          out["fp64"]["peak_stress_MPa"] - 1.0)
    out["workload"] = (f"uniaxial cylinder D{D:g}xH{H:g} mm, quasi-static load steps of strain "
                       f"{dstrain:g} to past the peak, FIRE to |F|max < 1e-6 E pi r_mean^2")
    return out


def bench_c1(steps: int = 5000, cpu_seconds: float = 5.0, cpu: bool = True) -> dict:
    """The CPU-runnable Brazilian disc (D20 x T5 mm, random packing at
    phi = 0.6, bonds at gap < 0.05 r_mean, strengths N(20,4)/N(40,8) MPa,
    platens closing at 0.5 m/s): `steps` explicit steps on the FP64 GPU path
    and on the FP32 path; beside them the oracle's step loop (single thread,
    reference slot layout) on the reference's monodisperse disc of the same
    size, for a bounded time."""
    import numpy as np

    from rfd import native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    out = {}
    for name, fp64 in (("gpu_fp64", True), ("gpu_fp32", False)):
        nc = native_cases.packed_brazil_disc(ctx, fp64=fp64)
        s = nc.sys
        s.step(10)
        ctx.sync()
        t0 = time.perf_counter()
        s.step(steps)
        ctx.sync()
        sec = time.perf_counter() - t0
        h = s.history()
        active = s.bond_counts()[0]
        out[name] = {"particles": s.N, "bonds": s.bonds.n, "steps": steps,
                     "ms_per_step": sec * 1e3 / steps,
                     "bond_updates_per_s": active * steps / sec,
                     "bonds_active_end": active,
                     "peak_stress_MPa": float(np.abs(h[:, 1:]).max(1).max()) * 1e3}
        assert ctx.flags() == 0, "device flags raised in the disc run"
        s.close()
    if cpu:
        from oracle.oracle import OracleUniaxialDriver
        from rfd import cases

        c = cases.reference_brazil(20.0, 5.0)
        d = OracleUniaxialDriver(c.R, c.mat, c.box, c.cmp, c.grid(), c.K)
        d.setup()
        d.step()
        nb = int((d.s.ILtype == 1).sum()) // 2
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds and n < steps:
            d.step()
            n += 1
        sec = time.perf_counter() - t0
        out["cpu_port"] = {"particles": c.N, "bonds": nb, "steps": n,
                           "ms_per_step": sec * 1e3 / n, "bond_updates_per_s": nb * n / sec,
                           "cores": 1, "kind": "port",
                           "sample": "oracle step loop on the reference's monodisperse "
                                     "D20xT5 mm disc, single thread"}
    return {"workload": "Brazilian disc D20xT5 mm, random packing phi=0.6, r~U[0.4,0.6] mm, "
                        f"{steps} explicit steps", **out}


def bench_peak_precision(D: float = 10.0, H: float = 20.0, max_steps: int = 20000) -> dict:
    """Peak strength of one seeded uniaxial cylinder (explicit loading at 10x
    the reference platen speed, capped ends) on the FP32 path and on the FP64
    GPU path, and their relative difference."""
    from rfd import native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    out = {}
    for name, fp64 in (("fp32", False), ("fp64", True)):
        nc = native_cases.uniaxial_cylinder(ctx, D, H, seed=31, fp64=fp64, flat_ends=True,
                                            v_load=-100.0 * 0.02 / 1e3)
        s = nc.sys
        s.seat_platens(preload=0.0)
        done, t0 = 0, time.perf_counter()
        while done < max_steps:
            s.step(500)
            done += 500
            if s.loading().fracture:
                break
        L = s.loading()
        out[name] = {"peak_stress_MPa": L.fracture_stress * 1e3, "peak_strain": L.fracture_strain,
                     "steps": done, "fractured": bool(L.fracture),
                     "seconds": time.perf_counter() - t0, "particles": s.N}
        print(f"[bench] peak {name}: {out[name]}", file=sys.stderr, flush=True)
        del s, nc
    out["relative_difference"] = abs(out["fp32"]["peak_stress_MPa"] / out["fp64"]["peak_stress_MPa"]
                                     - 1.0)
    out["workload"] = f"uniaxial cylinder D{D:g}xH{H:g} mm, explicit steps to fracture"
    return out


def bench_packed(D: float = 100.0, H: float = 200.0, steps: int = 50, warmup: int = 5) -> dict:
    """The uniaxial cylinder on a seeded random packing at solid fraction 0.6
    (rfd_system_pack) instead of the lattice: bonds join pairs with gap below
    0.05 r_mean.  Same step and roofline accounting as the main line."""
    from rfd import native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    pk = native_cases.packed_cylinder(ctx, D, H, 0.6)
    pack_s = time.perf_counter() - t0
    nc = native_cases.packed_uniaxial_cylinder(ctx, D, H, packing=pk)
    s = nc.sys
    N, Nb = s.N, s.bonds.n
    print(f"[bench] packed: N={N} Nb={Nb} phi={pk.phi:.4f} packed in {pack_s:.1f}s "
          f"({pk.result['iterations']} iterations)", file=sys.stderr, flush=True)
    s.step(warmup)
    ctx.sync()
    s.timing(True)
    t1 = time.perf_counter()
    s.step(steps)
    ctx.sync()
    t2 = time.perf_counter()
    kern_ms, launches = s.timing_read()
    kern_ms /= max(launches, 1)
    active, broken, removed = s.bond_counts()
    kernel = s.force_kernel()
    contacts = s.contact_entries()
    alg = force_alg_bytes(kernel, Nb, N)
    alg_c = force_alg_bytes(kernel, Nb, N, contacts)
    out = {
        "workload": f"uniaxial cylinder D{D:g}xH{H:g} mm, seeded random packing phi=0.6, "
                    "r~U[0.4,0.6] mm, bonds at gap < 0.05 r_mean",
        "particles": N, "bonds": Nb, "bonds_per_particle": Nb / N, "phi": pk.phi,
        "pack": {"seconds": pack_s, **pk.result},
        "bond_updates_per_s": active * steps / (t2 - t1),
        "ms_per_step": (t2 - t1) * 1e3 / steps,
        "force_kernel": kernel, "force_ms": kern_ms,
        "force_roofline_frac": alg / (kern_ms * 1e-3) / PEAK_HBM,
        "contact_entries_per_particle": contacts / N,
        "force_roofline_frac_with_contacts": alg_c / (kern_ms * 1e-3) / PEAK_HBM,
    }
    assert ctx.flags() == 0, "device flags raised in the packed run"
    s.close()
    return out


def bench_calibrate(steps: int = 2) -> dict:
    """The calibration outer loop on the reference's test geometry (uniaxial
    D25 x H50, Brazilian D50 x T20, packed at solid fraction 0.61): CG steps
    from the granite start, each evaluating 6 parameter sets x 2 tests in
    batches; then the same 12 runs of one step one at a time."""
    import numpy as np

    from rfd import calibrate as cal
    from rfd.params import STRESS
    from rfd.reflayout import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    ev = cal.Evaluator(ctx)
    setup = time.perf_counter() - t0
    st = cal.CGState(params=cal.GRANITE_START.copy(), goals=cal.GRANITE_GOALS.copy())
    times = []
    for k in range(steps):
        t = time.perf_counter()
        R = ev.evaluate(st.parameter_sets())
        times.append(time.perf_counter() - t)
        st.update(R)
        print(f"[bench] CG step {k}: {times[-1]:.1f}s, functional "
              f"{st.history[-1]['functional']:.4g}", file=sys.stderr, flush=True)
    sets = st.parameter_sets()
    dt = min(cal._material(p).dt for p in sets)
    t = time.perf_counter()
    for p in sets:
        ev.run(p[None, :], "uniaxial", dt)
        ev.run(p[None, :], "brazil", dt)
    seq = time.perf_counter() - t
    return {
        "workload": "granite CG calibration, uniaxial D25xH50 and Brazilian D50xT20 mm packed "
                    "at phi 0.61 (monodisperse r 0.5 mm), 6 parameter sets x 2 tests per step",
        "particles": {"uniaxial": int(len(ev.uni.radius)), "brazil": int(len(ev.bra.radius))},
        "setup_seconds": setup,
        "seconds_per_cg_step_batched": float(np.mean(times)),
        "seconds_12_runs_one_at_a_time": seq,
        "steps": [{"functional": h["functional"],
                   "uniaxial_peak_strain": float(h["results"][0]),
                   "uniaxial_peak_MPa": float(h["results"][1] * STRESS / 1e6),
                   "brazil_peak_MPa": float(h["results"][2] * STRESS / 1e6),
                   "params": [float(x) for x in h["params"]]} for h in st.history],
        "goals": {"uniaxial_peak_strain": float(cal.GRANITE_GOALS[0]),
                  "uniaxial_peak_MPa": float(cal.GRANITE_GOALS[1] * STRESS / 1e6),
                  "brazil_peak_MPa": float(cal.GRANITE_GOALS[2] * STRESS / 1e6)},
    }


def _quantiles(x) -> dict:
    import numpy as np

    q = np.percentile(x, [0, 10, 50, 90, 100])
    return {"min": q[0], "p10": q[1], "median": q[2], "p90": q[3], "max": q[4],
            "mean": float(np.mean(x)), "std": float(np.std(x))}


def bench_batch(S: int = 256, max_steps: int = 48000, chunk: int = 400) -> dict:
    """Batched Brazilian discs: S specimens (particle counts log-uniform in
    [1k, 40k], own seeds, own strength means) in one system, launch shapes
    from the autotuner, stepped with plain launches until every disc has
    fractured or passed its peak (stress below half of it) or max_steps.
    Reports the step rate of the batch and the per-specimen peak loads (a
    disc that has not passed its peak when the run stops is listed in
    `unfinished`, and its entry in `peaks_MPa` is a lower bound; the
    quantiles cover the finished discs); a 2k-particle disc alone measures
    what graph replays and multi-step launches save on small work."""
    import numpy as np

    from rfd import native_cases
    from rfd.reflayout import Context

    ctx = Context(0)
    t0 = time.perf_counter()
    case = native_cases.brazil_batch(ctx, S=S)
    s = case.sys
    setup_s = time.perf_counter() - t0
    tune = s.autotune(use_cache=True).as_dict()
    print(f"[bench] batch: {S} discs, N={s.N}, bonds={s.bonds.n}, setup {setup_s:.1f}s, "
          f"tiles {tune['force_tile']}/{tune['nbr_block']}", file=sys.stderr, flush=True)

    def sync():
        return s.specimen_loadings()

    sync()
    # plain launches and graph replays over the first steps of the run
    t = time.perf_counter()
    s.step(chunk)
    sync()
    ms_plain = (time.perf_counter() - t) / chunk * 1e3
    t = time.perf_counter()
    s.step_graph(chunk)
    Ls = sync()
    ms_graph = (time.perf_counter() - t) / chunk * 1e3
    done, t_run = 2 * chunk, time.perf_counter()
    while done < max_steps:
        s.step(chunk)  # plain launches: graph replays measured no faster above
        done += chunk
        Ls = sync()
        st = np.array([L.stress for L in Ls])
        pk = np.array([L.fracture_stress for L in Ls])
        fin = np.array([L.fracture for L in Ls], bool) | (st < 0.5 * pk)
        if done % (10 * chunk) == 0:
            print(f"[bench] batch step {done}: {fin.sum()} of {S} past peak",
                  file=sys.stderr, flush=True)
        if fin.all():
            break
    run_s = time.perf_counter() - t_run
    Ls = s.specimen_loadings()
    pk = np.array([L.fracture_stress for L in Ls])
    st = np.array([L.stress for L in Ls])
    fr = np.array([L.fracture for L in Ls], bool)
    past = fr | (st < 0.5 * pk)
    peak_mpa = pk * 1e3
    load_kn = pk / np.array([L.inv_area for L in Ls])  # force unit: 1 kN
    n = np.diff(case.first)
    counts = s.bond_counts()
    steps_timed = done - 2 * chunk
    ms_run = run_s / max(steps_timed, 1) * 1e3
    out = {
        "workload": f"{S} Brazilian discs, N log-uniform in [1k, 40k] "
                    f"(total {s.N}), r~U[0.4,0.6] mm, bond strength N(mu,0.2mu), "
                    "mu~U[10,30] MPa per disc, platens 0.5 m/s, explicit FP32 steps",
        "specimens": S,
        "particles": int(s.N),
        "bonds": int(s.bonds.n),
        "particles_per_specimen": _quantiles(n),
        "autotune": tune,
        "setup_seconds": setup_s,
        "ms_per_step_plain": ms_plain,
        "ms_per_step_graph": ms_graph,
        "ms_per_step_run": ms_run,
        "bond_updates_per_s": s.bonds.n / (ms_run * 1e-3),
        "specimen_steps_per_s": S / (ms_run * 1e-3),
        "steps": done,
        "run_seconds": run_s,
        "past_peak": int(past.sum()),
        "fractured": int(fr.sum()),
        "unfinished": [int(k) for k in np.flatnonzero(~past)],
        "bonds_intact_broken_removed": list(counts),
        "peak_stress_MPa": _quantiles(peak_mpa[past]),
        "peak_load_kN": _quantiles(load_kn[past]),
        "corr_peak_vs_mu": float(np.corrcoef(peak_mpa[past], case.mu_t[past])[0, 1]),
        "corr_peak_vs_log_n": float(np.corrcoef(peak_mpa[past], np.log(n[past]))[0, 1]),
        "peaks_MPa": [round(float(x), 3) for x in peak_mpa],
    }
    s.close()
    # a single small disc: launch latency bounds its step
    one = native_cases.brazil_batch(ctx, S=1, n_lo=2000, n_hi=2000, seed=7).sys
    one.step(40)
    one.specimen_loadings()
    small = {}
    for name, fn in (("plain", one.step), ("graph", one.step_graph),
                     ("multi_step_launch", one.step)):
        one.set_persistent(name == "multi_step_launch")
        fn(400)
        one.specimen_loadings()
        t = time.perf_counter()
        fn(2000)
        one.specimen_loadings()
        small[name] = (time.perf_counter() - t) / 2000 * 1e6
    out["small_disc_us_per_step"] = {"particles": int(one.N), **small}
    one.close()
    print(f"[bench] batch: {out['steps']} steps, {out['past_peak']} past peak, "
          f"{ms_run:.3f} ms/step; small disc {small}", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the path runs on one GPU; how it would shard is described, not run",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        print(json.dumps({"copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES,
                          **stream_copy()}))
        return
    if args.warmstart_only:
        print(json.dumps({"warm_start": bench_warmstart(),
                          "warm_start_damage": bench_warmstart_damage()}))
        return
    if args.batch_only:
        print(json.dumps({"batch": bench_batch()}))
        return
    if args.c2_peak_only:
        print(json.dumps({"c2_peak": bench_c2_peak()}))
        return
    if args.packed_only:
        print(json.dumps({"packed": bench_packed()}))
        return
    if args.disc_only:
        print(json.dumps({"disc_c1": bench_c1(cpu=not args.no_cpu)}))
        return
    if args.fire_only:
        print(json.dumps({"fire_fp64": bench_fire_fp64(),
                          "fire_large": bench_fire_large(args.D, args.H)}))
        return
    if args.calibrate:
        print(json.dumps({"calibrate": bench_calibrate()}))
        return
    res = bench_ref(args) if args.tier == "ref" else bench_native(args)
    res["roofline"]["traffic"] = load_traffic(res["roofline"]["kernel"],
                                              res["config"]["particles"],
                                              res["config"]["bonds"])
    res["roofline"].update(stream_copy())
    res["roofline"]["clock"] = ("kernel_ms: mean of HIP-event pairs recorded around each "
                                "force launch on its stream in this run")
    out = {
        "metric": "bond-updates/s",
        "value": res["value"],
        "particle_updates_per_s": res["config"]["particles"] * 1e3 / res["ms_per_step"],
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": res["dtype"],
        "data": "synthetic",
        "config": res["config"],
        "roofline": res["roofline"],
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ref(args.cpu_seconds)
        out["cpu_baseline"]["host_cores_available"] = os.cpu_count()
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        out["cpu_baseline_threads"] = cpu_baseline_ref(args.cpu_seconds, omp=True)
        try:
            out["cylinder_c2_gpu"] = bench_c2()
        except Exception as e:  # reported, never hides the main line
            out["cylinder_c2_gpu"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_warmstart and args.tier == "native":
        try:
            out["peak_strength"] = bench_peak_precision()
        except Exception as e:  # reported, never hides the main line
            out["peak_strength"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["disc_c1"] = bench_c1(cpu=not args.no_cpu)
        except Exception as e:  # reported, never hides the main line
            out["disc_c1"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["fire_fp64"] = bench_fire_fp64()
        except Exception as e:  # reported, never hides the main line
            out["fire_fp64"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            c2 = bench_c2_peak()
            for k in ("fp32", "fp64"):  # every 4th point of the curves in the line
                c2[k]["curve_strain_stressMPa"] = c2[k]["curve_strain_stressMPa"][::4]
            out["c2_peak"] = c2
        except Exception as e:  # reported, never hides the main line
            out["c2_peak"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["warm_start"] = bench_warmstart()
        except Exception as e:  # reported, never hides the main line
            out["warm_start"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["warm_start_damage"] = bench_warmstart_damage()
        except Exception as e:  # reported, never hides the main line
            out["warm_start_damage"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["fire_large"] = bench_fire_large(args.D, args.H)
        except Exception as e:  # reported, never hides the main line
            out["fire_large"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_batch and args.tier == "native":
        try:
            out["batch"] = bench_batch()
        except Exception as e:  # reported, never hides the main line
            out["batch"] = {"error": f"{type(e).__name__}: {e}"}
        try:
            out["packed"] = bench_packed()
        except Exception as e:  # reported, never hides the main line
            out["packed"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
