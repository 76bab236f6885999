#!/usr/bin/env python3
"""Throughput of the PIMC sampling path on one MI355X.

One step = one sweep of every chain of the batched trapped ideal-gas workload
(4096 chains, N ~ U[4,64], M ~ U{32,64,128,256}, beta ~ U[0.5,4], omega = 1,
FP32 paths and action differences, FP64 estimators): per particle, staging
segments of length 8 tiling the ring at a random offset, one centre-of-mass
try, then the energy and radial-density estimators.  Inputs are resident in
HBM before the timed region.  The metric counts proposed beads: m-1 per
staging segment of m links, M per centre-of-mass try.

Printed (one JSON line): the batched value, the plain single-chain GPU form
(N=10, M=64, beta=1, FP64), the energy check against the closed-form
primitive-action energies, the roofline of the sweep kernel and the CPU
baseline (the oracle's FP64 restatement of the reference schedule on host
cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "pimc.jl_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_MOVE_FP32 = 16  # read + write of one 2D FP32 bead
BYTES_PER_MOVE_FP64 = 32
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4 cycles
# (16 lanes per cycle, packed FP32 off) at the 2.4 GHz peak engine clock
VALU_PEAK_LANE_OPS = 1024 * 16 * 2.4e9
MIN_TIMED_LAUNCHES = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--chains", type=int, default=4096)
    ap.add_argument("--spl", type=int, default=50, help="sweeps per launch (one estimator block)")
    ap.add_argument("--plain-steps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-beads", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep staging length 8 and the default tile")
    ap.add_argument("--ell", type=int, default=0,
                    help="staging length of every chain (capped at M); with --tile-beads and --threads this "
                         "fixes the configuration the autotuner would otherwise pick")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a hipGraph")
    ap.add_argument("--measure-kernel", type=int, default=0, choices=[0, 1, 2],
                    help="estimators of whole chains: 0 automatic, 1 in the sweep kernel, 2 a wide kernel")
    ap.add_argument("--pair-kernel", type=int, default=0, choices=[0, 1, 2],
                    help="Gaussian-pair chains: 0 automatic, 1 the whole-chain kernel, 2 the window kernels")
    ap.add_argument("--workload", default="trapped_batch",
                    choices=["trapped_batch", "trapped_plain", "gauss_box", "trapped_gauss", "lattice", "contact_box"],
                    help="workload of the timed sweeps (the default is the headline batched form)")
    ap.add_argument("--check-sweeps", type=int, default=2000,
                    help="untimed measuring sweeps after the timed ones (at least 8 blocks), for the energy check")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rounded_steps(requested: int, spl: int, least: int) -> int:
    """Sweeps actually run: whole launches of spl sweeps, at least `least` launches."""
    return max(least, -(-max(0, requested) // spl)) * spl


def roofline_from_profile(name: str, moves_per_launch: float, avg_launch_s: float, bytes_per_move: int):
    """The roofline of the dominant kernel from its committed counter summary
    (profiles/traffic_<name>.json): VALU issue, the bound of these LDS-resident
    sweeps, with the instruction count per bead-move from SQ_INSTS_VALU and the
    launch time measured live; HBM traffic from FETCH_SIZE / WRITE_SIZE, both
    scaled to this launch's bead-moves.  The algorithmic-bytes model (16 B per
    bead-move in FP32) is kept as an informational field: the beads stay in LDS
    for the sweeps of a launch, so those bytes never reach HBM."""
    hbm_alg = bytes_per_move * moves_per_launch / avg_launch_s / 1e9
    info = {"achieved": hbm_alg, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_alg / HBM_PEAK_GBS,
            "note": f"{bytes_per_move} B per bead-move (read + write of its coordinates); LDS-resident, not "
                    "HBM traffic"}
    tp = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    try:
        with open(tp) as f:
            tj = json.load(f)
        per_move = tj["valu_wave_instructions_per_launch"] * 64.0 / tj["bead_moves_per_launch"]
        scale = moves_per_launch / tj["bead_moves_per_launch"]
        achieved = per_move * moves_per_launch / avg_launch_s
        return {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK_LANE_OPS / 1e12,
                "unit": "T VALU lane-instructions/s", "frac": achieved / VALU_PEAK_LANE_OPS,
                "traffic": tj["hbm_bytes_per_launch"] * scale,
                "kernel": tj.get("kernel"),
                "valu_lane_instructions_per_bead_move": per_move,
                "counters_from": os.path.relpath(tp, ROOT),
                "note": "achieved = SQ_INSTS_VALU x 64 per bead-move (counter run of the same configuration) x "
                        "bead-moves per launch / live launch time; peak = 1024 SIMDs x 16 lanes x 2.4 GHz; "
                        "traffic = scaled FETCH_SIZE + WRITE_SIZE bytes per launch",
                "hbm_algorithmic": info}
    except (OSError, ValueError, KeyError, ZeroDivisionError):
        return {"bound": "valu", "achieved": None, "peak": VALU_PEAK_LANE_OPS / 1e12,
                "unit": "T VALU lane-instructions/s", "frac": None, "traffic": None,
                "note": f"no counter summary {os.path.relpath(tp, ROOT)} for this workload",
                "hbm_algorithmic": info}


def energy_check(batch, wl):
    from analytic import energy_boltzmann_primitive
    z = []
    for c, d in enumerate(wl.descs):
        e = batch.estimates(c)
        exact = energy_boltzmann_primitive(d.N, d.M, d.beta)
        if e.E_err > 0:
            z.append((e.E - exact) / e.E_err)
    z = np.array(z)
    return {
        "chains": int(len(z)),
        "frac_within_2sigma": float(np.mean(np.abs(z) < 2.0)) if len(z) else None,
        "chi2_per_chain": float(np.mean(z ** 2)) if len(z) else None,
        "mean_pull": float(np.mean(z)) if len(z) else None,
        "reference": "closed-form primitive-action energy per chain (normal modes of the discrete ring)",
    }


def cpu_baseline(wl, seconds: float):
    """The oracle's FP64 restatement of run!(s, n, [(1, SingleCenterOfMass(s, 1.0)),
    (1, ReshapeLinear(s, 8))]) on one chain per host core, chains taken from
    the same batched workload.  ctypes releases the GIL, so threads run in parallel."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib()
    cores = max(1, min(16, os.cpu_count() or 1))
    systems = []
    for c in range(cores):
        d = wl.descs[c % len(wl.descs)]
        s = O.System(wl.init[c % len(wl.init)], d.lam, d.L, d.beta, O.make_pot(O.POT_HARMONIC, d.omega))
        systems.append((s, O.RefRunner(s, cap=8, step=1.0, seed=1000 + c)))
    # each thread runs chunks of attempts until the shared deadline
    for s, r in systems:
        r.run(2000)  # thermalise the adaptive cap and step a little
        r.st.bead_moves = 0
    attempts = [0] * len(systems)
    t0 = time.perf_counter()
    deadline = t0 + seconds

    def worker(i, r):
        while time.perf_counter() < deadline:
            r.run(2000)
            attempts[i] += 2000

    th = [threading.Thread(target=worker, args=(i, r)) for i, (_, r) in enumerate(systems)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    dt = time.perf_counter() - t0
    moves = sum(r.st.bead_moves for _, r in systems)
    iters = int(np.mean(attempts))
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": moves / dt,
        "unit": "bead-moves/s",
        "cores": cores,
        "kind": "port",
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model,
        "sample": (f"{cores} chains of the batched workload (one per thread), {iters} attempts each of the "
                   f"reference schedule (COM + ReshapeLinear cap 8, adaptive), FP64 C restatement, {dt:.1f} s; "
                   f"CPU: {cpu_model}"),
    }


WORKLOAD_TEXT = {
    "trapped_plain": "one trapped 2D ideal-boson chain, N=10, M=64, beta=1, omega=1, FP64",
    "gauss_box": "256 chains of N=256 bosons in a periodic 16x16 box, Gaussian pair g=1 sigma=0.5 (cut-off 4 sigma, "
                 "cell lists), M=128, beta=2, FP32 pair arithmetic, FP64 energy sums",
    "trapped_gauss": "512 chains of N=64 trapped 2D bosons with Gaussian pair g=0.5 sigma=0.5, M=256, beta=2, "
                     "FP32, reweighting sums for d<E>/d omega and d<E>/d g",
    "lattice": "64 chains of N=400 bosons on a 20x20 periodic square lattice V0=4 a=1, Gaussian pair g=1 "
               "sigma=0.3, M=512, beta=4, FP32, FP64 256x256 density histogram and energies",
    "contact_box": "256 chains of N=256 bosons in a periodic 16x16 box with the tabulated pair-product propagator "
                   "of a contact coupling g0=2 (600x600 table built on the GPU) and hard spheres a=exp(-2pi/g0), "
                   "M=128, beta=2, FP32 paths, FP64 lnU and energy sums",
}


def other_workload(args, A):
    """Throughput of one of the other workloads (no CPU leg)."""
    from pimc import batch as B
    build = {"trapped_plain": B.trapped_plain, "gauss_box": B.gauss_box, "trapped_gauss": B.trapped_gauss,
             "lattice": B.lattice, "contact_box": B.contact_box}[args.workload]
    wl = build()
    spl = args.spl
    bt = Batch = B.Batch
    t = time.perf_counter()
    b = Batch(wl, sweeps_per_launch=spl, seed=99, tile_beads=args.tile_beads, block_threads=args.threads,
              measure_kernel=args.measure_kernel, pair_kernel=args.pair_kernel)
    log(f"{args.workload}: {len(wl.descs)} chains, setup {time.perf_counter() - t:.1f} s")
    warm = rounded_steps(args.warmup, spl, 1)
    steps = rounded_steps(args.steps, spl, MIN_TIMED_LAUNCHES)
    cw = b.run(warm, measure=False)
    t0 = time.perf_counter()
    cnt = b.run(steps, measure=True)
    wall = time.perf_counter() - t0
    value = cnt.bead_moves / wall
    launches = max(1, cnt.kernel_launches)
    bpm = BYTES_PER_MOVE_FP64 if wl.precision == A.FP64 else BYTES_PER_MOVE_FP32
    roof = roofline_from_profile(args.workload, cnt.bead_moves / launches, cnt.kernel_ms / 1000.0 / launches, bpm)
    est = [b.estimates(c) for c in range(len(wl.descs))]
    E = np.array([e.E for e in est])
    Ev = np.array([e.Ev for e in est])
    out = {
        "metric": "bead-moves/sec at 1 GPU, batched vs plain, E within 2σ of FP64 ref; % HBM peak",
        "value": value, "unit": "bead-moves/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "steps_requested": args.steps, "warmup_requested": args.warmup, "launches": cnt.kernel_launches,
        "ms_per_step": 1000.0 * wall / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if wl.precision == A.FP64 else "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOAD_TEXT[args.workload]}", "chains": len(wl.descs),
                   "sweeps_per_launch": spl, "bead_moves_per_sweep": int(cnt.bead_moves // steps),
                   "pair_kernel": args.pair_kernel},
        "roofline": roof,
        "energy": {"E_mean_over_chains": float(E.mean()), "E_sem": float(E.std(ddof=1) / math.sqrt(len(E)))
                   if len(E) > 1 else float(est[0].E_err),
                   "Ev_mean_over_chains": float(Ev.mean()),
                   "Epair_mean_over_chains": float(np.mean([e.Epair for e in est]))},
        "acceptance": {"levy": cnt.levy_acc / max(1, cnt.levy_tries), "com": cnt.com_acc / max(1, cnt.com_tries)},
        "kernel_avg_launch_ms": cnt.kernel_ms / launches,
        "kernel_ms_per_sweep": {"thermalising": cw.kernel_ms / max(1, warm), "measuring": cnt.kernel_ms / steps},
    }
    if wl.derivatives:
        out["derivatives"] = {"dE_domega_mean": float(np.mean([e.dE_domega for e in est])),
                              "dE_dg_mean": float(np.mean([e.dE_dg for e in est]))}
    b.close()
    del bt
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: chains are independent and would shard across "
                                   "GPUs with one final FP64 reduction of the estimator sums, which is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from pimc import _capi as A
    from pimc.batch import Batch, trapped_batch, trapped_plain

    name, cus = A.device_info()
    log(f"device {name}, {cus} CUs")
    if args.calib_copy_mib > 0:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = ctypes.c_size_t(args.calib_copy_mib << 20)
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        ok = hip.hipMalloc(ctypes.byref(src), nbytes) == 0 and hip.hipMalloc(ctypes.byref(dst), nbytes) == 0
        ok = ok and hip.hipMemset(src, 1, nbytes) == 0 and hip.hipDeviceSynchronize() == 0
        ok = ok and hip.hipMemcpy(dst, src, nbytes, 3) == 0 and hip.hipDeviceSynchronize() == 0  # device to device
        hip.hipFree(src)
        hip.hipFree(dst)
        log(f"calibration copy of {args.calib_copy_mib} MiB {'done' if ok else 'FAILED'}")

    if args.workload != "trapped_batch":
        print(json.dumps(other_workload(args, A)), flush=True)
        return

    # ---- batched workload
    spl = args.spl
    wl = trapped_batch(n_chains=args.chains, seed=2, precision=A.FP32)
    if args.ell > 0:
        for d in wl.descs:
            d.ell = min(args.ell, d.M)
    t = time.perf_counter()
    tune = not args.no_autotune and args.tile_beads == 0 and args.threads == 0 and args.ell == 0
    batch = Batch(wl, sweeps_per_launch=spl, seed=20240601, tile_beads=args.tile_beads, block_threads=args.threads,
                  autotune=tune, use_graph=not args.no_graph)
    log(f"batched context: {len(wl.descs)} chains, {sum(d.N * d.M for d in wl.descs)} beads, "
        f"setup {time.perf_counter() - t:.1f} s")
    if batch.tune:
        ells, counts = np.unique(batch.tune.ell, return_counts=True)
        log(f"autotune ({batch.tune.seconds:.1f} s): tile_beads {batch.tune.tile_beads}, threads "
            f"{batch.tune.block_threads}, staging lengths {dict(zip(ells.tolist(), counts.tolist()))}")
    warm = rounded_steps(args.warmup, spl, 1)
    steps = rounded_steps(args.steps, spl, MIN_TIMED_LAUNCHES)
    batch.run(warm, measure=False)
    # two untimed measuring launches: the second captures the measuring launch step into a
    # hipGraph, so every timed launch is a graph replay
    batch.run(2 * spl, measure=True)
    log(f"warmup {warm} sweeps and 2 measuring launches done")
    t0 = time.perf_counter()
    cnt = batch.run(steps, measure=True)
    wall = time.perf_counter() - t0
    log(f"timed {steps} sweeps in {cnt.kernel_launches} launches: wall {wall:.3f} s, device {cnt.wall_ms:.1f} ms, "
        f"kernels {cnt.kernel_ms:.1f} ms")
    value = cnt.bead_moves / wall
    ms_per_step = 1000.0 * wall / steps
    launches = max(1, cnt.kernel_launches)
    avg_launch_s = cnt.kernel_ms / 1000.0 / launches
    roof = roofline_from_profile("trapped_batch", cnt.bead_moves / launches, avg_launch_s, BYTES_PER_MOVE_FP32)
    # untimed measuring sweeps, so that every chain has enough blocks for its error bar
    check = rounded_steps(args.check_sweeps, spl, 8)
    batch.run(check, measure=True)
    echeck = energy_check(batch, wl)
    echeck["blocks_per_chain"] = (2 * spl + steps + check) // spl
    acc = {"levy": cnt.levy_acc / max(1, cnt.levy_tries), "com": cnt.com_acc / max(1, cnt.com_tries)}
    batch.close()

    # ---- plain single-chain form
    pw = trapped_plain(seed=1)
    plain = Batch(pw, sweeps_per_launch=spl, seed=7, use_graph=not args.no_graph)
    plain.run(max(spl, 500 // spl * spl), measure=False)
    psteps = max(spl, (args.plain_steps // spl) * spl)
    t0 = time.perf_counter()
    pc = plain.run(psteps, measure=True)
    pwall = time.perf_counter() - t0
    plain_value = pc.bead_moves / pwall
    pe = plain.estimates(0)
    from analytic import energy_boltzmann_primitive
    pex = energy_boltzmann_primitive(10, 64, 1.0)
    plain.close()
    log(f"plain: {plain_value:.3e} bead-moves/s, E = {pe.E:.4f} +- {pe.E_err:.4f} (exact {pex:.4f})")

    cpu = None if args.no_cpu else cpu_baseline(wl, args.cpu_seconds)
    if cpu:
        log(f"cpu baseline {cpu['value']:.3e} bead-moves/s on {cpu['cores']} cores")

    out = {
        "metric": "bead-moves/sec at 1 GPU, batched vs plain, E within 2σ of FP64 ref; % HBM peak",
        "value": value,
        "unit": "bead-moves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "launches": cnt.kernel_launches,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "trapped_batch: 4096 independent trapped 2D ideal-boson chains, N~U[4,64], "
                        "M~U{32,64,128,256}, beta~U[0.5,4], omega=1, FP32 paths, FP64 estimators",
            "chains": len(wl.descs),
            "beads": int(sum(d.N * d.M for d in wl.descs)),
            "staging_length": ({str(k): int(v) for k, v in zip(*np.unique(batch.tune.ell, return_counts=True))}
                               if batch.tune else (args.ell or 8)),
            "tile_beads": batch.tune.tile_beads if batch.tune else (args.tile_beads or 4096),
            "block_threads": batch.tune.block_threads if batch.tune else (args.threads or 256),
            "hipgraph": not args.no_graph,
            "timed_launches_replayed": (not args.no_graph),
            "sweeps_per_launch": spl,
            "bead_moves_per_sweep": int(cnt.bead_moves // steps),
            "parallelism": "single GPU, one workgroup per tile of particles",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_gpu": {
            "workload": "trapped_plain: one chain N=10, M=64, beta=1, omega=1, FP64",
            "value": plain_value,
            "unit": "bead-moves/s",
            "E": pe.E, "E_err": pe.E_err, "E_exact_primitive": pex,
        },
        "batched_vs_plain": value / plain_value,
        "batched_vs_cpu": (value / cpu["value"]) if cpu else None,
        "energy_check": echeck,
        "acceptance": acc,
        "kernel_avg_launch_ms": avg_launch_s * 1000.0,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
