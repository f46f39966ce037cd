"""Benchmark: atom-steps/s of the energy+force path on one MI355X.

Workload: the 12,288-atom water box (4,096 H2O on a seeded 16^3 lattice,
49.68 A cubic cell, cosine cutoff 6.35 A, 50 SFs per element, 50-25-25-1 tanh
NN, random-init weights).  One step = one full evaluation (neighbour list,
symmetry functions, NNs, energy, forces) of the resident structure.

Prints one JSON line.  value = atom-steps/s of the mixed-precision form (the
BASELINE metric); the plain FP64 form is timed beside it, forces of both are
compared with the CPU FP64 path (force RMS error), and that CPU path -- n2p2's
own compute classes where they were built, else the from-scratch oracle -- is
timed on a bounded sample on the host cores as the cpu_baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--form mixed|plain]
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (E+F) mixed vs FP64 GPU form at force RMS err <=1e-3 eV/A"
PEAK_FP64 = 78.6    # TFLOP/s, MI355X FP64 vector
PEAK_FP32 = 157.3   # TFLOP/s, MI355X FP32 vector
PEAK_HBM = 8000.0   # GB/s


def flop_counts(w, nn_macs_per_atom, natoms):
    """Algorithmic FLOPs per evaluation: 13P + 15P_m + 60T + 41T_m + 3T_e
    (symmetry functions and their derivatives) + 6C per atom (NN forward and
    adjoint)."""
    sf = 13 * w["pairs"] + 15 * w["pair_members"] + 60 * w["triplets"] + 41 * w["triplet_members"] \
        + 3 * w["triplet_exps"]
    nn = 6 * nn_macs_per_atom * natoms
    return sf, nn


def time_form(m, form, steps, warmup):
    import n2p2_amd
    for _ in range(warmup):
        m.evaluate(form, forces=True)
    m.synchronize()
    m.set_option(n2p2_amd.OPT_TIMING, 1)
    m.reset_kernel_times()
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.evaluate(form, forces=True)
    m.synchronize()
    dt = time.perf_counter() - t0
    times = {}
    for k in ("nbr", "rev", "sf64", "nn64", "energy", "force64", "sf32", "nnbf16", "force32"):
        ms, n = m.kernel_time(k)
        if n:
            times[k] = (ms, n)
    m.set_option(n2p2_amd.OPT_TIMING, 0)
    return dt, times


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` as recorded by the FETCH_SIZE /
    WRITE_SIZE counter runs (scaled by a copy-kernel calibration) in
    profiles/hbm_traffic_water12288.txt; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_water12288.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and f[0] == kernel and f[1] == "traffic_bytes_per_launch":
            return float(f[2])
    return None


def calibrate_copy(mib):
    """A plain device copy of known size (torch's copy kernel), five
    launches: the FETCH_SIZE / WRITE_SIZE counters of this run give the
    scale factor between counter readings and bytes actually moved."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "copy_GBps": 2 * n * 4 / dt / 1e9}))


def cpu_baseline(model_dir, pos, elem, box, budget_s=15.0):
    """n2p2's compute path on the host cores: reference classes (SF groups,
    NN, force contraction compiled from n2p2's sources) when built, else the
    from-scratch FP64 oracle; O(N^2) neighbour list as the reference does."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import BRUTE, Oracle, RefLib, ref_available
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    if ref_available():
        c = RefLib(model_dir)
        kind = "reference"
        run = c.predict
    else:
        c = Oracle(model_dir)
        kind = "port"
        run = lambda: c.predict(0, BRUTE, threads)  # noqa: E731
    c.set_structure(pos, elem, box)
    t0 = time.perf_counter()
    run()
    first = time.perf_counter() - t0
    F = c.forces()
    E = c.energy()
    reps, tot, phases = 0, 0.0, np.zeros(5)
    while tot < budget_s and reps < 20:
        t0 = time.perf_counter()
        run()
        tot += time.perf_counter() - t0
        phases += np.array(c.phase_times())
        reps += 1
    per = tot / reps
    ph = phases / reps
    n = len(elem)
    sample = ("%d evaluations of the %d-atom box (first %.2f s untimed), %d OpenMP threads; per evaluation "
              "%.3f s: neighbour list %.3f, SF %.3f, NN %.3f, forces %.3f; without the O(N^2) neighbour "
              "list %.1f atom-steps/s" % (reps, n, first, threads, per, ph[0], ph[1], ph[2], ph[4],
                                          n / max(per - ph[0], 1e-12)))
    return {"value": n / per, "unit": "atom-steps/s", "cores": threads, "kind": kind,
            "sample": sample}, E, F


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--form", choices=["mixed", "plain"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import n2p2_amd
    from n2p2_amd import remap, synth

    mdir = synth.water_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_water_model"))
    pos, elem, box = synth.water_box(16)
    gpos = remap(pos, box)
    n = len(elem)
    m = n2p2_amd.Mode(mdir)
    m.set_structures(gpos, elem, box[None])
    w = m.work_counts()
    nn_macs = 50 * 25 + 25 * 25 + 25 * 1

    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    dt_plain, t_plain = time_form(m, n2p2_amd.PLAIN_FP64, max(5, args.steps // 5), 2)
    F_plain = m.forces()
    E_plain = m.energies()[0]
    if form == n2p2_amd.MIXED:
        dt, times = time_form(m, form, args.steps, args.warmup)
        F_form = m.forces()
        E_form = m.energies()[0]
        steps = args.steps
    else:
        dt, times = time_form(m, form, args.steps, args.warmup)
        F_form, E_form, steps = m.forces(), m.energies()[0], args.steps
    value = n * steps / dt
    plain_rate = n * max(5, args.steps // 5) / dt_plain

    sf_flop, nn_flop = flop_counts(w, nn_macs, n)
    dom = max((k for k in times if k not in ("nbr", "rev")), key=lambda k: times[k][0])
    ms, launches = times[dom]
    avg_s = ms / launches / 1e3
    if dom in ("sf64", "sf32", "force32"):
        achieved = sf_flop / avg_s / 1e12
        peak = PEAK_FP64 if dom == "sf64" else PEAK_FP32
        roof = {"bound": "valu", "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic_from_profile("k_" + dom),
                "flop_per_launch": sf_flop, "avg_launch_ms": avg_s * 1e3}
    else:
        achieved = nn_flop / avg_s / 1e12
        roof = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32, "traffic": traffic_from_profile("k_" + dom),
                "flop_per_launch": nn_flop,
                "avg_launch_ms": avg_s * 1e3}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": dt / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+fp32 (fp64 reductions)" if form == n2p2_amd.MIXED else "f64",
        "data": "synthetic (seeded water box, random-init weights)",
        "config": {"workload": "water-12288: 4096 H2O, 12288 atoms, 49.68 A periodic cell, rc 6.35 A, "
                               "50 SF/element, NN 50-25-25-1",
                   "form": args.form, "atoms": n, "pairs": w["pairs"], "triplets": w["triplets"]},
        "plain_fp64_atom_steps_per_s": plain_rate,
        "form_over_plain": value / plain_rate,
        "kernel_ms_per_step": {k: v[0] / v[1] for k, v in times.items()},
        "roofline": roof,
    }
    if not args.no_cpu:
        cb, E_cpu, F_cpu = cpu_baseline(mdir, pos, elem, box, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["force_rms_err_vs_fp64_cpu"] = float(np.sqrt(np.mean((F_form - F_cpu) ** 2)))
        out["energy_rel_err_vs_fp64_cpu"] = float(abs(E_form - E_cpu) / abs(E_cpu))
        out["plain_force_max_err_vs_fp64_cpu"] = float(np.abs(F_plain - F_cpu).max())
        out["plain_energy_rel_err_vs_fp64_cpu"] = float(abs(E_plain - E_cpu) / abs(E_cpu))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
