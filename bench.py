"""Benchmark: atom-steps/s of the energy+force path on one MI355X.

Default workload (`water-12288`): the 12,288-atom water box (4,096 H2O on a
seeded 16^3 lattice, 49.68 A cubic cell, cosine cutoff 6.35 A, 50 SFs per
element, 50-25-25-1 tanh NN, random-init weights).  One step = one full
evaluation (neighbour list, symmetry functions, NNs, energy, forces) of the
resident structure.

Other workloads (`--workload`):
  batch-512  512 periodic structures of 16-512 atoms (141,122 atoms, two
             elements 1:2, density 0.1/A^3), same model, one launch per
             kernel over the packed batch; a step evaluates every structure.
  fcc-md     131,072-atom four-species FCC crystal (32^3 cells, a = 3.6 A),
             cutoff 6.0 A, 80 SFs per element, 80-30-30-1 NN; 100
             velocity-Verlet steps (dt 1 fs, 300 K) with a 1 A skin and a
             device-side rebuild trigger, replayed from HIP graphs of 10 steps.

Prints one JSON line.  value = atom-steps/s of the mixed-precision form (the
metric); the plain FP64 form is timed beside it, forces of both are compared
with the CPU FP64 path, and that CPU path -- n2p2's own compute classes where
they were built, else the from-scratch oracle -- is timed on a bounded sample
on the host cores as the cpu_baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W]
                       [--form mixed|plain] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (E+F) mixed vs FP64 GPU form at force RMS err <=1e-3 eV/A"
PEAK_FP64 = 78.6    # TFLOP/s, MI355X FP64 vector
PEAK_FP32 = 157.3   # TFLOP/s, MI355X FP32 vector
KERNELS = ("nbr", "rev", "sf64", "sf64v", "nn64", "energy", "force64", "force64f", "sf32", "nnbf16", "force32",
           "residual", "tangent64", "tangent32", "grad64", "grad32")
NN_MACS = {"water": 50 * 25 + 25 * 25 + 25 * 1, "fcc": 80 * 30 + 30 * 30 + 30 * 1}


def flop_counts(w, nn_macs_per_atom, natoms):
    """Algorithmic FLOPs per evaluation: 13P + 15P_m + 60T + 41T_m + 3T_e
    (symmetry functions and their derivatives) and 6C per atom (NN forward
    and adjoint)."""
    sf = 13 * w["pairs"] + 15 * w["pair_members"] + 60 * w["triplets"] + 41 * w["triplet_members"] \
        + 3 * w["triplet_exps"]
    return sf, 6 * nn_macs_per_atom * natoms


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch of `kernel` as recorded by the FETCH_SIZE /
    WRITE_SIZE counter runs (scaled by a copy-kernel calibration) in
    profiles/hbm_traffic_<workload>.txt; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.txt" % workload.replace("-", ""))
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


# Per-kernel HIP events are recorded on every TIMING_EVERY-th timed step: each
# event is a marker packet between two dispatches, which delays the next
# dispatch by several microseconds (five stages per evaluation: ~40 us of a
# 0.65 ms step); sampling keeps most of that out of the measured throughput
# while the kernel averages still come from launches inside the timed region.
TIMING_EVERY = 4


def time_form(m, form, steps, warmup):
    """Wall time of `steps` evaluations after `warmup`, and the per-kernel
    HIP-event times (events on the library's stream) of every
    TIMING_EVERY-th timed evaluation."""
    import n2p2_amd
    for _ in range(warmup):
        m.evaluate(form, forces=True)
    m.synchronize()
    m.reset_kernel_times()
    m.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        m.set_option(n2p2_amd.OPT_TIMING, int(i % TIMING_EVERY == 0))
        m.evaluate(form, forces=True)
    m.synchronize()
    dt = time.perf_counter() - t0
    times = {}
    for k in KERNELS:
        ms, n = m.kernel_time(k)
        if n:
            times[k] = (ms, n)
    m.set_option(n2p2_amd.OPT_TIMING, 0)
    return dt, times


def time_plain(m, n, steps, warmup):
    """The plain FP64 form both ways: fused (derivatives recomputed in the
    force kernel, nnpx_evaluate's default) and stored (dG/dr written in the
    reference layout and contracted).  Returns the two rates (atom-steps/s),
    the fused path's forces and energies and its kernel times."""
    import n2p2_amd
    m.set_option(n2p2_amd.OPT_PLAIN_FORCES, 1)
    dt_s, _ = time_form(m, n2p2_amd.PLAIN_FP64, steps, warmup)
    m.set_option(n2p2_amd.OPT_PLAIN_FORCES, 0)
    dt_f, times = time_form(m, n2p2_amd.PLAIN_FP64, steps, warmup)
    return n * steps / dt_f, n * steps / dt_s, m.forces(), m.energies(), {k: v[0] / v[1] for k, v in times.items()}


def tune(m):
    """Launch shapes of the mixed kernels from the autotuner, timed on the
    resident structures before the measured region."""
    return m.autotune()


def plain_fields(out, value, rate_f, rate_s, times_p):
    out["plain_fp64_atom_steps_per_s"] = rate_f
    out["plain_fp64_stored_atom_steps_per_s"] = rate_s
    out["form_over_plain"] = value / rate_f
    out["form_over_plain_stored"] = value / rate_s
    out["plain_kernel_ms_per_step"] = times_p


# the kernel computing SF values only, paired with the fused force kernel of its form
VALUE_PAIR = {"sf32": "force32", "sf64v": "force64f"}


def roofline(times, w, nn_macs, natoms, workload):
    """Roofline object of the dominant kernel: algorithmic FLOPs per launch
    over its average HIP-event duration.  The fused force kernel recomputes
    every SF term with its derivative, so its algorithmic work is the SF
    part of FLOP_EF; the NN kernel's is 6C per atom.  FLOP_EF has no split
    into values and derivatives, so where the SF value kernel dominates it
    is reported together with its force kernel: the SF part over the two
    kernels' summed durations."""
    sf_flop, nn_flop = flop_counts(w, nn_macs, natoms)
    dom = max((k for k in times if k not in ("nbr", "rev")), key=lambda k: times[k][0])
    ms, launches = times[dom]
    avg_s = ms / launches / 1e3
    if dom in VALUE_PAIR and VALUE_PAIR[dom] in times:
        f = VALUE_PAIR[dom]
        avg_s += times[f][0] / times[f][1] / 1e3
        dom = dom + "+" + f
    if dom in ("nn64", "nnbf16"):
        flop, peak, bound = nn_flop, PEAK_FP32, "mfma"
    elif dom in ("grad64", "grad32"):
        # forward pass with the tangent and the reverse pass: 8 C per atom
        flop, peak, bound = nn_flop * 8 / 6, (PEAK_FP64 if dom.endswith("64") else PEAK_FP32), "valu"
    else:
        flop, peak, bound = sf_flop, (PEAK_FP64 if dom.endswith("64") else PEAK_FP32), "valu"
    achieved = flop / avg_s / 1e12
    traffic = None
    if "+" not in dom:
        traffic = traffic_from_profile("k_" + {"force64f": "force64f", "sf64v": "values64"}.get(dom, dom), workload)
    else:
        parts = [traffic_from_profile("k_" + {"sf64v": "values64"}.get(k, k), workload) for k in dom.split("+")]
        traffic = sum(parts) if all(x is not None for x in parts) else None
    return {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "flop_per_launch": flop, "avg_launch_ms": avg_s * 1e3}


def cpu_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS where the job
    environment sets it (the CPU share of a one-GPU job on the GPU machine),
    else every CPU this process may run on."""
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))


def host_cpu():
    """Host CPU as lscpu reports it (model, sockets x cores per socket x
    threads per core) and the CPUs this process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_checker(model_dir, cell=False):
    """n2p2's compute path on the host cores: the reference classes (SF
    groups, NN, force contraction compiled from n2p2's sources) when built,
    else the from-scratch FP64 oracle.  `cell` selects the oracle's
    set-identical cell list (for systems where the O(N^2) list is
    infeasible).  Returns the checker and run(threads)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import BRUTE, CELL, Oracle, RefLib, ref_available
    if ref_available() and not cell:
        c = RefLib(model_dir)

        def run(threads):
            c.set_threads(threads)
            c.predict()
        return c, run, "reference"
    c = Oracle(model_dir)
    return c, (lambda threads: c.predict(0, CELL if cell else BRUTE, threads)), "port"


def one_thread_rate(model_dir, pos, elem, box, cell):
    """Atom-steps/s of one evaluation on one thread (after one untimed)."""
    c, run, _ = cpu_checker(model_dir, cell)
    c.set_structure(pos, elem, box)
    run(1)
    t0 = time.perf_counter()
    run(1)
    return len(elem) / (time.perf_counter() - t0)


def cpu_baseline_single(model_dir, pos, elem, box, budget_s, cell=False, max_reps=20, one_thread=None):
    """The multi-thread rate on the structure, and a one-thread rate on
    `one_thread` = (pos, elem, box, label) or on the structure itself."""
    threads = cpu_threads()
    c, run, kind = cpu_checker(model_dir, cell)
    c.set_structure(pos, elem, box)
    t0 = time.perf_counter()
    run(threads)
    first = time.perf_counter() - t0
    F, E = c.forces(), c.energy()
    reps, tot, phases = 0, 0.0, np.zeros(5)
    while tot < budget_s and reps < max_reps:
        t0 = time.perf_counter()
        run(threads)
        tot += time.perf_counter() - t0
        phases += np.array(c.phase_times())
        reps += 1
    per, ph, n = tot / reps, phases / reps, len(elem)
    p1, e1, b1, label = one_thread if one_thread else (pos, elem, box, "the same structure")
    r1 = one_thread_rate(model_dir, p1, e1, b1, cell)
    host = host_cpu()
    sample = ("%d evaluations of the %d-atom structure (first %.2f s untimed), %d OpenMP threads, %s "
              "neighbour list; per evaluation %.3f s: neighbour list %.3f, SF %.3f, NN %.3f, forces %.3f; "
              "without the neighbour list %.1f atom-steps/s; one thread (%s): %.1f atom-steps/s; host %s, "
              "%s sockets x %s cores x %s threads, %d CPUs in this process's affinity"
              % (reps, n, first, threads, "cell" if cell else "O(N^2)", per, ph[0], ph[1], ph[2], ph[4],
                 n / max(per - ph[0], 1e-12), label, r1, host["model"], host["sockets"], host["cores_per_socket"],
                 host["threads_per_core"], host["affinity_cpus"]))
    return {"value": n / per, "unit": "atom-steps/s", "cores": threads, "kind": kind, "sample": sample,
            "one_thread": r1, "host": host}, E, F


def cpu_baseline_batch(model_dir, pos, elem, boxes, off, budget_s):
    """Loop over structures as nnp-dataset does (nnp-dataset.cpp:210-215),
    until the budget is spent; returns the rate and the energies/forces of
    the structures it covered."""
    threads = cpu_threads()
    c, run, kind = cpu_checker(model_dir)
    atoms, tot, done, E, F = 0, 0.0, 0, [], []
    for s in range(len(off) - 1):
        a, b = off[s], off[s + 1]
        c.set_structure(pos[a:b], elem[a:b], boxes[s])
        t0 = time.perf_counter()
        run(threads)
        tot += time.perf_counter() - t0
        atoms += b - a
        done += 1
        E.append(c.energy())
        F.append(c.forces())
        if tot >= budget_s:
            break
    # one thread on the first structures, about a fifth of the budget
    a1, t1 = 0, 0.0
    for s in range(done):
        a, b = off[s], off[s + 1]
        c.set_structure(pos[a:b], elem[a:b], boxes[s])
        t0 = time.perf_counter()
        run(1)
        t1 += time.perf_counter() - t0
        a1 += b - a
        if t1 >= budget_s / 5:
            break
    host = host_cpu()
    sample = ("first %d of %d structures (%d atoms), one evaluation each, %d OpenMP threads, %.2f s; one thread: "
              "%.1f atom-steps/s over the first %d atoms; host %s, %s sockets x %s cores x %s threads, %d CPUs in "
              "this process's affinity" % (done, len(off) - 1, atoms, threads, tot, a1 / t1, a1, host["model"],
                                           host["sockets"], host["cores_per_socket"], host["threads_per_core"],
                                           host["affinity_cpus"]))
    return {"value": atoms / tot, "unit": "atom-steps/s", "cores": threads, "kind": kind, "sample": sample,
            "one_thread": a1 / t1, "host": host}, np.array(E), np.vstack(F), done


def base_record(value, steps, warmup, ms_per_step, dtype, config, data):
    return {"metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": data, "config": config}


DTYPE = {"mixed": "bf16+fp32 (fp64 reductions)", "plain": "f64"}


WATER = {"water-12288": (16, "water-12288: 4096 H2O, 12288 atoms, 49.68 A periodic cell"),
         "water-192": (4, "water-192: 64 H2O, 192 atoms, 12.42 A periodic cell")}


def run_water(args):
    import n2p2_amd
    from n2p2_amd import remap, synth
    mdir = synth.water_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_water_model"))
    n_side, label = WATER[args.workload]
    pos, elem, box = synth.water_box(n_side)
    n = len(elem)
    m = n2p2_amd.Mode(mdir)
    m.set_structures(remap(pos, box), elem, box[None])
    w = m.work_counts()
    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    plain_steps = max(5, args.steps // 5)
    rate_f, rate_s, F_plain, E_plain, times_p = time_plain(m, n, plain_steps, 2)
    E_plain = E_plain[0]
    tuning = tune(m)
    dt, times = time_form(m, form, args.steps, args.warmup)
    F_form, E_form = m.forces(), m.energies()[0]
    value = n * args.steps / dt
    out = base_record(value, args.steps, args.warmup, dt / args.steps * 1e3, DTYPE[args.form],
                      {"workload": label + ", rc 6.35 A, 50 SF/element, NN 50-25-25-1", "form": args.form,
                       "atoms": n, "pairs": w["pairs"], "triplets": w["triplets"]},
                      "synthetic (seeded water box, random-init weights)")
    plain_fields(out, value, rate_f, rate_s, times_p)
    out["tuning"] = tuning
    out["kernel_ms_per_step"] = {k: v[0] / v[1] for k, v in times.items()}
    out["roofline"] = roofline(times, w, NN_MACS["water"], n, args.workload)
    if not args.no_cpu:
        cb, E_cpu, F_cpu = cpu_baseline_single(mdir, pos, elem, box, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["force_rms_err_vs_fp64_cpu"] = float(np.sqrt(np.mean((F_form - F_cpu) ** 2)))
        out["energy_rel_err_vs_fp64_cpu"] = float(abs(E_form - E_cpu) / abs(E_cpu))
        out["plain_force_max_err_vs_fp64_cpu"] = float(np.abs(F_plain - F_cpu).max())
        out["plain_energy_rel_err_vs_fp64_cpu"] = float(abs(E_plain - E_cpu) / abs(E_cpu))
    else:
        out["cpu_baseline"] = None
    return out


def run_batch(args):
    import n2p2_amd
    from n2p2_amd import synth
    mdir = synth.water_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_water_model"))
    pos, elem, boxes, off = synth.random_batch()
    n = len(elem)
    m = n2p2_amd.Mode(mdir)
    m.set_structures(pos, elem, boxes, off)
    w = m.work_counts()
    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    plain_steps = max(3, args.steps // 5)
    rate_f, rate_s, F_plain, E_plain, times_p = time_plain(m, n, plain_steps, 1)
    tuning = tune(m)
    dt, times = time_form(m, form, args.steps, args.warmup)
    F_form, E_form = m.forces(), m.energies()
    value = n * args.steps / dt
    out = base_record(value, args.steps, args.warmup, dt / args.steps * 1e3, DTYPE[args.form],
                      {"workload": "batch-512: 512 periodic structures, 16-512 atoms each (141,122 atoms), "
                                   "2 elements 1:2, density 0.1/A^3, rc 6.35 A, 50 SF/element, NN 50-25-25-1",
                       "form": args.form, "atoms": n, "structures": len(off) - 1, "pairs": w["pairs"],
                       "triplets": w["triplets"]},
                      "synthetic (seeded random structures, random-init weights)")
    plain_fields(out, value, rate_f, rate_s, times_p)
    out["tuning"] = tuning
    out["kernel_ms_per_step"] = {k: v[0] / v[1] for k, v in times.items()}
    out["roofline"] = roofline(times, w, NN_MACS["water"], n, "batch-512")
    if not args.no_cpu:
        cb, E_cpu, F_cpu, done = cpu_baseline_batch(mdir, pos, elem, boxes, off, args.cpu_budget)
        out["cpu_baseline"] = cb
        k = off[done]
        out["force_rms_err_vs_fp64_cpu"] = float(np.sqrt(np.mean((F_form[:k] - F_cpu) ** 2)))
        out["energy_max_rel_err_vs_fp64_cpu"] = float(np.max(np.abs(E_form[:done] - E_cpu) / np.abs(E_cpu)))
        out["plain_force_max_err_vs_fp64_cpu"] = float(np.abs(F_plain[:k] - F_cpu).max())
        out["plain_energy_max_rel_err_vs_fp64_cpu"] = float(np.max(np.abs(E_plain[:done] - E_cpu)
                                                                   / np.abs(E_cpu)))
    else:
        out["cpu_baseline"] = None
    return out


def run_md(args):
    import n2p2_amd
    from n2p2_amd import remap, synth
    mdir = synth.fcc_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_fcc_model"))
    pos, elem, box = synth.fcc_lattice(n_cells=32)
    n = len(elem)
    mass = np.array([synth.MASS[s] for s in synth.FCC_SYMBOLS])
    vel = synth.maxwell_velocities(mass[elem], 300.0)
    steps = args.steps if args.steps_given else 100
    graph = 10
    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    m = n2p2_amd.Mode(mdir)
    m.set_structures(remap(pos, box), elem, box[None])
    w = m.work_counts()
    # the plain form's full evaluation (fused and stored) of the initial positions
    _, _, F_plain, E_plain, times_p = time_plain(m, n, 3, 1)
    tuning = tune(m)
    # per-kernel durations of the same kernels in a full evaluation
    _, times = time_form(m, form, 5, 1)
    F_form, E_form = m.forces(), m.energies()[0]
    m.md_init(mass, vel, 1.0, 1.0, form)
    m.md_run(graph, graph_steps=graph)   # capture and one warm-up replay
    m.synchronize()
    t0 = time.perf_counter()
    m.md_run(steps, graph_steps=graph)
    m.synchronize()
    dt = time.perf_counter() - t0
    ep, ek, rebuilds = m.md_energies()
    et = ep + ek
    value = n * steps / dt
    # the plain FP64 form: the same MD in the same HIP graphs, fused forces;
    # then a few eager steps with the stored dG/dr
    rates = []
    for stored in (0, 1):
        mp = n2p2_amd.Mode(mdir)
        mp.set_option(n2p2_amd.OPT_PLAIN_FORCES, stored)
        mp.set_structures(remap(pos, box), elem, box[None])
        mp.md_init(mass, vel, 1.0, 1.0, n2p2_amd.PLAIN_FP64)
        gs = 0 if stored else graph
        pst = max(3, steps // 20) if stored else graph * max(1, steps // (5 * graph))
        mp.md_run(gs or 1, graph_steps=gs)
        mp.synchronize()
        t0 = time.perf_counter()
        mp.md_run(pst, graph_steps=gs)
        mp.synchronize()
        rates.append(n * pst / (time.perf_counter() - t0))
        del mp
    out = base_record(value, steps, graph, dt / steps * 1e3, DTYPE[args.form],
                      {"workload": "fcc-md: 131072-atom 4-species FCC (32^3 cells, a 3.6 A), rc 6.0 A, 80 "
                                   "SF/element, NN 80-30-30-1; velocity Verlet dt 1 fs, 300 K, skin 1 A, "
                                   "HIP graphs of 10 steps", "form": args.form, "atoms": n,
                       "pairs": w["pairs"], "triplets": w["triplets"]},
                      "synthetic (seeded FCC lattice and velocities, random-init weights)")
    plain_fields(out, value, rates[0], rates[1], times_p)
    out["tuning"] = tuning
    out["md"] = {"rebuilds": int(rebuilds), "steps_total": int(len(ep)),
                 "etot_drift_eV": float(np.abs(et - et[0]).max()), "ekin_first": float(ek[0]),
                 "ekin_last": float(ek[-1])}
    out["kernel_ms_per_step"] = {k: v[0] / v[1] for k, v in times.items()}
    out["roofline"] = roofline(times, w, NN_MACS["fcc"], n, "fcc-md")
    if not args.no_cpu:
        # The reference's O(N^2) list (~4.6e11 pair tests) is infeasible at this
        # size: the oracle with its set-identical cell list, one evaluation per step.
        # One thread: a 10^3-cell piece of the same crystal (4,000 atoms,
        # identical per-atom work).
        p1, e1, b1 = synth.fcc_lattice(n_cells=10)
        cb, E_cpu, F_cpu = cpu_baseline_single(mdir, pos, elem, box, args.cpu_budget, cell=True, max_reps=3,
                                               one_thread=(p1, e1, b1, "a 4,000-atom 10^3-cell piece"))
        out["cpu_baseline"] = cb
        out["force_rms_err_vs_fp64_cpu"] = float(np.sqrt(np.mean((F_form - F_cpu) ** 2)))
        out["energy_rel_err_vs_fp64_cpu"] = float(abs(E_form - E_cpu) / abs(E_cpu))
        out["plain_force_max_err_vs_fp64_cpu"] = float(np.abs(F_plain - F_cpu).max())
        out["plain_energy_rel_err_vs_fp64_cpu"] = float(abs(E_plain[0] - E_cpu) / abs(E_cpu))
    else:
        out["cpu_baseline"] = None
    return out


def run_train(args):
    """Loss gradient of 256 water structures (64-256 atoms) with energy and
    force targets from a teacher network (Training's error function,
    beta = 0.7), both forms; a step is one gradient of the whole batch."""
    import n2p2_amd
    from n2p2_amd import synth
    base = tempfile.gettempdir()
    student = synth.water_model(os.path.join(base, "n2p2_amd_train_student"))
    teacher = synth.water_model(os.path.join(base, "n2p2_amd_train_teacher"), seed=synth.SEED_TEACHER)
    pos, elem, boxes, off = synth.water_set()
    n, ns, beta = len(elem), len(off) - 1, 0.7
    t = n2p2_amd.Mode(teacher)
    t.set_structures(pos, elem, boxes, off)
    t.evaluate(n2p2_amd.PLAIN_FP64, forces=True)
    E_ref, F_ref = t.energies(), t.forces()
    del t
    m = n2p2_amd.Mode(student)
    m.set_structures(pos, elem, boxes, off)
    w = m.work_counts()

    def time_grad(prec, steps, warmup):
        for _ in range(warmup):
            m.loss_gradient(E_ref, F_ref, beta, prec)
        m.synchronize()
        m.reset_kernel_times()
        t0 = time.perf_counter()
        for i in range(steps):
            m.set_option(n2p2_amd.OPT_TIMING, int(i % TIMING_EVERY == 0))
            loss, g = m.loss_gradient(E_ref, F_ref, beta, prec)
        m.synchronize()
        dt = time.perf_counter() - t0
        times = {}
        for k in KERNELS:
            ms, cnt = m.kernel_time(k)
            if cnt:
                times[k] = (ms, cnt)
        m.set_option(n2p2_amd.OPT_TIMING, 0)
        return dt, times, loss, g
    dt_p, times_p, loss_p, g_p = time_grad(n2p2_amd.PLAIN_FP64, max(3, args.steps // 5), 1)
    tuning = tune(m)
    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    dt, times, loss, g = time_grad(form, args.steps, args.warmup)
    value = n * args.steps / dt
    rate_p = n * max(3, args.steps // 5) / dt_p
    out = base_record(value, args.steps, args.warmup, dt / args.steps * 1e3, DTYPE[args.form],
                      {"workload": "train-256: loss gradient (energy + force terms, beta 0.7) of 256 water "
                                   "structures, 66-255 atoms (%d atoms), rc 6.35 A, 50 SF/element, NN 50-25-25-1"
                                   % n, "form": args.form, "atoms": n, "structures": ns, "pairs": w["pairs"],
                       "triplets": w["triplets"]},
                      "synthetic (seeded water structures, random-init student and teacher weights)")
    out["plain_fp64_atom_steps_per_s"] = rate_p
    out["form_over_plain"] = value / rate_p
    out["plain_kernel_ms_per_step"] = {k: v[0] / v[1] for k, v in times_p.items()}
    out["gradient_max_rel_err_vs_plain"] = float(np.abs(g - g_p).max() / np.abs(g_p).max())
    out["loss_rel_err_vs_plain"] = [float(x) for x in np.abs(loss - loss_p) / np.maximum(np.abs(loss_p), 1e-300)]
    out["tuning"] = tuning
    out["kernel_ms_per_step"] = {k: v[0] / v[1] for k, v in times.items()}
    out["roofline"] = roofline(times, w, NN_MACS["water"], n, args.workload)
    if not args.no_cpu:
        # the oracle's Training-style gradient (calculateDEdc / calculateDFdc
        # with the contracted dG/dx Jacobian), structure by structure until
        # the budget is spent
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import BRUTE, PULL, Oracle
        o = Oracle(student)
        threads = cpu_threads()
        atoms, tot, done = 0, 0.0, 0
        g_cpu = None
        for s_ in range(ns):
            a, b = off[s_], off[s_ + 1]
            o.set_structure(pos[a:b], elem[a:b], boxes[s_])
            t0 = time.perf_counter()
            o.predict(PULL, BRUTE, threads)
            _, gs = o.loss_gradient(E_ref[s_], F_ref[a:b], 1.0 / (b - a), beta / (3.0 * n))
            tot += time.perf_counter() - t0
            g_cpu = gs if g_cpu is None else g_cpu + gs
            atoms += b - a
            done += 1
            if tot >= args.cpu_budget:
                break
        host = host_cpu()
        out["cpu_baseline"] = {"value": atoms / tot, "unit": "atom-steps/s", "cores": threads, "kind": "port",
                               "sample": "loss gradient of the first %d of %d structures (%d atoms), %.2f s, "
                                         "%d OpenMP threads (evaluation; the gradient pass is serial); host %s, "
                                         "%s sockets x %s cores x %s threads"
                                         % (done, ns, atoms, tot, threads, host["model"], host["sockets"],
                                            host["cores_per_socket"], host["threads_per_core"]),
                               "host": host}
    else:
        out["cpu_baseline"] = None
    return out


def run_profile(args):
    """Only the timed work of one form on one workload (after the autotuner
    for the mixed form), for kernel-trace and counter runs: no plain/mixed
    comparison and no CPU leg.  fcc-md runs its MD steps (HIP graphs of 10)."""
    import n2p2_amd
    from n2p2_amd import remap, synth
    form = n2p2_amd.MIXED if args.form == "mixed" else n2p2_amd.PLAIN_FP64
    if args.workload == "fcc-md":
        mdir = synth.fcc_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_fcc_model"))
        pos, elem, box = synth.fcc_lattice(n_cells=32)
        m = n2p2_amd.Mode(mdir)
        m.set_option(n2p2_amd.OPT_PLAIN_FORCES, args.plain_forces)
        m.set_structures(remap(pos, box), elem, box[None])
        if form == n2p2_amd.MIXED:
            m.autotune()
        mass = np.array([synth.MASS[s] for s in synth.FCC_SYMBOLS])
        m.md_init(mass, synth.maxwell_velocities(mass[elem], 300.0), 1.0, 1.0, form)
        m.md_run(10, graph_steps=10)
        m.synchronize()
        t0 = time.perf_counter()
        m.md_run(args.steps, graph_steps=10)
        m.synchronize()
        dt = time.perf_counter() - t0
        n = len(elem)
    else:
        mdir = synth.water_model(os.path.join(tempfile.gettempdir(), "n2p2_amd_water_model"))
        if args.workload == "water-12288":
            pos, elem, box = synth.water_box(16)
            boxes, off = box[None], None
            pos = remap(pos, box)
        else:
            pos, elem, boxes, off = synth.random_batch()
        m = n2p2_amd.Mode(mdir)
        m.set_option(n2p2_amd.OPT_PLAIN_FORCES, args.plain_forces)
        m.set_structures(pos, elem, boxes, off)
        if form == n2p2_amd.MIXED:
            m.autotune()
        dt, _ = time_form(m, form, args.steps, args.warmup)
        n = len(elem)
    return {"profile": args.workload, "form": args.form, "plain_forces": args.plain_forces, "steps": args.steps,
            "atom_steps_per_s": n * args.steps / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["water-12288", "water-192", "batch-512", "fcc-md", "train-256"],
                    default="water-12288")
    ap.add_argument("--form", choices=["mixed", "plain"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration)")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed work of --form on --workload (kernel traces, counters)")
    ap.add_argument("--plain-forces", type=int, default=0, choices=[0, 1],
                    help="plain form forces in --profile runs: 0 fused recompute, 1 stored dG/dr")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 50 if args.workload == "water-12288" else 10
    if args.profile:
        print(json.dumps(run_profile(args)))
        return
    run = {"water-12288": run_water, "water-192": run_water, "batch-512": run_batch, "fcc-md": run_md,
           "train-256": run_train}[args.workload]
    print(json.dumps(run(args)))


if __name__ == "__main__":
    main()
