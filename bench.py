"""Benchmark of the MI355X DFSPH step: particle-substeps/s at 1,048,576
particles (a block filling 40% of the unit box, sinusoidal tilt gravity,
4,096 tracers), the reference's fixed iteration schedule.

A bench step is one full DFSPH substep (grid build, density and alpha,
divergence solve, forces, density solve with R refine iterations,
integrate) over the whole scene, inputs resident in HBM. Per-frame work
(stats, tracers) runs once per frame outside the timed substeps and is
reported on its own.

Prints one JSON line. `roofline` is priced on the canonical algorithmic
bytes per particle of each kernel (BYTES_PER_PARTICLE) over the kernel's average
launch time measured with HIP events on the sim's own stream during the
timed region. `cpu_baseline` times the CPU oracle (a C++ restatement of the
reference's step, OpenMP over the host cores) on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "fluid-simulation-app_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Canonical algorithmic bytes per particle of each kernel class (per launch).
# The plain step's integrate also takes the next grid build's count (the cell
# and provisional rank it writes, 8 bytes a particle, from the positions it
# already holds), so that step runs no grid_count; a batch counts in
# grid_count (24 bytes: the position read, cell and rank written) and its
# integrate moves 120.
BYTES_PER_PARTICLE = {
    "grid_count": 24, "grid_scatter": 124, "density_div": 64, "div_apply": 68, "forces": 124,
    "den_kappa": 68, "den_apply": 68, "integrate": 128,
}
BYTES_PER_CELL = {"grid_scan": 16}
INTEGRATE_UNFUSED = 120


def canonical_bytes(n, cells, refine, count_fused=True):
    """A substep's canonical bytes: N (508 + 136 R) + 16 C with the count in
    integrate (the plain step), N (524 + 136 R) + 16 C with grid_count."""
    return n * ((508 if count_fused else 524) + 136 * refine) + 16 * cells


def kernel_bytes(name, n, cells):
    if name in BYTES_PER_PARTICLE:
        return BYTES_PER_PARTICLE[name] * n
    if name in BYTES_PER_CELL:
        return BYTES_PER_CELL[name] * cells
    return None


# The committed counter summary `roofline.traffic` comes from (written by
# --pmc-summary from rocprofv3 runs of this bench; see its header).
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_block_1m.md")
COPY_BYTES = 268435456


def _kernel_classes(names):
    """Labels a substep's dispatches (in dispatch order) with the bench's
    kernel classes. kappa_apply serves div_apply (after density_div) and
    den_apply (after den_kappa); grid_scatter is place + gather and
    density_div includes its tile_regions launch, as the sim's own timer
    groups them. Returns (class, primary) per dispatch: `primary` marks the
    launch that closes one occurrence of the class."""
    import re
    out, after = [], None
    for name in names:
        m = re.search(r"\b(k_\w+?)(?:<|\(|$)", name)
        k = m.group(1) if m else name
        cls = {"k_grid_count": "grid_count", "k_grid_scan": "grid_scan", "k_grid_place": "grid_scatter",
               "k_grid_gather": "grid_scatter", "k_tile_regions": "density_div", "k_density_div": "density_div",
               "k_forces": "forces",
               "k_den_kappa": "den_kappa", "k_integrate": "integrate"}.get(k)
        if k == "k_kappa_apply":
            cls = "div_apply" if after == "k_density_div" else "den_apply"
        primary = k not in ("k_grid_place", "k_tile_regions")
        if k in ("k_density_div", "k_den_kappa"):
            after = k
        out.append((cls, primary))
    return out


def pmc_summary(base, workload="block_1m", steps=0, warmup=0):
    """Prints the markdown summary of four rocprofv3 --pmc runs under `base`:
    pmc_F / pmc_W (FETCH_SIZE, WRITE_SIZE of a short bench run of
    `workload`) and pmc_copyF / pmc_copyW (the same counters over
    --copy-calibration). Each counter is scaled by what it reads for the
    plain copy of known size. With a kernel trace of the same workload under
    base/trace (rocprofv3 --kernel-trace), adds each kernel class's average
    launch time and its achieved fraction of the HBM roofline."""
    import csv
    import glob

    def rows(d, name="pmc_counter_collection.csv"):
        f = os.path.join(base, d, name)
        return sorted(csv.DictReader(open(f)), key=lambda r: int(r["Dispatch_Id"]))

    def copy_factor(d):
        vals = [float(r["Counter_Value"]) for r in rows(d) if "copyBuffer" in r["Kernel_Name"]]
        return (COPY_BYTES / 1024.0) / (sum(vals) / len(vals)), sum(vals) / len(vals)

    ff, fcopy = copy_factor("pmc_copyF")
    fw, wcopy = copy_factor("pmc_copyW")
    per = {}
    for d, key in (("pmc_F", "fetch"), ("pmc_W", "write")):
        rs = [r for r in rows(d) if "k_" in r["Kernel_Name"]]
        for r, (cls, primary) in zip(rs, _kernel_classes([r["Kernel_Name"] for r in rs])):
            if cls is None:
                continue
            e = per.setdefault(cls, {"fetch": 0.0, "write": 0.0, "launches_fetch": 0, "launches_write": 0})
            e[key] += float(r["Counter_Value"])
            if primary:
                e["launches_" + key] += 1
    import fluid_amd as fa
    sc = {"block_1m": fa.scenes.block_1m, "block_8m": fa.scenes.block_8m,
          "dam_break_16k": fa.scenes.dam_break}[workload](seed=0)
    n, cells = sc.n, sc.grid.cell_count()
    refine = sc.schedule()[2]
    window = (f" --steps {steps} --warmup {warmup}" if steps else "")
    cmd = "python3 -u bench.py" + ("" if workload == "block_1m" else f" --workload {workload}") + window + \
        " --no-cpu --no-batched --no-adjoint --no-cfl --no-tolerance"
    print(f"# HBM traffic per launch, {workload} ({n:,} particles, {cells:,} cells, R = {refine})\n")
    print("Commands, from the repository root on the MI355X box (each counter in a run of its own):\n")
    print("```")
    for c, d in (("FETCH_SIZE", "pmc_F"), ("WRITE_SIZE", "pmc_W")):
        print(f"rocprofv3 --pmc {c} --output-format csv -d {base}/{d} -o pmc -- {cmd}")
    for c, d in (("FETCH_SIZE", "pmc_copyF"), ("WRITE_SIZE", "pmc_copyW")):
        print(f"rocprofv3 --pmc {c} --output-format csv -d {base}/{d} -o pmc -- "
              f"python3 -u bench.py --copy-calibration {COPY_BYTES}")
    traces = glob.glob(os.path.join(base, "trace", "*kernel_trace.csv"))
    if traces:
        print(f"rocprofv3 --kernel-trace --stats --output-format csv -d {base}/trace -o trace -- {cmd}")
    out = "profiles/pmc_block_1m.md" if workload == "block_1m" else f"profiles/roofline_{workload}.md"
    print(f"python3 bench.py --pmc-summary {base}" + ("" if workload == "block_1m" else f" --workload {workload}") +
          (f" --steps {steps} --warmup {warmup}" if steps else "") + f" > {out}")
    print("```\n")
    print(f"Calibration: a {COPY_BYTES >> 20} MiB device copy (HIP's copy kernel) reads FETCH_SIZE = "
          f"{fcopy:.0f} KiB and WRITE_SIZE = {wcopy:.0f} KiB, so traffic = {ff:.3f} x FETCH_SIZE + "
          f"{fw:.3f} x WRITE_SIZE (KiB x 1024 bytes).\n")
    print("| kernel | launches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic MB | algorithmic MB |")
    print("|---|---|---|---|---|---|")
    classes = ["grid_count", "grid_scan", "grid_scatter", "density_div", "div_apply", "forces", "den_kappa",
               "den_apply", "integrate"]
    traffic = {}
    for cls in classes:
        e = per.get(cls)
        if not e or not e["launches_fetch"] or not e["launches_write"]:
            continue
        fk = e["fetch"] / e["launches_fetch"]
        wk = e["write"] / e["launches_write"]
        traffic[cls] = (ff * fk + fw * wk) * 1024.0
        alg = kernel_bytes(cls, n, cells)
        print(f"| {cls} | {e['launches_fetch']} | {fk:.0f} | {wk:.0f} | {traffic[cls] / 1e6:.1f} | "
              f"{alg / 1e6:.1f} |")
    if not traces:
        return
    rs = sorted(csv.DictReader(open(traces[0])), key=lambda r: int(r["Start_Timestamp"]))
    rs = [r for r in rs if "k_" in r["Kernel_Name"]]
    tot, occ = {}, {}
    for r, (cls, primary) in zip(rs, _kernel_classes([r["Kernel_Name"] for r in rs])):
        if cls is None:
            continue
        tot[cls] = tot.get(cls, 0.0) + (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        occ[cls] = occ.get(cls, 0) + (1 if primary else 0)
    print("\nPer-kernel roofline, from the kernel trace of the same command: average launch time (a class's"
          " launches summed, place + gather as grid_scatter, tile_regions with density_div where it runs: the"
          " plain step's gather writes the tile regions, and its integrate takes the next grid count, so"
          " grid_count runs only in the first substep), algorithmic"
          f" bytes over that time, and its fraction of MI355X's {HBM_PEAK_GBS / 1000:.0f} TB/s.\n")
    print("| kernel | launches | avg µs | algorithmic GB/s | frac of HBM peak | measured traffic GB/s |")
    print("|---|---|---|---|---|---|")
    sub = 0.0
    for cls in classes:
        if not occ.get(cls):
            continue
        us = tot[cls] / occ[cls]
        alg = kernel_bytes(cls, n, cells) / (us * 1e-6) / 1e9
        meas = traffic[cls] / (us * 1e-6) / 1e9 if cls in traffic else float("nan")
        print(f"| {cls} | {occ[cls]} | {us:.1f} | {alg:.0f} | {alg / HBM_PEAK_GBS:.3f} | {meas:.0f} |")
        sub += us * occ[cls]
    substeps = occ.get("integrate", 1)
    canon = canonical_bytes(n, cells, refine)
    print(f"\nA substep's kernels: {sub / substeps:.1f} µs of kernel time; canonical bytes "
          f"{canon / 1e6:.1f} MB, {canon / (sub / substeps * 1e-6) / 1e9:.0f} GB/s "
          f"({canon / (sub / substeps * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f} of peak).")


def _short_name(name):
    import re
    m = re.search(r"\b(k_\w+?)(<[^()]*>)?\(", name)
    if not m:
        return name.split("(")[0]
    args = m.group(2) or ""
    return m.group(1) + args.replace("fluid::", "")


def trace_summary(base, workload="block_1m", steps=0, warmup=0):
    """Prints the markdown table of a rocprofv3 --kernel-trace --stats run
    under base/trace (trace_kernel_stats.csv): calls, average, min and max
    duration of each kernel."""
    import csv
    f = os.path.join(base, "trace", "trace_kernel_stats.csv")
    rs = list(csv.DictReader(open(f)))
    window = (f" --steps {steps} --warmup {warmup}" if steps else "")
    cmd = "python3 -u bench.py" + ("" if workload == "block_1m" else f" --workload {workload}") + window + \
        " --no-cpu --no-batched --no-adjoint --no-cfl --no-tolerance"
    if workload.startswith("ensemble"):
        cmd = "python3 -u bench.py --batched-only"
    print(f"# DFSPH substep kernels, {workload}, kernel trace\n")
    print("Commands, from the repository root on the MI355X box:\n")
    print("```")
    print(f"rocprofv3 --kernel-trace --stats --output-format csv -d {base}/trace -o trace -- {cmd}")
    print(f"python3 bench.py --trace-summary {base}" + ("" if workload == "block_1m" else f" --workload {workload}") +
          (f" --steps {steps} --warmup {warmup}" if steps else "") + f" > profiles/kernels_{workload}.md")
    print("```\n")
    if workload.startswith("ensemble"):
        print("The run holds the autotuner's trials, the batched substeps under graph replay, the same batch at the"
              " configs' tolerances (work-list passes: `_list` kernels), a profiled substep, and the plain loop"
              " over the same scenes, fixed and at the tolerances (`OneScene` kernels).\n")
    else:
        print("`k_kappa_apply` serves div_apply (1 a substep) and den_apply (R a substep); the solve's kernels are"
              " the 128-slot tile build (namespace t128).\n")
    print("| kernel | calls | avg µs | min µs | max µs | total ms | % |")
    print("|---|---|---|---|---|---|---|")
    for r in sorted(rs, key=lambda r: -float(r["TotalDurationNs"])):
        print(f"| {_short_name(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f} |")


def tolerance_summary(base):
    """Prints the kernel table of the tolerance leg from a rocprofv3
    --kernel-trace run of the default bench under base/trace
    (trace_kernel_trace.csv): the substeps that ran convergence checks, less
    the leg's untimed first one, each from the end of the previous
    substep's integrate to the end of its own."""
    import csv
    import re
    f = os.path.join(base, "trace", "trace_kernel_trace.csv")
    ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short_name(r["Kernel_Name"]))
                for r in csv.DictReader(open(f)))
    subs, cur, t_prev = [], [], None
    for e in ev:
        cur.append(e)
        if e[2].startswith("k_integrate"):
            subs.append((t_prev, cur))
            t_prev, cur = e[1], []
    tol = [(t0, s) for t0, s in subs if t0 is not None and any(re.match(r"k_converge", e[2]) for e in s)][1:]
    if not tol:
        print("no tolerance-mode substeps in the trace")
        return
    n = len(tol)
    wall = sum(s[-1][1] - t0 for t0, s in tol) / n / 1e3
    acc = {}
    for _, s in tol:
        for a, z, k in s:
            c = acc.setdefault(k, [0, 0.0])
            c[0] += 1
            c[1] += (z - a) / 1e3
    busy = sum(v[1] for v in acc.values()) / n
    cmd = "python3 -u bench.py --no-cpu --no-batched --no-adjoint --no-cfl"
    print("# DFSPH substep at the configs' tolerances, block_1m, kernel trace\n")
    print("Commands, from the repository root on the MI355X box:\n")
    print("```")
    print(f"rocprofv3 --kernel-trace --stats --output-format csv -d {base}/trace -o trace -- {cmd}")
    print(f"python3 bench.py --tolerance-summary {base} > profiles/kernels_tolerance_1m.md")
    print("```\n")
    print(f"The bench's tolerance leg (divergence {TOL_DIVERGENCE:g}, density {TOL_DENSITY:g}, caps of {MAX_PASSES} "
          f"passes) continues the developed 1,048,576-particle scene; {n} substeps after its untimed first one. "
          f"A substep spans {wall:.0f} µs from the previous substep's integrate to its own, {busy:.0f} µs of it "
          "in kernels. The host issues no pass once a check says the solve has converged, so the launches "
          "are the passes the scene ran plus at most four no-op passes a solve.\n")
    print("| kernel | launches a substep | avg µs | µs a substep | % of kernel time |")
    print("|---|---|---|---|---|")
    for k, (c, t) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
        print(f"| {k} | {c / n:.1f} | {t / c:.1f} | {t / n:.0f} | {100 * t / n / busy:.1f} |")


def sq_summary(base, workload="block_1m"):
    """Prints per-wave shader counters of two rocprofv3 --pmc runs under
    base/sq_A and base/sq_B, averaged over every launch of each kernel."""
    import csv
    acc = {}
    for d in ("sq_A", "sq_B"):
        f = os.path.join(base, d, "pmc_counter_collection.csv")
        for r in csv.DictReader(open(f)):
            if "k_" not in r["Kernel_Name"]:
                continue
            k = _short_name(r["Kernel_Name"])
            disp = acc.setdefault(k, {}).setdefault((d, r["Dispatch_Id"]), {})
            disp[r["Counter_Name"]] = float(r["Counter_Value"])
    cmd = "python3 -u bench.py" + ("" if workload == "block_1m" else f" --workload {workload}") + \
        " --no-cpu --no-batched --no-adjoint --no-cfl --no-tolerance"
    print(f"# Shader counters per wave, {workload}\n")
    print("Commands, from the repository root on the MI355X box (two passes within the per-block counter limits):\n")
    print("```")
    print("rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR "
          f"SQ_WAVE_CYCLES SQ_BUSY_CYCLES --output-format csv -d {base}/sq_A -o pmc -- {cmd}")
    print("rocprofv3 --pmc SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_ACTIVE_INST_ANY "
          f"SQ_LDS_BANK_CONFLICT SQ_WAIT_ANY GRBM_GUI_ACTIVE --output-format csv -d {base}/sq_B -o pmc -- {cmd}")
    print(f"python3 bench.py --sq-summary {base} > profiles/sq_{workload}.md")
    print("```\n")
    print("Per wave, averaged over every launch of the run (the counters' own cycle units).\n")
    cols = [("SQ_INSTS_VALU", "VALU"), ("SQ_INSTS_SALU", "SALU"), ("SQ_INSTS_LDS", "LDS"),
            ("SQ_INSTS_VMEM_RD", "VMEM rd"), ("SQ_INSTS_VMEM_WR", "VMEM wr"), ("SQ_WAVE_CYCLES", "wave cycles"),
            ("SQ_WAIT_ANY", "wait cycles"), ("SQ_LDS_BANK_CONFLICT", "LDS bank conflicts")]
    print("| kernel | waves a launch | " + " | ".join(c[1] for c in cols) + " |")
    print("|---|---|" + "---|" * len(cols))
    for k, disps in sorted(acc.items()):
        waves = [v["SQ_WAVES"] for (d, _), v in disps.items() if d == "sq_A" and v.get("SQ_WAVES")]
        if not waves:
            continue
        wave_n = sum(waves) / len(waves)
        row = []
        for c, _ in cols:
            vals = [v[c] for v in disps.values() if c in v]
            row.append(f"{sum(vals) / len(vals) / wave_n:.1f}" if vals else "-")
        print(f"| {k} | {wave_n:.0f} | " + " | ".join(row) + " |")


def load_traffic(path=PMC_SUMMARY):
    """kernel class -> measured HBM bytes per launch from the committed
    counter summary's traffic table, or {} when there is none."""
    out = {}
    if not os.path.exists(path):
        return out
    in_table = False
    for line in open(path):
        if line.startswith("|"):
            in_table = in_table or "traffic MB" in line
        else:
            in_table = False
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_table and len(cells) == 6 and (cells[0] in BYTES_PER_PARTICLE or cells[0] in BYTES_PER_CELL):
            try:
                out[cells[0]] = float(cells[4]) * 1e6
            except ValueError:
                pass
    return out


def cpu_baseline(scene, step, refine, budget_s=10.0, max_substeps=64):
    """Times the oracle's FP32 step on this host's cores on a bounded sample."""
    import oracle
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "1024")))
    oracle.set_threads(cores)
    o = oracle.OracleSim(scene.params, precision=32)
    o.set_state(scene.pos4)
    o.substep(step, refine)  # warm-up substep (untimed)
    done, t0 = 0, time.perf_counter()
    while done < max_substeps:
        o.substep(step, refine)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    o.close()
    return {"value": scene.n * done / dt, "unit": "particle-substeps/s", "cores": cores, "kind": "port",
            "sample": f"{done} substeps (R={refine}) of the same {scene.n}-particle scene after 1 warm-up "
                      f"substep, {dt:.1f} s"}


# The configs' solver tolerances: mean divergence error 0.1%, mean density
# error 0.01%; each solve stops at MAX_PASSES passes.
# The timed region's dominant-kernel events time one launch in every 7 (7
# and the 5 density passes are coprime, so every pass is sampled).
DOMINANT_STRIDE = 7
TOL_DIVERGENCE = 1e-3
TOL_DENSITY = 1e-4
MAX_PASSES = 256
# Substeps of the plain scene's tolerance leg (the first untimed).
TOL_SUBSTEPS = 11


def cpu_configs(fa, runs=5):
    """The CPU baseline plan's runs of the oracle's FP32 step on this host's
    cores: the 16,384-particle dam break for 200 frames in full (median of
    `runs` runs, after 1 warm-up frame) with the time of each kernel class,
    10 frames of block_1m and 1 frame of block_8m, each reported per
    particle-substep."""
    import oracle
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "1024")))
    oracle.set_threads(cores)
    out = {"cores": cores, "kind": "port", "timer": "time.perf_counter around whole substeps"}

    def frames_of(sc, frames, first=0):
        n, dt, refine, vc = sc.schedule()
        return [(sc.step_block((first * n + k) * dt, dt, vc, seed=first * n + k), refine)
                for k in range(frames * n)], n, refine

    # config 1 in full
    sc = fa.scenes.dam_break(seed=0)
    warm, n, refine = frames_of(sc, 1)
    plan, _, _ = frames_of(sc, 200, first=1)
    times = []
    for _ in range(runs):
        o = oracle.OracleSim(sc.params, precision=32)
        o.set_state(sc.pos4)
        for st, r in warm:
            o.substep(st, r)
        t0 = time.perf_counter()
        for st, r in plan:
            o.substep(st, r)
        times.append(time.perf_counter() - t0)
        o.close()
    med = sorted(times)[len(times) // 2]
    # the kernel classes of one frame, launch by launch
    o = oracle.OracleSim(sc.params, precision=32)
    o.set_state(sc.pos4)
    per = {}
    for st, r in plan[:n]:
        for name in ("grid", "density_div", "div_apply", "forces_eval", "forces_den_apply") + \
                ("den_kappa", "den_apply") * r + ("integrate",):
            t0 = time.perf_counter()
            o.kernel(name, st)
            per[name] = per.get(name, 0.0) + (time.perf_counter() - t0) * 1e3 / n
    o.close()
    out["dam_break_16k"] = {"particles": sc.n, "frames": 200, "substeps": len(plan), "refine": refine,
                            "runs": runs, "median_s": med, "value": sc.n * len(plan) / med,
                            "unit": "particle-substeps/s", "ms_per_substep": med * 1e3 / len(plan),
                            "per_kernel_ms_per_substep": {k: round(v, 4) for k, v in per.items()}}
    # configs 2 and 3, subsets
    for name, maker, frames in (("block_1m", fa.scenes.block_1m, 10), ("block_8m", fa.scenes.block_8m, 1)):
        sc = maker(seed=0)
        plan, n, refine = frames_of(sc, frames)
        o = oracle.OracleSim(sc.params, precision=32)
        o.set_state(sc.pos4)
        t0 = time.perf_counter()
        for st, r in plan:
            o.substep(st, r)
        dt_ = time.perf_counter() - t0
        o.close()
        out[name] = {"particles": sc.n, "frames": frames, "substeps": len(plan), "refine": refine, "seconds": dt_,
                     "value": sc.n * len(plan) / dt_, "unit": "particle-substeps/s",
                     "ms_per_substep": dt_ * 1e3 / len(plan), "subset": f"the first {frames} frame(s)"}
        print(f"cpu {name}: {out[name]['ms_per_substep']:.1f} ms a substep", file=sys.stderr, flush=True)
    return out


def hist_summary(h):
    """A pass-count histogram as {passes: count}, with its mean and max."""
    cnt = {int(p): int(c) for p, c in enumerate(h) if c}
    n_ = sum(cnt.values())
    return {"histogram": cnt, "mean": sum(p * c for p, c in cnt.items()) / max(n_, 1),
            "max": max(cnt) if cnt else 0, "at_cap": cnt.get(MAX_PASSES, 0)}


def batched_ensemble(fa, scenes=256, steps=20, warmup=3, plain_rounds=10, seed=0, autotune=True,
                     tol_plain_rounds=2):
    """The batched configuration: `scenes` ensemble boxes stepped as one
    batch (one launch per kernel per substep, replayed from a hipGraph),
    against looping the plain step over the same scenes. Each scene takes its
    own reference schedule (dt_sub, R) and tilt; both legs run the same
    substeps, from the same initial state, inputs resident on the GPU."""
    ens = fa.scenes.ensemble(scenes, seed=seed)
    total = sum(sc.n for sc in ens)
    sched = [sc.schedule() for sc in ens]
    plan = []
    for k in range(warmup + max(steps, plain_rounds) + 1):
        plan.append(([sc.step_block(k * sd[1], sd[1], sd[3], seed=k) for sc, sd in zip(ens, sched)],
                     [sd[2] for sd in sched]))
    b = fa.Batch([sc.params for sc in ens])
    for i, sc in enumerate(ens):
        b.upload_state(i, sc.pos4)
    tuned = b.autotune(plan[0][0], plan[0][1], trials=2) if autotune else None
    b.use_graph(True)
    for k in range(warmup):
        b.substep(*plan[k])
    b.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        b.substep(*plan[k])
    b.synchronize()
    t_batch = time.perf_counter() - t0
    # the same batch with each scene's divergence and density solves iterated
    # to the configs' tolerances (0.1% and 0.01% mean error, at most 256
    # passes each), graph replay; then the plain loop in the same mode
    b.set_tolerance(TOL_DENSITY, 1, MAX_PASSES)
    b.set_div_tolerance(TOL_DIVERGENCE, MAX_PASSES)
    for k in range(warmup):
        b.substep(*plan[k])
    b.synchronize()
    tol_start = [b.state(i) for i in range(scenes)]
    b.pass_histogram(reset=True)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        b.substep(*plan[k])
    b.synchronize()
    t_tol = time.perf_counter() - t0
    hist = b.pass_histogram(reset=True)
    b.set_tolerance(0.0)
    b.set_div_tolerance(0.0)
    b.use_graph(False)
    b.profile(fa.KERNELS)
    b.substep(*plan[warmup + steps])
    prof = b.profile_read()
    b.profile([])
    desc = b.describe()
    b.close()
    sims = []
    for sc in ens:
        s = fa.Sim(sc.params)
        s.upload_state(sc.pos4)
        sims.append(s)
    for s, st, r in zip(sims, *plan[0]):
        s.substeps(st, 1, r)
    for s in sims:
        s.synchronize()
    t0 = time.perf_counter()
    for k in range(1, 1 + plain_rounds):
        for s, st, r in zip(sims, *plan[k]):
            s.substeps(st, 1, r)
    for s in sims:
        s.synchronize()
    t_plain = time.perf_counter() - t0
    # the plain loop in tolerance mode from the state the batch's timed
    # tolerance substeps started from, over the first of those substeps (each
    # scene runs the passes it ran in the batch)
    for s, st0 in zip(sims, tol_start):
        s.upload_state(st0["pos"], st0["vel"], st0["prev_vel"], st0["prev_pressure"], st0["temperature"])
        s.set_tolerance(TOL_DENSITY, 1, MAX_PASSES, TOL_DIVERGENCE, MAX_PASSES)
    del tol_start
    for s in sims:
        s.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + tol_plain_rounds):
        for s, st, r in zip(sims, *plan[k]):
            s.substeps(st, 1, r)
    for s in sims:
        s.synchronize()
    t_tol_plain = time.perf_counter() - t0
    for s in sims:
        s.close()
    batched = total * steps / t_batch
    plain = total * plain_rounds / t_plain
    tol_batched = total * steps / t_tol
    tol_plain = total * tol_plain_rounds / t_tol_plain

    canon = sum(canonical_bytes(sc.n, sc.grid.cell_count(), sd[2], count_fused=False) for sc, sd in zip(ens, sched))
    return {
        "metric": "DFSPH particle-steps/s, batched ensemble",
        "value": batched, "unit": "particle-substeps/s",
        "ms_per_step": t_batch * 1e3 / steps, "steps": steps, "warmup": warmup,
        "config": {"workload": f"ensemble_{scenes}", "scenes": scenes, "particles": total,
                   "counts": "log-uniform [2048, 65536]", "fill": "U(0.2, 0.5)", "seed": seed,
                   "refine": sorted(set(sd[2] for sd in sched)), "graph": True},
        "plain_loop": {"value": plain, "ms_per_step": t_plain * 1e3 / plain_rounds, "rounds": plain_rounds,
                       "what": "fluid_sim_substeps over each scene in turn, one stream, same substeps"},
        "batched_over_plain": batched / plain,
        "tolerance_mode": {"density_tol": TOL_DENSITY, "divergence_tol": TOL_DIVERGENCE, "max_passes": MAX_PASSES,
                           "value": tol_batched, "ms_per_step": t_tol * 1e3 / steps,
                           "density_passes": hist_summary(hist[0]), "divergence_passes": hist_summary(hist[1]),
                           "plain_loop": {"value": tol_plain, "ms_per_step": t_tol_plain * 1e3 / tol_plain_rounds,
                                          "rounds": tol_plain_rounds},
                           "batched_over_plain": tol_batched / tol_plain,
                           "what": "each scene's divergence and density solves iterated until its mean error is "
                                   "at most the tolerance; pass counts over every (scene, substep) timed"},
        "step_roofline": {"canonical_bytes": canon, "achieved_gbs": canon / (t_batch / steps) / 1e9,
                          "frac": canon / (t_batch / steps) / 1e9 / HBM_PEAK_GBS},
        "per_kernel_ms": {k: v[0] for k, v in prof.items()},
        "launch": desc, "autotuned_stage_cap": tuned,
    }


def cfl_leg(fa, sc, develop_s, frames=40):
    """The same scene on the reference's adaptive schedule: each 1/120 s
    frame takes substeps_for(dt, v_max of the frame before) substeps
    (render.rs:62-106), each at R = refine_passes(dt_sub) density passes
    (render.rs:77-79), from the seeded block run for `develop_s` seconds of
    simulated time first. Times `frames` whole frames, each with its stats
    reduction and the v_max readback the next frame's count needs."""
    sim = fa.Sim(sc.params)
    sim.upload_state(sc.pos4)
    state = {"t": 0.0, "v_max": 0.0}

    def frame():
        n, dt, r, vc = sc.schedule(state["v_max"])
        for _ in range(n):
            sim.substeps(sc.step_block(state["t"], dt, vc), 1, r)
            state["t"] += dt
        sim.reduce_stats()
        state["v_max"] = float(sim.read_stats()[6])
        return n, r

    while state["t"] < develop_s:
        frame()
    clamps0, _ = sim.counters()
    t0 = time.perf_counter()
    substeps, passes = 0, {}
    for _ in range(frames):
        n, r = frame()
        substeps += n
        passes[r] = passes.get(r, 0) + n
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    clamps1, _ = sim.counters()
    sim.close()
    return {"schedule": "reference adaptive: substeps_for(1/120 s, v_max of the previous frame), R = refine_passes(dt_sub)",
            "value": sc.n * substeps / elapsed, "unit": "particle-steps/s", "ms_per_step": elapsed * 1e3 / substeps,
            "frames": frames, "substeps": substeps, "substeps_by_refine": {str(k): v for k, v in sorted(passes.items())},
            "developed_for_s": develop_s, "speed_clamps": clamps1 - clamps0,
            "what": "wall time of whole frames (substeps, stats reduction, v_max readback)"}


def adjoint_leg(fa, n=65536, steps=64, every=8, seed=0):
    """The adjoint configuration: one 65,536-particle scene, 64 substeps with
    a seeded random-walk tilt of gravity, a checkpoint every 8 substeps; the
    forward run and the backward run (segment replays + reverse sweeps),
    each timed with HIP events on the adjoint's stream, for the reference's
    whole step and for the viscous profile (near pressure, cohesion,
    adhesion and the XSPH blend off). The loss seed is the centroid term's
    gradient (the cost does not depend on the seed)."""
    from fluid_amd import adjoint as fad
    sc = fa.scenes.fill_block("adjoint_64k", n, 0.3, seed=seed)
    _, dt, refine, vc = sc.schedule()
    rng = np.random.default_rng(seed)
    th = np.cumsum(rng.normal(0.0, 0.05, steps))
    grav = np.stack([9.81 * np.sin(th), -9.81 * np.cos(th), np.zeros(steps)], 1)
    v0 = np.zeros((sc.n, 4), np.float32)
    v0[:, :3] = rng.normal(0.0, 0.05, (sc.n, 3))
    blocks = fad.steps_for(grav, dt, vc)

    def timed(physics):
        adj = fad.Adjoint(sc.params, steps, every, physics=physics)
        adj.set_mu(1.002e-3)
        pos, _ = adj.forward(sc.pos4, v0, blocks, [refine] * steps)  # warm-up
        c_t = pos[:, :3].mean(0) + 0.01
        seed4 = np.zeros((sc.n, 4), np.float32)
        seed4[:, :3] = 2.0 * (pos[:, :3].mean(0) - c_t) / sc.n
        adj.backward(seed4)
        adj.forward(sc.pos4, v0, blocks, [refine] * steps)
        adj.backward(seed4)
        fwd_ms, bwd_ms = adj.timing()
        adj.close()
        return fwd_ms, bwd_ms

    fwd_ms, bwd_ms = timed(fad.FULL)
    vf, vb = timed(fad.VISCOUS)
    return {"config": {"workload": "adjoint_64k", "particles": sc.n, "substeps": steps, "checkpoint_every": every,
                       "refine": refine, "profile": "the whole step (viscosity, near pressure, cohesion, adhesion, "
                                                    "XSPH); gradients of v0, mu and the gravity vectors"},
            "forward_ms": fwd_ms, "backward_ms": bwd_ms, "adjoint_over_forward": bwd_ms / fwd_ms,
            "forward_particle_substeps_per_s": sc.n * steps / (fwd_ms * 1e-3),
            "viscous": {"forward_ms": vf, "backward_ms": vb, "adjoint_over_forward": vb / vf}}


def copy_calibration(nbytes: int, reps: int = 4) -> None:
    """Runs `reps` device-to-device copies of `nbytes` (HIP's own copy kernel)
    so that a counter run can read what FETCH_SIZE / WRITE_SIZE report for a
    plain copy of known size; the ratio scales the solver kernels' counts."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--workload", default="block_1m",
                    choices=["block_1m", "block_8m", "dam_break_16k", "ensemble_256"],
                    help="the plain step's scene (ensemble_256 only names the batched leg in summaries)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device copies of BYTES (counter calibration) and exit")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batched-only", action="store_true", help="print only the batched-ensemble line")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-cfl", action="store_true", help="skip the adaptive-schedule leg")
    ap.add_argument("--no-tolerance", action="store_true", help="skip the configs' tolerance leg of the scene")
    ap.add_argument("--adjoint-only", action="store_true", help="print only the adjoint line")
    ap.add_argument("--cpu-configs", action="store_true",
                    help="print only the oracle's CPU timings of configs 1-3 (the CPU baseline plan)")
    ap.add_argument("--trace-summary", metavar="DIR", default="",
                    help="print the kernel table of a rocprofv3 --kernel-trace --stats run under DIR/trace and exit")
    ap.add_argument("--tolerance-summary", metavar="DIR", default="",
                    help="print the tolerance leg's kernel table of a rocprofv3 --kernel-trace run under DIR/trace "
                         "and exit")
    ap.add_argument("--sq-summary", metavar="DIR", default="",
                    help="print the per-wave shader counters of the runs under DIR/sq_A, DIR/sq_B and exit")
    ap.add_argument("--pmc-summary", metavar="DIR", default="",
                    help="print the counter summary of rocprofv3 --pmc runs under DIR and exit")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary, args.workload, args.steps if args.steps != 300 else 0,
                      args.warmup if args.steps != 300 else 0)
        return
    if args.sq_summary:
        sq_summary(args.sq_summary, args.workload)
        return
    if args.tolerance_summary:
        tolerance_summary(args.tolerance_summary)
        return
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.workload, args.steps if args.steps != 300 else 0,
                    args.warmup if args.steps != 300 else 0)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the DFSPH step runs on a single MI355X, and no "
              "multi-GPU path is built.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import fluid_amd as fa
    if args.cpu_configs:
        print(json.dumps({"cpu_configs": cpu_configs(fa)}), flush=True)
        return
    if fa.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(fa, seed=args.seed)), flush=True)
        return
    if args.batched_only or args.workload == "ensemble_256":
        print(json.dumps(batched_ensemble(fa, scenes=args.scenes, seed=args.seed)), flush=True)
        return
    maker = {"block_1m": fa.scenes.block_1m, "block_8m": fa.scenes.block_8m,
             "dam_break_16k": fa.scenes.dam_break}[args.workload]
    sc = maker(seed=args.seed)
    n_sub, dt_sub, refine, v_clamp = sc.schedule()
    ntr = len(sc.tracers)
    sim = fa.Sim(sc.params, tracers=ntr)
    sim.upload_state(sc.pos4)
    if ntr:
        sim.upload_tracers(sc.tracers)

    # A bench step is one substep of the reference schedule from the seeded
    # block: `warmup` untimed substeps, `steps` timed ones, then 3 profiled.
    total = args.warmup + args.steps + 3 + TOL_SUBSTEPS
    steps = [sc.step_block(k * dt_sub, dt_sub, v_clamp) for k in range(total)]
    frame_blocks = [fa.pack_step(force=sc.gravity(k * dt_sub), dt=dt_sub * n_sub, seed=k) for k in range(total)]

    # warm-up substeps; the last 3 of them profile every kernel class to name
    # the dominant one
    for k in range(args.warmup):
        if k == args.warmup - 3:
            sim.synchronize()
            sim.profile(fa.KERNELS)
        sim.substeps(steps[k], 1, refine)
    pre = sim.profile_read()
    sim.profile([])
    dominant = max((k for k in pre if kernel_bytes(k, sc.n, sim.cells)), key=lambda k: pre[k][0])

    # the timed region: exactly `steps` substeps, the dominant kernel timed by
    # start/stop events carried on its own dispatches, one launch in every
    # DOMINANT_STRIDE (the events slow the stream: about 2% of a 1M substep
    # with every den_kappa launch timed)
    clamps0, overflow0 = sim.counters()
    sim.profile([dominant], stride=DOMINANT_STRIDE)
    sim.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        sim.substeps(steps[args.warmup + k], 1, refine)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    dom = sim.profile_read()[dominant]
    sim.profile([])
    dom_ms = dom[0] / max(dom[1], 1)
    clamps1, overflow1 = sim.counters()

    # every kernel class over 3 substeps just after the timed region
    sim.profile(fa.KERNELS)
    for k in range(3):
        sim.substeps(steps[args.warmup + args.steps + k], 1, refine)
    prof = sim.profile_read()
    sim.profile([])
    per_kernel_ms = {k: v[0] / 3.0 for k, v in prof.items()}
    launch_us = {k: v[0] * 1e3 / max(v[1], 1) for k, v in prof.items()}

    # per-frame work, timed on its own: reduce_stats + tracers
    sim.synchronize()
    t1 = time.perf_counter()
    frames = 5
    for k in range(frames):
        sim.reduce_stats()
        if ntr:
            sim.tracers(frame_blocks[k])
    sim.synchronize()
    frame_extra_ms = (time.perf_counter() - t1) * 1e3 / frames
    stats = sim.read_stats()

    # the same scene from where the timed run left it, both solves iterated
    # to the configs' tolerances (the configs' own reading of the step)
    tol_leg = None
    if not args.no_tolerance:
        sim.set_tolerance(TOL_DENSITY, 1, MAX_PASSES, TOL_DIVERGENCE, MAX_PASSES)
        sim.pass_histogram(reset=True)
        first = args.warmup + args.steps + 3
        sim.substeps(steps[first], 1, refine)  # untimed
        sim.synchronize()
        t1 = time.perf_counter()
        for k in range(1, TOL_SUBSTEPS):
            sim.substeps(steps[first + k], 1, refine)
        sim.synchronize()
        dt_tol = (time.perf_counter() - t1) / (TOL_SUBSTEPS - 1)
        h = sim.pass_histogram()
        sim.set_tolerance(0.0)
        tol_leg = {"density_tol": TOL_DENSITY, "divergence_tol": TOL_DIVERGENCE, "max_passes": MAX_PASSES,
                   "substeps": TOL_SUBSTEPS - 1, "ms_per_step": dt_tol * 1e3, "value": sc.n / dt_tol,
                   "density_passes": hist_summary(h[0]), "divergence_passes": hist_summary(h[1]),
                   "what": "the timed scene continued with both solves iterated until the mean error is at "
                           "most the tolerance (the histograms include the untimed first substep)"}

    ms_per_step = elapsed * 1e3 / args.steps
    value = sc.n * args.steps / elapsed
    b_dom = kernel_bytes(dominant, sc.n, sim.cells)
    achieved = b_dom / (dom_ms * 1e-3) / 1e9
    step_bytes = canonical_bytes(sc.n, sim.cells, refine)
    traffic = load_traffic() if sc.name == "block_1m" else {}
    kernels = {}
    for k, us in launch_us.items():
        kb = kernel_bytes(k, sc.n, sim.cells)
        kernels[k] = {"avg_launch_us": us, "launches_per_substep": prof[k][1] / 3.0,
                      "algorithmic_bytes": kb, "achieved_gbs": kb / (us * 1e-6) / 1e9 if kb else None,
                      "traffic_bytes": traffic.get(k),
                      "traffic_gbs": traffic[k] / (us * 1e-6) / 1e9 if k in traffic else None}
    out = {
        "metric": "DFSPH particle-steps/s @1M (plain vs batched) + adjoint/forward cost ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded jittered lattice block, tilt gravity",
        "config": {"workload": sc.name, "particles": sc.n, "cells": sim.cells, "refine": refine,
                   "substep_dt_s": dt_sub, "substeps_per_frame": n_sub, "h_over_spacing": sc.h / sc.spacing,
                   "tracers": ntr, "schedule": "reference fixed (1 divergence + warm start + R density)"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic.get(dominant),
                     "traffic_gbs": traffic[dominant] / (dom_ms * 1e-3) / 1e9 if dominant in traffic else None,
                     "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) if dominant in traffic else None,
                     "avg_launch_us": dom_ms * 1e3, "algorithmic_bytes": b_dom},
        "kernels": kernels,
        "step_roofline": {"canonical_bytes": step_bytes, "achieved_gbs": step_bytes / (ms_per_step * 1e-3) / 1e9,
                          "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "per_kernel_ms": per_kernel_ms,
        "launch": sim.describe(),
        "frame_extra_ms": frame_extra_ms,
        "timed_counters": {"speed_clamps": clamps1 - clamps0, "list_overflow_pairs": overflow1 - overflow0},
        "stats": {"v_max": float(stats[6]), "rho_min": float(stats[2]), "rho_max": float(stats[3]),
                  "compression_avg": float(stats[0])},
    }
    if tol_leg:
        out["tolerance_mode"] = tol_leg
    if not args.no_cfl:
        out["cfl_schedule"] = cfl_leg(fa, sc, develop_s=(args.warmup + args.steps) * dt_sub)
    if not args.no_batched:
        out["batched"] = batched_ensemble(fa, scenes=args.scenes, seed=args.seed)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(fa, seed=args.seed)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, steps[0], refine)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
