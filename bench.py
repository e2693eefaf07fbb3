"""Benchmark of the MI355X TMz stepper.

Default workload: the 8192x8192 single grid (CPML 16, 16-element steered
array, parabolic plasma with 2% Gaussian fluctuations, 64 detectors), the
configuration the Gcell-updates/s metric is quoted on. A "step" is one FDTD
time step over the whole grid; inputs are resident in HBM before the timed
region, which brackets exactly --steps steps with device synchronisation on
both sides.

Prints ONE JSON line (rank 0). Besides the metric it carries
  roofline      the interior step kernel against the 8 TB/s HBM peak. A
                launch advances its cells T steps in registers but must move
                each cell's state through HBM once: 36 B per cell (read and
                write Ez, Hx, Hy, Jz; read the 4-byte medium word).
                `achieved` = 36 B x interior cells x launches / summed kernel
                time (HIP events on the plan's stream), `frac` = achieved /
                peak, `traffic` = counter-measured HBM bytes per launch.
                `step_equiv_gbs` = 36 B x interior cells x steps / the same
                time: the rate a kernel moving every step's state through
                HBM would need (above the peak: temporal blocking);
  cpu_baseline  the single-threaded FP32 CPU oracle on a bounded sample of the
                same grid (a few steps), timed on this host; "all_cores"
                gives the same sample on up to 16 threads.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--tblock T]
                  [--cpu-steps S] [--no-cpu] [--traffic-file F]

Profiling helpers (no bench line):
  --copy-probe MIB      copy a MIB-MiB buffer 5 times with torch (the plain
                        copy kernel that calibrates rocprofv3's byte counters)
  --pmc-summary FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV
                        turn four `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE`
                        counter_collection.csv files (bench run, copy probe)
                        into the traffic file and a text summary on stdout
  --stats-summary KERNEL_STATS_CSV
                        print a `rocprofv3 --stats` kernel table as text
  --pmc-total FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV
                        HBM bytes summed over every pool and step kernel
                        dispatch of two counter runs (a serving window, say),
                        calibrated like --pmc-summary, over the summed
                        durations of those dispatches
  --sq-summary CSV [CSV ...]
                        per-dispatch means of the SQ_* / GRBM_* counters of
                        `rocprofv3 --pmc` runs for the single-grid step
                        kernel, with the VALU issue fraction derived
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
# launch shape of a serving pool whose stream has no cached shape near its
# mix (pool options): the pool's defaults, the batch mix's tuned shape
SERVE_SHAPE = {"fused": 0, "tblock": 8, "cols_per_lane": 1, "band": 128, "wg_per_cu": 3, "band_cpml": 64,
               "wg_per_cu_cpml": 2, "prefetch": 1}
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 36  # r+w of Ez, Hx, Hy, Jz + one 4-byte medium word


def host_cpu() -> dict:
    """CPU model and logical CPU count of this host (for the CPU baseline)."""
    model = "unknown"
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": affinity, "cpu_count": os.cpu_count()}


def trace_check(det) -> tuple:
    """(ok, message): every detector trace finite and not all zero."""
    import numpy as np

    finite = bool(np.isfinite(det).all())
    zero = int((np.abs(det[1:]).max(axis=0) == 0).sum()) if det.size else 0
    return finite and zero == 0, f"finite: {finite}, all-zero detector traces: {zero} of {det.shape[1]}"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def copy_probe(mib: int) -> int:
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: {mib} MiB x 5")
    return 0


def _read_counter(path: str, counter: str) -> dict:
    """Mean counter value (KiB per dispatch) per kernel name."""
    import csv

    acc: dict = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            v = acc.setdefault(row["Kernel_Name"], [0.0, 0])
            v[0] += float(row["Counter_Value"])
            v[1] += 1
    return {k: v[0] / v[1] for k, v in acc.items()}


def _pick(d: dict, pred) -> tuple:
    names = [k for k in d if pred(k)]
    if not names:
        raise SystemExit("no matching kernel in counter file")
    k = max(names, key=lambda n: d[n])
    return k, d[k]


def _step_kernel_mode(name: str) -> int:
    """Last template argument of a single-grid step kernel's name
    (step_kernel<T, C, D, medium, mode>), -1 for other kernels."""
    if "step_kernel<" not in name:
        return -1
    args = name.split("step_kernel<", 1)[1].split(">", 1)[0].split(",")
    try:
        return int(args[-1])
    except ValueError:
        return -1


def pmc_summary(paths: list, copy_mib: int, traffic_file: str) -> int:
    fetch = _read_counter(paths[0], "FETCH_SIZE")
    write = _read_counter(paths[1], "WRITE_SIZE")
    cfetch = _read_counter(paths[2], "FETCH_SIZE")
    cwrite = _read_counter(paths[3], "WRITE_SIZE")
    # the single-grid kernel that covers the interior: the fused one (mode 2),
    # or the interior one of a two-kernel launch (mode 0), not the CPML-class
    # kernel beside it (mode 1)
    kname, fkib = _pick(fetch, lambda n: _step_kernel_mode(n) in (0, 2))
    wkib = write[kname]
    cname, cf = _pick(cfetch, lambda n: "copy" in n.lower() or "elementwise" in n.lower())
    _, cw = _pick(cwrite, lambda n: "copy" in n.lower() or "elementwise" in n.lower())
    copy_bytes = copy_mib * (1 << 20)
    fscale = copy_bytes / (cf * 1024.0)
    wscale = copy_bytes / (cw * 1024.0)
    fb = fkib * 1024.0 * fscale
    wb = wkib * 1024.0 * wscale
    rec = {
        "kernel": kname,
        "fetch_kib_raw": round(fkib, 1),
        "write_kib_raw": round(wkib, 1),
        "copy_kernel": cname,
        "copy_bytes": copy_bytes,
        "copy_fetch_kib_raw": round(cf, 1),
        "copy_write_kib_raw": round(cw, 1),
        "fetch_scale": round(fscale, 4),
        "write_scale": round(wscale, 4),
        "fetch_bytes_per_launch": round(fb),
        "write_bytes_per_launch": round(wb),
        "bytes_per_launch": round(fb + wb),
    }
    Path(traffic_file).write_text(json.dumps(rec, indent=1) + "\n")
    print(f"kernel: {kname}")
    print(f"FETCH_SIZE {fkib:.1f} KiB/dispatch x {fscale:.3f} = {fb / 1e6:.2f} MB")
    print(f"WRITE_SIZE {wkib:.1f} KiB/dispatch x {wscale:.3f} = {wb / 1e6:.2f} MB")
    print(f"HBM traffic per launch: {(fb + wb) / 1e6:.2f} MB")
    print(f"calibration: {cname} copying {copy_bytes} B reads FETCH {cf:.1f} KiB, WRITE {cw:.1f} KiB")
    return 0


def pmc_total(paths: list, copy_mib: int) -> int:
    """Sum of calibrated FETCH_SIZE and WRITE_SIZE bytes over every dispatch
    of the stepper's kernels, per kernel and in total, with the summed
    dispatch durations (under --pmc rocprofv3 runs dispatches one at a time,
    so the sum is the time the device spent in them)."""
    import csv

    def per_dispatch(path, counter):
        out = {}
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                name = row["Kernel_Name"]
                if "fdtd2d::" not in name:
                    continue
                short = name.replace("(anonymous namespace)::", "").split("(")[0].split("::")[-1]
                v = out.setdefault(short, [0.0, 0, 0])
                v[0] += float(row["Counter_Value"])
                v[1] += 1
                v[2] += int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
        return out

    copy_bytes = copy_mib * (1 << 20)
    _, cf = _pick(_read_counter(paths[2], "FETCH_SIZE"), lambda n: "copy" in n.lower() or "elementwise" in n.lower())
    _, cw = _pick(_read_counter(paths[3], "WRITE_SIZE"), lambda n: "copy" in n.lower() or "elementwise" in n.lower())
    fscale, wscale = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    fetch, write = per_dispatch(paths[0], "FETCH_SIZE"), per_dispatch(paths[1], "WRITE_SIZE")
    tot_b = tot_ns = 0.0
    print(f"{'dispatches':>10} {'fetch GB':>9} {'write GB':>9} {'ms':>9} {'GB/s':>7}  kernel")
    for k in sorted(set(fetch) | set(write)):
        fb = fetch.get(k, [0, 0, 0])[0] * 1024 * fscale
        wb = write.get(k, [0, 0, 0])[0] * 1024 * wscale
        n = max(fetch.get(k, [0, 0, 0])[1], write.get(k, [0, 0, 0])[1])
        ns = (fetch.get(k, [0, 0, 0])[2] + write.get(k, [0, 0, 0])[2]) / 2
        tot_b += fb + wb
        tot_ns += ns
        print(f"{n:>10} {fb / 1e9:>9.2f} {wb / 1e9:>9.2f} {ns / 1e6:>9.1f} {(fb + wb) / max(ns, 1):>7.0f}  {k}")
    print(f"total: {tot_b / 1e9:.2f} GB in {tot_ns / 1e6:.1f} ms of dispatches: {tot_b / max(tot_ns, 1):.0f} GB/s "
          f"(calibration: {fscale:.3f} x FETCH_SIZE, {wscale:.3f} x WRITE_SIZE)")
    return 0


def sq_summary(paths: list, simds: int = 1024, xcds: int = 8, match: str | None = None) -> int:
    """SQ_* and GRBM_* counter means per dispatch of the single-grid step
    kernel (mode 0 or 2; or of each kernel whose name contains `match`) from
    one or more counter_collection.csv files, and derived figures: VALU
    issue = SQ_INSTS_VALU over the one wave64 VALU instruction per SIMD every
    4 cycles that `simds` SIMDs can issue in GRBM_GUI_ACTIVE / xcds cycles
    (GRBM_GUI_ACTIVE counts on every XCD); the fraction of wave time waiting
    (SQ_WAIT_ANY, SQ_WAIT_INST_ANY over SQ_WAVE_CYCLES, all in the same
    4-cycle units)."""
    import csv

    per: dict = {}
    for path in paths:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                name = row["Kernel_Name"]
                if match is None:
                    if _step_kernel_mode(name) not in (0, 2):
                        continue
                elif match not in name:
                    continue
                acc = per.setdefault(name, {})
                v = acc.setdefault(row["Counter_Name"], [0.0, 0, 0])
                v[0] += float(row["Counter_Value"])
                v[1] += 1
                v[2] += int(row["End_Timestamp"]) - int(row["Start_Timestamp"])
    if not per:
        raise SystemExit("no matching kernel dispatch in the counter files")
    for kname in sorted(per):
        _sq_print(kname, per[kname], simds, xcds)
    return 0


def _sq_print(kname: str, acc: dict, simds: int, xcds: int) -> None:
    mean = {k: v[0] / v[1] for k, v in acc.items()}
    total = {k: v[0] for k, v in acc.items()}
    n = max(v[1] for v in acc.values())
    dur_ms = max(v[2] for v in acc.values()) / max(n, 1) / 1e6
    print(f"kernel: {kname}")
    print(f"  dispatches {n}, mean duration {dur_ms:.4f} ms")
    for k in sorted(mean):
        print(f"  {k:<22} {mean[k]:.4g} per dispatch, {total[k]:.4g} in all")
    if "GRBM_GUI_ACTIVE" in mean:
        cyc = mean["GRBM_GUI_ACTIVE"] / xcds
        print(f"  cycles per XCD        {cyc:.4g}")
        if "SQ_INSTS_VALU" in mean:
            print(f"  VALU issue            {mean['SQ_INSTS_VALU'] / (simds * cyc / 4):.3f} "
                  f"(of {simds} SIMDs x 1 per 4 cycles)")
    if "SQ_WAVE_CYCLES" in mean:
        for w in ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY"):
            if w in mean:
                print(f"  {w} / SQ_WAVE_CYCLES {mean[w] / mean['SQ_WAVE_CYCLES']:.3f}")


def stats_summary(path: str) -> int:
    import csv

    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    print(f"{'calls':>7} {'avg_us':>10} {'min_us':>10} {'max_us':>10} {'pct':>6}  kernel")
    for r in rows:
        print(f"{r['Calls']:>7} {float(r['AverageNs']) / 1e3:>10.1f} {float(r['MinNs']) / 1e3:>10.1f} "
              f"{float(r['MaxNs']) / 1e3:>10.1f} {float(r['Percentage']):>6.2f}  {r['Name']}")
    return 0


def busy_summary(path: str, match: str = "pool_kernel", gap_ms: float = 1.0) -> dict:
    """From a rocprofv3 kernel trace (kernel_trace.csv): the union of the
    intervals of the kernels whose name contains `match` over the span from
    the first to the last of them, the idle share of that span, and the gaps
    between them longer than `gap_ms`."""
    import csv

    iv = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if match in r["Kernel_Name"]:
                iv.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    iv.sort()
    merged = []
    for a, b in iv:
        if merged and a <= merged[-1][1]:
            merged[-1][1] = max(merged[-1][1], b)
        else:
            merged.append([a, b])
    if not merged:
        return {"kernels": 0}
    busy = sum(b - a for a, b in merged)
    span = merged[-1][1] - merged[0][0]
    gaps = sorted((merged[k + 1][0] - merged[k][1] for k in range(len(merged) - 1)), reverse=True)
    long = [g for g in gaps if g > gap_ms * 1e6]
    out = {"kernels": len(iv), "span_s": span / 1e9, "busy_s": busy / 1e9,
           "idle_frac": 1 - busy / span if span > 0 else 0.0,
           "gaps_over": {"ms": gap_ms, "count": len(long), "sum_s": sum(long) / 1e9,
                         "largest_ms": [round(g / 1e6, 1) for g in long[:5]]}}
    print(json.dumps(out))
    return out


def bench_batch(args) -> int:
    """The 256-shot uneven mix (seed 3): every shot through one pool (persistent
    kernels, graph chunks, autotuned shape) against looping the plain
    single-grid stepper over the same shots, with the batched detector traces
    checked bitwise against the plain ones."""
    import numpy as np

    import fdtd2d_amd as F
    from fdtd2d_amd.batch import build_setups, run_pool
    from fdtd2d_amd.tuning import ShapeCache

    t0 = time.time()
    setups = build_setups(3, 3, range(args.shots))
    probs = [s.c_problem.contents for s in setups]
    cells = sum(float(p.nx) * p.ny * p.nsteps for p in probs)
    log(f"batch: {len(setups)} shots, {cells / 1e12:.3f} T cell updates, set up in {time.time() - t0:.1f}s")

    cache = None if args.shape_cache == "off" else ShapeCache()
    pool, ids, st = run_pool(setups, 0, tune=not args.no_tune, log=log, cache=cache,
                             lookup=args.shape_cache == "use", warmup=True)
    det_b = [pool.read_detectors(i)[0] for i in ids]
    pool.close()
    tb = st["seconds"]
    log(f"batched: {tb:.3f}s, {cells / tb / 1e9:.1f} Gcell/s, shape {st['tuned']['shape'] if st['tuned'] else None}, "
        f"graph replays {st['graph_replays']:.0f}, eager {st['eager_launches']:.0f}, tails {st['tail_launches']:.0f}")

    tp = None
    same = None
    if not args.no_plain:
        tp = 0.0
        same = True
        nonfinite = 0
        for k, s in enumerate(setups):
            with F.Plan(s, 0) as plan:
                plan.sync()
                c0 = time.perf_counter()
                plan.run(probs[k].nsteps)
                plan.sync()
                tp += time.perf_counter() - c0
                d, _ = plan.read_detectors()
            # bit patterns, so that equal NaNs compare equal too
            same = same and bool(np.array_equal(d.view(np.uint32), det_b[k].view(np.uint32)))
            nonfinite += int((~np.isfinite(d)).sum())
        log(f"plain loop: {tp:.3f}s, {cells / tp / 1e9:.1f} Gcell/s; batched traces bitwise equal: {same}; "
            f"non-finite trace values: {nonfinite}")
    bad = [k for k, d in enumerate(det_b) if not trace_check(d)[0]]
    if bad:
        log(f"batch: shots {bad[:10]} have non-finite or all-zero detector traces")

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle

        # a bounded sample (about 10 s single-threaded): the first 32 shots of
        # the mix for 64 steps each, scaled to whole shots of the mix
        nshot, nstep = min(32, len(setups)), 64
        ssum, csum = 0.0, 0.0
        for s in setups[:nshot]:
            ref = Oracle(s, precision=32, threads=1)
            c0 = time.perf_counter()
            ref.run(nstep)
            ssum += time.perf_counter() - c0
            csum += float(s.c_problem.contents.nx) * s.c_problem.contents.ny * nstep
            ref.close()
        rate = csum / ssum
        cpu = {"value": round(rate / (cells / len(setups)), 6), "unit": "sims/s", "cores": 1, "kind": "port",
               "sample": f"FP32 CPU oracle, shots 0-{nshot - 1} of the mix x {nstep} steps ({ssum:.1f} s, single "
                         f"thread, {rate / 1e9:.3f} Gcell/s), scaled to whole shots"}
        # the batched detector traces of the smallest shots, whole runs,
        # against the FP64 oracle (all host cores)
        small = sorted(range(len(setups)), key=lambda k: probs[k].nx * probs[k].ny)[: args.fp64_shots]
        if small:
            c0 = time.perf_counter()
            worst = 0.0
            for k in small:
                ref = Oracle(setups[k], precision=64, threads=max(1, host_cpu()["nproc"]))
                ref.run(probs[k].nsteps)
                worst = max(worst, rel_l2(det_b[k], ref.read_detectors()[0]))
                ref.close()
            cpu["fp64_check"] = {"shots": small, "steps": "all", "worst_det_rel_l2": float(f"{worst:.3e}"),
                                 "bound": 1e-5, "seconds": round(time.perf_counter() - c0, 1)}
            log(f"FP64 oracle, shots {small}: worst batched detector rel L2 {worst:.3e}")
    line = {
        "metric": "batched sims/s",
        "value": round(len(setups) / tb, 3),
        "unit": "sims/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(setups)} uneven shots (seed 3): 256-2048 per side, 8-40 GHz, 1-16 elements, "
                               "linear/parabolic plasma, 0-5% fluctuation, CPML 12, 4000 steps",
                   "cell_updates": cells, "parallelism": "single GPU"},
        "batched": {"seconds": round(tb, 4), "gcell_per_s": round(cells / tb / 1e9, 2),
                    "shape": st["tuned"]["shape"] if st["tuned"] else None,
                    "shape_source": st["tuned"]["source"] if st["tuned"] else "pool defaults",
                    "mix_key": st["tuned"]["key"] if st["tuned"] else None,
                    "tune_trials": len(st["tuned"]["trials"]) if st["tuned"] else 0,
                    "warmup": "shot 0 alone for 16 launches at the same shape, untimed" if st["warmup"] else None,
                    "graph_replays": st["graph_replays"], "eager_launches": st["eager_launches"]},
        "plain": None if tp is None else {"seconds": round(tp, 4), "sims_per_s": round(len(setups) / tp, 3),
                                          "gcell_per_s": round(cells / tp / 1e9, 2)},
        "speedup_vs_plain": None if tp is None else round(tp / tb, 3),
        "traces_bitwise_equal_to_plain": same,
        "traces_checked": {"shots": len(det_b), "failed": len(bad)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0 if same is not False and not bad else 1


def rel_l2(a, b) -> float:
    """Relative L2 error of a against b, in float64 (0 when both are 0)."""
    import numpy as np

    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float(np.sqrt((b * b).sum()))
    num = float(np.sqrt(((a - b) ** 2).sum()))
    return num / den if den > 0 else (0.0 if num == 0 else float("inf"))


def slicing_options(args) -> dict:
    """Pool time-slicing options of a serving run: the 16384^2 shots of the
    mixed stream run only in every 16th launch by default, so that the
    1024^2 shots do not wait for a full large launch at every step, and the
    small shots take shorter bands meanwhile; the sweep stream, whose shots
    are alike, runs unsliced unless asked."""
    every = args.slice_every if args.slice_every >= 0 else (16 if args.mode == "serve-mixed" else 1)
    on = every > 1
    return {"slice_every": every, "slice_cells": args.slice_cells if on else 0,
            "band_small": args.band_small if on else 0, "band_cpml_small": args.band_cpml_small if on else 0}


def bench_serve(args) -> int:
    """Served streams through the continuous-batching loop: frequency sweeps
    (--mode serve-sweep) or the mixed 1024^2 / 16384^2 stream (--mode
    serve-mixed), at --rate requests/s for --duration seconds of arrivals."""
    from fdtd2d_amd.serve import Server, latency_summary, mixed_requests, request_mix_key, sweep_requests
    from fdtd2d_amd.tuning import ShapeCache

    t0 = time.time()
    if args.mode == "serve-sweep":
        rate = args.rate or 200.0
        reqs = sweep_requests(rate, args.duration, seed=4)
        budget = int(160e9)
        classes = ["sweep"]
    else:
        rate = args.rate or 0.3
        reqs = mixed_requests(rate, args.duration, seed=5)
        budget = int(240e9)
        classes = ["1024", "16384"]
    nshots = sum(len(r.shots) for r in reqs)
    log(f"{args.mode}: {len(reqs)} requests ({nshots} shots) at {rate} req/s over {args.duration}s; "
        f"drawn in {time.time() - t0:.1f}s")
    srv = Server(0, budget_bytes=budget)
    key = request_mix_key(reqs)
    if args.pool_shape:
        shape, source = json.loads(args.pool_shape), "--pool-shape"
    else:
        # the shape cached for the nearest shot mix, else the batch mix's
        shape, source = (None, None) if args.shape_cache == "off" else \
            ShapeCache().lookup("pool", key, int(srv.pool.info("cus")), nearest=True)
        if shape is None:
            shape, source = dict(SERVE_SHAPE), "built-in default"
    for name, value in shape.items():
        srv.pool.set_option(name, value)
    log(f"pool shape {shape} ({source}; stream mix {key})")
    slicing = slicing_options(args)
    for name, value in slicing.items():
        srv.pool.set_option(name, value)
    done, st = srv.run(reqs, max_seconds=args.duration + args.drain, log=log)
    srv.close()
    lat = latency_summary(done, classes)
    wall = st["wall"]
    line = {
        "metric": "p50/p99 request latency",
        "value": None if lat["all"] is None else round(lat["all"]["p50"], 4),
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.mode, "rate_req_per_s": rate, "arrival_window_s": args.duration,
                   "requests": len(reqs), "shots": nshots, "budget_bytes": budget, "parallelism": "single GPU"},
        "pool_shape": shape,
        "shape_source": source,
        "mix_key": key,
        "time_slicing": slicing,
        "latency_s": lat,
        "completed_requests": len(done),
        "unfinished_requests": st["unfinished_requests"],
        "sims_per_s": round(st["shots"] / wall, 3),
        "gcell_per_s": round(st["cell_updates"] / wall / 1e9, 2),
        "hbm_gbs_algorithmic": round(36 * st["cell_updates"] / wall / 1e9, 1),
        "wall_s": round(wall, 2),
        "host_s": st["host_s"],
        "pool_host_s": st["pool_host_s"],
        "pool_counts": st["pool_counts"],
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    ap.add_argument("--pmc-summary", nargs=4, default=None)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--stats-summary", default=None)
    ap.add_argument("--busy-summary", default=None, metavar="KERNEL_TRACE_CSV",
                    help="busy and idle time of the pool kernels in a rocprofv3 kernel trace")
    ap.add_argument("--busy-kernel", default="pool_kernel")
    ap.add_argument("--sq-summary", nargs="+", default=None)
    ap.add_argument("--sq-kernel", default=None, help="with --sq-summary: every kernel whose name contains this")
    ap.add_argument("--pmc-total", nargs=4, default=None)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--tblock", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--fp64-steps", type=int, default=512,
                    help="single grid, CPU leg: steps of the FP64 oracle the GPU traces are checked against")
    ap.add_argument("--fp64-shots", type=int, default=4,
                    help="batch, CPU leg: smallest shots checked whole against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "traffic_8192.json"))
    ap.add_argument("--mode", choices=["single", "batch", "serve-sweep", "serve-mixed"], default="single")
    ap.add_argument("--rate", type=float, default=0.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--drain", type=float, default=120.0)
    ap.add_argument("--slice-every", type=int, default=-1,
                    help="serving: shots of at least --slice-cells cells run in every Nth launch only "
                         "(1 = off; default 16 for serve-mixed, off for serve-sweep)")
    ap.add_argument("--slice-cells", type=float, default=16e6)
    ap.add_argument("--band-small", type=int, default=64,
                    help="serving with time slicing: band rows of the shots that are not sliced")
    ap.add_argument("--band-cpml-small", type=int, default=32,
                    help="serving with time slicing: CPML sub-band rows of the shots that are not sliced")
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--plan-shape", choices=["cache", "tune", "defaults"], default="cache",
                    help="single grid: take the launch shape cached for the grid (cache; the library's defaults "
                         "on a miss), tune it on the live plan on a miss and cache it (tune), or keep the "
                         "library's defaults (defaults)")
    ap.add_argument("--shape-cache", choices=["use", "retune", "off"], default="use",
                    help="batch and serving pools: take the launch shape cached for the shot mix (use), run the "
                         "autotuner and cache its result (retune), or run it without the cache (off)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--pool-shape", default=None,
                    help="launch shape of the serving pool, as JSON pool options (default: the shape cached for "
                         "the nearest shot mix, else a built-in one)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary, args.copy_mib, args.traffic_file)
    if args.stats_summary:
        return stats_summary(args.stats_summary)
    if args.busy_summary:
        busy_summary(args.busy_summary, match=args.busy_kernel)
        return 0
    if args.sq_summary:
        return sq_summary(args.sq_summary, match=args.sq_kernel)
    if args.pmc_total:
        return pmc_total(args.pmc_total, args.copy_mib)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(a single grid would shard in row slabs with halo exchange, batched and served "
              "shots by shot; neither is built)", file=sys.stderr)
        return 2
    if args.mode == "batch":
        return bench_batch(args)
    if args.mode.startswith("serve"):
        return bench_serve(args)

    import fdtd2d_amd as F
    from fdtd2d_amd.tuning import ShapeCache, plan_key, plan_shape

    total = args.warmup + args.steps
    # plan trials (--plan-shape tune on a cache miss) advance the real grid
    trial_steps = 1600 if args.plan_shape == "tune" else 0
    t0 = time.time()
    setup = F.Setup.workload(2, seed=args.seed, steps=max(total + trial_steps, args.cpu_steps))
    p = setup.c_problem.contents
    nx, ny = p.nx, p.ny
    log(f"setup {nx}x{ny} P={p.npml} nsrc={p.nsrc} ndet={p.ndet} in {time.time() - t0:.1f}s")

    plan = F.Plan(setup, device=0, tblock=args.tblock)
    shape_rec = plan_shape(plan, plan_key(nx, ny, p.npml, p.medium_kind), int(plan.info("cus")),
                           None if args.plan_shape == "defaults" else ShapeCache(), tune=args.plan_shape == "tune",
                           log=log)
    log(f"plan shape {shape_rec['shape']} ({shape_rec['source']}, key {shape_rec['key']})")
    T = int(plan.info("tblock"))
    log(f"plan: pitch={plan.info('pitch'):.0f} tblock={T} band={plan.info('band'):.0f} "
        f"workgroups={plan.info('workgroups'):.0f} bytes={plan.info('bytes') / 1e9:.2f} GB")
    plan.run(args.warmup)
    plan.sync()
    plan.time_kernels(True)
    plan.sync()
    t_start = time.perf_counter()
    plan.run(args.steps)
    plan.sync()
    elapsed = time.perf_counter() - t_start
    avg_ms, nlaunch = plan.kernel_ms(0)
    cpml_ms, _ = plan.kernel_ms(1)
    fused = int(plan.info("fused"))
    cells_cpml = plan.info("cells_cpml")
    # the timed kernel: the fused one covers every cell, else the interior one
    cells_int = plan.info("cells_interior") + (cells_cpml if fused else 0)
    plan_c = int(plan.info("cols_per_lane"))
    plan_d = int(plan.info("prefetch"))
    band = int(plan.info("band"))
    kernel_sig = f"step_kernel<{T}, {plan_c}, {plan_d}, 1, {2 if fused else 0}>"
    det, _ = plan.read_detectors()
    plan.close()
    # only the rows of the steps run so far can hold samples
    ok, check_msg = trace_check(det[: args.warmup + args.steps + 1])
    cells = nx * ny
    value = cells * args.steps / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    # The run splits --steps into launches of min(T, steps left); each
    # launch moves the state of its cells through HBM once, whatever its
    # step count.
    kernel_s = avg_ms * nlaunch * 1e-3
    launch_bytes = BYTES_PER_CELL_UPDATE * cells_int
    achieved = launch_bytes * nlaunch / kernel_s / 1e9 if kernel_s > 0 else 0.0
    step_equiv = BYTES_PER_CELL_UPDATE * cells_int * args.steps / kernel_s / 1e9 if kernel_s > 0 else 0.0
    traffic = None
    tf = Path(args.traffic_file)
    if tf.exists():
        try:
            rec = json.loads(tf.read_text())
            # only a record taken on this kernel describes it
            if kernel_sig in rec.get("kernel", ""):
                traffic = rec.get("bytes_per_launch")
        except Exception:
            traffic = None
    log(f"GPU: {value:.1f} Gcell/s, {ms_per_step:.3f} ms/step, {kernel_sig} avg {avg_ms:.4f} ms over {nlaunch} "
        f"launches ({cells_int / cells:.3f} of the grid): {achieved:.0f} GB/s of one pass per launch "
        f"(frac {achieved / HBM_PEAK_GBS:.3f}), single-step equivalent {step_equiv:.0f} GB/s; traces {check_msg}")

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle

        ref = Oracle(setup, precision=32, threads=1)
        c0 = time.perf_counter()
        ref.run(args.cpu_steps)
        cs = time.perf_counter() - c0
        hc = host_cpu()
        cpu = {
            "value": round(cells * args.cpu_steps / cs / 1e9, 4),
            "unit": "Gcell-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"FP32 CPU oracle, {nx}x{ny}, {args.cpu_steps} steps ({cs:.1f} s, single thread)",
            "cpu_model": hc["model"],
            "nproc": hc["nproc"],
        }
        ref.close()
        # the same sample on the host cores a one-GPU job gets (rows split
        # over threads, bitwise the single-threaded result)
        # every CPU this process may run on (the box gives a one-GPU job 16)
        nthr = max(1, hc["nproc"])
        ref = Oracle(setup, precision=32, threads=nthr)
        c0 = time.perf_counter()
        ref.run(args.cpu_steps)
        cs_all = time.perf_counter() - c0
        ref.close()
        cpu["all_cores"] = {"value": round(cells * args.cpu_steps / cs_all / 1e9, 4), "cores": nthr}
        log(f"CPU oracle: {cpu['value']} Gcell/s single-threaded, {cpu['all_cores']['value']} on {nthr} threads")
        # the GPU's detector traces against the FP64 oracle of the same grid
        # over its first --fp64-steps steps (all host cores)
        k = min(args.fp64_steps, total)
        if k > 0:
            ref = Oracle(setup, precision=64, threads=nthr)
            c0 = time.perf_counter()
            ref.run(k)
            det64 = ref.read_detectors()[0][: k + 1]
            ref.close()
            err = rel_l2(det[: k + 1], det64)
            cpu["fp64_check"] = {"steps": k, "det_rel_l2": float(f"{err:.3e}"), "bound": 1e-5,
                                 "seconds": round(time.perf_counter() - c0, 1)}
            log(f"FP64 oracle, first {k} steps: GPU detector traces rel L2 {err:.3e}")

    line = {
        "metric": "Gcell-updates/s",
        "value": round(value, 2),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "8192x8192 TMz FDTD single grid, CPML 16, 16-element 15-degree steered array, "
                        "parabolic plasma x(1+0.02g), 64 detectors",
            "nx": nx,
            "ny": ny,
            "steps_per_launch": T,
            "launch_shape": {"cols_per_lane": plan_c, "band": band, "prefetch": plan_d, "fused": fused},
            "shape_source": shape_rec["source"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_frac": None if traffic is None else round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernel": f"{'fused' if fused else 'interior'} step kernel {kernel_sig}",
            "kernel_avg_ms": round(avg_ms, 4),
            "launches": nlaunch,
            "cells_per_launch": int(cells_int),
            "bytes_per_launch": int(launch_bytes),
            "step_equiv_gbs": round(step_equiv, 1),
            "cpml_cells_per_launch": int(cells_cpml),
            "cpml_kernel": None if fused else {
                "avg_ms": round(cpml_ms, 4),
                "gbs_one_pass": round(BYTES_PER_CELL_UPDATE * cells_cpml / (cpml_ms * 1e-3) / 1e9, 1)
                if cpml_ms > 0 else None},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if not ok:
        log(f"bench: detector traces failed the check ({check_msg})")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
