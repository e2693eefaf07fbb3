"""Benchmark of the MI355X TMz stepper.

Default workload: the 8192x8192 single grid (CPML 16, 16-element steered
array, parabolic plasma with 2% Gaussian fluctuations, 64 detectors), the
configuration the Gcell-updates/s metric is quoted on. A "step" is one FDTD
time step over the whole grid; inputs are resident in HBM before the timed
region, which brackets exactly --steps steps with device synchronisation on
both sides.

Prints ONE JSON line (rank 0). Besides the metric it carries
  roofline      the step kernel's algorithmic bytes per launch (36 B per cell
                update) over its average launch duration, timed with HIP
                events on the plan's stream, against the 8 TB/s HBM peak;
  cpu_baseline  the single-threaded FP32 CPU oracle on a bounded sample of the
                same grid (a few steps), timed on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--tblock T]
                  [--cpu-steps S] [--no-cpu] [--traffic-file F]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 36  # r+w of Ez, Hx, Hy, Jz + one 4-byte medium word


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--tblock", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-file", default=str(ROOT / "profiles" / "traffic_8192.json"))
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(shots shard across GPUs by construction, see the design notes)", file=sys.stderr)
        return 2

    import fdtd2d_amd as F

    total = args.warmup + args.steps
    t0 = time.time()
    setup = F.Setup.workload(2, seed=args.seed, steps=max(total, args.cpu_steps))
    p = setup.c_problem.contents
    nx, ny = p.nx, p.ny
    log(f"setup {nx}x{ny} P={p.npml} nsrc={p.nsrc} ndet={p.ndet} in {time.time() - t0:.1f}s")

    plan = F.Plan(setup, device=0, tblock=args.tblock)
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
    avg_ms, nlaunch = plan.kernel_ms()
    det, _ = plan.read_detectors()
    plan.close()
    ok = bool((det == det).all())  # no NaN in the traces
    cells = nx * ny
    value = cells * args.steps / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    # launches of the timed region: all T steps except possibly the last
    full_launch_updates = cells * T
    achieved = BYTES_PER_CELL_UPDATE * full_launch_updates / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tf = Path(args.traffic_file)
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            traffic = None
    log(f"GPU: {value:.1f} Gcell/s, {ms_per_step:.3f} ms/step, kernel avg {avg_ms:.3f} ms over {nlaunch} "
        f"launches of T={T}, finite traces: {ok}")

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle

        ref = Oracle(setup, precision=32, threads=1)
        c0 = time.perf_counter()
        ref.run(args.cpu_steps)
        cs = time.perf_counter() - c0
        cpu = {
            "value": round(cells * args.cpu_steps / cs / 1e9, 4),
            "unit": "Gcell-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"FP32 CPU oracle, {nx}x{ny}, {args.cpu_steps} steps ({cs:.1f} s, single thread)",
        }
        ref.close()
        log(f"CPU oracle: {cpu['value']} Gcell/s")

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
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_avg_ms": round(avg_ms, 4),
            "bytes_per_launch": BYTES_PER_CELL_UPDATE * full_launch_updates,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
