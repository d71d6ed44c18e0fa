"""Benchmark: AVBD body-iterations/s on the 51,039-cube pyramid (one MI355X).

A step is one full AVBD time step of the whole scene (broad phase, narrow
phase, colouring, 20 primal/dual iterations, velocity update); inputs are
resident in HBM. body-iterations/s = dynamic bodies x iterations x steps /
seconds, over exactly --steps steps after --warmup untimed steps.

Also reported:
  roofline      the primal kernel (the dominant kernel): algorithmic bytes per
                launch / average launch duration (events stamped by each
                dispatch on the library's stream, over a second region of the
                same length) against 8 TB/s HBM; traffic = HBM bytes per launch
                from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary;
  cpu_baseline  the f64 CPU oracle (a single-threaded restatement of the
                reference engine) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "avbd-sandbox_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "body-iterations/s on 51k-body pyramid vs host f64 CPU; adjoint/forward time"
HBM_PEAK_GBS = 8000.0


def measured_traffic(precision, team):
    """HBM bytes per k_primal launch (read + written) from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary of the same workload,
    scaled by the copy-probe calibration; None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"k_primal_pyramid51k_{precision}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        doc = json.load(f)
    k = doc["kernels"].get(f"k_primal<{'float' if precision == 'mixed' else 'double'}, "
                           f"{'true' if precision == 'exact' else 'false'}, {team}>")
    if doc.get("team") != team or k is None:
        return None, None
    return k["read_bytes"] + k["write_bytes"], os.path.relpath(path, ROOT)


def copy_probe(mib):
    """Three launches of one elementwise kernel that reads MIB MiB and writes
    MIB MiB: what a byte counter reads for it calibrates the counter."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "f64", "exact"])
    ap.add_argument("--team", type=int, default=0, help="lanes per body; 0 = autotune")
    ap.add_argument("--order", default="priority", choices=["priority", "index"])
    ap.add_argument("--layers", type=int, default=53)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the known-size reference "
                         "for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    import avbd
    from avbd import _lib as L

    prec = {"mixed": avbd.PREC_MIXED, "f64": avbd.PREC_F64, "exact": avbd.PREC_EXACT}[args.precision]
    scene, params, meta = avbd.scenes.pyramid(layers=args.layers)
    params = dict(params, colorOrder=1 if args.order == "priority" else 0)
    gpu = avbd.HipBackend(precision=prec, team=args.team)
    gpu.pack(scene, params)
    iters = params["iterations"]
    t0 = time.perf_counter()
    gpu.step(max(args.warmup - 1, 0))
    tune = None
    if args.team == 0 and prec != avbd.PREC_EXACT:
        tune = gpu.autotune()  # one warm-up step that also picks the team shape
    elif args.warmup > 0:
        gpu.step(1)
    gpu.idle()
    print(f"[bench] warmup {args.warmup} steps: {time.perf_counter() - t0:.2f} s, team {tune}", file=sys.stderr,
          flush=True)
    c = gpu.counters()
    nd = c["dynamic_bodies"]

    # timed region: plain stepping (the solve loop replays a captured graph)
    t0 = time.perf_counter()
    gpu.step(args.steps)
    gpu.idle()
    secs = time.perf_counter() - t0

    # the same number of further steps with per-kernel HIP events, for the
    # roofline and the phase split (direct launches, so each is bracketed)
    gpu.set_profiling(True)
    t1 = time.perf_counter()
    gpu.step(args.steps)
    gpu.idle()
    psecs = time.perf_counter() - t1
    phases, psteps = gpu.profile()
    pms, plaunch, pbytes = gpu.kernel_stats(L.K_PRIMAL)
    dms, dlaunch, dbytes = gpu.kernel_stats(L.K_DUAL)
    nms, nlaunch, _ = gpu.kernel_stats(L.K_NARROW)
    c = gpu.counters()
    gpu.set_profiling(False)

    value = nd * iters * args.steps / secs
    avg_ms = pms / plaunch
    achieved = (pbytes / plaunch) / (avg_ms * 1e-3) / 1e9
    team = tune[0] if tune else args.team
    traffic, traffic_src = measured_traffic(args.precision, team)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32+f64", "f64": "f64", "exact": "f64"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"pyramid of {meta['cubes']} unit cubes ({args.layers} layers, 1 m pitch, mu 0.6) "
                        f"on a static ground, dt 1/60, {iters} iterations, steps {args.warmup + 1}-"
                        f"{args.warmup + args.steps} after pack",
            "bodies": len(scene["bodies"]), "dynamic_bodies": nd, "iterations": iters,
            "precision": args.precision, "team": team, "color_order": args.order,
            "autotune_ms_per_sweep": tune[1] if tune else None,
            "color_rounds": c["color_rounds"],
            "manifolds": c["manifolds"], "contacts": c["contacts"], "colors": c["colors"],
        },
        "roofline": {
            "kernel": "k_primal", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "avg_launch_us": avg_ms * 1e3, "launches": plaunch,
            "alg_bytes_per_launch": pbytes / plaunch,
        },
        "kernels": {
            "primal_ms_per_step": pms / args.steps, "dual_ms_per_step": dms / args.steps,
            "dual_GBs": (dbytes / dms * 1e-6) if dms > 0 else None,
            "narrow_ms_per_step": nms / args.steps,
        },
        "phases_ms": phases,
        "profiled_ms_per_step": psecs / args.steps * 1e3,
    }
    if not args.no_cpu:
        import pyoracle
        t0 = time.perf_counter()
        orc = pyoracle.Oracle(scene, params, schedule=pyoracle.SCHED_COLOR, color_order=params["colorOrder"])
        build_s = time.perf_counter() - t0
        cpu_secs = orc.step(args.cpu_steps)
        cpu_rate = orc.n_dynamic * iters * args.cpu_steps / cpu_secs
        line["cpu_baseline"] = {
            "value": cpu_rate, "unit": "body-iterations/s", "cores": 1, "kind": "port",
            "sample": f"steps 1-{args.cpu_steps} of the same {meta['cubes']}-cube pyramid, {iters} iterations, "
                      f"f64 oracle (C++ -O2, single thread); {cpu_secs:.1f} s (+{build_s:.1f} s scene build)",
        }
        line["vs_cpu"] = value / cpu_rate
    print(json.dumps(line))


if __name__ == "__main__":
    main()
