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
                reference engine) on a bounded sample of the same workload;
  adjoint       (pyramid runs) forward and reverse time of the 64-box ramp
                gradient benchmark.

--workload rain / batched measure the 50,000-body mixed-shape rain and the 256
independent chain and stack scenes the same way (not the headline line).
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


POSE_LAUNCHES = 100


def measured_traffic(workload, precision, kernel_prefixes):
    """HBM bytes per launch (read + written) of the dominant kernel from the
    committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary of the same
    workload (profiles/<workload>_<precision>_kernels.md, already scaled by the
    copy-probe calibration): the dispatch-weighted mean over the rows whose
    kernel name starts with one of `kernel_prefixes` (the team shapes a sweep
    used). None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_kernels.md")
    if not os.path.exists(path):
        return None, None
    n, total = 0, 0.0
    with open(path) as f:
        for ln in f:
            if not ln.startswith("| ") or ln.startswith("| kernel"):
                continue
            cells = [c.strip() for c in ln.strip().strip("|").split("|")]
            name = cells[0].replace("void ", "").replace("avbd::", "")
            if not any(name.startswith(p) for p in kernel_prefixes) or cells[5] == "-":
                continue
            calls = int(cells[1])
            n += calls
            total += calls * (float(cells[5]) + float(cells[6])) * 1e6
    if n == 0:
        return None, None
    return total / n, os.path.relpath(path, ROOT)


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


WORKLOADS = ("pyramid", "rain", "batched", "fabric")


def build_workload(args):
    """(scene, params, description, cpu sample runner) of the chosen workload."""
    import avbd
    if args.workload == "pyramid":
        scene, params, meta = avbd.scenes.pyramid(layers=args.layers)
        desc = (f"pyramid of {meta['cubes']} unit cubes ({args.layers} layers, 1 m pitch, mu 0.6) on a static ground, "
                f"dt 1/60, {params['iterations']} iterations")
    elif args.workload == "rain":
        scene, params, meta = avbd.scenes.rain(n=args.bodies)
        desc = (f"rain of {args.bodies} bodies (70% boxes 0.2-1 m, 30% spheres r 0.1-0.5 m, density log-uniform "
                f"1-1000, mu 0.2-0.9) into a 40 x 40 m walled column, dt 1/60, {params['iterations']} iterations")
    elif args.workload == "fabric":
        scene, params, meta = avbd.scenes.fabric(count=args.bodies if args.bodies != 50000 else 15000)
        desc = (f"hanging fabric of {meta['rows']} x {meta['cols']} sphere particles and {meta['springs']} springs "
                f"(structural, shear, bend; tear strain 0.75) with 16 heavy spheres thrown through it, dt 1/60, "
                f"{params['iterations']} iterations")
    else:
        scene, params, meta = avbd.scenes.batched(num_scenes=args.scenes)
        desc = (f"{args.scenes} independent scenes (chains of 10-200 hard links with a 26,000:1 end mass, stacks of "
                f"5-50 cubes), dt 1/60, {params['iterations']} iterations")
    params = dict(params, colorOrder=1 if args.order == "priority" else 0)
    return scene, params, meta, desc


def cpu_sample(args, scene, params, meta, start):
    """The f64 oracle (one thread) on a bounded sample of the same workload,
    from the bodies' state where the timed region starts (`start`: the GPU's
    poses and velocities after the set-up and warm-up; the oracle's contacts
    then start cold, as a pack's do): steps one at a time until about
    `args.cpu_seconds` of CPU work. Returns (body-iterations/s, description)."""
    import pyoracle
    iters = params["iterations"]
    budget = args.cpu_seconds

    def resumed(part, lo):
        bodies = []
        for i, b in enumerate(part["bodies"]):
            j = lo + i
            bodies.append(dict(b, pos=[float(x) for x in start["pos"][j]], quat=[float(x) for x in start["quat"][j]],
                               vel=[float(x) for x in start["vel"][j]], angvel=[float(x) for x in start["angvel"][j]]))
        return dict(part, bodies=bodies)

    if args.workload == "batched":
        # scene by scene (the oracle has no notion of independent scenes):
        # every scene one step per round, rounds until the budget is spent
        worlds, lo = [], 0
        for part in meta["parts"]:
            worlds.append(pyoracle.Oracle(resumed(part, lo), params, schedule=pyoracle.SCHED_COLOR,
                                          color_order=params["colorOrder"]))
            lo += len(part["bodies"])
        nd = sum(o.n_dynamic for o in worlds)
        secs, rounds = 0.0, 0
        while rounds < 1000 and (rounds < 1 or secs < budget):
            secs += sum(o.step(1) for o in worlds)
            rounds += 1
        return nd * iters * rounds / secs, (f"all {len(worlds)} scenes from their state at the timed region's "
                                            f"start (the oracle's contacts starting cold), {rounds} steps each, "
                                            f"f64 oracle (C++ -O2, single thread); {secs:.1f} s")
    t0 = time.perf_counter()
    o = pyoracle.Oracle(resumed(scene, 0), params, schedule=pyoracle.SCHED_COLOR, color_order=params["colorOrder"])
    build_s = time.perf_counter() - t0
    secs, steps = 0.0, 0
    while steps < 1000 and (steps < 1 or secs < budget):
        secs += o.step(1)
        steps += 1
    return (o.n_dynamic * iters * steps / secs,
            f"{steps} steps of the same scene from its state at the timed region's start (the oracle's contacts "
            f"starting cold), {iters} iterations, f64 oracle (C++ -O2, single thread); {secs:.1f} s "
            f"(+{build_s:.1f} s scene build)")


def settle_and_strike(gpu, spec):
    """The pyramid's set-up: `spec['step']` settling steps, then the sphere
    launched at the apex (position, velocity and previous velocity set,
    angular velocity 0)."""
    gpu.step(spec["step"])
    st = gpu.read_state()
    b = spec["body"]
    st["pos"][b] = spec["pos"]
    st["vel"][b] = spec["vel"]
    st["prev_vel"][b] = spec["vel"]
    st["angvel"][b] = 0.0
    gpu.write_state(st["pos"], st["quat"], st["vel"], st["angvel"], st["prev_vel"])


def graphs_cfg(args):
    """avbd_config.graphs for the command line's choice."""
    import avbd
    if args.no_graphs:
        return avbd.GRAPHS_NONE
    return {"step": avbd.GRAPHS_STEP, "solve": avbd.GRAPHS_SOLVE, "none": avbd.GRAPHS_NONE}[args.graphs]


def exact_leg(args, scene, params, spec, timed_end):
    """The same workload in the exact precision (f64, bit-identical to the
    oracle: the error-free line) over the same steps: its throughput, and the
    pose error of the measured precision against it at the end of the timed
    region (the largest position and quaternion-component difference over the
    dynamic bodies)."""
    import avbd
    import numpy as np
    ex = avbd.HipBackend(precision=avbd.PREC_EXACT, graphs=graphs_cfg(args))
    ex.pack(scene, params)
    if spec is not None:
        settle_and_strike(ex, spec)
    ex.step(args.warmup)
    ex.idle()
    t0 = time.perf_counter()
    ex.step(args.steps)
    ex.idle()
    secs = time.perf_counter() - t0
    st = ex.read_state()
    nd = ex.counters()["dynamic_bodies"]
    dyn = np.array([b["density"] > 0 for b in scene["bodies"]])
    qa, qb = st["quat"][dyn], timed_end["quat"][dyn]
    sgn = np.where(np.sum(qa * qb, axis=1) < 0, -1.0, 1.0)[:, None]
    err_p = float(np.abs(st["pos"][dyn] - timed_end["pos"][dyn]).max())
    err_q = float(np.abs(qa - sgn * qb).max())
    return {"value": nd * params["iterations"] * args.steps / secs, "unit": "body-iterations/s",
            "ms_per_step": secs / args.steps * 1e3, "dtype": "f64",
            "pose_error_of_measured_precision_m": err_p, "quat_error_of_measured_precision": err_q,
            "note": "exact precision: f64, every body's system summed in the reference's order, bit-identical "
                    "to the f64 oracle; same steps as the timed region"}


def contact_law_leg(args, scene, params, spec, prec):
    """The same workload under the reference GPU backend's contact law
    (contactModel 1: static/dynamic friction, restitution targets, a 5 mm
    contact skin per body, rest offsets, geometric warm-start persistence;
    include/avbd.h AVBD_CONTACTS_GPU) in the measured precision, over the same
    steps after the same set-up: its throughput and the contact counts it
    carries (the skin adds speculative contacts)."""
    import avbd
    g = avbd.HipBackend(precision=prec, team=args.team, graphs=graphs_cfg(args))
    g.pack(scene, dict(params, contactModel=1))
    if spec is not None:
        settle_and_strike(g, spec)
    tune = None
    if args.team == 0 and prec != avbd.PREC_EXACT and args.warmup >= 3:
        after = max(2, min(12, args.warmup // 2))
        g.step(args.warmup - after - 1)
        tune = g.autotune()
        g.step(after)
    else:
        g.step(args.warmup)
    g.idle()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.idle()
    secs = time.perf_counter() - t0
    c = g.counters()
    return {"value": c["dynamic_bodies"] * params["iterations"] * args.steps / secs, "unit": "body-iterations/s",
            "ms_per_step": secs / args.steps * 1e3, "team": tune[0] if tune else args.team,
            "manifolds": c["manifolds"], "contacts": c["contacts"], "pairs": c["pairs"]}


def host_cpu():
    """The host CPU's model name and the logical CPUs this process may run
    on (the CPU baseline uses one of them)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def adjoint_leg(graphs=0):
    """Forward and reverse timing of the ramp benchmark: 64 boxes (0.5 m, rho
    500) on a 30-degree ramp, 120 steps, 10 iterations, f64, checkpoint every
    10 steps; gradients of a loss of the final positions with respect to each
    box's friction and initial velocity. Timed with the tape kept resident
    (recorded as the forward runs; the default when it fits its budget) and
    with every checkpoint segment re-run to record it in the backward."""
    import avbd
    import numpy as np
    scene, params, meta = avbd.scenes.ramp(64)
    target = np.array([b["pos"] for b in scene["bodies"]])

    ctxs = {0: avbd.HipBackend(precision=avbd.PREC_F64, adjoint_tape_mb=0, graphs=graphs),
            -1: avbd.HipBackend(precision=avbd.PREC_F64, adjoint_tape_mb=-1, graphs=graphs)}

    def adjoint(tape_mb):
        # one context per mode across the rounds, as an optimisation loop
        # keeps one: the second round reuses the tape's buffers
        a = ctxs[tape_mb]
        a.pack(scene, params)
        t0 = time.perf_counter()
        a.adjoint_forward(meta["steps"], 10)
        a.idle()
        cfwd = time.perf_counter() - t0
        dL = 2 * (a.read_state()["pos"] - target)
        dL[:2] = 0
        t0 = time.perf_counter()
        a.adjoint_backward(dL)
        return cfwd, time.perf_counter() - t0

    out = None
    for _ in range(2):  # the first round warms up the code paths
        f = avbd.HipBackend(precision=avbd.PREC_F64, graphs=graphs)
        f.pack(scene, params)
        t0 = time.perf_counter()
        f.step(meta["steps"])
        f.idle()
        fwd = time.perf_counter() - t0
        cfwd, bwd = adjoint(0)
        rfwd, rbwd = adjoint(-1)
        out = {"workload": "ramp: 64 boxes (0.5 m, rho 500, mu 0.1-0.9, v0 0-2 m/s) on a 30-degree ramp, 120 steps, "
                           "10 iterations, f64, checkpoint every 10 steps; d loss / d mu and d loss / d v0",
               "forward_ms": fwd * 1e3, "checkpointed_forward_ms": cfwd * 1e3, "backward_ms": bwd * 1e3,
               "backward_over_forward": bwd / fwd, "adjoint_total_over_forward": (cfwd + bwd) / fwd,
               "tape": "resident (recorded during the checkpointed forward)",
               "segment_rerun": {"checkpointed_forward_ms": rfwd * 1e3, "backward_ms": rbwd * 1e3,
                                 "backward_over_forward": rbwd / fwd,
                                 "adjoint_total_over_forward": (rfwd + rbwd) / fwd}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="pyramid", choices=WORKLOADS,
                    help="pyramid (the headline), rain (50k mixed shapes), batched (256 scenes) or fabric "
                         "(15k-particle spring cloth)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "f64", "exact"])
    ap.add_argument("--team", type=int, default=0, help="lanes per body; 0 = autotune")
    ap.add_argument("--workgroup", type=int, default=0,
                    help="with --team: the sweep's workgroup (64, 128 or 256; batched scenes 128, 256 or 512); "
                         "0 = the library's default")
    ap.add_argument("--order", default="priority", choices=["priority", "index"])
    ap.add_argument("--layers", type=int, default=53)
    ap.add_argument("--bodies", type=int, default=50000, help="rain: body count")
    ap.add_argument("--scenes", type=int, default=256, help="batched: scene count")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline: oracle steps until about this many seconds of CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--graphs", default="step", choices=["step", "solve", "none"],
                    help="graph replay: whole steps with no host read inside a step (step), each step's solve loop "
                         "after host reads of its counts (solve), or direct launches (none)")
    ap.add_argument("--no-graphs", action="store_true", help="the same as --graphs none")
    ap.add_argument("--no-exact", action="store_true",
                    help="pyramid: skip the exact-precision leg (its throughput and the pose error against it)")
    ap.add_argument("--no-launch", action="store_true",
                    help="pyramid: time the first steps after pack (the settling pile) instead of the strike")
    ap.add_argument("--contact-model", default="cpu", choices=["cpu", "gpu"],
                    help="contact law of the timed region: the CPU engine's (the headline) or the reference GPU "
                         "backend's (static/dynamic friction, restitution, contact skin and rest offsets)")
    ap.add_argument("--no-contact-law", action="store_true",
                    help="pyramid: skip the leg timing the same steps under the GPU backend's contact law")
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
    scene, params, meta, desc = build_workload(args)
    if args.contact_model == "gpu":
        params = dict(params, contactModel=1)
        desc += ", the GPU backend's contact law"
    gpu = avbd.HipBackend(precision=prec, team=args.team, graphs=graphs_cfg(args))
    gpu.pack(scene, params)
    iters = params["iterations"]
    # The pyramid configuration settles for its first steps and is then struck
    # at the apex by the sphere: both are part of the workload's set-up, so
    # the warm-up and timed steps come after the strike and cover the impact
    # and the pile's response.
    spec = meta.get("launch") if args.workload == "pyramid" and not args.no_launch else None
    settle = spec["step"] if spec is not None else 0
    if spec is not None:
        t0 = time.perf_counter()
        settle_and_strike(gpu, spec)
        print(f"[bench] settled {settle} steps and launched the sphere at the apex: "
              f"{time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tune = None
    if args.team == 0 and prec != avbd.PREC_EXACT and args.warmup >= 3:
        # a warm-up step picks the team shape; the next reads the counts the
        # replayed graphs are shaped for under it, and the rest capture the
        # step graphs (one per colour count and buffer parity the steps meet)
        # before the timed steps
        after = max(2, min(12, args.warmup // 2))
        gpu.step(args.warmup - after - 1)
        tune = gpu.autotune()
        gpu.step(after)
    else:
        if args.team and args.workgroup:
            gpu.set_launch_shape(args.team, args.workgroup)
        gpu.step(args.warmup)
    gpu.idle()
    print(f"[bench] warmup {args.warmup} steps: {time.perf_counter() - t0:.2f} s, team {tune}", file=sys.stderr,
          flush=True)
    c = gpu.counters()
    nd = c["dynamic_bodies"]

    start = gpu.read_state() if not args.no_cpu else None  # where the CPU baseline resumes
    # timed region: plain stepping (replayed graphs: whole steps by default)
    t0 = time.perf_counter()
    gpu.step(args.steps)
    gpu.idle()
    secs = time.perf_counter() - t0
    timed_end = gpu.read_state() if args.workload == "pyramid" and prec != avbd.PREC_EXACT else None

    # the same number of further steps with per-kernel timing, for the
    # roofline and the phase split (direct launches)
    gpu.set_profiling(True)
    t1 = time.perf_counter()
    gpu.step(args.steps)
    gpu.idle()
    psecs = time.perf_counter() - t1
    phases, psteps = gpu.profile()
    pms, plaunch, pbytes = gpu.kernel_stats(L.K_PRIMAL)
    dms, dlaunch, dbytes = gpu.kernel_stats(L.K_DUAL)
    nms, nlaunch, _ = gpu.kernel_stats(L.K_NARROW)
    # the renderer's pose export (blend at alpha 0.5), timed by the same
    # dispatch-stamped events
    poses = None
    for _ in range(POSE_LAUNCHES):
        poses = gpu.export_poses(0.5, out=poses)
    xms, xlaunch, xbytes = gpu.kernel_stats(L.K_POSES)
    c = gpu.counters()
    gpu.set_profiling(False)

    value = nd * iters * args.steps / secs
    avg_ms = pms / max(plaunch, 1)
    achieved = (pbytes / max(plaunch, 1)) / (avg_ms * 1e-3) / 1e9 if pms > 0 else 0.0
    team = tune[0] if tune else args.team
    shape = gpu.launch_shape()
    kern = {"exact": ("k_primal_xt",), "f64": ("k_primal<double",), "mixed": ("k_primal<float",)}[args.precision]
    if args.workload == "batched":
        kern = ("k_solve_scenes<" + ("double" if args.precision != "mixed" else "float"),)
    traffic, traffic_src = measured_traffic(args.workload, args.precision, kern)
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
            "workload": (f"{desc}, " + (f"settled {settle} steps, then the 1 m sphere (10x density) launched at the "
                                         f"apex at 20 m/s; " if settle else "") +
                         f"timed steps {settle + args.warmup + 1}-{settle + args.warmup + args.steps} after pack"),
            "bodies": len(scene["bodies"]), "dynamic_bodies": nd, "iterations": iters,
            "precision": args.precision, "team": team, "workgroup": shape[1], "color_order": args.order,
            "contact_law": args.contact_model,
            "graphs": "none" if args.no_graphs else args.graphs,
            "autotune_ms_per_sweep": tune[1] if tune else None,
            "color_rounds": c["color_rounds"],
            "manifolds": c["manifolds"], "contacts": c["contacts"], "joints": c["joints"], "colors": c["colors"],
        },
        "roofline": {
            "kernel": "k_primal", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "avg_launch_us": avg_ms * 1e3, "launches": plaunch,
            "alg_bytes_per_launch": pbytes / max(plaunch, 1),
        },
        "kernels": {
            "primal_ms_per_step": pms / args.steps, "dual_ms_per_step": dms / args.steps,
            "dual_GBs": (dbytes / dms * 1e-6) if dms > 0 else None,
            "narrow_ms_per_step": nms / args.steps,
            "export_poses_us": xms / max(xlaunch, 1) * 1e3,
            "export_poses_GBs": (xbytes / xms * 1e-6) if xms > 0 else None,
        },
        "phases_ms": phases,
        "profiled_ms_per_step": psecs / args.steps * 1e3,
    }
    # a topology edit that keeps state: erase one cube, append one
    if args.workload == "pyramid":
        gpu.idle()
        n = len(scene["bodies"])
        old_index = [i for i in range(n) if i != 1] + [-1]
        scene2 = dict(bodies=[b for i, b in enumerate(scene["bodies"]) if i != 1] +
                      [dict(scene["bodies"][1], pos=[0.0, 0.0, params.get("layers", 53) + 5.0])], joints=[])
        # the description's arrays are built first: the time is the library's
        # (avbd_repack), not the Python dict conversion
        import ctypes
        import numpy as np
        s2 = gpu._scene_struct(scene2)
        oi = np.ascontiguousarray(old_index, dtype=np.int32)
        t2 = time.perf_counter()
        rc = gpu.L.avbd_repack(gpu.ctx, ctypes.byref(s2), oi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        line["kernels"]["repack_ms"] = (time.perf_counter() - t2) * 1e3
        if rc != 0:
            raise RuntimeError(f"avbd_repack failed: {rc}")
        gpu.n = len(scene2["bodies"])
        c2 = gpu.counters()
        line["kernels"]["repack_kept_manifolds"] = c2["manifolds"]
    if timed_end is not None and not args.no_exact:
        line["exact"] = exact_leg(args, scene, params, spec, timed_end)
    if args.workload == "pyramid" and args.contact_model == "cpu" and not args.no_contact_law:
        line["contact_law_gpu"] = contact_law_leg(args, scene, params, spec, prec)
    if not args.no_adjoint and args.workload == "pyramid":
        # direct launches there too when the command asks for them (as under a profiler)
        line["adjoint"] = adjoint_leg(avbd.GRAPHS_NONE if graphs_cfg(args) == avbd.GRAPHS_NONE else 0)
    if not args.no_cpu:
        cpu_rate, sample = cpu_sample(args, scene, params, meta, start)
        line["cpu_baseline"] = {"value": cpu_rate, "unit": "body-iterations/s", "cores": 1, "kind": "port",
                                "sample": sample, "host": host_cpu()}
        line["vs_cpu"] = value / cpu_rate
    print(json.dumps(line))


if __name__ == "__main__":
    main()
