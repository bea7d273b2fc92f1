"""Benchmark of the plain GPU VBD hot path.

Workload (one "step" = one frame): the 256x256 grid cloth (65,536 vertices,
195,585 springs, 194,565 bending pairs, 7 colours), top row pinned, seeded
per-frame wind, 100 VBD iterations per frame, FP32.  Inputs live in HBM
before the timed region; the timed region replays the captured frame graph
K times.  value = vertex updates per second over the whole job
(V * iterations * K / wall time).

Also reported:
  roofline      the per-colour VBD kernel (k_vbd_colour): algorithmic bytes
                per launch (53 V + 20 E + 72 B + 8 bytes per full sweep, spread
                over the colour launches of one sweep) / mean launch duration
                (HIP events around a graph of 20 sweeps on the solver's
                stream), vs 8 TB/s HBM.
  cpu_baseline  the CPU oracle (C++ restatement of the reference's CPU VBD,
                per-colour parallel passes) on this host's cores, over a bounded
                sample of the same workload (whole frames, >= 10 s of work).
  warm_start    the learned warm start on the same cloth (trained in the run).
  other_workloads
                short runs of the other configured workloads -- batch_c2 (512
                cloths), sphere_c4 (1024x1024 on a sphere), xpbd_c1 (XPBD on the
                256x256 cloth) -- and of square8k (the reference's own example
                mesh, ms per iteration beside the reference's published task
                costs), each this script in a child process, its line
                summarised (--no-other-workloads skips them).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import subprocess
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "VBD vertex-updates/s (256² cloth); ms/frame to fixed residual, learned vs plain"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_vbd_colour_grid256_traffic.txt")
PATCH_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_vbd_patch_grid256_traffic.txt")


def read_traffic(path, key="hbm_bytes_per_launch"):
    """HBM bytes per launch from a committed PMC summary (its `<key>: N` line)."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def read_colour_traffic(path, lanes, regions, threads):
    """HBM bytes per colour launch from a committed PMC summary (its
    `hbm_bytes_per_launch L R T: N` lines): measured at this lanes / bands /
    block-threads configuration, else at the same lanes and bands with
    another block size (a lane's loads do not depend on the block size),
    else None."""
    b = read_colour_traffic_shape(path, lanes, regions, threads)
    return b[0] if b else None


def read_colour_traffic_shape(path, lanes, regions, threads):
    """(bytes, "L R T" shape measured) for read_colour_traffic, or None."""
    if not os.path.exists(path):
        return None
    exact = f"hbm_bytes_per_launch {lanes} {regions} {threads}:"
    near = f"hbm_bytes_per_launch {lanes} {regions} "
    fallback = None
    for line in open(path):
        if line.startswith(exact):
            return float(line.split(":", 1)[1]), f"{lanes} {regions} {threads}"
        if fallback is None and line.startswith(near):
            fallback = float(line.split(":", 1)[1]), line[len("hbm_bytes_per_launch "):].split(":", 1)[0]
    return fallback


def algorithmic_bytes_per_sweep(nv, ne, nb):
    """Compulsory bytes of one sweep over all colours (compact layout)."""
    return 53 * nv + 20 * ne + 72 * nb + 8


def _host_cpu():
    """(CPU model, logical CPUs) of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def _oracle_vbd_rate(cloth, params, wind, threads, min_seconds, max_frames):
    """(vertex updates per second, frames, seconds) of the oracle's FP32 VBD frames."""
    from oracle import oracle as O
    m = O.Mesh(cloth.x, cloth.faces)
    s = O.Sim(m, cloth.mass, cloth.fixed, implicit_dt=params.implicit_dt, num_substep=params.num_substep,
              iter_count=params.constraint_iter_count, k_stretch=params.stiffness_stretch_spring,
              k_bending=params.stiffness_quadratic_bending, damping=params.damping_cloth, threads=threads)
    frames = 0
    t0 = time.perf_counter()
    while True:
        s.set_wind(wind[frames % len(wind)])
        s.step()
        frames += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or frames >= max_frames:
            break
    vu = cloth.num_verts * params.constraint_iter_count * frames
    return vu / el, frames, el


def cpu_baseline(cloth, params, wind, min_seconds=10.0, max_frames=120):
    """The oracle on this host's cores (at most 16 threads), and on one thread
    over a shorter sample."""
    from oracle import oracle as O
    threads = min(O.max_threads(), 16)
    rate, frames, el = _oracle_vbd_rate(cloth, params, wind, threads, min_seconds, max_frames)
    rate1, frames1, el1 = _oracle_vbd_rate(cloth, params, wind, 1, min_seconds / 2, max_frames)
    model, ncpu = _host_cpu()
    vu = cloth.num_verts * params.constraint_iter_count * frames
    return {
        "value": rate,
        "unit": "vertex-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} frames of the same 256x256 workload ({vu} vertex updates, {el:.1f} s), "
                  f"oracle/vbd_oracle.cpp FP32, OpenMP per-colour passes",
        "one_thread": {"value": rate1, "sample": f"{frames1} frames, {el1:.1f} s"},
        "host_cpu": model,
        "host_logical_cpus": ncpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="grid_c1", choices=["grid_c1", "batch_c2", "sphere_c4", "copy_calibration",
                                                               "mlp_kernels", "comm_matrix", "launch_floor",
                                                               "xpbd_c1", "square8k"],
                    help="grid_c1 (default, the metric's workload), batch_c2 (512 cloths) or "
                         "copy_calibration (a plain 256 MiB device copy, the reference for HBM counters)")
    ap.add_argument("--strategy", default="auto", choices=["auto", "graph", "persistent", "flow", "patch"])
    ap.add_argument("--patches", type=int, default=0, help="patch strategy: patches (multiple of 8; 0 = derived)")
    ap.add_argument("--patch-colours", type=int, default=0, help="patch strategy: colours per launch (0 = all)")
    ap.add_argument("--patch-threads", type=int, default=1024, help="patch strategy: threads per workgroup")
    ap.add_argument("--tiles", type=int, default=0, help="flow strategy: vertex tiles (0 = default)")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per vertex (1, 4, 8, 16); 0 = autotuned")
    ap.add_argument("--regions", type=int, default=0, help="colour-pass cloth bands (1..8); 0 = autotuned")
    ap.add_argument("--colour-block", type=int, default=0, help="colour-pass block threads (64, 128, 256); "
                                                                 "0 = autotuned")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned warm-start measurement")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="grid_c1 only: skip the short batch_c2 / sphere_c4 / xpbd_c1 runs reported beside it")
    ap.add_argument("--warm-scenarios", type=int, default=8, help="training scenarios of the warm start")
    ap.add_argument("--warm-frames", type=int, default=60, help="frames per training scenario")
    ap.add_argument("--warm-record-every", type=int, default=1, help="record every k-th training frame")
    ap.add_argument("--warm-steps", type=int, default=2000, help="Adam steps per training phase of the warm start")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the VBD hot path runs on a single MI355X "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    # the other workloads run first, each in a child process, before this
    # process uses the GPU (a device shared with a live context refuses
    # cooperative launches, which the children's autotuners time)
    others = None
    if args.workload == "grid_c1" and not args.no_other_workloads:
        others = other_workloads()

    from libatssim_amd import GpuVBD, configs, _lib

    if args.workload == "batch_c2":
        return bench_batch(args)
    if args.workload == "copy_calibration":
        return copy_calibration(args)
    if args.workload == "mlp_kernels":
        return bench_mlp(args)
    if args.workload == "sphere_c4":
        return bench_sphere(args)
    if args.workload == "comm_matrix":
        return bench_comm_matrix(args)
    if args.workload == "launch_floor":
        return bench_launch_floor(args)
    if args.workload == "xpbd_c1":
        return bench_xpbd(args)
    if args.workload == "square8k":
        return bench_square8k(args)

    K, W = args.steps, args.warmup
    cloth, params, _, wind = configs.grid_c1(frames=W + K)
    t = cloth.topology
    g = GpuVBD(cloth, params)
    if args.strategy == "auto":
        strategy, blocks = g.autotune(frames=2)
    else:
        strategy = {"graph": _lib.LATS_STRATEGY_GRAPH, "persistent": _lib.LATS_STRATEGY_PERSISTENT,
                    "flow": _lib.LATS_STRATEGY_FLOW, "patch": _lib.LATS_STRATEGY_PATCH}[args.strategy]
        if args.strategy == "patch":
            g.set_patches(args.patches, args.patch_colours, args.patch_threads)
        g.set_strategy(strategy, persistent_blocks=args.tiles if args.strategy == "flow" else None)
        blocks = g.strategy()[1]
    if args.lanes:
        g.set_lanes_per_vertex(args.lanes)
        strategy, blocks = g.strategy()
    if args.regions:
        g.set_regions(args.regions)
    if args.colour_block:
        g.set_colour_block(args.colour_block)
    if W:
        g.run_frames(W, wind[:W])
    g.synchronize()
    t0 = time.perf_counter()
    g.run_frames(K, wind[W:W + K])
    g.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = g.last_elapsed_ms()
    iters = params.constraint_iter_count
    vu = cloth.num_verts * iters * K
    value = vu / wall

    b_sweep = algorithmic_bytes_per_sweep(t.num_verts, t.num_edges, t.num_bending)
    if strategy in (_lib.LATS_STRATEGY_PERSISTENT, _lib.LATS_STRATEGY_FLOW):
        # one cooperative launch runs the K timed frames: its duration is the
        # HIP-event time of the timed region on the solver's stream
        kernel = "k_vbd_persistent" if strategy == _lib.LATS_STRATEGY_PERSISTENT else "k_vbd_flow"
        mean_ms = gpu_ms
        bytes_per_launch = b_sweep * iters * K
    elif strategy == _lib.LATS_STRATEGY_PATCH:
        # launches_per_sweep patch launches per sweep, each running its
        # group's colours: the timed region's HIP events over K frames
        # (slightly overstated by one predict and one update per frame)
        kernel = "k_vbd_patch"
        per_sweep = g.patch_info()["launches_per_sweep"]
        mean_ms = gpu_ms / (K * iters * per_sweep)
        bytes_per_launch = b_sweep / per_sweep
    else:
        # The timed region's own HIP events (on the solver's stream, around
        # the K frame graphs): K x iterations x colours colour launches plus
        # one predict and one update per frame, so the mean launch duration
        # is slightly overstated.  Cross-check: 20 sweeps of colour launches
        # alone, captured in a graph and timed the same way.
        kernel = "k_vbd_colour"
        mean_ms = gpu_ms / (K * iters * t.num_clusters)
        graph_ms, launches, kvu = g.time_colour_kernels(20)
        bytes_per_launch = b_sweep / t.num_clusters
    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9
    traffic = None
    if kernel == "k_vbd_colour":
        traffic = read_colour_traffic(TRAFFIC_FILE, g.lanes_per_vertex(), g.regions(), g.colour_block())
    elif kernel == "k_vbd_patch":
        pi = g.patch_info()
        traffic = read_traffic(PATCH_TRAFFIC_FILE, f"hbm_bytes_per_launch_patch {pi['patches']} "
                                                   f"{pi['colours_per_launch'] % t.num_clusters} {pi['threads']} {pi['lanes']}")
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "kernel": kernel,
        "mean_launch_us": round(mean_ms * 1e3, 3),
        "algorithmic_bytes_per_launch": round(bytes_per_launch, 1),
        "bytes_per_vertex_update": round(b_sweep / t.num_verts, 2),
    }
    if kernel == "k_vbd_colour":
        roofline["colour_only_graph_launch_us"] = round(graph_ms * 1e3, 3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "grid_c1: 256x256 cloth (65,536 verts, 195,585 springs, 194,565 bending pairs), "
                        "top row pinned, seeded wind, 100 VBD iterations/frame, one frame per step",
            "colours": t.num_clusters,
            "iterations": iters,
            "strategy": {_lib.LATS_STRATEGY_PERSISTENT: "persistent", _lib.LATS_STRATEGY_FLOW: "flow",
                         _lib.LATS_STRATEGY_PATCH: "patch"}.get(strategy, "graph"),
            "persistent_blocks_or_tiles": blocks,
            "lanes_per_vertex": g.lanes_per_vertex(),
            "regions": g.regions(),
            "colour_block_threads": g.colour_block(),
            "gpu_ms_per_step": gpu_ms / K,
        },
        "roofline": roofline,
    }
    if strategy == _lib.LATS_STRATEGY_PATCH:
        out["config"]["patch"] = {k: v for k, v in g.patch_info().items() if k != "build_ms"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cloth, params, wind)
    if not args.no_warm_start:
        patch = None
        if strategy == _lib.LATS_STRATEGY_PATCH:
            pi = g.patch_info()
            patch = (pi["patches"], pi["colours_per_launch"], pi["threads"])
        out["warm_start"] = bench_warm_start(args, cloth, params, wind, plain_ms=gpu_ms / K,
                                             lanes=g.lanes_per_vertex(), regions=g.regions(),
                                             colour_block=g.colour_block(), patch=patch)
    if others is not None:
        out["other_workloads"] = others
    print(json.dumps(out), flush=True)


# The other configured workloads, each a short run of this script in a child
# process (its own GPU context), summarised beside the grid_c1 line.
OTHER_WORKLOADS = [
    ("batch_c2", ["--workload", "batch_c2", "--steps", "5", "--warmup", "1", "--no-cpu-baseline"]),
    ("sphere_c4", ["--workload", "sphere_c4", "--steps", "5", "--warmup", "1", "--no-warm-start"]),
    ("xpbd_c1", ["--workload", "xpbd_c1", "--steps", "5", "--warmup", "1", "--no-cpu-baseline"]),
    ("square8k", ["--workload", "square8k", "--steps", "20", "--warmup", "3"]),
]


def other_workloads():
    res = {}
    for name, extra in OTHER_WORKLOADS:
        cmd = [sys.executable, os.path.abspath(__file__)] + extra
        print(f"[bench] {name}: python3 bench.py {' '.join(extra)}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
            line = [x for x in r.stdout.splitlines() if x.startswith("{")]
            if r.returncode != 0 or not line:
                res[name] = {"error": f"exit {r.returncode}: {r.stderr[-300:]}"}
                continue
            d = json.loads(line[-1])
        except (subprocess.TimeoutExpired, ValueError) as e:
            res[name] = {"error": str(e)[:300]}
            continue
        keep = {k: d[k] for k in ("metric", "value", "unit", "ms_per_step", "steps", "warmup", "ms_per_iteration",
                                  "reference_published_ms_per_iteration", "speedup_vs_reference_published")
                if k in d}
        keep["command"] = "python3 bench.py " + " ".join(extra)
        keep["seconds"] = round(time.perf_counter() - t0, 1)
        print(f"[bench] {name}: {keep['seconds']} s", file=sys.stderr, flush=True)
        keep["config"] = {k: v for k, v in d.get("config", {}).items()}
        if "roofline" in d:
            keep["roofline"] = d["roofline"]
        res[name] = keep
    return res


BF16_DENSE_PEAK_TFLOPS = 2500.0


def _mlp_roofline(rows, ms):
    """The MFMA forward kernel (k_mlp_forward): algorithmic FLOPs of `rows`
    rows (unpadded 48-128-128-128-3) / its mean duration, vs dense BF16 peak."""
    from libatssim_amd.warm_start import FLOPS_PER_ROW
    tflops = rows * FLOPS_PER_ROW / (ms * 1e-3) / 1e12
    return {"kernel": "k_mlp_forward", "bound": "mfma", "rows": rows, "mean_launch_us": round(ms * 1e3, 3),
            "achieved": round(tflops, 2), "peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / BF16_DENSE_PEAK_TFLOPS, 5)}


def bench_mlp(args):
    """The warm-start network's kernels alone on 65,536 random rows (the
    256x256 cloth's vertex count): inference forward and a training step."""
    import numpy as np
    from libatssim_amd.warm_start import WarmStartMLP
    rows = 65536
    m = WarmStartMLP(2024)
    rng = np.random.default_rng(0)
    m.append_rows(rng.standard_normal((rows, 48)).astype(np.float32),
                  rng.standard_normal((rows, 3)).astype(np.float32) * 0.05)
    fwd_ms, step_ms = m.time_kernels(rows, args.steps)
    print(json.dumps({"workload": "mlp_kernels", "rows": rows, "forward": _mlp_roofline(rows, fwd_ms),
                      "train_step_ms": round(step_ms, 4), "iters": args.steps}), flush=True)


def bench_warm_start(args, cloth, params, wind, plain_ms, lanes=8, regions=1, colour_block=256, patch=None):
    """Learned warm start on the same cloth and wind, against the plain frame
    graph, per frame, measured two ways:

    oracle_stopped: sweeps until the energy reaches the plain solver's
      100-iteration energy from the same frame start, which a second context
      computes beforehand (outside the timed call).  It shows how soon the
      warm-started solve gets there; a deployed solver does not know that
      energy.
    deployable: the target-free stopping rule (energy decrease per sweep <=
      tol * |E|) with tol and the check schedule calibrated on four training
      scenarios; the plain energy is computed only to report how many frames
      reach it.

    The warm-started frames run the plain frame's configuration: its lanes,
    bands and block size, and with patch = (patches, colours per launch,
    threads) its patch strategy."""
    import numpy as np
    from libatssim_amd import configs
    from libatssim_amd.warm_start import WarmStartMLP, calibrate_stop_schedule, evaluate, evaluate_rule, train_recipe
    t0 = time.perf_counter()
    scen = [configs.warm_c3_scenario(k, n=256, frames=args.warm_frames) for k in range(args.warm_scenarios)]
    m = WarmStartMLP(2024)
    steps = args.warm_steps
    # rounds of rows from warm-started trajectories: 3 -> 13.7 sweeps per
    # frame, 5 -> 11.6, 8 -> 11.1 (more rounds, states closer to the ones the
    # warm-started solver visits)
    # (the dataset holds every pass; at most ~400 M rows = 58 GB, so larger
    # training sets get fewer rounds, never fewer than 3)
    rows_per_pass = args.warm_scenarios * (-(-args.warm_frames // args.warm_record_every)) * cloth.num_verts
    rounds = max(3, min(8, 400_000_000 // rows_per_pass - 1))
    train_recipe(m, scen, params, args.warm_frames, plain_steps=steps, guided_rounds=rounds, guided_steps=steps,
                 record_every=args.warm_record_every)
    cal_frames = min(30, args.warm_frames)
    cal_scen = min(4, args.warm_scenarios)
    from libatssim_amd import _lib

    def setup(g):
        if patch is not None:
            g.set_patches(*patch)
            g.set_strategy(_lib.LATS_STRATEGY_PATCH)

    cal = calibrate_stop_schedule(m, [configs.warm_c3_scenario(k, n=256, frames=cal_frames) for k in range(cal_scen)],
                                  params, cal_frames, lanes=lanes, regions=regions, colour_block=colour_block,
                                  setup=setup)
    tol, fc, ce = cal["tol"], cal["first_check"], cal["check_every"]
    train_s = time.perf_counter() - t0
    full_wind = configs.grid_c1(frames=60)[3]
    frames = len(full_wind)

    def summary(res):
        wall = np.array([r["wall_ms"] for r in res])
        sweeps = np.array([r["sweeps"] for r in res])
        excess = np.array([(r["e_final"] - r["target"]) / abs(r["target"]) for r in res])
        return {"ms_per_frame": round(float(wall.mean()), 4), "speedup": round(plain_ms / float(wall.mean()), 3),
                "mean_sweeps": round(float(sweeps.mean()), 2), "max_sweeps": int(sweeps.max()),
                "kept_fraction": round(float(np.mean([r["kept"] for r in res])), 3),
                "frames_reaching_plain_energy": round(float(np.mean(excess <= 0)), 3),
                "max_energy_excess": float(f"{excess.max():.3g}")}

    # the warm-started frames run the plain frame's colour-pass configuration
    oracle = summary(evaluate(m, cloth, params, full_wind, frames, lanes=lanes, regions=regions,
                              colour_block=colour_block, setup=setup))
    oracle["stop"] = ("first check at or below the plain 100-iteration energy of the same frame start, computed "
                      "beforehand on a second context (not known to a deployed solver)")
    deploy = summary(evaluate_rule(m, cloth, params, full_wind, frames, tol, first_check=fc, check_every=ce,
                                   lanes=lanes, regions=regions, colour_block=colour_block, setup=setup))
    deploy["stop"] = "energy decrease per sweep <= tol * |E| since the previous check (no reference solve)"
    deploy["tol"] = tol
    deploy["first_check"] = fc
    deploy["check_every"] = ce
    deploy["calibration"] = {"scenarios": cal_scen, "frames": cal_frames,
                             "rule": "fastest (schedule, tol) within 1 point of the best share of frames reaching "
                                     "the plain energy, on training scenarios",
                             "candidates": [{k: (round(v, 4) if isinstance(v, float) and k != "tol" else v)
                                             for k, v in t.items()} for t in cal["tried"]]}
    fwd_ms, step_ms = m.time_kernels(cloth.num_verts, 20)   # after the evaluation: timing steps train
    return {
        "ms_per_frame_plain": round(plain_ms, 4),
        "strategy": "patch" if patch is not None else "graph",
        "frames": frames,
        "deployable": deploy,
        "oracle_stopped": oracle,
        "mlp_forward": _mlp_roofline(cloth.num_verts, fwd_ms),
        "mlp_train_step_ms": round(step_ms, 4),
        "training": {"scenarios": args.warm_scenarios, "frames": args.warm_frames,
                     "record_every": args.warm_record_every, "guided_rounds": rounds,
                     "adam_steps": (1 + rounds) * steps,
                     "batch_rows": 16384, "rows": m.dataset_rows(), "seconds": round(train_s, 1)},
    }


def bench_sphere(args):
    """1024x1024 cloth dropped on the sphere (contact detection every frame,
    50 iterations per frame): plain GPU VBD throughput and the colour kernel's
    roofline at a million vertices."""
    from libatssim_amd import GpuVBD, configs, _lib
    cloth, params, (ox, of), ct = configs.sphere_scene()
    t = cloth.topology
    g = GpuVBD(cloth, params)
    g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])
    K, W = args.steps, args.warmup
    # the cloth reaches the sphere after ~9 frames; before that a frame has no
    # contacts and its cost does not rank the candidates as the contact
    # frames do, so the tuning (and the timed frames) start after 12 frames
    g.set_strategy(_lib.LATS_STRATEGY_GRAPH)
    if args.lanes:
        g.set_lanes_per_vertex(args.lanes)
    if args.regions:
        g.set_regions(args.regions)
    if args.colour_block:
        g.set_colour_block(args.colour_block)
    g.run_frames(12)
    if args.strategy == "auto":
        # three timed frames per candidate: one frame (~10 ms) is too short to
        # separate candidates a few per cent apart
        strategy, blocks = g.autotune(frames=3)
        if args.lanes:
            g.set_lanes_per_vertex(args.lanes)
        if args.regions:
            g.set_regions(args.regions)
        if args.colour_block:
            g.set_colour_block(args.colour_block)
    else:
        strategy, blocks = g.strategy()
    if W:
        g.run_frames(W)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_frames(K)
    wall = time.perf_counter() - t0
    gpu_ms = g.last_elapsed_ms()
    iters = params.constraint_iter_count
    counts, _ = g.contacts()
    b_sweep = algorithmic_bytes_per_sweep(t.num_verts, t.num_edges, t.num_bending)
    mean_ms, launches, _ = g.time_colour_kernels(10)
    achieved = b_sweep / t.num_clusters / (mean_ms * 1e-3) / 1e9
    out = {
        "metric": "VBD vertex-updates/s (1024x1024 cloth on a sphere)",
        "value": t.num_verts * iters * K / wall, "unit": "vertex-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"sphere_c4: {t.num_verts} verts, {t.num_clusters} colours, {iters} iterations/frame, "
                               f"contact detection every frame", "gpu_ms_per_step": gpu_ms / K,
                   "strategy": {_lib.LATS_STRATEGY_PERSISTENT: "persistent", _lib.LATS_STRATEGY_FLOW: "flow"}.get(
                       strategy, "graph"),
                   "lanes_per_vertex": g.lanes_per_vertex(), "regions": g.regions(),
                   "colour_block_threads": g.colour_block(), "contact_vertices": int((counts > 0).sum())},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": read_colour_traffic(os.path.join(ROOT, "profiles", "k_vbd_colour_sphere1024_traffic.txt"),
                                                    g.lanes_per_vertex(), g.regions(), g.colour_block()),
                     "traffic_measured_at_lanes_bands_threads": (read_colour_traffic_shape(
                         os.path.join(ROOT, "profiles", "k_vbd_colour_sphere1024_traffic.txt"),
                         g.lanes_per_vertex(), g.regions(), g.colour_block()) or (None, None))[1],
                     "kernel": "k_vbd_colour", "mean_launch_us": round(mean_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": round(b_sweep / t.num_clusters, 1)},
    }
    if not args.no_warm_start:
        out["warm_start"] = bench_sphere_warm_start(args, cloth, params, (ox, of), ct, plain_ms=gpu_ms / K,
                                                    lanes=g.lanes_per_vertex(), regions=g.regions())
    print(json.dumps(out), flush=True)


def bench_sphere_warm_start(args, cloth, params, obstacle, ct, plain_ms, scenarios=4, frames=120, lanes=8,
                            regions=1):
    """Warm start retrained on the sphere regime (plain and guided rollouts of
    smaller drops), evaluated on the 1024x1024 drop: per frame, wall time to
    the plain 50-iteration energy from the same frame start."""
    import numpy as np
    from libatssim_amd import configs
    from libatssim_amd.warm_start import WarmStartMLP, evaluate, train_recipe
    ox, of = obstacle

    def setup(g):
        g.set_obstacle(ox, of, ct["thickness"], ct["stiffness"], ct["radius"], ct["cell"])

    t0 = time.perf_counter()
    scen = [configs.sphere_train_scenario(k) for k in range(scenarios)]
    m = WarmStartMLP(2025)
    train_recipe(m, scen, params, frames, plain_steps=2000, guided_rounds=3, guided_steps=2000, record_every=2,
                 setup=setup)
    train_s = time.perf_counter() - t0
    n_eval = min(frames, args.warm_frames * 2)
    res = evaluate(m, cloth, params, None, n_eval, setup=setup, lanes=lanes, regions=regions)
    wall = np.array([r["wall_ms"] for r in res])
    sweeps = np.array([r["sweeps"] for r in res])
    # a frame's cost depends on its contacts: the plain reference is the mean
    # plain frame (graph, same lanes and bands) over the same frame range
    from libatssim_amd import GpuVBD, _lib
    p = GpuVBD(cloth, params, _lib.LATS_STRATEGY_GRAPH)
    setup(p)
    p.set_lanes_per_vertex(lanes)
    p.set_regions(regions)
    p.run_frames(1)                      # graph capture
    p.set_state(cloth.x, np.zeros_like(cloth.x))
    plain = []
    for _ in range(n_eval):
        p.run_frames(1)
        plain.append(p.last_elapsed_ms())
    p.close()
    plain_range = float(np.mean(plain))
    return {
        "ms_per_frame_learned": round(float(wall.mean()), 4), "ms_per_frame_plain": round(plain_range, 4),
        "speedup": round(plain_range / float(wall.mean()), 3), "frames": n_eval,
        "ms_per_frame_plain_timed_window": round(plain_ms, 4),
        "mean_sweeps": round(float(sweeps.mean()), 2), "max_sweeps": int(sweeps.max()),
        "kept_fraction": round(float(np.mean([r["kept"] for r in res])), 3),
        "stop": ("oracle-stopped: first check at or below the plain 50-iteration energy of the same frame start, "
                 "computed beforehand on a second context (not known to a deployed solver)"),
        "training": {"scenarios": scenarios, "cloth": "384x384, 2 mm", "frames": frames, "record_every": 2,
                     "guided_rounds": 3, "adam_steps": 8000, "rows": m.dataset_rows(), "seconds": round(train_s, 1)},
    }


def bench_batch(args):
    """512 uneven cloths (configs batch_c2), one persistent launch per frame."""
    from libatssim_amd import configs
    from libatssim_amd.vbd import GpuBatch
    cloths, ks, kb, params = configs.batch_c2()
    b = GpuBatch(cloths, params, ks, kb)
    n, nv = b.counts()
    threads, blocks = b.autotune(frames=1)
    K, W = args.steps, args.warmup
    if W:
        b.run_frames(W)
    t0 = time.perf_counter()
    b.run_frames(K)
    wall = time.perf_counter() - t0
    gpu_ms = b.last_elapsed_ms()
    iters = params.constraint_iter_count
    vu = nv * iters * K
    nb_sweep = sum(algorithmic_bytes_per_sweep(c.topology.num_verts, c.topology.num_edges, c.topology.num_bending)
                   for c in cloths)
    achieved = nb_sweep * iters * K / (gpu_ms * 1e-3) / 1e9
    out = {
        "metric": "batched VBD vertex-updates/s (512 cloths)",
        "value": vu / wall,
        "unit": "vertex-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batch_c2: {n} cloths, {nv} verts, 100 iterations/frame", "block_threads": threads,
                   "blocks": blocks, "lanes_per_vertex": b.lanes_per_vertex(),
                   "lds_staged_max_verts": b.lds_slots()[0],
                   "lds_staged_cloths": sum(c.topology.num_verts <= b.lds_slots()[0] for c in cloths),
                   "gpu_ms_per_step": gpu_ms / K},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     # one launch runs a frame (100 sweeps of every cloth): PMC bytes of one launch
                     "traffic": read_traffic(os.path.join(ROOT, "profiles", "k_vbd_batch_c2_profile.txt")),
                     "kernel": "k_vbd_batch", "mean_launch_us": round(gpu_ms / K * 1e3, 3),
                     "algorithmic_bytes_per_launch": round(nb_sweep * iters, 1)},
    }
    print(json.dumps(out), flush=True)


# The reference's published per-iteration cost of its own example on its own
# mesh (square8K.obj, 8,282 vertices, 10 colours): the sum over the 10
# colour clusters of the pre-profiled VBD task costs, Apple GPU (Metal) and
# Apple CPU (12 threads), examples/example3_hybrid_cpu_gpu.cpp:2297-2306.
REF_SQUARE8K_MS_PER_ITER = {"gpu_metal": 0.6527, "cpu_12_threads": 0.7306}


def bench_square8k(args):
    """The reference's own example: square8K.obj with example2's parameters
    (pins, mass 0.01, k_s 1e4, k_b 2.5e-3, damping 2, 100 iterations per
    frame), autotuned; ms per VBD iteration beside the reference's published
    per-iteration task costs of the same mesh."""
    import numpy as np
    from libatssim_amd import GpuVBD, configs, _lib
    d = np.load(os.path.join(ROOT, "tests", "golden", "square8k_mesh.npz"))
    cloth = configs.square8k_cloth(d["x"], d["faces"])
    params = configs.square8k_params()
    t = cloth.topology
    g = GpuVBD(cloth, params)
    if args.strategy == "auto":
        strategy, blocks = g.autotune(frames=3)
    else:
        # a fixed strategy (no autotune: e.g. under a profiler)
        if args.strategy == "patch":
            g.set_patches(args.patches, args.patch_colours, args.patch_threads)
        if args.lanes:
            g.set_lanes_per_vertex(args.lanes)
        g.set_strategy({"graph": _lib.LATS_STRATEGY_GRAPH, "persistent": _lib.LATS_STRATEGY_PERSISTENT,
                        "flow": _lib.LATS_STRATEGY_FLOW, "patch": _lib.LATS_STRATEGY_PATCH}[args.strategy])
        strategy, blocks = g.strategy()
    K, W = args.steps, args.warmup
    if W:
        g.run_frames(W)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_frames(K)
    wall = time.perf_counter() - t0
    gpu_ms = g.last_elapsed_ms()
    iters = params.constraint_iter_count
    ms_iter = gpu_ms / K / iters
    out = {
        "metric": "VBD vertex-updates/s (square8K.obj, the reference's example mesh)",
        "value": t.num_verts * iters * K / wall, "unit": "vertex-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "the reference's square8K.obj (as committed under tests/golden)",
        "config": {"workload": f"square8k: {t.num_verts} verts, {t.num_clusters} colours, {iters} iterations/frame, "
                               "example2 parameters",
                   "strategy": {_lib.LATS_STRATEGY_PERSISTENT: "persistent", _lib.LATS_STRATEGY_FLOW: "flow",
                                _lib.LATS_STRATEGY_PATCH: "patch"}.get(strategy, "graph"),
                   "blocks_or_patches": blocks, "lanes_per_vertex": g.lanes_per_vertex(),
                   "gpu_ms_per_step": gpu_ms / K},
        "ms_per_iteration": round(ms_iter, 5),
        "reference_published_ms_per_iteration": REF_SQUARE8K_MS_PER_ITER,
        "speedup_vs_reference_published": {k: round(v / ms_iter, 1) for k, v in REF_SQUARE8K_MS_PER_ITER.items()},
    }
    print(json.dumps(out), flush=True)


# Compulsory bytes of one projection (xpbd.h): a spring reads its slot pair,
# multiplier, rest length, two inverse masses and two positions, and writes
# the multiplier and two positions; a bending pair reads its four slots, Q
# (16 floats), multiplier, four inverse masses and four positions, and
# writes the multiplier and four positions.
XPBD_SPRING_BYTES = 8 + 4 + 4 + 8 + 24 + 4 + 24
XPBD_BENDING_BYTES = 16 + 64 + 4 + 16 + 48 + 4 + 48


def bench_xpbd(args):
    """The XPBD solver on the metric's cloth: 256x256, 100 iterations per
    frame, every spring cluster then every bending cluster per iteration,
    one kernel per constraint cluster in a frame graph.  value = constraint
    projections per second; roofline of the cluster kernel with the larger
    share of the frame."""
    from libatssim_amd import GpuVBD, configs
    K, W = args.steps, args.warmup
    cloth, params, _, _ = configs.grid_c1(frames=W + K)
    t = cloth.topology
    g = GpuVBD(cloth, params)
    g.set_solver("xpbd")
    ns, nb = len(g.xpbd_clusters(0)), len(g.xpbd_clusters(1))
    if W:
        g.run_frames(W)
    g.synchronize()
    t0 = time.perf_counter()
    g.run_frames(K)
    g.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = g.last_elapsed_ms()
    iters = params.constraint_iter_count
    projections = (t.num_edges + t.num_bending) * iters * K
    ms_s, ms_b, ls, lb = g.time_xpbd_kernels(10)
    per_s = t.num_edges * XPBD_SPRING_BYTES / ns
    per_b = t.num_bending * XPBD_BENDING_BYTES / nb
    share_s, share_b = ms_s * ns, ms_b * nb
    kern, per, ms = ("k_xpbd_bending", per_b, ms_b) if share_b >= share_s else ("k_xpbd_stretch", per_s, ms_s)
    achieved = per / (ms * 1e-3) / 1e9
    out = {
        "metric": "XPBD constraint projections/s (256x256 cloth)",
        "value": projections / wall, "unit": "projections/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"xpbd_c1: 256x256 cloth ({t.num_edges} springs, {t.num_bending} bending pairs), "
                               f"top row pinned, {iters} XPBD iterations/frame, one frame per step",
                   "spring_clusters": ns, "bending_clusters": nb, "gpu_ms_per_step": gpu_ms / K},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": read_traffic(os.path.join(ROOT, "profiles", "k_xpbd_grid256_traffic.txt"),
                                             "hbm_bytes_per_launch_" + kern), "kernel": kern,
                     "mean_launch_us": round(ms * 1e3, 3), "algorithmic_bytes_per_launch": round(per, 1),
                     "stretch_launch_us": round(ms_s * 1e3, 3), "bending_launch_us": round(ms_b * 1e3, 3),
                     "bytes_per_spring": XPBD_SPRING_BYTES, "bytes_per_bending_pair": XPBD_BENDING_BYTES},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_xpbd(cloth, params)
    print(json.dumps(out), flush=True)


def cpu_baseline_xpbd(cloth, params, min_seconds=10.0, max_frames=40):
    from oracle import oracle as O
    threads = min(O.max_threads(), 16)
    m = O.Mesh(cloth.x, cloth.faces)
    s = O.Sim(m, cloth.mass, cloth.fixed, implicit_dt=params.implicit_dt, num_substep=params.num_substep,
              iter_count=params.constraint_iter_count, k_stretch=params.stiffness_stretch_spring,
              k_bending=params.stiffness_quadratic_bending, damping=params.damping_cloth, threads=threads)
    t = cloth.topology
    frames = 0
    t0 = time.perf_counter()
    while True:
        s.step_xpbd()
        frames += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or frames >= max_frames:
            break
    n = (t.num_edges + t.num_bending) * params.constraint_iter_count * frames
    return {"value": n / el, "unit": "projections/s", "cores": threads, "kind": "port",
            "sample": f"{frames} XPBD frames of the same 256x256 workload ({n} projections, {el:.1f} s), "
                      f"oracle/vbd_oracle.cpp FP32, OpenMP per-cluster passes",
            "host_cpu": _host_cpu()[0], "host_logical_cpus": _host_cpu()[1]}


def bench_launch_floor(args):
    """Latency floors at the colour-pass shape, microseconds: an empty kernel
    and a dependent-gather kernel per launch (graphs of 800 launches, as a
    frame graph of 7 colours x 100 sweeps replays its passes), and one grid
    barrier inside a cooperative kernel (800 barriers), at the thread counts
    of the 256x256 colour passes (8,192-vertex colours x 8 lanes = 65,536
    threads) and of 256 / 64 blocks for the barrier; launches that re-read
    one 2 MiB buffer (for the FETCH_SIZE counter: does a launch find the
    previous launch's lines in L2?); and the hand-off between neighbouring
    blocks of one persistent kernel (per phase, with and without fences)."""
    import ctypes
    from libatssim_amd import _lib
    L = _lib.lib()

    def floor(kind, threads, launches=800):
        us = ctypes.c_float()
        _lib.check(_lib.measure_lib().lats_time_floor(kind, threads, launches, ctypes.byref(us)))
        return round(us.value, 3)

    out = {"workload": "launch_floor", "launches_per_graph": 800, "unit": "us",
           "empty_65536_threads": floor(_lib.LATS_FLOOR_EMPTY, 65536),
           "empty_64_threads": floor(_lib.LATS_FLOOR_EMPTY, 64),
           "gather_65536_threads": floor(_lib.LATS_FLOOR_GATHER, 65536),
           "gather_8192_threads": floor(_lib.LATS_FLOOR_GATHER, 8192),
           "grid_barrier_256_blocks": floor(_lib.LATS_FLOOR_GRID_BARRIER, 256 * 256),
           "grid_barrier_64_blocks": floor(_lib.LATS_FLOOR_GRID_BARRIER, 64 * 256),
           "reread_2MiB_buffer": floor(_lib.LATS_FLOOR_REREAD, 131072),
           "handoff_fenced_256_blocks": floor(_lib.LATS_FLOOR_HANDOFF_FENCED, 256 * 256),
           "handoff_relaxed_256_blocks": floor(_lib.LATS_FLOOR_HANDOFF_RELAXED, 256 * 256),
           "handoff_relaxed_same_xcd_256_blocks": floor(_lib.LATS_FLOOR_HANDOFF_XCD, 256 * 256)}
    print(json.dumps(out), flush=True)


def bench_comm_matrix(args):
    """The scheduler's cost inputs measured on this GPU for a square8K-sized
    cloth (91x91 grid, 8,281 vertices), as the reference hard-codes them for
    its Apple machine (communication_cost_matrix_uma, example3_hybrid_cpu_gpu.cpp:
    2341-2345; GPU column of the computation matrix, :2276-2311), in ms:
      cpu_to_gpu  host signal -> the stream passes its wait -> iterate + begin
                  buffers uploaded (pinned, vertex order) -> marker seen by the host
      gpu_to_cpu  iterate + begin buffers downloaded -> marker seen by the host
      gpu_gpu     one dispatch of the smallest cluster (the per-task floor)
      gpu_cluster per-cluster all-in-one task (mean of repeated dispatches)."""
    import ctypes
    import numpy as np
    from libatssim_amd import GpuVBD, Params, configs, _lib
    L = _lib.lib()
    cloth = configs.grid_cloth(91, "top", seed=3)
    g = GpuVBD(cloth, Params(constraint_iter_count=100))
    g.physics_step()
    V = cloth.num_verts
    nbytes = 3 * V * 4
    ptrs = []
    for _ in range(2):
        p = ctypes.c_void_p()
        _lib.check(L.lats_host_alloc(nbytes, ctypes.byref(p)))
        ptrs.append(p)
    h = g.handle
    reps = max(args.steps, 20)

    def wait(marker):
        done = ctypes.c_int(0)
        while not done.value:
            _lib.check(L.lats_vbd_query_marker(h, marker, ctypes.byref(done)))

    # predict once so the first cluster task can copy sa_x into buffer 0
    g.dispatch(_lib.LATS_FN_PREDICT_POSITION)
    g.dispatch(_lib.LATS_FN_VBD_ALL_IN_ONE, iter_idx=0, cluster_idx=0, buffer_idx=0,
               left_buffer_idx=_lib.LATS_BUFFER_DEFAULT)
    # the staging starts with the buffer's own contents (uploads then rewrite the same positions)
    _lib.check(L.lats_vbd_buffer_download(h, _lib.LATS_BUF_ITER, 0, ptrs[0]))
    _lib.check(L.lats_vbd_buffer_download(h, _lib.LATS_BUF_BEGIN, 0, ptrs[1]))
    g.synchronize()
    c2g, g2c = [], []
    value = 0
    for r in range(reps + 3):
        value += 1
        _lib.check(L.lats_vbd_stream_wait_host(h, value))
        _lib.check(L.lats_vbd_buffer_upload(h, _lib.LATS_BUF_ITER, 0, ptrs[0]))
        _lib.check(L.lats_vbd_buffer_upload(h, _lib.LATS_BUF_BEGIN, 0, ptrs[1]))
        _lib.check(L.lats_vbd_record_marker(h, 1))
        time.sleep(0.002)                       # the stream is parked on its wait
        t0 = time.perf_counter()
        _lib.check(L.lats_vbd_host_signal(h, value))
        wait(1)
        t1 = time.perf_counter()
        _lib.check(L.lats_vbd_buffer_download(h, _lib.LATS_BUF_ITER, 0, ptrs[0]))
        _lib.check(L.lats_vbd_buffer_download(h, _lib.LATS_BUF_BEGIN, 0, ptrs[1]))
        _lib.check(L.lats_vbd_record_marker(h, 2))
        wait(2)
        t2 = time.perf_counter()
        if r >= 3:
            c2g.append((t1 - t0) * 1e3)
            g2c.append((t2 - t1) * 1e3)
    sizes = [len(c) for c in cloth.topology.clusters()]
    per_cluster = []
    for c in range(len(sizes)):
        _lib.check(L.lats_vbd_record_marker(h, 10))
        for _ in range(reps):
            g.dispatch(_lib.LATS_FN_VBD_ALL_IN_ONE, iter_idx=1, cluster_idx=c, buffer_idx=0)
        _lib.check(L.lats_vbd_record_marker(h, 11))
        g.synchronize()
        ms = ctypes.c_float()
        _lib.check(L.lats_vbd_marker_elapsed_ms(h, 10, 11, ctypes.byref(ms)))
        per_cluster.append(ms.value / reps)
    for p in ptrs:
        L.lats_host_free(p)
    out = {
        "workload": "comm_matrix", "cloth": "91x91 grid", "verts": V, "clusters": sizes,
        "cpu_to_gpu_ms": round(float(np.median(c2g)), 4), "gpu_to_cpu_ms": round(float(np.median(g2c)), 4),
        "gpu_gpu_ms": round(min(per_cluster), 4), "gpu_cluster_ms": [round(v, 4) for v in per_cluster],
        "reference_apple_uma_ms": {"cpu_to_gpu": 0.220, "gpu_to_cpu": 0.145, "cpu_cpu": 0.002, "gpu_gpu": 0.01},
        "repeats": reps,
    }
    print(json.dumps(out), flush=True)


def copy_calibration(args):
    """A plain device-to-device copy of 256 MiB: the HBM-counter calibration
    (FETCH_SIZE / WRITE_SIZE per dispatch of this copy vs its 256 MiB)."""
    import torch
    n = 64 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup + args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy_calibration", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                      "copies": args.warmup + args.steps}), flush=True)


if __name__ == "__main__":
    try:
        main()
    finally:
        # every context, batch and network is destroyed before the interpreter
        # (and the HIP runtime) shuts down
        from libatssim_amd import _lib as _l
        _l.close_all()
