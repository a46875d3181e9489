#!/usr/bin/env python3
"""SPH steps/s of the two-body impact at 1e6 particles on one MI355X.

One step = Headless::step of the reference (Headless.cpp:37-55): octree
hand-off, sort, FMM gravity, density, acceleration, drift/kick, then a wait
for the GPU -- the reference waits once per frame.  Inputs are generated
on the host, uploaded once and stay resident in HBM; the timed region
contains only steps.

Prints one JSON line with the headline metric, the per-pass times (HIP events
on the context's stream), the roofline of the dominant kernel and the oracle
(OpenMP CPU restatement) timed on the host's cores from the same warmed
state.
"""
import argparse
import importlib
import json
import os
import sys
import time

# OpenMP placement of the CPU baseline (the oracle): one thread per core,
# packed, fixed before any OpenMP runtime starts
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "real-time-planetary-impacts-mps_amd"

METRIC = "SPH steps/s at 1e6 particles; density-pass ms, learned-h vs Newton-Raphson"

# MI355X vendor peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3

# Algorithmic work of the density pass (per unit), counted by the kernel's
# work counters:
#   9 FLOP per neighbour candidate examined (distance + compare), every scan,
#   +9 per candidate inside 2h of an NR scan (sqrt, kernel, accumulate),
#   +54 per neighbour of the final pass (sqrt, dW/dh, grad W, div, curl).
DENSITY_BYTES_PER_PARTICLE = 69.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="particles (default 1e6; 4e6 with --learned-precision)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--tree", choices=["async", "every", "fixed", "gpu"], default="async",
                    help="octree: CPU thread (async, as the reference), CPU before every step, the first one kept, "
                         "or built on the GPU before every step")
    ap.add_argument("--tree-threads", type=int, default=8)
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--c1-steps", type=int, default=100,
                    help="steps of the 2e4-particle run on the CPU path and on the GPU (0 = skip)")
    ap.add_argument("--no-timing", action="store_true", help="do not record per-pass HIP events")
    ap.add_argument("--density-probe", action="store_true",
                    help="after the warm-up steps, one sort and two density passes back to back, then exit "
                         "(for profiler counters of a density launch that follows another)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes 5 times (calibrates the "
                         "HBM byte counters of a profiler against a kernel of known traffic) and exit")
    ap.add_argument("--learned-steps", type=int, default=2000,
                    help="Adam steps of the learned-h network (0 = skip the learned-h comparison)")
    ap.add_argument("--learned-eta", type=float, default=2.5e-6,
                    help="Newton-Raphson tolerance of the learned-h comparison: distance of h from the fixed point "
                         "(density_stop = 1), so the two runs end within about twice this of each other")
    ap.add_argument("--density-lanes", type=int, default=None,
                    help="lanes per particle in the density pass (0, 8, 16, 32, 64; default: the library's)")
    ap.add_argument("--accel-lanes", type=int, default=None,
                    help="lanes per particle in the acceleration pass (0, 8, 16, 32, 64; default: the library's)")
    ap.add_argument("--param", action="append", default=[], metavar="NAME=VALUE",
                    help="override one sph_params field (repeatable), e.g. --param gravity_eta=2e-5")
    ap.add_argument("--headon", action="store_true",
                    help="head-on impact variant (impact parameter 0.3, 12 km/s, u0 5e6 J/kg)")
    ap.add_argument("--fixes", type=lambda v: int(v, 0), default=0,
                    help="SPH_FIX_* bit mask (0 = the reference step as written, 0x3f = every correction)")
    ap.add_argument("--learned-precision", action="store_true",
                    help="instead of the headline: the learned-h + precision form against the plain FP32 form on "
                         "a head-on impact (--n, default 4e6 here), README activity rule with up to 8 rungs; "
                         "prints one JSON line with both forms' steps/s and per-pass times")
    args = ap.parse_args()
    if args.n is None:
        args.n = 4_000_000 if args.learned_precision else 1_000_000
    return args


def apply_overrides(prm, args):
    """--density-lanes, --accel-lanes and --param NAME=VALUE on an sph_params
    (lanes given on the command line are used at any particle count)."""
    if args.density_lanes is not None:
        prm.density_lanes = args.density_lanes
        prm.small_n_lanes = 0
    if args.accel_lanes is not None:
        prm.accel_lanes = args.accel_lanes
        prm.small_n_lanes = 0
    for kv in args.param:
        name, value = kv.split("=", 1)
        cur = getattr(prm, name)
        setattr(prm, name, type(cur)(float(value)) if isinstance(cur, float) else int(value, 0))
    return prm


def profiled_traffic():
    """HBM bytes per density-pass launch at the default workload, measured by
    two rocprofv3 counter passes and recorded in profiles/."""
    path = os.path.join(ROOT, "profiles", "k_density_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def density_flops(work):
    w = work.astype(np.float64).sum(axis=0)
    return 9.0 * w[0] + 9.0 * w[1] + 9.0 * w[2] + 54.0 * w[3]


def host_cpu():
    """CPU model, physical cores and logical CPUs of this host (lscpu's
    sources: /proc/cpuinfo and the process's CPU affinity).  Called before
    the OpenMP runtime starts: with OMP_PROC_BIND set it binds this thread
    to one place, and the affinity would then read 1."""
    model, phys = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            cur = {}
            for line in f:
                if ":" in line:
                    k, v = [x.strip() for x in line.split(":", 1)]
                    cur[k] = v
                    if k == "model name":
                        model = v
                elif cur:
                    phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
                    cur = {}
            if cur:
                phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "physical_cores_host": len(phys) or None, "logical_cpus_usable": usable}


CPU_PASSES = ("sort", "gravity", "density", "acceleration", "step")


def oracle_steps(oracle, st, prm, steps, rebuild_tree):
    """Runs the oracle for `steps` steps with per-pass wall time; the octree
    (serial CPU build, as in the reference) is timed apart."""
    ms = {k: 0.0 for k in CPU_PASSES + ("octree",)}
    t_all = time.perf_counter()
    for i in range(steps):
        if rebuild_tree or i == 0:
            t0 = time.perf_counter()
            st.set_tree(oracle.build_octree(st["pos"], prm.max_children_in_leaf, prm.max_tree_depth))
            ms["octree"] += 1000 * (time.perf_counter() - t0)
        for name in CPU_PASSES:
            t0 = time.perf_counter()
            getattr(st, name)()
            ms[name] += 1000 * (time.perf_counter() - t0)
    wall = time.perf_counter() - t_all
    step_ms = sum(ms[k] for k in CPU_PASSES)
    return {"steps": steps, "steps_per_s": steps / (step_ms / 1000), "steps_per_s_with_octree": steps / wall,
            "pass_ms": {k: round(v / steps, 3) for k, v in ms.items()}}


def cpu_baseline(sph, oracle, hl, prm, tables, steps):
    """Times the oracle on the warmed GPU state (same particles, same tree rule)."""
    ctx = hl.ctx
    ic = {k: ctx.download(k) for k in ("pos", "vel", "rho", "u", "mass", "mat", "pressure", "h")}
    st = oracle.State(prm, ic, tables)
    for f in ("acc", "gradterm", "soundspeed", "du", "balsara", "dhdt", "active", "next_active", "gravnorm",
              "global_time", "dt"):
        st[f] = ctx.download(f)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r = oracle_steps(oracle, st, prm, steps, rebuild_tree=False)
    cpu = HOST_CPU
    return {
        "value": r["steps_per_s"],
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{steps} steps of the same {st.n}-particle impact from the warmed GPU state "
                   f"(sort, gravity, density, acceleration, step), OpenMP over {threads} threads "
                   f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, OMP_PLACES={os.environ.get('OMP_PLACES')}) "
                   f"on {cpu['cpu_model']}; the serial octree build ({r['pass_ms']['octree']:.0f} ms) is excluded "
                   "as it runs on a separate thread in the GPU path"),
        "pass_ms": r["pass_ms"],
        **cpu,
    }


def c1_run(sph, oracle, args):
    """The 2e4-particle impact (89% / 11%, seed 1), 100 global steps, on the
    CPU path (the oracle, octree rebuilt serially before every step) and on
    the GPU (Headless, octree rebuilt before every step): steps/s and
    per-pass ms of both."""
    prm = sph.default_params()
    tables = sph.make_eos_tables(prm, 256)
    cfg = sph.impact_config(20000, 1)
    ic = sph.make_impact(prm, cfg, tables)
    st = oracle.State(prm, ic, tables)
    cpu = oracle_steps(oracle, st, prm, args.c1_steps, rebuild_tree=True)
    cpu["threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    hl = sph.Headless(prm, cfg, eos_res=256, tree_mode=sph.TREE_EVERY, tree_threads=args.tree_threads)
    ctx = hl.ctx
    ctx.reset_pass_times()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.c1_steps):
        hl.step()
    ctx.sync()
    el = time.perf_counter() - t0
    passes = ctx.pass_times()
    ctx.set_timing(False)
    gpu = {"steps": args.c1_steps, "steps_per_s_with_octree": args.c1_steps / el,
           "pass_ms": {k: round(v["ms"] / v["calls"], 4) for k, v in passes.items() if v["calls"]}}
    d = float(np.abs(ctx.download("pos")[:, :3] - st["pos"][:, :3]).max())
    hl.close()
    return {"workload": f"two-body 45-degree impact, 20000 particles, {args.c1_steps} steps, reference step as "
                        "written, octree rebuilt before every step", "cpu": cpu, "gpu": gpu,
            "final_positions_max_abs_diff": d}


def learned_comparison(sph, args):
    """Learned-h against plain Newton-Raphson at a tightened tolerance.

    Samples (features, converged h) are dumped by plain runs of seeded 1e5
    (seed 11) and N-particle (seed 12) impacts, the network is trained on the
    GPU (Adam, batch 65536), then the density pass of the benchmark's own
    impact (seed --seed, after 3 steps) runs twice from the same state:
    plain, and from the network's guess.  Both stop once the estimated
    distance of h from the fixed point of the update is within --learned-eta
    (density_stop = 1, at most 400 iterations), so the two converged h agree
    to about twice that; the pass time includes the feature and inference
    kernels.
    """
    prm = apply_overrides(sph.default_params(fixes=args.fixes), args)
    prm.density_eta = args.learned_eta
    prm.density_stop = 1
    prm.max_nr_iterations = 400
    net, train = train_guess(sph, args, prm, ((100_000, 11), (args.n, 12)), args.headon)
    hl = sph.Headless(prm, sph.impact_config(args.n, args.seed, headon=args.headon), eos_res=256, tree_mode=sph.TREE_EVERY,
                      tree_threads=args.tree_threads)
    for _ in range(3):
        hl.step()
    ctx = hl.ctx
    ctx.sort()
    ctx.sync()
    keep = {k: ctx.download(k) for k in ("h", "rho", "global_time", "dt")}
    reps = 5
    out = {}
    for mode in ("plain", "learned"):
        ctx.set_learned_guess(net if mode == "learned" else None)
        ctx.reset_pass_times()
        ctx.set_timing(True)
        for _ in range(reps):
            for k, v in keep.items():
                ctx.upload(k, v)
            ctx.density()
        t = ctx.pass_times()["density"]
        ctx.set_timing(False)
        out[mode] = {"ms": t["ms"] / max(t["calls"], 1), "h": ctx.download("h").astype(np.float64),
                     "it": float(ctx.download("nr_iters").mean()), "it_max": int(ctx.download("nr_iters").max())}
    ctx.set_learned_guess(None)
    hl.close()
    dh = np.abs(out["learned"]["h"] / out["plain"]["h"] - 1)
    res = {
        "nr_tolerance": args.learned_eta,
        "nr_stop_rule": "estimated distance of h from the fixed point <= nr_tolerance (density_stop = 1), "
                        "at most 400 iterations, in both runs",
        "nr_iterations_max_plain": int(out["plain"]["it_max"]),
        "nr_iterations_max_learned": int(out["learned"]["it_max"]),
        "density_ms_plain": round(out["plain"]["ms"], 4),
        "density_ms_learned": round(out["learned"]["ms"], 4),
        "density_speedup": round(out["plain"]["ms"] / out["learned"]["ms"], 4),
        "nr_iterations_plain": round(out["plain"]["it"], 4),
        "nr_iterations_learned": round(out["learned"]["it"], 4),
        "converged_h_rel_diff_max": float(dh.max()),
        "converged_h_rel_diff_q999": float(np.quantile(dh, 0.999)),
        "train": train,
    }
    net.close()
    return res


def train_guess(sph, args, prm, sizes, headon, dump_steps=8):
    """Trains the learned-h network on (features, converged h) dumped by plain
    runs of the given (n, seed) impacts.  Returns (net, info)."""
    smp = sph.Samples(sum(n for n, _ in sizes) * dump_steps)
    t0 = time.perf_counter()
    for n, seed in sizes:
        hl = sph.Headless(prm, sph.impact_config(n, seed, headon=headon), eos_res=256, tree_mode=sph.TREE_EVERY,
                          tree_threads=args.tree_threads)
        hl.step()
        hl.ctx.set_sample_dump(smp)
        for _ in range(dump_steps):
            hl.step()
        hl.ctx.set_sample_dump(None)
        hl.close()
    t_dump = time.perf_counter() - t0
    log(f"learned: {len(smp)} samples dumped in {t_dump:.1f} s")
    net = sph.Mlp(seed=13)
    net.fit_norm(smp)
    t0 = time.perf_counter()
    losses = net.train(smp, steps=args.learned_steps, batch=65536, lr=2e-3, seed=14)
    t_train = time.perf_counter() - t0
    log(f"learned: {args.learned_steps} Adam steps in {t_train:.2f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    info = {"samples": len(smp), "adam_steps": args.learned_steps, "batch": 65536, "lr": 2e-3,
            "seconds": round(t_train, 3), "loss_first": float(losses[0]),
            "loss_last10_mean": float(np.mean(losses[-10:])), "dump_seconds": round(t_dump, 2)}
    smp.close()
    return net, info


# Algorithmic bytes per particle of the sort phase (hash 20 B, three 8-bit
# radix passes of 16 B) and the compulsory bytes of the neighbour passes.
SORT_BYTES_PER_PARTICLE = 68.0
ACCEL_BYTES_PER_PARTICLE = 93.0


def learned_precision_run(sph, args):
    """The learned-h + precision form against the plain FP32 form, same GPU.

    Workload: head-on impact (impact parameter 0.3 (R_t + R_i), 12 km/s,
    u0 5e6 J/kg) of --n particles, individual timesteps (README rule, up to 8
    rungs, leaf/branch activation syncing kept), octree rebuilt on a CPU
    thread.  Plain: FP32, the reference step as written except for the
    inactive-particle drift (SPH_FIX_INACTIVE_DRIFT: as written, the README
    rule drives h of inactive particles to 0 and rho to infinity within a few
    steps at 1e6).  Learned-h + precision: every correction, compensated sums,
    FP64 time and conservation sums, third-order gravity at GRAVITY_ETA 1e-3,
    and the network (trained first on dumps of 1e5 and 1e6 head-on runs of the
    plain form) as the Newton-Raphson guess.  Both run --warmup untimed then
    --steps timed steps."""
    base = dict(activity_rule=sph.SPH_ACTIVE_README, max_rungs=8)
    forms = {
        "plain_fp32": sph.default_params(fixes=sph.SPH_FIX_INACTIVE_DRIFT, **base),
        "learned_precision": sph.default_params(fixes=sph.SPH_FIX_ALL, precision=1, expansion_order=3, mac_order=2,
                                                gravity_eta=1e-3, max_unchecked_pointers=8192, **base),
    }
    for prm in forms.values():
        apply_overrides(prm, args)
    # dumps with every particle active and the iteration stopped within
    # 2.5e-6 of its fixed point: the targets are the fixed points
    dump_prm = apply_overrides(sph.default_params(density_eta=2.5e-6, density_stop=1, max_nr_iterations=400), args)
    net, train = train_guess(sph, args, dump_prm, ((100_000, 11), (1_000_000, 12)), True)
    res = {}
    for name, prm in forms.items():
        hl = sph.Headless(prm, sph.impact_config(args.n, 5, headon=True), eos_res=256, tree_mode=sph.TREE_ASYNC,
                          tree_threads=args.tree_threads)
        ctx = hl.ctx
        if name == "learned_precision":
            ctx.set_learned_guess(net)
        for i in range(args.warmup):
            hl.step()
        ctx.sync()
        c0 = ctx.conservation() if prm.precision else None
        ctx.reset_pass_times()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        for i in range(args.steps):
            hl.step()
            if i % 20 == 0:
                log(f"{name}: step {i}")
        ctx.sync()
        el = time.perf_counter() - t0
        passes = ctx.pass_times()
        ctx.set_timing(False)
        per = {k: v["ms"] / max(v["calls"], 1) for k, v in passes.items()}
        r = {"steps_per_s": args.steps / el, "ms_per_step": 1000 * el / args.steps,
             "pass_ms": {k: round(v, 4) for k, v in per.items()},
             "active_fraction_last": float(ctx.download("active").mean()),
             "nr_iterations_mean_last": float(ctx.download("nr_iters").mean()),
             "sort_gbs_algorithmic": SORT_BYTES_PER_PARTICLE * args.n / (per["sort"] * 1e-3) / 1e9,
             "density_gbs_algorithmic": DENSITY_BYTES_PER_PARTICLE * args.n / (per["density"] * 1e-3) / 1e9,
             "acceleration_gbs_algorithmic": ACCEL_BYTES_PER_PARTICLE * args.n / (per["acceleration"] * 1e-3) / 1e9}
        if c0 is not None:
            c1 = ctx.conservation()
            e0 = c0["kinetic"] + c0["internal"] + c0["potential"]
            e1 = c1["kinetic"] + c1["internal"] + c1["potential"]
            r["energy_drift"] = abs(e1 - e0) / abs(e0)
            r["angular_momentum_drift"] = float(np.linalg.norm(c1["angular_momentum"] - c0["angular_momentum"]) /
                                                np.linalg.norm(c0["angular_momentum"]))
            r["time_s"] = ctx.time()
        ctx.set_learned_guess(None)
        hl.close()
        log(f"{name}: {r['steps_per_s']:.2f} steps/s")
        res[name] = r
    net.close()
    out = {
        "metric": "steps/s, learned-h + precision form vs plain FP32 form",
        "value": res["learned_precision"]["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "dtype": "fp32 (fp64 time and conservation sums in the precision form)",
        "data": "synthetic",
        "config": {"workload": f"two-body head-on impact, {args.n} particles, README activity rule, up to 8 rungs, "
                               "CPU octree rebuilt on a thread (async)", "seed": 5},
        "forms": res,
        "training": train,
    }
    print(json.dumps(out))


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: reads nbytes, writes nbytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for i in range(5):
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy {i} done")
    print(json.dumps({"copy_probe_bytes": src.numel() * 4, "launches": 5}))


HOST_CPU = {}


def main():
    HOST_CPU.update(host_cpu())
    args = parse()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "the step is built and measured for a single MI355X; its multi-GPU decomposition is a design, "
              "not a code path", file=sys.stderr)
        sys.exit(2)
    sph = importlib.import_module(PKG)
    if args.learned_precision:
        learned_precision_run(sph, args)
        return
    prm = apply_overrides(sph.default_params(fixes=args.fixes), args)
    tables = sph.make_eos_tables(prm, 256)
    cfg = sph.impact_config(args.n, args.seed, headon=args.headon)
    tree_mode = {"async": sph.TREE_ASYNC, "every": sph.TREE_EVERY, "fixed": sph.TREE_FIXED,
                 "gpu": sph.TREE_GPU}[args.tree]
    t_setup = time.perf_counter()
    hl = sph.Headless(prm, cfg, eos_res=256, tree_mode=tree_mode, tree_threads=args.tree_threads)
    ctx = hl.ctx
    log(f"setup {args.n} particles in {time.perf_counter() - t_setup:.1f} s")
    for i in range(args.warmup):
        t_w = time.perf_counter()
        hl.step()
        log(f"warmup step {i} {1000 * (time.perf_counter() - t_w):.1f} ms")
    ctx.sync()
    if args.density_probe:
        ctx.sort()
        ctx.density()
        ctx.density()
        ctx.sync()
        print(json.dumps({"density_probe": args.n, "launches": 2}))
        return
    trees0, _ = hl.tree_stats()
    if not args.no_timing:
        ctx.reset_pass_times()
        ctx.set_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hl.step()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    trees1, build_ms = hl.tree_stats()
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    passes = ctx.pass_times() if not args.no_timing else {}
    ctx.set_timing(False)
    steps_per_s = args.steps / elapsed

    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"two-body {'head-on' if args.headon else '45-degree'} impact, {args.n} particles "
                         "(89% target / 11% impactor), plain "
                         "Newton-Raphson smoothing length, FP32, "
                         + ("reference step as written" if args.fixes == 0 else f"corrections 0x{args.fixes:x}")
                         + (", octree rebuilt on the GPU before every step (Morton keys)" if args.tree == "gpu"
                            else f", CPU octree rebuilt on a thread ({args.tree})")),
            "fixes": args.fixes,
            "param_overrides": args.param,
            "n_particles": args.n,
            "eos_table": "256x256 Tillotson per material",
            "tree_mode": args.tree,
            "tree_threads": args.tree_threads,
            "density_lanes": int(prm.density_lanes) if args.n >= prm.small_n_lanes else 32,
            "accel_lanes": int(prm.accel_lanes) if args.n >= prm.small_n_lanes else (16 if args.n < 65536 else 8),
            "small_n_lanes": int(prm.small_n_lanes),
            "lane_split": int(prm.lane_split),
            "trees_swapped_in_timed_region": trees1 - trees0,
            "last_tree_build_ms": round(build_ms, 2),
        },
    }
    if passes:
        per_pass = {k: v["ms"] / v["calls"] for k, v in passes.items() if v["calls"]}
        out["pass_ms"] = {k: round(v, 4) for k, v in per_pass.items()}
        out["pass_ms_note"] = ("gravity_down runs on its own stream, concurrently with density; its events span "
                               "that overlap, so the passes do not add up to ms_per_step")
        out["density_ms"] = round(per_pass["density"], 4)
        # dominant kernel: the density pass is one kernel
        dlanes = int(prm.density_lanes) if args.n >= prm.small_n_lanes else 32
        kname = (f"k_density_group<{dlanes}, {int(prm.lane_split)}, {int(prm.density_flat)}, "
                 "false, false>" if dlanes > 0 else "k_density")
        # the reference algorithm's work, counted by one more sort + density
        # pass of the state after the timed steps (untimed: counting is a
        # measurement aid)
        nr = ctx.download("nr_iters")
        ctx.set_work_counters(True)
        ctx.sort()
        ctx.density()
        ctx.sync()
        ctx.set_work_counters(False)
        work = ctx.download("density_work")
        out["nr_iterations_mean"] = round(float(nr.mean()), 4)
        out["density_work_per_particle"] = {
            "nr_scan_candidates": round(float(work[:, 0].mean()), 1),
            "nr_scan_inside_2h": round(float(work[:, 1].mean()), 1),
            "final_candidates": round(float(work[:, 2].mean()), 1),
            "final_neighbours": round(float(work[:, 3].mean()), 1),
        }
        flops = density_flops(work)
        # the recorded counters cover the default workload and launch shape only
        default_shape = (args.n == 1_000_000 and args.fixes == 0 and not args.param and not args.headon
                         and dlanes == 8 and prm.lane_split == 0 and prm.density_flat == 0)
        traffic_bytes = profiled_traffic() if default_shape else None
        t = per_pass["density"] / 1000.0
        achieved = flops / t / 1e12
        bytes_alg = DENSITY_BYTES_PER_PARTICLE * args.n
        out["roofline"] = {
            "kernel": kname,
            "bound": "mfma",
            "compute_unit": "fp32 VALU (no MFMA in this kernel; peak = MI355X FP32 157.3 TFLOP/s)",
            "achieved": round(achieved, 4),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
            "traffic": traffic_bytes,
            "traffic_source": "profiles/k_density_hbm_traffic.md (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, "
                              "copy-calibrated)" if traffic_bytes else None,
            "algorithmic_flop_per_launch": flops,
            "algorithmic_bytes_per_launch": bytes_alg,
            "hbm_frac_of_algorithmic_bytes": round(bytes_alg / t / 1e9 / PEAK_HBM_GBS, 6),
            "avg_launch_ms": round(per_pass["density"], 4),
        }
    if args.learned_steps > 0:
        log("learned-h comparison")
        out["learned_h"] = learned_comparison(sph, args)
        log("learned-h comparison done")
    if args.cpu_steps > 0:
        import oracle
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(sph, oracle, hl, prm, tables, args.cpu_steps)
        log("cpu baseline done")
    hl.close()
    if args.c1_steps > 0:
        import oracle
        log("2e4-particle run, CPU path and GPU")
        out["c1_2e4"] = c1_run(sph, oracle, args)
        log("2e4-particle run done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
