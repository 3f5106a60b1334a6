#!/usr/bin/env python3
"""Benchmark of the sparse-block-grid smoother path on one MI355X.

Workload (one step): the 512^3, 4-level noisy-sphere multires grid; one
19-tap mean-curvature sweep and one 25-tap bi-Laplacian sweep over every
level-0 block, 8-colour in place, with cross-level (fill_multires) halos.
Inputs are resident in HBM before timing starts.

Prints one JSON line: value = cell-updates/s over the timed steps, plus the
roofline of the dominant kernel (HIP events around each launch) and the CPU
baseline (the oracle's C++ restatement of the reference path, OpenMP on the
host cores, on the same grid).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s & % HBM peak; time-to-residual 1e-6, learned vs plain"


def load_pkg():
    if "msbg_rs_amd" in sys.modules:
        return sys.modules["msbg_rs_amd"]
    path = os.path.join(ROOT, "msbg-rs_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("msbg_rs_amd", path,
                                                  submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["msbg_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_scene(pkg, lib):
    from msbg_rs_amd import scenes

    def regularize_on_device(levels, nx, ny, nz, n_levels):
        d = lib.to_device(levels)
        pkg.regularize_refinement_map(d, nx, ny, nz, n_levels)
        levels[:] = d.get()

    t = time.time()
    sc = scenes.noisy_sphere_scene(S=512, n_levels=4, regularize=regularize_on_device)
    log("scene: %.1fs, blocks per level %s" % (time.time() - t, [int((sc.levels == l).sum()) for l in range(4)]))
    return sc


def cpu_baseline(sc, dt_mc, dt_bl, reps):
    """The oracle's C++ restatement of the reference sweeper (thread-local halo
    tiles, OpenMP over the blocks of a colour), FTZ/DAZ as the reference benches."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi

    orc = oracle_ffi.load()
    fine, coarse = sc.grids[0], sc.grids[1]
    active = sc.level_blocks(0)
    threads = max(1, min(16, os.cpu_count() or 1))
    slab = np.ascontiguousarray(fine.slab).copy()
    best = None
    for r in range(reps + 1):
        t = time.perf_counter()
        for st, dt in ((1, dt_mc), (2, dt_bl)):
            orc.orc_sweep_multires(fine.sx, fine.sy, fine.sz, 4, slab, fine.b2s, coarse.slab, coarse.b2s, sc.levels,
                                   active, active.size, st, dt, 1, 0, threads)
        el = time.perf_counter() - t
        if r > 0:
            best = el if best is None else min(best, el)
    cells = active.size * 4096 * 2
    return {"value": cells / best, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "full workload: one MC19 + one BiLap25 8-colour multires sweep over %d level-0 blocks, "
                      "best of %d after 1 warm-up, FTZ/DAZ off" % (active.size, reps)}


def particle_leg(pkg, lib, sc, n, steps, warmup):
    """16M particles uniform in the level-0 band: stable 30-bit Morton sort,
    then quadratic B-spline P2G (sort + voxel ranges + gather) and G2P, each
    timed with HIP events around its call (inputs resident in HBM)."""
    import ctypes as C
    from msbg_rs_amd import scenes
    from msbg_rs_amd import transfer as T

    t = time.time()
    pos, vel = scenes.band_particles(sc, n)
    log("particles: %d generated in %.1fs" % (n, time.time() - t))
    dpos, dvel = lib.to_device(pos), lib.to_device(vel)
    cap = 3 * int((sc.levels == 0).sum()) + 4096
    tg = T.TransferGrid(sc.S, sc.S, sc.S, capacity_blocks=cap, lib=lib)
    keys, perm = lib.empty(n, np.uint32), lib.empty(n, np.uint32)
    out = lib.empty((n, 3), np.float32)
    handles = (C.c_void_p * 3)(*[g.h for g in tg.velocity])
    k = C.c_int64()

    def p2g():
        lib.call("msbg_p2g_bspline", 2, tg.mass.h, handles, dpos.ptr, dvel.ptr, None, n, C.byref(k))

    def g2p():
        lib.call("msbg_g2p_bspline", 2, handles, dpos.ptr, n, out.ptr)

    def sort():
        lib.call("msbg_morton_sort", dpos.ptr, n, keys.ptr, perm.ptr)

    for _ in range(warmup):
        sort()
        p2g()
        g2p()
    ms = {"sort": 0.0, "p2g": 0.0, "g2p": 0.0}
    e0, e1 = lib.event(), lib.event()
    for _ in range(steps):
        for name, fn in (("sort", sort), ("p2g", p2g), ("g2p", g2p)):
            lib.call("msbg_event_record", e0)
            fn()
            lib.call("msbg_event_record", e1)
            ms[name] += lib.elapsed_ms(e0, e1)
    for key in ms:
        ms[key] /= steps
    step_ms = ms["p2g"] + ms["g2p"]
    return {
        "particles": n,
        "active_blocks": k.value,
        "ms": ms,
        "morton_sort_particles_per_s": n / (ms["sort"] * 1e-3),
        "p2g_g2p_particles_per_s": n / (step_ms * 1e-3),
        "note": "quadratic B-spline P2G (includes its own Morton sort, active-block build and voxel ranges) "
                "+ G2P, cell-centred 16^3-block grid over the 512^3 domain",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--particles", type=int, default=16 * 1024 * 1024)
    ap.add_argument("--particle-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the sparse-grid path is not sharded across GPUs "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    pkg = load_pkg()
    lib = pkg.lib.load()
    if lib.device_count() < 1:
        raise RuntimeError("bench.py needs a HIP device")
    sc = build_scene(pkg, lib)
    fine = pkg.DeviceGrid.from_host(sc.grids[0], lib=lib)
    coarse = pkg.DeviceGrid.from_host(sc.grids[1], lib=lib)
    levels = lib.to_device(sc.levels)
    active = sc.level_blocks(0)
    plan = pkg.SweepPlan(fine, active, coarse=coarse, levels=levels)
    cells = active.size * 4096
    dt_mc, dt_bl = 0.02, 0.005

    # Each stencil's 8 colour passes replay as one HIP graph (captured during
    # warm-up); an event pair brackets each graph inside the timed region.
    for _ in range(args.warmup):
        plan.run_graph_async(pkg.Stencil.MeanCurvature, dt_mc, 1)
        plan.run_graph_async(pkg.Stencil.BiLaplacian, dt_bl, 1)
    lib.synchronize()
    log("warmup done")
    launches_per_graph = int(np.count_nonzero(pkg.color_buckets(fine, active)[1]))

    ev = [lib.event() for _ in range(2 * args.steps + 1)]
    lib.call("msbg_event_record", ev[0])
    for s in range(args.steps):
        plan.run_graph_async(pkg.Stencil.MeanCurvature, dt_mc, 1)
        lib.call("msbg_event_record", ev[2 * s + 1])
        plan.run_graph_async(pkg.Stencil.BiLaplacian, dt_bl, 1)
        lib.call("msbg_event_record", ev[2 * s + 2])
    lib.synchronize()
    k_ms = {1: 0.0, 2: 0.0}
    for s in range(args.steps):
        k_ms[1] += lib.elapsed_ms(ev[2 * s], ev[2 * s + 1])
        k_ms[2] += lib.elapsed_ms(ev[2 * s + 1], ev[2 * s + 2])
    total_ms = lib.elapsed_ms(ev[0], ev[-1])
    for e in ev:
        lib.call("msbg_event_destroy", e)
    log("timed %d steps: %.3f ms" % (args.steps, total_ms))
    ms_per_step = total_ms / args.steps
    value = 2 * cells / (ms_per_step * 1e-3)

    dom = max(k_ms, key=lambda s: k_ms[s])
    name = {1: "sweep_kernel<MeanCurvature, multires>", 2: "sweep_kernel<BiLaplacian, multires>"}[dom]
    # Average launch duration = graph time / launches in it (inter-kernel gaps
    # inside the graph are charged to the kernel, so `achieved` is a floor).
    launches = launches_per_graph * args.steps
    bytes_total = 8.0 * cells * args.steps  # 4 B read + 4 B written per cell-update, in place
    avg_launch_ms = k_ms[dom] / launches
    achieved = bytes_total / launches / (avg_launch_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(name.split("<")[0] + "_" + str(dom))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "512^3 4-level noisy sphere (R=0.35*512, seed 4), %d level-0 16^3 blocks; per step one "
                        "MC19 + one BiLap25 8-colour in-place sweep with cross-level halos" % active.size,
            "level0_cells": cells,
            "blocks_per_level": [int((sc.levels == l).sum()) for l in range(4)],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_us": avg_launch_ms * 1e3,
            "launches": launches,
            "algorithmic_bytes_per_launch": bytes_total / launches,
        },
        "kernel_ms": {"mean_curvature": k_ms[1], "bilaplacian": k_ms[2]},
    }
    if args.particles > 0:
        log("particle leg ...")
        out["particles"] = particle_leg(pkg, lib, sc, args.particles, args.particle_steps, 2)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(sc, dt_mc, dt_bl, args.cpu_reps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
