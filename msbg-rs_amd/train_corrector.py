#!/usr/bin/env python3
"""Trains the Poisson-band error corrector on the GPU from pairs the solver
generates, and writes its weights (npz) for msbg_rs_amd.CorrectorNet.load.

Pairs: on seeded 128^3 noisy-sphere bands (16^3 blocks within 8 cells of the
surface, zero Dirichlet outside the band) with an LCG right-hand side, the
reference solution x* comes from a converged solve; plain red-black SOR
(omega 1.9) then runs from x = 0 and every 4 sweeps each block contributes
(input: residual / 2^k and the phase field on its 20^3 halo tile, target:
(x* - x) / 2^k / out_scale), 2^k being the power of two the solver uses for
rms(r).  Training: Adam on batches of 64 pairs, by the library's own BF16
MFMA forward / backward kernels.

Usage: python train_corrector.py --out corrector.npz [--scenes 12 --steps 2000]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    import importlib.util
    if "msbg_rs_amd" in sys.modules:
        return sys.modules["msbg_rs_amd"]
    spec = importlib.util.spec_from_file_location("msbg_rs_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["msbg_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pow2_inv(r2, n_cells):
    """2^-k with rms / 2^k in [1, 2) (the solver's refinement scale)."""
    rms = math.sqrt(r2 / n_cells) if n_cells else 0.0
    if rms <= 0.0 or not math.isfinite(rms):
        return 1.0
    _, e = math.frexp(rms)
    e = max(-120, min(120, e - 1))
    return math.ldexp(1.0, -e)


def band_scene(pkg, lib, seed, S=128):
    from msbg_rs_amd import scenes
    sc = scenes.noisy_sphere_scene(S=S, n_levels=2, band=8.0, seed=seed, amp=6.0)
    g = sc.grids[0]
    bids = sc.level_blocks(0)
    phi = pkg.DeviceGrid.from_host(g, lib=lib)
    band = pkg.PoissonBand(phi, bids, lib=lib)
    b = lib.to_device(scenes.poisson_rhs(g, bids, seed=1000 + seed))
    return band, phi, b


def make_pairs(pkg, lib, seeds, states, every, out_scale, log):
    """Device arrays (x0 bf16 [n][20^3][2], target f32 [n][4096]) from the
    solver trajectories of the given scenes."""
    chunks_x, chunks_t = [], []
    total = 0
    for s in seeds:
        band, phi, b = band_scene(pkg, lib, s)
        xs = band.field(0)
        st = band.solve(xs, b, storage=0, refine=True, omega=1.9, sweeps_per_check=10, max_sweeps=20000, tol=2e-7)
        x = band.field(0)
        r = band.field(0)
        n = band.n_active
        X = lib.empty((states * n, 20, 20, 20, 2), np.uint16)
        T = lib.empty((states * n, 4096), np.float32)
        for k in range(states):
            for _ in range(every):
                band.rbgs(0, x, b, 0, 1.9)
                band.rbgs(0, x, b, 1, 1.9)
            r2, _ = band.residual(0, x, b, r)
            inv = pow2_inv(r2, band.n_cells)
            x0 = band.net_input(phi, r, inv)
            lib.call("msbg_memcpy_dtod", X.ptr + k * n * 32000, x0.ptr, n * 32000)
            t = band.error_target(x, xs, inv / out_scale)
            lib.call("msbg_memcpy_dtod", T.ptr + k * n * 16384, t.ptr, n * 16384)
        total += states * n
        chunks_x.append(X)
        chunks_t.append(T)
        log("scene %d: %d blocks, x* rel residual %.2e, %d pairs so far" % (s, n, st["rel_residual"], total))
    X = lib.empty((total, 20, 20, 20, 2), np.uint16)
    T = lib.empty((total, 4096), np.float32)
    off = 0
    for cx, ct in zip(chunks_x, chunks_t):
        m = cx.shape[0]
        lib.call("msbg_memcpy_dtod", X.ptr + off * 32000, cx.ptr, m * 32000)
        lib.call("msbg_memcpy_dtod", T.ptr + off * 16384, ct.ptr, m * 16384)
        off += m
        cx.free()
        ct.free()
    return X, T, total


def make_pairs_mg(pkg, lib, seeds, rhs_per_scene, cycles, out_scale, log):
    """Pairs from multigrid trajectories: for each scene and right-hand side,
    the iterate after each of `cycles` V(2, 2)-cycles (FP32 refinement, FP16
    band storage) from x = 0 gives (residual / 2^k + phi, (x* - x) / 2^k /
    out_scale) per block."""
    from msbg_rs_amd import scenes
    xs_, ts_ = [], []
    total = 0
    for s in seeds:
        sc = scenes.noisy_sphere_scene(S=128, n_levels=2, band=8.0, seed=s, amp=6.0)
        g_h, bids = sc.grids[0], sc.level_blocks(0)
        phi = pkg.DeviceGrid.from_host(g_h, lib=lib)
        band = pkg.PoissonBand(phi, bids, lib=lib)
        for j in range(rhs_per_scene):
            b = lib.to_device(scenes.poisson_rhs(g_h, bids, seed=2000 + 97 * s + j))
            xs = band.field(0)
            band.solve(xs, b, storage=0, refine=True, omega=1.9, sweeps_per_check=10, max_sweeps=20000, tol=2e-7)
            x, r = band.field(0), band.field(0)
            n = band.n_active
            X = lib.empty((cycles * n, 20, 20, 20, 2), np.uint16)
            T = lib.empty((cycles * n, 4096), np.float32)
            for k in range(cycles):
                band.solve(x, b, storage=1, multigrid=True, omega=1.0, sweeps_per_check=2, max_sweeps=4, tol=0.0)
                r2, _ = band.residual(0, x, b, r)
                inv = pow2_inv(r2, band.n_cells)
                x0 = band.net_input(phi, r, inv)
                lib.call("msbg_memcpy_dtod", X.ptr + k * n * 32000, x0.ptr, n * 32000)
                t = band.error_target(x, xs, inv / out_scale)
                lib.call("msbg_memcpy_dtod", T.ptr + k * n * 16384, t.ptr, n * 16384)
            total += cycles * n
            xs_.append(X)
            ts_.append(T)
        log("scene %d: %d pairs so far" % (s, total))
    X = lib.empty((total, 20, 20, 20, 2), np.uint16)
    T = lib.empty((total, 4096), np.float32)
    off = 0
    for cx, ct in zip(xs_, ts_):
        m = cx.shape[0]
        lib.call("msbg_memcpy_dtod", X.ptr + off * 32000, cx.ptr, m * 32000)
        lib.call("msbg_memcpy_dtod", T.ptr + off * 16384, ct.ptr, m * 16384)
        off += m
        cx.free()
        ct.free()
    return X, T, total


def evaluate_mg(pkg, lib, net, tol=1e-6):
    """V-cycles and wall time to tol on the 128^3 shell, multigrid (FP16 band
    storage) with and without a learned correction after every cycle."""
    from msbg_rs_amd import scenes
    g, bids, _ = scenes.shell_scene()
    phi = pkg.DeviceGrid.from_host(g, lib=lib)
    band = pkg.PoissonBand(phi, bids, lib=lib)
    b = lib.to_device(scenes.poisson_rhs(g, bids))
    out = {}
    for learned in (False, True):
        best = None
        for _ in range(3):
            x = band.field(0)
            st = band.solve(x, b, storage=1, multigrid=True, omega=1.0, sweeps_per_check=2, max_sweeps=400, tol=tol,
                            learned=net if learned else None, phi=phi if learned else None)
            if best is None or st["ms"] < best["ms"]:
                best = st
        out["f16_mg" + ("_learned" if learned else "")] = {
            k: (float(v) if isinstance(v, float) else v) for k, v in best.items()}
    return out


def evaluate(pkg, lib, net, tol=1e-6, every=4):
    """Sweeps and wall time to tol on the 128^3 shell (not in the training
    scenes): plain vs learned, FP32 and FP16 / BF16 refinement."""
    from msbg_rs_amd import scenes
    g, bids, _ = scenes.shell_scene()
    phi = pkg.DeviceGrid.from_host(g, lib=lib)
    band = pkg.PoissonBand(phi, bids, lib=lib)
    b = lib.to_device(scenes.poisson_rhs(g, bids))
    out = {}
    for name, storage, refine in (("f32", 0, False), ("f16", 1, True), ("bf16", 2, True)):
        for learned in (False, True):
            best = None
            for _ in range(3):
                x = band.field(0)
                st = band.solve(x, b, storage=storage, refine=refine, omega=1.9, sweeps_per_check=every,
                                max_sweeps=5000, tol=tol, learned=net if learned else None,
                                phi=phi if learned else None)
                if best is None or st["ms"] < best["ms"]:
                    best = st
            out[name + ("_learned" if learned else "_plain")] = {
                k: (float(v) if isinstance(v, float) else v) for k, v in best.items()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="corrector.npz")
    ap.add_argument("--scenes", type=int, default=12)
    ap.add_argument("--states", type=int, default=60)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out-scale", type=float, default=64.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--solver", choices=("sor", "mg"), default="sor",
                    help="trajectories the pairs come from: red-black SOR every --every sweeps, or multigrid "
                         "V-cycles (--rhs right-hand sides per scene, --cycles iterates each)")
    ap.add_argument("--rhs", type=int, default=8)
    ap.add_argument("--cycles", type=int, default=6)
    args = ap.parse_args()

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    pkg = load_pkg()
    lib = pkg.lib.load()
    from msbg_rs_amd.learned import init_params
    t0 = time.time()
    if args.solver == "mg":
        X, T, n = make_pairs_mg(pkg, lib, range(100, 100 + args.scenes), args.rhs, args.cycles, args.out_scale, log)
    else:
        X, T, n = make_pairs(pkg, lib, range(100, 100 + args.scenes), args.states, args.every, args.out_scale, log)
    lib.synchronize()
    t_gen = time.time() - t0
    log("pairs: %d in %.1fs" % (n, t_gen))
    net = pkg.CorrectorNet(lib, init_params(args.seed), args.out_scale)
    losses = []
    t0 = time.time()
    segs = 4
    for k in range(segs):  # step-wise decay of the learning rate
        lr = args.lr * (0.3 ** k)
        losses.append(net.train(X, T, n, args.steps // segs, args.batch, lr, seed=args.seed * 131 + k))
        log("segment %d (lr %.1e): loss %.4g -> %.4g" % (k, lr, losses[-1][:20].mean(), losses[-1][-20:].mean()))
    t_train = time.time() - t0
    losses = np.concatenate(losses)
    net.save(args.out)
    ev = evaluate_mg(pkg, lib, net) if args.solver == "mg" else evaluate(pkg, lib, net)
    print(json.dumps({"pairs": n, "gen_s": t_gen, "train_s": t_train, "steps": int(losses.size),
                      "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-20:].mean()),
                      "eval": ev}), flush=True)


if __name__ == "__main__":
    main()
