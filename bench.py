#!/usr/bin/env python3
"""Benchmark of the LBVH build + SPH line-integrated interpolation hot path on one MI355X.

Workload (default, the 16M-particle configuration the headline metric is quoted on):
  warped disc, 1.6e7 particles stored FP32, Wendland C4 kernel, face-on 1024 x 1024 column
  map (lines along +z), scalars (rho: column density, not Shepard-normalised; vz: column-
  averaged, Shepard-normalised), itpScatter.
A step is one LBVH build from device-resident particle columns followed by one
interpolation of all map lines into device-resident outputs (the work of
LineSamples_interpolation with the data already in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f32c|f64|both]

Prints one JSON line.  `value` is map lines (grid points) per second over whole steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="best", choices=["best", "f32c", "f64", "both"])
    ap.add_argument("--particles", type=int, default=16_000_000)
    ap.add_argument("--map", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-lines", type=int, default=16384, help="lines in the CPU baseline sample")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (no multi-GPU path)",
                          "n_gpus": a.gpus}))
        return 2
    import torch

    from distorteddustydiscanalysistoolkits_amd import _lib, disc

    dev = torch.device("cuda:0")
    n, M = a.particles, a.map
    t0 = time.time()
    d = disc.make_disc(n, seed=0, warped=True)
    names = ("x", "y", "z", "m", "h", "rho")
    host32 = {k: np.ascontiguousarray(d[k], np.float32) for k in names + ("vz",)}
    log(f"[bench] disc generated n={n} in {time.time() - t0:.1f}s")
    dcols = [torch.from_numpy(host32[k]).to(dev) for k in names]
    dquant = [torch.from_numpy(host32["rho"]).to(dev), torch.from_numpy(host32["vz"]).to(dev)]
    o, dvec = disc.faceon_lines(M)
    do = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o]
    dd = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in dvec]
    P = M * M
    cat = _lib.catalog_struct(scalar_slots=[0, 1], scalar_shepard=[0, 1])
    KERNEL_C4 = 4
    ctx = _lib.Context(0)

    def run(prec, steps, warmup):
        outs = [torch.empty(P, dtype=torch.float32 if prec == _lib.PMX_PREC_F32C else torch.float64, device=dev)
                for _ in range(2)]

        def step():
            ctx.build(*dcols, quant=dquant, kernel=KERNEL_C4)
            tb = ctx.timings()[0]
            ctx.interpolate_lines(_lib.PMX_SCATTER, prec, cat, do, dd, outs)
            return tb, ctx.timings()[5]

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        tb_list, ti_list = [], []
        t = time.perf_counter()
        for _ in range(steps):
            tb, ti = step()
            tb_list.append(tb)
            ti_list.append(ti)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, build_ms=float(np.median(tb_list)),
                    interp_ms=float(np.median(ti_list)), outs=[o_.double().cpu().numpy() for o_ in outs])

    precs = {"f64": [_lib.PMX_PREC_F64], "f32c": [_lib.PMX_PREC_F32C],
             "both": [_lib.PMX_PREC_F32C, _lib.PMX_PREC_F64], "best": [_lib.PMX_PREC_F32C, _lib.PMX_PREC_F64]}[a.precision]
    results = {}
    for prec in precs:
        key = "f32c" if prec == _lib.PMX_PREC_F32C else "f64"
        try:
            log(f"[bench] timing precision={key} steps={a.steps} warmup={a.warmup}")
            results[key] = run(prec, a.steps, a.warmup)
            log(f"[bench] {key}: {results[key]['ms_per_step']:.2f} ms/step "
                f"(build {results[key]['build_ms']:.2f} ms, interp {results[key]['interp_ms']:.2f} ms)")
        except _lib.PmxError as e:
            log(f"[bench] precision {key} unavailable: {e}")
    main_key = "f32c" if "f32c" in results else "f64"
    r = results[main_key]

    # traversal statistics -> algorithmic bytes of the line kernel (per launch)
    st = ctx.line_stats(do, dd)
    n_int, n_leaf, n_acc = int(st[2]), int(st[3]), int(st[4])
    ncolq = 2
    if main_key == "f64":
        # node record 64 B, leaf record 32 B + escape 4 B, accepted: m, rho, 2 quant (8 B each),
        # line in 48 B, outputs 2 x 8 B
        bytes_launch = 64 * n_int + 36 * n_leaf + 8 * (2 + ncolq) * n_acc + P * (48 + 8 * 2)
    else:
        bytes_launch = 32 * n_int + 20 * n_leaf + 4 * (1 + ncolq) * n_acc + P * (48 + 4 * 2)
    achieved = bytes_launch / (r["interp_ms"] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"line_{main_key}")
        except Exception:
            traffic = None

    errors = {}
    if "f32c" in results and "f64" in results:
        from tests.helpers import rel_norms
        for j, nm in enumerate(("rho_column", "vz_column")):
            linf, l2 = rel_norms(results["f32c"]["outs"][j], results["f64"]["outs"][j])
            errors[nm] = {"rel_linf": linf, "rel_l2": l2}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(d, o, dvec, M, a.cpu_lines)

    line = {
        "metric": "interp grid points/s + LBVH build particles/s at 16M parts, FP32c vs FP64",
        "value": P / (r["ms_per_step"] * 1e-3),
        "unit": "grid points/s (1024^2 column-map lines per second, build + interpolation per step)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_key == "f32c" else "f64",
        "data": "synthetic (deterministic SplitMix64 warped disc, seed 0)",
        "config": {"workload": f"warped disc {n} particles (FP32 storage), Wendland C4, face-on {M}x{M} "
                               "column map: column density + column-averaged vz, itpScatter",
                   "particles": n, "lines": P, "path": main_key},
        "build_particles_per_s": n / (r["build_ms"] * 1e-3),
        "interp_points_per_s": P / (r["interp_ms"] * 1e-3),
        "build_ms": r["build_ms"],
        "interp_ms": r["interp_ms"],
        "paths": {k: {"ms_per_step": v["ms_per_step"], "build_ms": v["build_ms"], "interp_ms": v["interp_ms"],
                      "grid_points_per_s": P / (v["ms_per_step"] * 1e-3),
                      "interp_points_per_s": P / (v["interp_ms"] * 1e-3),
                      "build_particles_per_s": n / (v["build_ms"] * 1e-3)} for k, v in results.items()},
        "field_error_f32c_vs_f64": errors or None,
        "traversal": {"internal_tested": n_int, "leaves_tested": n_leaf, "leaves_accepted": n_acc},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"line interpolation ({main_key})", "algorithmic_bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()
    return 0


def cpu_baseline(d, o, dvec, M, nsample):
    """The oracle (C FP64 restatement of the reference CPU algorithm, OpenMP where the reference
    threads) on this host: the full LBVH build of the same particle set, then the line kernel
    on an evenly strided sample of the map lines; whole-step throughput is extrapolated as
    P / (t_build + t_sample * P / sample)."""
    from oracle import oracle as O
    nthreads = min(os.cpu_count() or 1, 16)
    cols = [np.array(d[k], np.float32).astype(np.float64) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [np.array(d["rho"], np.float32).astype(np.float64), np.array(d["vz"], np.float32).astype(np.float64)]
    log(f"[bench] cpu baseline: oracle build n={cols[0].size} threads={nthreads}")
    t = time.perf_counter()
    tree = O.LBVH(*cols, quant=quant, nthreads=nthreads)
    tb = time.perf_counter() - t
    P = M * M
    step = max(1, P // nsample)
    sel = np.arange(0, P, step)[:nsample]
    oo = [v[sel] for v in o]
    dd = [v[sel] for v in dvec]
    log(f"[bench] cpu baseline: oracle lines sample={sel.size}")
    t = time.perf_counter()
    tree.line_samples("C4", "scatter", [0, 1], [0, 1], oo, dd, nthreads=nthreads)
    ti = time.perf_counter() - t
    total = tb + ti * P / sel.size
    return {"value": P / total, "unit": "grid points/s", "cores": nthreads, "kind": "port",
            "sample": f"full LBVH build of {cols[0].size} particles ({tb:.2f} s) + {sel.size} of {P} lines "
                      f"({ti:.2f} s), extrapolated to the full map",
            "build_particles_per_s": cols[0].size / tb, "interp_points_per_s": sel.size / ti}


if __name__ == "__main__":
    sys.exit(main())
