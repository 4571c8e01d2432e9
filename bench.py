#!/usr/bin/env python3
"""Benchmark of the LBVH build + SPH interpolation hot path on one MI355X.

Workloads
  lines  (default; the 16M-particle configuration the headline metric is quoted on):
         warped disc, 1.6e7 particles stored FP32, Wendland C4 kernel, face-on 1024 x 1024
         column map (lines along +z), scalars rho (column density, not Shepard-normalised)
         and vz (column-averaged, Shepard-normalised), itpScatter.
  points flat disc, 4e6 particles stored FP32, M4 cubic spline, 256 x 256 x 64 cylindrical
         grid, scalar rho, grad rho, div v, curl v (8 outputs), itpSymmetric.
A step is one LBVH build from device-resident particle columns followed by one
interpolation of all samples into device-resident outputs (the work of
LineSamples_interpolation / StructuredGrid_interpolation with the data already in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lines|points]
                    [--precision best|f32c|f64|both]

Prints one JSON line.  `value` is grid points (map lines or grid nodes) per second over
whole steps (build + interpolation).
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
FP32_PEAK_TFLOPS = 157.3
KERNEL_M4, KERNEL_C4 = 0, 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lines", choices=["lines", "points", "batch", "dusty", "phantom"])
    ap.add_argument("--snapshots", type=int, default=64, help="batch workload: number of snapshots")
    ap.add_argument("--no-tune", action="store_true", help="batch workload: skip the autotuner")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch workload: skip the per-call FP32c / FP64 comparison")
    ap.add_argument("--precision", default="best", choices=["best", "f32c", "f64", "both"])
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--map", type=int, default=1024, help="lines workload: map side")
    ap.add_argument("--geometry", default="faceon", choices=["faceon", "perspective"],
                    help="lines workload: parallel face-on lines, or the rays of a camera 500 above the "
                         "disc through the same image points (slopes up to 0.2)")
    ap.add_argument("--grid", default="256,256,64", help="points workload: ns,nphi,nz")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-samples", type=int, default=16384, help="samples in the CPU baseline leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (reference kernel for HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_norms(got, ref):
    """Normwise relative L-inf and L2 difference of one output field (FP32c against FP64) over
    the finite entries of the reference field."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.isfinite(ref)
    if not m.any():
        return 0.0, 0.0
    e = got[m] - ref[m]
    linf = np.abs(e).max() / max(np.abs(ref[m]).max(), 1e-300)
    l2 = np.linalg.norm(e) / max(np.linalg.norm(ref[m]), 1e-300)
    return float(linf), float(l2)


class Workload:
    """Device-resident particles + samples of one benchmark configuration."""

    def __init__(self, a, torch, dev):
        from distorteddustydiscanalysistoolkits_amd import _lib, disc
        self.kind = a.workload
        self.geometry = a.geometry if a.workload == "lines" else "faceon"
        t0 = time.time()
        if self.kind == "lines":
            self.n = a.particles or 16_000_000
            self.d = disc.make_disc(self.n, seed=0, warped=True)
            self.qnames = ("rho", "vz")
            self.kernel = KERNEL_C4
            self.kname = "C4"
            o, dv = disc.faceon_lines(a.map) if a.geometry == "faceon" else disc.perspective_lines(a.map)
            self.samples = (o, dv)
            self.P = a.map * a.map
            self.dsamples = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o],
                             [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in dv])
            self.cat_d = dict(scalar_slots=[0, 1], scalar_shepard=[0, 1])
            self.L = 2
            self.strategy = _lib.PMX_SCATTER
            self.fields = ("rho_column", "vz_column")
            view = "face-on" if a.geometry == "faceon" else "perspective (camera at z = 500)"
            self.desc = (f"warped disc {self.n} particles (FP32 storage), Wendland C4, {view} {a.map}x{a.map} "
                         "column map: column density + column-averaged vz, itpScatter")
        else:
            self.n = a.particles or 4_000_000
            self.d = disc.make_disc(self.n, seed=0, warped=False)
            self.qnames = ("rho", "vx", "vy", "vz")
            self.kernel = KERNEL_M4
            self.kname = "M4"
            ns, nphi, nz = (int(v) for v in a.grid.split(","))
            sa, pa, za = disc.cylindrical_axes((1.0, 100.0, ns), (0.0, 2 * np.pi, nphi), (-5.0, 5.0, nz))
            q = disc.flatten_cylindrical(sa, pa, za)
            self.samples = q
            self.P = q[0].size
            self.dsamples = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in q]
            self.cat_d = dict(scalar_slots=[0], scalar_shepard=[0], grad_slots=[0], div_slots=[1, 2, 3],
                              curl_slots=[1, 2, 3])
            self.L = 8
            self.strategy = _lib.PMX_SYMMETRIC
            self.fields = ("rho", "grad_rho_x", "grad_rho_y", "grad_rho_z", "div_v", "curl_v_x", "curl_v_y",
                           "curl_v_z")
            self.desc = (f"flat disc {self.n} particles (FP32 storage), M4 cubic spline, cylindrical "
                         f"{ns}x{nphi}x{nz} grid: rho, grad rho, div v, curl v, itpSymmetric")
        self.cat = _lib.catalog_struct(**self.cat_d)
        names = ("x", "y", "z", "m", "h", "rho")
        self.host32 = {k: np.ascontiguousarray(self.d[k], np.float32) for k in names + self.qnames}
        self.dcols = [torch.from_numpy(self.host32[k]).to(dev) for k in names]
        self.dquant = [torch.from_numpy(self.host32[k]).to(dev) for k in self.qnames]
        log(f"[bench] workload {self.kind}: n={self.n} samples={self.P} prepared in {time.time() - t0:.1f}s")

    def interpolate(self, ctx, prec, outs):
        if self.kind == "lines":
            ctx.interpolate_lines(self.strategy, prec, self.cat, self.dsamples[0], self.dsamples[1], outs)
        else:
            ctx.interpolate_points(self.strategy, prec, self.cat, *self.dsamples, outs)

    def stats(self, ctx):
        if self.kind == "lines":
            return ctx.line_stats(*self.dsamples)
        return ctx.point_stats(self.strategy, *self.dsamples)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (no multi-GPU path)",
                          "n_gpus": a.gpus}))
        return 2
    import torch

    from distorteddustydiscanalysistoolkits_amd import _lib

    dev = torch.device("cuda:0")
    if a.workload == "batch":
        return run_batch(a, torch, dev)
    if a.workload == "dusty":
        return run_dusty(a, torch, dev)
    if a.workload == "phantom":
        return run_phantom(a, torch, dev)
    if a.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 1 GiB read + 1 GiB written")
        del src, dst
    w = Workload(a, torch, dev)
    ctx = _lib.Context(0)

    def run(prec, steps, warmup):
        outs = [torch.empty(w.P, dtype=torch.float32 if prec == _lib.PMX_PREC_F32C else torch.float64, device=dev)
                for _ in range(w.L)]

        def step():
            # the build returns once its inputs are consumed; its phase times are read after the
            # interpolation, which is ordered after it on the context's stream
            ctx.build(*w.dcols, quant=w.dquant, kernel=w.kernel)
            w.interpolate(ctx, prec, outs)
            t = ctx.timings()
            return t[0], t[5], t[7]

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        tb_list, ti_list, tk_list = [], [], []
        t = time.perf_counter()
        for _ in range(steps):
            tb, ti, tk = step()
            tb_list.append(tb)
            ti_list.append(ti)
            tk_list.append(tk)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        return dict(wall_s=wall, ms_per_step=1e3 * wall / steps, build_ms=float(np.median(tb_list)),
                    interp_ms=float(np.median(ti_list)), kernel_ms=float(np.mean(tk_list)),
                    outs=[o_.double().cpu().numpy() for o_ in outs])

    precs = {"f64": [_lib.PMX_PREC_F64], "f32c": [_lib.PMX_PREC_F32C],
             "both": [_lib.PMX_PREC_F32C, _lib.PMX_PREC_F64],
             "best": [_lib.PMX_PREC_F32C, _lib.PMX_PREC_F64]}[a.precision]
    results = {}
    for prec in precs:
        key = "f32c" if prec == _lib.PMX_PREC_F32C else "f64"
        try:
            log(f"[bench] timing precision={key} steps={a.steps} warmup={a.warmup}")
            results[key] = run(prec, a.steps, a.warmup)
            r = results[key]
            log(f"[bench] {key}: {r['ms_per_step']:.2f} ms/step (build {r['build_ms']:.2f} ms, interp "
                f"{r['interp_ms']:.2f} ms, kernels {r['kernel_ms']:.2f} ms)")
        except _lib.PmxError as e:
            log(f"[bench] precision {key} unavailable: {e}")
    main_key = "f32c" if "f32c" in results else "f64"
    r = results[main_key]

    # Algorithmic bytes of the interpolation kernels per launch (pair-stream model): one
    # particle record of (6 + k) FP32 columns per accepted (sample, particle) pair, 8 B columns
    # on the FP64 path.  The per-sample traversal model (FP32 node 36 B, leaf 20 B per visit,
    # FP64 64 B / 36 B) is reported beside it; HBM traffic is the measured FETCH + WRITE bytes
    # per launch (profiles/pmc_traffic.json) when present.
    st = w.stats(ctx)
    n_int, n_leaf, n_acc = int(st[2]), int(st[3]), int(st[4])
    n_near_int, n_near_leaf = int(st[0]), int(st[1])
    k = len(w.qnames)
    if main_key == "f64":
        rec, node, leaf = 8 * (6 + k), 64, 36
    else:
        rec, node, leaf = 4 * (6 + k), 36, 20
    trav_bytes = node * (n_int + n_near_int) + leaf * (n_leaf + n_near_leaf)
    pair_bytes = rec * n_acc
    t_kernel = r["kernel_ms"] * 1e-3
    achieved = pair_bytes / t_kernel / 1e9
    flops_pair = 103 if w.kind == "points" else 30
    flops = flops_pair * n_acc
    geo = "" if getattr(w, "geometry", "faceon") == "faceon" else f"_{w.geometry}"
    traffic, traffic_src = measured_traffic(f"{w.kind}_{main_key}{geo}", n=w.n, P=w.P)

    errors = {}
    if "f32c" in results and "f64" in results:
        for j, nm in enumerate(w.fields):
            linf, l2 = rel_norms(results["f32c"]["outs"][j], results["f64"]["outs"][j])
            errors[nm] = {"rel_linf": linf, "rel_l2": l2}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(w, a.cpu_samples)

    P, n = w.P, w.n
    line = {
        "metric": "interp grid points/s + LBVH build particles/s at 16M parts, FP32c vs FP64",
        "value": P / (r["ms_per_step"] * 1e-3),
        "unit": ("grid points/s (column-map lines per second, build + interpolation per step)" if w.kind == "lines"
                 else "grid points/s (grid nodes per second, build + interpolation per step)"),
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_key == "f32c" else "f64",
        "data": f"synthetic (deterministic SplitMix64 {'warped' if w.kind == 'lines' else 'flat'} disc, seed 0)",
        "config": {"workload": w.desc, "particles": n, "samples": P, "path": main_key},
        "build_particles_per_s": n / (r["build_ms"] * 1e-3),
        "build_roofline": build_roofline(n, k, r["build_ms"], w.kind),
        "interp_points_per_s": P / (r["interp_ms"] * 1e-3),
        "build_ms": r["build_ms"],
        "interp_ms": r["interp_ms"],
        "kernel_ms": r["kernel_ms"],
        "paths": {kk: {"ms_per_step": v["ms_per_step"], "build_ms": v["build_ms"], "interp_ms": v["interp_ms"],
                       "kernel_ms": v["kernel_ms"],
                       "grid_points_per_s": P / (v["ms_per_step"] * 1e-3),
                       "interp_points_per_s": P / (v["interp_ms"] * 1e-3),
                       "build_particles_per_s": n / (v["build_ms"] * 1e-3)} for kk, v in results.items()},
        "speedup_f32c_vs_f64": ({"step": results["f64"]["ms_per_step"] / results["f32c"]["ms_per_step"],
                                 "interpolation": results["f64"]["interp_ms"] / results["f32c"]["interp_ms"],
                                 "kernel": results["f64"]["kernel_ms"] / results["f32c"]["kernel_ms"]}
                                if "f32c" in results and "f64" in results else None),
        "field_error_f32c_vs_f64": errors or None,
        "traversal": {"nearest_internal": n_near_int, "nearest_leaves": n_near_leaf, "internal_tested": n_int,
                      "leaves_tested": n_leaf, "leaves_accepted": n_acc},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"{w.kind} interpolation ({main_key})",
                     "algorithmic_bytes_per_launch": pair_bytes, "bytes_per_unit": rec, "units": n_acc,
                     "unit_name": "accepted (sample, particle) pairs",
                     "measured_GBs": (traffic / t_kernel / 1e9) if traffic else None,
                     "traffic_source": traffic_src,
                     "traversal_model_bytes": trav_bytes,
                     "valu_TFLOPs": flops / t_kernel / 1e12,
                     "valu_frac": flops / t_kernel / 1e12 / FP32_PEAK_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()
    return 0


def batch_specs(nsnap):
    """Snapshot i (seed i): N log-uniform in [2e5, 8e6], flat or warped with probability 1/2,
    and one of {cylindrical 128x128x32, Cartesian z = 0 slice 512^2, face-on map 768^2}."""
    from distorteddustydiscanalysistoolkits_amd import disc
    specs = []
    for i in range(nsnap):
        u = disc.uniform(i, 900, 3)
        n = int(round(np.exp(np.log(2e5) + u[0] * (np.log(8e6) - np.log(2e5)))))
        specs.append(dict(seed=i, n=n, warped=bool(u[1] < 0.5), grid=("cyl", "slice", "map")[min(2, int(u[2] * 3))]))
    return specs


def run_batch(a, torch, dev):
    """The batched workload: every snapshot's LBVH build and all interpolations in one
    hipGraph replay (persistent kernel over all tiles).  Points: scalar density (rho column,
    Shepard-normalised), itpSymmetric; maps: column density, itpScatter; M4 kernel."""
    from distorteddustydiscanalysistoolkits_amd import _lib, disc
    t0 = time.time()
    specs = batch_specs(a.snapshots)
    cyl = disc.flatten_cylindrical(*disc.cylindrical_axes((1.0, 100.0, 128), (0.0, 2 * np.pi, 128), (-5.0, 5.0, 32)))
    ax = disc.linrange(-100.0, 100.0, 512)
    sl = disc.flatten_cartesian(ax, ax, np.array([0.0]))
    lo, ld = disc.faceon_lines(768)
    dcyl = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in cyl]
    dsl = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in sl]
    dmap = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in lo],
            [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in ld])
    jobs, P_total, N_total = [], 0, 0
    for sp in specs:
        d = disc.make_disc(sp["n"], seed=sp["seed"], warped=sp["warped"])
        cols = [torch.from_numpy(np.ascontiguousarray(d[k], np.float32)).to(dev) for k in ("x", "y", "z", "m", "h", "rho")]
        quant = [torch.from_numpy(np.ascontiguousarray(d["rho"], np.float32)).to(dev)]
        del d
        if sp["grid"] == "map":
            P = dmap[0][0].shape[0]
            job = dict(kind="lines", samples=dmap, scalar_slots=[0], scalar_shepard=[0])
        else:
            samp = dcyl if sp["grid"] == "cyl" else dsl
            P = samp[0].shape[0]
            job = dict(kind="points", strategy=_lib.PMX_SYMMETRIC, samples=samp, scalar_slots=[0], scalar_shepard=[1])
        job.update(cols=cols, quant=quant, kernel=KERNEL_M4, out=[torch.empty(P, dtype=torch.float32, device=dev)])
        jobs.append(job)
        P_total += P
        N_total += sp["n"]
    log(f"[bench] batch: {len(jobs)} snapshots, {N_total} particles, {P_total} samples prepared in {time.time() - t0:.1f}s")
    t0 = time.time()
    b = _lib.Batch(jobs)
    log(f"[bench] batch planned in {time.time() - t0:.1f}s")
    if not a.no_tune:
        b.tune()
        log(f"[bench] autotuned: {b.info()}")
    b.run(graph=False)
    eager = b.info()
    log(f"[bench] eager run {eager['run_ms']:.1f} ms (builds {eager['build_ms']:.1f}, interpolation {eager['interp_ms']:.1f})")
    b.run(graph=True)                       # capture
    for _ in range(a.warmup):
        b.run(graph=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        b.run(graph=True)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / a.steps
    info = b.info()
    log(f"[bench] graph replay {ms:.1f} ms/step")
    b.close()
    compare = None if a.no_compare else per_call_compare(jobs, torch, dev)
    roof = None
    if compare is not None:
        # pair-stream model of the batch's interpolation (persistent kernel + split items + finish):
        # one (6 + k)-column FP32 record per accepted (sample, particle) pair, over the eager run's
        # interpolation phase
        t_i = eager["interp_ms"] * 1e-3
        pb = compare["f32c"]["pair_bytes"]
        roof = {"bound": "hbm", "achieved": pb / t_i / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": pb / t_i / 1e9 / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_batch + k_batch_items + k_batch_finish (eager interpolation phase)",
                "algorithmic_bytes_per_launch": pb, "bytes_per_unit": 28,
                "units": compare["f32c"]["accepted_pairs"], "unit_name": "accepted (sample, particle) pairs"}
    line = {
        "metric": "interp grid points/s + LBVH build particles/s at 16M parts, FP32c vs FP64",
        "value": P_total / (ms * 1e-3),
        "unit": "grid points/s (all snapshots' samples per second, builds + interpolation per step)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (deterministic SplitMix64 discs, seed = snapshot index)",
        "config": {"workload": f"{len(jobs)} snapshots, N log-uniform in [2e5, 8e6], flat or warped, each with a "
                               "cylindrical 128x128x32 grid, a z = 0 slice 512^2 or a face-on 768^2 map; M4; one "
                               "hipGraph of all builds + one persistent interpolation kernel",
                   "particles": N_total, "samples": P_total,
                   "grids": {g: sum(1 for sp in specs if sp["grid"] == g) for g in ("cyl", "slice", "map")}},
        "build_particles_per_s": N_total / (eager["build_ms"] * 1e-3),
        "interp_points_per_s": P_total / (eager["interp_ms"] * 1e-3),
        "eager_ms": eager["run_ms"], "eager_build_ms": eager["build_ms"], "eager_interp_ms": eager["interp_ms"],
        "tuning": {k: info[k] for k in ("waves_per_cu", "bucket_points", "bucket_lines", "heavy_first",
                                        "split_budget")},
        "per_call": compare,
        "speedup_vs_f64_per_call": None if compare is None else (P_total / (ms * 1e-3)) / compare["f64"]["grid_points_per_s"],
        "roofline": roof, "cpu_baseline": None,
    }
    print(json.dumps(line))
    return 0


def run_phantom(a, torch, dev):
    """The input side of the path (ParticleIO's read_phantom + add_rho! + build_input): a
    synthetic warped-disc Phantom dump (default 16M particles; x y z h as default reals FP32,
    vx vy vz real*4, itype int*1: 29 B per particle) is opened, every gas column is streamed from
    the file straight into HBM (pinned double-buffered copies), rho is formed on the device from
    h and massoftype, and the InterpolationInput + catalog (vz; M4) are built on the device.  A
    step is all of that from the file name; the file is page-cache resident after the first
    read.  The achieved host-to-device rate is set against a plain pinned 1 GiB copy measured in
    the same run, and against the host-memory path (the same reader into numpy arrays)."""
    import tempfile

    from distorteddustydiscanalysistoolkits_amd import disc
    from distorteddustydiscanalysistoolkits_amd import partia as P
    from distorteddustydiscanalysistoolkits_amd import particleio as PIO
    from distorteddustydiscanalysistoolkits_amd import phantomdump as W
    n = a.particles or 16_000_000
    t0 = time.time()
    d = disc.make_disc(n, seed=0, warped=True)
    path = os.path.join(tempfile.gettempdir(), f"pmx_bench_disc_{n}_00000")
    W.disc_dump(path, d)
    del d
    fbytes = os.path.getsize(path)
    log(f"[bench] phantom: wrote {fbytes / 1e6:.0f} MB dump of {n} particles in {time.time() - t0:.1f}s")

    def step(device):
        (gas,) = PIO.read_phantom(path, device=device)
        gas.params["mass"] = gas.params["massoftype"]
        PIO.add_rho(gas)
        return PIO.build_input(gas, PIO.MassFromParams("massoftype"), scalars=("vz",), smoothed_kernel=P.M4_spline)

    for _ in range(a.warmup):
        step(dev)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        inp, _cat = step(dev)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / a.steps
    assert inp.coord[0].is_cuda and inp.Npart == n
    # plain pinned host-to-device copy of 1 GiB: the link's achievable rate in this process
    src = torch.empty(1 << 28, dtype=torch.float32).pin_memory()
    dst = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    h2d = 3 * (1 << 30) / (time.perf_counter() - t) / 1e9
    del src, dst
    t = time.perf_counter()
    step(None)
    host_ms = (time.perf_counter() - t) * 1e3
    gbs = fbytes / (ms * 1e-3) / 1e9
    # the whole pipeline from the file: input side, LBVH build (C4 Wendland) and a face-on map of
    # column density and column-averaged vz (itpScatter, FP32c), outputs left in HBM
    from distorteddustydiscanalysistoolkits_amd import _lib
    o, dv = disc.faceon_lines(a.map)
    do_ = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o]
    dd_ = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in dv]
    outs = [torch.empty(a.map * a.map, dtype=torch.float32, device=dev) for _ in range(2)]
    ctx = _lib.Context(0)
    cat_s = _lib.catalog_struct(scalar_slots=[0, 1], scalar_shepard=[0, 1])

    def pipeline():
        (gas,) = PIO.read_phantom(path, device=dev)
        gas.params["mass"] = gas.params["massoftype"]
        PIO.add_rho(gas)
        inp, _ = PIO.build_input(gas, PIO.MassFromParams("massoftype"), scalars=("rho", "vz"),
                                 smoothed_kernel=P.C4_Wendland)
        ctx.build(*inp.coord, inp.m, inp.h, inp.ρ, quant=list(inp.quant), kernel=P.C4_Wendland.id)
        ctx.interpolate_lines(_lib.PMX_SCATTER, _lib.PMX_PREC_F32C, cat_s, do_, dd_, outs)

    pipeline()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        pipeline()
    torch.cuda.synchronize()
    pipe_ms = (time.perf_counter() - t) * 1e3 / a.steps
    assert bool(torch.isfinite(outs[0]).any())
    ctx.close()
    line = {"metric": "Phantom dump -> device-resident InterpolationInput (particles/s)", "value": n / (ms * 1e-3),
            "unit": "particles/s (file read, H2D, rho, build_input)", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (deterministic SplitMix64 warped disc written as a Phantom dump)",
            "config": {"workload": f"Phantom dump of a warped disc, {n} particles, {fbytes} bytes; read_phantom "
                                   "(device) + add_rho + build_input(MassFromParams, vz, M4)", "particles": n},
            "file_GBs": gbs, "h2d_copy_GBs": h2d, "frac_of_h2d_copy": gbs / h2d,
            "host_path_ms": host_ms,
            "pipeline_ms": pipe_ms,
            "pipeline": f"dump -> InterpolationInput in HBM -> LBVH build (C4) -> face-on {a.map}x{a.map} map of "
                        "column density and column-averaged vz (FP32c)"}
    print(json.dumps(line))
    return 0


def run_dusty(a, torch, dev):
    """Two-species dusty disc: warped gas (default 2.4e7) and dust (a third of it: H_d = 0.3 H,
    dust-to-gas mass 0.01, v_phi x 0.995).  Per species and step: LBVH build, rho, div v and
    curl v (itpSymmetric, M4) on a 512x512x64 cylindrical grid, and a face-on 1024^2 column
    density map; then the dust-to-gas ratio rho_d / rho_g on the grid (device epilogue; the
    reference has no operator for it).  FP32c timed over the steps; the plain FP64 path timed
    over one step and used for the per-field errors."""
    from distorteddustydiscanalysistoolkits_amd import _lib, disc
    t0 = time.time()
    ng = a.particles or 24_000_000
    nd = ng // 3
    ns, nphi, nz = (int(v) for v in (a.grid if a.grid != "256,256,64" else "512,512,64").split(","))
    grid = disc.flatten_cylindrical(*disc.cylindrical_axes((1.0, 100.0, ns), (0.0, 2 * np.pi, nphi), (-5.0, 5.0, nz)))
    dgrid = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in grid]
    P = grid[0].size
    del grid
    lo, ld = disc.faceon_lines(a.map)
    dlo = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in lo]
    dld = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in ld]
    M = lo[0].size
    species = []
    for name, d in (("gas", disc.make_disc(ng, seed=0, warped=True)), ("dust", disc.make_dust(nd, seed=0, warped=True))):
        cols = [torch.from_numpy(np.ascontiguousarray(d[k], np.float32)).to(dev) for k in ("x", "y", "z", "m", "h", "rho")]
        quant = [torch.from_numpy(np.ascontiguousarray(d[k], np.float32)).to(dev) for k in ("rho", "vx", "vy", "vz")]
        species.append(dict(name=name, n=d["x"].size, cols=cols, quant=quant, ctx=_lib.Context(0)))
        del d
    log(f"[bench] dusty: gas {ng} + dust {nd} particles, {P} grid nodes + {M} lines per species, "
        f"prepared in {time.time() - t0:.1f}s")
    cat_p = _lib.catalog_struct(scalar_slots=[0], scalar_shepard=[0], div_slots=[1, 2, 3], curl_slots=[1, 2, 3])
    cat_l = _lib.catalog_struct(scalar_slots=[0], scalar_shepard=[0])
    fields = ("rho", "div_v", "curl_v_x", "curl_v_y", "curl_v_z")

    def run(prec, steps, warmup):
        ft = torch.float32 if prec == _lib.PMX_PREC_F32C else torch.float64
        outs = [dict(p=[torch.empty(P, dtype=ft, device=dev) for _ in fields], l=[torch.empty(M, dtype=ft, device=dev)])
                for _ in species]
        ratio = torch.empty(P, dtype=ft, device=dev)

        kpts = [0.0] * len(species)     # point-interpolation kernel time of the last step, per species

        def step():
            tb = ti = 0.0
            for j, (sp, o) in enumerate(zip(species, outs)):
                ctx = sp["ctx"]
                ctx.build(*sp["cols"], quant=sp["quant"], kernel=KERNEL_M4)
                ctx.interpolate_points(_lib.PMX_SYMMETRIC, prec, cat_p, *dgrid, o["p"])
                t = ctx.timings()
                tb += t[0]
                ti += t[5]
                kpts[j] = t[7]
                ctx.interpolate_lines(_lib.PMX_SCATTER, prec, cat_l, dlo, dld, o["l"])
                ti += ctx.timings()[5]
            torch.div(outs[1]["p"][0], outs[0]["p"][0], out=ratio)
            return tb, ti

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        tbs, tis = [], []
        t = time.perf_counter()
        for _ in range(steps):
            tb, ti = step()
            tbs.append(tb)
            tis.append(ti)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t) / steps
        return dict(ms_per_step=ms, build_ms=float(np.median(tbs)), interp_ms=float(np.median(tis)), outs=outs,
                    ratio=ratio, point_kernel_ms=list(kpts))

    log(f"[bench] dusty: FP32c steps={a.steps} warmup={a.warmup}")
    r32 = run(_lib.PMX_PREC_F32C, a.steps, a.warmup)
    log(f"[bench] dusty FP32c: {r32['ms_per_step']:.1f} ms/step (builds {r32['build_ms']:.1f}, interpolation "
        f"{r32['interp_ms']:.1f})")
    points_per_step = len(species) * (P + M)
    # pair-stream model of the dominant kernels (the point tiles of rho, div v, curl v): one
    # 10-column FP32 record per accepted (sample, particle) pair; each context still holds its
    # species' tree
    n_acc = [int(sp["ctx"].point_stats(_lib.PMX_SYMMETRIC, *dgrid)[4]) for sp in species]
    pb = 4 * (6 + 4) * sum(n_acc)
    t_k = sum(r32["point_kernel_ms"]) * 1e-3
    roof = {"bound": "hbm", "achieved": pb / t_k / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": pb / t_k / 1e9 / HBM_PEAK_GBS, "traffic": None,
            "kernel": "point interpolation kernels (f32c), both species",
            "algorithmic_bytes_per_launch": pb, "bytes_per_unit": 40, "units": sum(n_acc),
            "unit_name": "accepted (sample, particle) pairs", "kernel_ms": r32["point_kernel_ms"]}
    res = {"f32c": {k: r32[k] for k in ("ms_per_step", "build_ms", "interp_ms")}}
    res["f32c"]["grid_points_per_s"] = points_per_step / (r32["ms_per_step"] * 1e-3)
    errors = None
    if a.precision in ("both", "best"):
        log("[bench] dusty: plain FP64 path, one step")
        r64 = run(_lib.PMX_PREC_F64, 1, 0)
        log(f"[bench] dusty FP64: {r64['ms_per_step']:.1f} ms/step")
        res["f64"] = {k: r64[k] for k in ("ms_per_step", "build_ms", "interp_ms")}
        res["f64"]["grid_points_per_s"] = points_per_step / (r64["ms_per_step"] * 1e-3)
        errors = {}
        for sp, o32, o64 in zip(species, r32["outs"], r64["outs"]):
            for nm, g, r in zip(fields, o32["p"], o64["p"]):
                linf, l2 = rel_norms(g.double().cpu().numpy(), r.cpu().numpy())
                errors[f"{sp['name']}_{nm}"] = {"rel_linf": linf, "rel_l2": l2}
            linf, l2 = rel_norms(o32["l"][0].double().cpu().numpy(), o64["l"][0].cpu().numpy())
            errors[f"{sp['name']}_rho_column"] = {"rel_linf": linf, "rel_l2": l2}
        linf, l2 = rel_norms(r32["ratio"].double().cpu().numpy(), r64["ratio"].cpu().numpy())
        errors["dust_to_gas"] = {"rel_linf": linf, "rel_l2": l2}
        del r64
    n_total = sum(sp["n"] for sp in species)
    line = {
        "metric": "interp grid points/s + LBVH build particles/s at 16M parts, FP32c vs FP64",
        "value": res["f32c"]["grid_points_per_s"],
        "unit": "grid points/s (both species' grid nodes and map lines per second, builds + interpolation per step)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (deterministic SplitMix64 warped gas + dust discs, seed 0)",
        "config": {"workload": f"two-species dusty disc: gas {ng} + dust {nd} particles (FP32 storage), M4; per "
                               f"species rho, div v, curl v on a cylindrical {ns}x{nphi}x{nz} grid (itpSymmetric) "
                               f"and a face-on {a.map}x{a.map} column-density map; dust-to-gas ratio on the grid",
                   "particles": n_total, "samples_per_step": points_per_step},
        "build_particles_per_s": n_total / (r32["build_ms"] * 1e-3),
        "paths": res,
        "speedup_vs_f64": (res["f64"]["ms_per_step"] / res["f32c"]["ms_per_step"]) if "f64" in res else None,
        "field_error_f32c_vs_f64": errors,
        "roofline": roof, "cpu_baseline": None,
    }
    print(json.dumps(line))
    for sp in species:
        sp["ctx"].close()
    return 0


def per_call_compare(jobs, torch, dev):
    """The same snapshots one call at a time on one context (largest first, after one warm-up
    call): build + interpolation device time per job, summed, for the FP32c path and for the
    plain FP64 path (FP64 outputs)."""
    from distorteddustydiscanalysistoolkits_amd import _lib
    order = sorted(range(len(jobs)), key=lambda j: -jobs[j]["cols"][0].shape[0])
    P_total = sum(jobs[j]["out"][0].shape[0] for j in order)
    res = {}
    for name, prec in (("f32c", _lib.PMX_PREC_F32C), ("f64", _lib.PMX_PREC_F64)):
        ctx = _lib.Context(0)
        tb = ti = 0.0
        n_acc = pair_bytes = 0
        t0 = time.time()
        for k, j in enumerate([order[0]] + order):
            job = jobs[j]
            ctx.build(*job["cols"], quant=job["quant"], kernel=KERNEL_M4)
            cat = _lib.catalog_struct(scalar_slots=job["scalar_slots"], scalar_shepard=job["scalar_shepard"])
            P = job["out"][0].shape[0]
            out = [torch.empty(P, dtype=torch.float32 if name == "f32c" else torch.float64, device=dev)]
            if job["kind"] == "points":
                ctx.interpolate_points(job["strategy"], prec, cat, *job["samples"], out)
            else:
                ctx.interpolate_lines(_lib.PMX_SCATTER, prec, cat, job["samples"][0], job["samples"][1], out)
            if k > 0:                      # the first call is the warm-up
                t = ctx.timings()
                tb += t[0]
                ti += t[6]
                if name == "f32c":         # accepted pairs of the same call (pair-stream model)
                    st = (ctx.point_stats(job["strategy"], *job["samples"]) if job["kind"] == "points"
                          else ctx.line_stats(*job["samples"]))
                    n_acc += int(st[4])
                    pair_bytes += 4 * (6 + len(job["quant"])) * int(st[4])
        ctx.close()
        log(f"[bench] per-call {name}: builds {tb:.1f} ms, interpolation {ti:.1f} ms ({time.time() - t0:.1f} s)")
        res[name] = {"build_ms": tb, "interp_ms": ti, "grid_points_per_s": P_total / ((tb + ti) * 1e-3)}
        if name == "f32c":
            res[name].update(accepted_pairs=n_acc, pair_bytes=pair_bytes)
    return res


def measured_traffic(entry, n=None, P=None):
    """HBM bytes per launch of a kernel group from profiles/pmc_traffic.json (calibrated rocprofv3
    FETCH_SIZE + WRITE_SIZE of an earlier profiled run), with where they come from.  The entry is
    used only when its workload has this run's particle and sample counts; otherwise (None,
    reason)."""
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        e = json.load(open(tfile)).get(entry)
    except Exception:
        e = None
    if not e:
        return None, {"file": "profiles/pmc_traffic.json", "entry": entry, "used": False, "reason": "no entry"}
    wl = e.get("workload", {})
    src = {"file": "profiles/pmc_traffic.json", "entry": entry, "commit": e.get("commit"), "workload": wl,
           "kernel": e.get("kernel")}
    if (n is not None and wl.get("particles") != n) or (P is not None and "samples" in wl and wl["samples"] != P):
        src.update(used=False, reason="profiled workload differs from this run")
        return None, src
    src["used"] = True
    return e.get("bytes"), src


def build_roofline(n, k, build_ms, kind=None, passes=8):
    """Algorithmic HBM bytes of the LBVH build per particle (FP32 storage, k quant columns,
    8-bit radix passes): bbox 12 + encode 24 + sort (8 + 24 P) + column gather (4 + 8 (6 + k))
    + BRT (8 + 32) + refit 56 = 192 + 8 k + 24 P, against the 8 TB/s peak.  Beside it, the
    measured bytes of one build (calibrated FETCH + WRITE of the build kernels, profiled on the
    line-map workload) when the particle count matches."""
    b = 192 + 8 * k + 24 * passes
    gbs = b * n / (build_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "bytes_per_particle": b, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": gbs / HBM_PEAK_GBS, "traffic": None, "measured_GBs": None}
    if kind == "lines":
        tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
        try:
            e = json.load(open(tfile)).get("build_lines")
        except Exception:
            e = None
        if e and e.get("workload", {}).get("particles") == n:
            # per-kernel entries are totals over the kernel's dispatches in one build
            tot = sum(v["fetch_bytes"] + v["write_bytes"] for v in e["kernels"].values())
            out["traffic"] = tot
            out["measured_GBs"] = tot / (build_ms * 1e-3) / 1e9
            out["traffic_source"] = {"file": "profiles/pmc_traffic.json", "entry": "build_lines",
                                     "commit": e.get("commit")}
    return out


def cpu_baseline(w, nsample):
    """The oracle (C FP64 restatement of the reference CPU algorithm, OpenMP where the reference
    threads) on this host: the full LBVH build of the same particle set, then the sample
    kernel on an evenly strided subset of the samples; whole-step throughput is extrapolated
    as P / (t_build + t_sample * P / sample)."""
    from oracle import oracle as O
    nthreads = min(os.cpu_count() or 1, 16)
    cols = [w.host32[k].astype(np.float64) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [w.host32[k].astype(np.float64) for k in w.qnames]
    log(f"[bench] cpu baseline: oracle build n={cols[0].size} threads={nthreads}")
    t = time.perf_counter()
    tree = O.LBVH(*cols, quant=quant, nthreads=nthreads)
    tb = time.perf_counter() - t
    P = w.P
    step = max(1, P // nsample)
    sel = np.arange(0, P, step)[:nsample]
    log(f"[bench] cpu baseline: oracle samples={sel.size}")
    t = time.perf_counter()
    if w.kind == "lines":
        o, dv = w.samples
        tree.line_samples(w.kname, "scatter", [0, 1], [0, 1], [v[sel] for v in o], [v[sel] for v in dv],
                          nthreads=nthreads)
    else:
        tree.point_samples(w.kname, "symmetric", w.cat_d, *[v[sel] for v in w.samples], nthreads=nthreads)
    ti = time.perf_counter() - t
    # the same sample kernel on one thread, over every 8th sample of the subset
    sel1 = sel[::8]
    t = time.perf_counter()
    if w.kind == "lines":
        tree.line_samples(w.kname, "scatter", [0, 1], [0, 1], [v[sel1] for v in o], [v[sel1] for v in dv],
                          nthreads=1)
    else:
        tree.point_samples(w.kname, "symmetric", w.cat_d, *[v[sel1] for v in w.samples], nthreads=1)
    t1 = time.perf_counter() - t
    total = tb + ti * P / sel.size
    return {"value": P / total, "unit": "grid points/s", "cores": nthreads, "kind": "port",
            "sample": f"full LBVH build of {cols[0].size} particles ({tb:.2f} s) + {sel.size} of {P} samples "
                      f"({ti:.2f} s), extrapolated to all samples",
            "build_particles_per_s": cols[0].size / tb, "interp_points_per_s": sel.size / ti,
            "interp_points_per_s_1thread": sel1.size / t1,
            "sample_1thread": f"{sel1.size} samples on one thread ({t1:.2f} s)"}


if __name__ == "__main__":
    sys.exit(main())
