"""Batches of snapshots (pmx_batch_*): one persistent kernel over the tiles of many jobs,
eager and replayed as a hipGraph, before and after the autotuner.  Every job must agree
with the per-call FP32-compensated path on the same job (same tree, same tile procedure:
1e-12 relative, line sums differ only in the order of FP64 additions), and one job is
checked against the oracle to the FP32c tolerance."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc
from oracle import oracle as O

from .helpers import disc_columns, oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
F32C_TOL = 1e-6


def _job_specs():
    cyl = disc.flatten_cylindrical(*disc.cylindrical_axes((1.0, 100.0, 12), (0.0, 2 * np.pi, 12), (-5.0, 5.0, 6)))
    ax = disc.linrange(-100.0, 100.0, 24)
    sl = disc.flatten_cartesian(ax, ax, np.array([0.0]))
    return [
        dict(n=6000, seed=1, warped=False, kind="points", samples=cyl, strategy=_lib.PMX_SYMMETRIC,
             slots=[0], shep=[1]),
        dict(n=9000, seed=2, warped=True, kind="points", samples=sl, strategy=_lib.PMX_GATHER, slots=[3], shep=[1]),
        dict(n=8000, seed=3, warped=True, kind="lines", samples=disc.faceon_lines(24), slots=[0, 3], shep=[0, 1]),
        dict(n=3000, seed=4, warped=False, kind="lines", samples=disc.faceon_lines(16), slots=[0], shep=[0]),
        dict(n=12000, seed=5, warped=False, kind="points", samples=cyl, strategy=_lib.PMX_SCATTER, slots=[0, 1, 2],
             shep=[1, 1, 1]),
        # a camera's rays (nearly parallel set, slopes up to 0.25)
        dict(n=7000, seed=6, warped=True, kind="lines", samples=disc.perspective_lines(20, height=400.0), slots=[0, 3],
             shep=[0, 1]),
    ]


@pytest.fixture(scope="module")
def batch_setup():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    specs = _job_specs()
    jobs, host = [], []
    for k, sp in enumerate(specs):
        cols, quant = disc_columns(sp["n"], seed=sp["seed"], warped=sp["warped"], dtype=np.float32)
        dcols = [torch.from_numpy(c).to(dev) for c in cols]
        dquant = [torch.from_numpy(q).to(dev) for q in quant]
        if sp["kind"] == "points":
            ds = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in sp["samples"]]
            P = ds[0].shape[0]
        else:
            o, d = sp["samples"]
            ds = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o],
                  [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in d])
            P = ds[0][0].shape[0]
        outs = [torch.full((P,), -1.0, dtype=torch.float32, device=dev) for _ in sp["slots"]]
        jobs.append(dict(cols=dcols, quant=dquant, kernel=0, kind=sp["kind"], strategy=sp.get("strategy", 1),
                         samples=ds, scalar_slots=sp["slots"], scalar_shepard=sp["shep"], out=outs))
        host.append((cols, quant))
    print(f"[batch] {len(jobs)} jobs", flush=True)
    b = _lib.Batch(jobs)
    yield b, jobs, specs, host
    b.close()


def _per_call(job, spec):
    """The same job through the per-call FP32c path."""
    import torch
    ctx = _lib.Context(0)
    try:
        ctx.build(*job["cols"], quant=job["quant"], kernel=0)
        cat = _lib.catalog_struct(scalar_slots=spec["slots"], scalar_shepard=spec["shep"])
        P = job["out"][0].shape[0]
        outs = [torch.empty(P, dtype=torch.float32, device=job["out"][0].device) for _ in spec["slots"]]
        if spec["kind"] == "points":
            ctx.interpolate_points(job["strategy"], _lib.PMX_PREC_F32C, cat, *job["samples"], outs)
        else:
            ctx.interpolate_lines(_lib.PMX_SCATTER, _lib.PMX_PREC_F32C, cat, job["samples"][0], job["samples"][1],
                                  outs)
        return [o.double().cpu().numpy() for o in outs]
    finally:
        ctx.close()


def _compare_all(jobs, specs, tol=1e-12):
    for k, (job, sp) in enumerate(zip(jobs, specs)):
        ref = _per_call(job, sp)
        for c, (g, r) in enumerate(zip(job["out"], ref)):
            g = g.double().cpu().numpy()
            assert np.array_equal(np.isnan(g), np.isnan(r)), f"job {k} col {c}: NaN pattern"
            assert rel_norms(g, r)[0] <= tol, f"job {k} col {c}: {rel_norms(g, r)}"


def test_batch_eager_matches_per_call(batch_setup):
    b, jobs, specs, _ = batch_setup
    b.run(graph=False)
    print("[batch] eager", b.info(), flush=True)
    _compare_all(jobs, specs)


def test_batch_graph_replays(batch_setup):
    b, jobs, specs, _ = batch_setup
    for rep in range(3):
        for job in jobs:
            for o in job["out"]:
                o.fill_(-1.0)
        b.run(graph=True)
        print(f"[batch] graph replay {rep}", b.info(), flush=True)
        _compare_all(jobs, specs)


def test_batch_tuned(batch_setup):
    b, jobs, specs, _ = batch_setup
    b.tune()
    info = b.info()
    print("[batch] tuned", info, flush=True)
    assert info["waves_per_cu"] in (4, 8, 16, 32)
    b.run(graph=True)
    _compare_all(jobs, specs)


def test_batch_against_oracle(batch_setup):
    b, jobs, specs, host = batch_setup
    b.run(graph=False)
    k = 2                                              # the two-column line map
    cols, quant = host[k]
    tree, _, _ = oracle_tree(cols, quant)
    o, d = specs[k]["samples"]
    ref = tree.line_samples("M4", "scatter", specs[k]["slots"], specs[k]["shep"], o, d)
    for g, r in zip(jobs[k]["out"], ref):
        g = g.double().cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(r))
        assert max(rel_norms(g, r)) <= F32C_TOL
    k = 5                                              # the camera's rays
    cols, quant = host[k]
    tree, _, _ = oracle_tree(cols, quant)
    o, d = specs[k]["samples"]
    ref = tree.line_samples("M4", "scatter", specs[k]["slots"], specs[k]["shep"], o, d)
    for g, r in zip(jobs[k]["out"], ref):
        g = g.double().cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(r))
        assert max(rel_norms(g, r)) <= F32C_TOL
    k = 0                                              # symmetric points on a cylindrical grid
    cols, quant = host[k]
    tree, _, _ = oracle_tree(cols, quant)
    ref = tree.point_samples("M4", "symmetric", dict(scalar_slots=[0], scalar_shepard=[1]), *specs[k]["samples"])
    g = jobs[k]["out"][0].double().cpu().numpy()
    assert np.array_equal(np.isnan(g), np.isnan(ref[0]))
    assert max(rel_norms(g, ref[0])) <= F32C_TOL


def test_batch_particles_grow_between_runs(batch_setup):
    """The particles of the line jobs spread to 4x their extent after the batch was planned: the
    frames' tolerances follow the rebuilt particle boxes (eager and replayed).  Against the
    per-call path to the FP32c tolerance: with the particles 4x deeper, the per-call path may
    send the camera's rays to the union walk, whose weights round differently."""
    b, jobs, specs, _ = batch_setup
    line_jobs = [k for k, sp in enumerate(specs) if sp["kind"] == "lines"]
    try:
        for k in line_jobs:
            for c in jobs[k]["cols"][:3] + [jobs[k]["cols"][4]]:
                c.mul_(4.0)                            # x, y, z, h (exact in FP32)
        for graph in (False, True):
            b.run(graph=graph)
            _compare_all([jobs[k] for k in line_jobs], [specs[k] for k in line_jobs], tol=F32C_TOL)
    finally:
        for k in line_jobs:
            for c in jobs[k]["cols"][:3] + [jobs[k]["cols"][4]]:
                c.mul_(0.25)
    b.run(graph=True)
    _compare_all(jobs, specs)


def test_batch_split_tiles(batch_setup, monkeypatch):
    """Every gather / symmetric point tile split (16 leaf tests, items of <= 64 particles):
    eager and graph replays agree with the per-call path to FP32 rounding."""
    _, jobs, specs, _ = batch_setup
    monkeypatch.setenv("PMX_SPLIT_BUDGET", "16")
    monkeypatch.setenv("PMX_SPLIT_ITEM", "64")
    b2 = _lib.Batch(jobs)
    try:
        assert b2.info()["split_budget"] == 16
        for graph in (False, True, True):
            for job in jobs:
                for o in job["out"]:
                    o.fill_(-1.0)
            b2.run(graph=graph)
            print(f"[batch split] graph={graph}", b2.info(), flush=True)
            _compare_all(jobs, specs, tol=1e-6)
    finally:
        b2.close()


def test_batch_rejects_bad_jobs():
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    cols, quant = disc_columns(500, seed=9, dtype=np.float32)
    dcols = [torch.from_numpy(c).to(dev) for c in cols]
    dquant = [torch.from_numpy(q).to(dev) for q in quant]
    o, d = disc.faceon_lines(8)
    d = [v.copy() for v in d]
    d[0][5] = 1e-3                                     # not one parallel set
    ds = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o],
          [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in d])
    out = [torch.empty(64, dtype=torch.float32, device=dev)]
    with pytest.raises(_lib.ArgumentError):
        _lib.Batch([dict(cols=dcols, quant=dquant, kind="lines", samples=ds, scalar_slots=[0], out=out)])
    with pytest.raises(_lib.ArgumentError):
        _lib.Batch([dict(cols=dcols, quant=dquant, kind="points", strategy=7, samples=[t for t in ds[0]],
                         scalar_slots=[0], out=out)])
