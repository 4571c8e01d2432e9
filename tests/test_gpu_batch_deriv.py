"""Batches with derivative catalogs (gradients, divergences, curls) on point jobs, mixed with a
line job: every output column agrees with the per-call FP32-compensated path on the same job
(same tile procedure, FP64 sums in another order: 1e-6 relative), eager, as graph replays and
with every heavy tile split; a multi-chunk catalog (5 scalars, 3 gradients, 3 divergences,
3 curls: 26 columns) is checked against the oracle to the FP32c tolerance."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc

from .helpers import CAT_RHO_V, disc_columns, oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
F32C_TOL = 1e-6

CAT_WIDE = dict(scalar_slots=[0, 1, 2, 3, 0], scalar_shepard=[0, 1, 1, 1, 1], grad_slots=[-1, 0, 2],
                div_slots=[1, 2, 3, 0, 1, 2, 3, 2, 1], curl_slots=[1, 2, 3, 2, 3, 0, 1, 1, 2])
CAT_DONLY = dict(scalar_slots=[], scalar_shepard=[], grad_slots=[-1], div_slots=[], curl_slots=[1, 2, 3])


def _ncols(cat):
    return (len(cat.get("scalar_slots", ())) + 3 * len(cat.get("grad_slots", ())) + len(cat.get("div_slots", ())) // 3
            + len(cat.get("curl_slots", ())))


def _specs():
    cyl = disc.flatten_cylindrical(*disc.cylindrical_axes((1.0, 100.0, 10), (0.0, 2 * np.pi, 10), (-5.0, 5.0, 6)))
    ax = disc.linrange(-100.0, 100.0, 20)
    sl = disc.flatten_cartesian(ax, ax, np.array([0.0]))
    return [
        dict(n=6000, seed=11, warped=False, kind="points", samples=cyl, strategy=_lib.PMX_SYMMETRIC, cat=CAT_RHO_V),
        dict(n=7000, seed=12, warped=True, kind="lines", samples=disc.faceon_lines(16),
             cat=dict(scalar_slots=[0, 3], scalar_shepard=[0, 1])),
        dict(n=9000, seed=13, warped=True, kind="points", samples=sl, strategy=_lib.PMX_GATHER, cat=CAT_WIDE),
        dict(n=5000, seed=14, warped=False, kind="points", samples=cyl, strategy=_lib.PMX_SCATTER, cat=CAT_DONLY),
    ]


def _jobs(specs):
    import torch
    dev = torch.device("cuda:0")
    jobs, host = [], []
    for sp in specs:
        cols, quant = disc_columns(sp["n"], seed=sp["seed"], warped=sp["warped"], dtype=np.float32)
        if sp["kind"] == "points":
            ds = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in sp["samples"]]
            P = ds[0].shape[0]
        else:
            o, d = sp["samples"]
            ds = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o],
                  [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in d])
            P = ds[0][0].shape[0]
        outs = [torch.full((P,), -1.0, dtype=torch.float32, device=dev) for _ in range(_ncols(sp["cat"]))]
        jobs.append(dict(cols=[torch.from_numpy(c).to(dev) for c in cols],
                         quant=[torch.from_numpy(q).to(dev) for q in quant], kernel=0, kind=sp["kind"],
                         strategy=sp.get("strategy", _lib.PMX_SCATTER), samples=ds, out=outs, **sp["cat"]))
        host.append((cols, quant))
    return jobs, host


@pytest.fixture(scope="module")
def deriv_setup():
    pytest.importorskip("torch")
    specs = _specs()
    jobs, host = _jobs(specs)
    print(f"[batch deriv] {len(jobs)} jobs, columns {[len(j['out']) for j in jobs]}", flush=True)
    b = _lib.Batch(jobs)
    yield b, jobs, specs, host
    b.close()


def _per_call(job, spec):
    import torch
    ctx = _lib.Context(0)
    try:
        ctx.build(*job["cols"], quant=job["quant"], kernel=0)
        cat = _lib.catalog_struct(**spec["cat"])
        outs = [torch.empty_like(o) for o in job["out"]]
        if spec["kind"] == "points":
            ctx.interpolate_points(job["strategy"], _lib.PMX_PREC_F32C, cat, *job["samples"], outs)
        else:
            ctx.interpolate_lines(_lib.PMX_SCATTER, _lib.PMX_PREC_F32C, cat, job["samples"][0], job["samples"][1],
                                  outs)
        return [o.double().cpu().numpy() for o in outs]
    finally:
        ctx.close()


def _compare(jobs, specs, tol=F32C_TOL):
    for k, (job, sp) in enumerate(zip(jobs, specs)):
        ref = _per_call(job, sp)
        for c, (g, r) in enumerate(zip(job["out"], ref)):
            g = g.double().cpu().numpy()
            assert np.array_equal(np.isnan(g), np.isnan(r)), f"job {k} col {c}: NaN pattern"
            assert np.isfinite(g).sum() > 0, f"job {k} col {c}: no finite values"
            assert max(rel_norms(g, r)) <= tol, f"job {k} col {c}: {rel_norms(g, r)}"


def _clear(jobs):
    for job in jobs:
        for o in job["out"]:
            o.fill_(-1.0)


def test_deriv_batch_eager_and_graph(deriv_setup):
    b, jobs, specs, _ = deriv_setup
    for graph in (False, True, True):
        _clear(jobs)
        b.run(graph=graph)
        print(f"[batch deriv] graph={graph}", b.info(), flush=True)
        _compare(jobs, specs)


def test_deriv_batch_tuned(deriv_setup):
    """The autotuner re-plans the units (derivative tile LDS included) for every candidate shape."""
    b, jobs, specs, _ = deriv_setup
    b.tune()
    print("[batch deriv] tuned", b.info(), flush=True)
    _clear(jobs)
    b.run(graph=True)
    _compare(jobs, specs)


def test_deriv_batch_against_oracle(deriv_setup):
    b, jobs, specs, host = deriv_setup
    _clear(jobs)
    b.run(graph=False)
    for k in (0, 2, 3):
        cols, quant = host[k]
        tree, _, _ = oracle_tree(cols, quant)
        strategy = {_lib.PMX_GATHER: "gather", _lib.PMX_SCATTER: "scatter", _lib.PMX_SYMMETRIC: "symmetric"}
        ref = tree.point_samples("M4", strategy[specs[k]["strategy"]], specs[k]["cat"], *specs[k]["samples"])
        assert len(ref) == len(jobs[k]["out"])
        for c, (g, r) in enumerate(zip(jobs[k]["out"], ref)):
            g = g.double().cpu().numpy()
            assert np.array_equal(np.isnan(g), np.isnan(r)), f"job {k} col {c}: NaN pattern"
            assert max(rel_norms(g, r)) <= F32C_TOL, f"job {k} col {c}: {rel_norms(g, r)}"


def test_deriv_batch_split_tiles(deriv_setup, monkeypatch):
    """Every gather / symmetric tile split (16 leaf tests, items of <= 64 particles): the split
    items and outputs of derivative units agree with the per-call path."""
    _, jobs, specs, _ = deriv_setup
    monkeypatch.setenv("PMX_SPLIT_BUDGET", "16")
    monkeypatch.setenv("PMX_SPLIT_ITEM", "64")
    b2 = _lib.Batch(jobs)
    try:
        assert b2.info()["split_budget"] == 16
        for graph in (False, True):
            _clear(jobs)
            b2.run(graph=graph)
            print(f"[batch deriv split] graph={graph}", b2.info(), flush=True)
            _compare(jobs, specs)
    finally:
        b2.close()


def test_deriv_batch_rejects_bad_catalogs():
    import torch
    dev = torch.device("cuda:0")
    cols, quant = disc_columns(500, seed=9, dtype=np.float32)
    dcols = [torch.from_numpy(c).to(dev) for c in cols]
    dquant = [torch.from_numpy(q).to(dev) for q in quant]
    o, d = disc.faceon_lines(8)
    ds = ([torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in o],
          [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in d])
    pts = [torch.zeros(64, dtype=torch.float64, device=dev) for _ in range(3)]
    out = [torch.empty(64, dtype=torch.float32, device=dev) for _ in range(4)]
    with pytest.raises(_lib.ArgumentError):            # lines take scalars only
        _lib.Batch([dict(cols=dcols, quant=dquant, kind="lines", samples=ds, scalar_slots=[0], grad_slots=[0],
                         out=out)])
    with pytest.raises(_lib.DimensionMismatch):        # gradient slot beyond the columns
        _lib.Batch([dict(cols=dcols, quant=dquant, kind="points", samples=pts, scalar_slots=[0], grad_slots=[9],
                         out=out)])
    with pytest.raises(_lib.DimensionMismatch):        # one output column short
        _lib.Batch([dict(cols=dcols, quant=dquant, kind="points", samples=pts, scalar_slots=[0], grad_slots=[0],
                         out=out[:3])])
