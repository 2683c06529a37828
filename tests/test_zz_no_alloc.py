"""pp_batch_run_static, the batch entry point that allocates nothing (needs a GPU).

Its results are checked against a batch plan of the same volumes (pp_batch_plan_create +
pp_batch_run, the same kernels over a hipMalloc'ed descriptor table), on two calls in a
row, the second with fewer and smaller volumes than the first (the fixed table keeps the
first call's descriptors beyond the second's): gathered outputs (pull, dL/dphi) bit for
bit, scattered ones (push, dL/df, the VJP's df) to rounding (their tiles meet in float
atomics in no fixed order), the loss to 1e-9. A HIP graph that captures the call replays
it, with eager work between the replays, to the eager results.
"""
import numpy as np
import pytest

from helpers import identity, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull import batch as ppb  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SCATTER_TOL = 1e-6


def _volumes(seed, n, lo, hi, op):
    rng = np.random.default_rng(seed)
    vols = []
    for i in range(n):
        d = tuple(int(v) for v in rng.integers(lo, hi, size=3))
        C = int(rng.integers(1, 4))
        deg = int(rng.choice([0, 1, 2, 3, 4]))
        phi = torch.from_numpy(identity(d) + smooth_field(rng, d, 3.0, 3.0)).to(DEV)
        img = torch.from_numpy(rng.standard_normal((C,) + d).astype(np.float32)).to(DEV)
        v = dict(phi=phi, deg=deg, bnd=tuple(int(b) for b in rng.integers(0, 3, size=3)), ext=int(i % 3 != 1))
        if op == "push":
            v.update(f=img, dim=tuple(int(x) for x in np.array(d) + rng.integers(-2, 3, size=3)))
        else:
            v.update(f=img, aux=torch.from_numpy(rng.standard_normal((C,) + d).astype(np.float32)).to(DEV))
        vols.append(v)
    return vols


def _with_outputs(op, vols):
    out = []
    for v in vols:
        C, d1 = v["f"].shape[0], tuple(v["phi"].shape[1:])
        d0 = tuple(v["dim"]) if op == "push" else tuple(v["f"].shape[1:])
        w = dict(v, out=torch.full({"pull": (C,) + d1, "push": (C,) + d0}.get(op, (3,) + d1), 7.0, device=DEV))
        if op in ("objective", "pull_vjp"):
            w["out2"] = torch.full((C,) + d0, 7.0, device=DEV)
        if op == "objective":
            w["loss"] = torch.full((), 7.0, dtype=torch.float64, device=DEV)
        out.append(w)
    return out


def _host(recs):
    return [{k: (t.detach().cpu().numpy() if t is not None else None) for k, t in r.items()
             if k in ("out", "out2", "loss")} for r in recs]


def _rel(a, b):
    a = np.nan_to_num(a.astype(np.float64))
    b = np.nan_to_num(b.astype(np.float64))
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def _compare(op, got, ref):
    for i, (g, r) in enumerate(zip(got, ref)):
        if op in ("pull", "objective", "pull_vjp"):  # gathered per sample
            assert np.array_equal(g["out"], r["out"], equal_nan=True), (op, i)
        else:
            assert _rel(g["out"], r["out"]) <= SCATTER_TOL, (op, i, _rel(g["out"], r["out"]))
        if op in ("objective", "pull_vjp"):
            assert _rel(g["out2"], r["out2"]) <= SCATTER_TOL, (op, i, _rel(g["out2"], r["out2"]))
        if op == "objective":
            assert abs(float(g["loss"]) - float(r["loss"])) <= 1e-9 * abs(float(r["loss"])), (i, g["loss"], r["loss"])


def _plan_results(op, vols):
    plan = ppb.BatchPlan(op, vols)
    recs = plan.run()
    torch.cuda.synchronize()
    res = _host(recs)
    plan.close()
    return res


@pytest.mark.parametrize("op", ["pull", "push", "objective", "pull_vjp"])
def test_static_batch_matches_plan_on_two_calls(op):
    big = _volumes(10, 13, 20, 56, op)   # 13 volumes: two upload chunks
    small = _volumes(11, 3, 8, 20, op)   # then fewer, smaller volumes
    for vols in (big, small):
        ref = _plan_results(op, vols)
        recs = ppb.run_static(op, _with_outputs(op, vols))
        torch.cuda.synchronize()
        _compare(op, _host(recs), ref)


def test_static_batch_under_graph_capture():
    op = "objective"
    vols = _volumes(12, 6, 16, 40, op)
    ref = _plan_results(op, vols)
    outs = _with_outputs(op, vols)
    graph = ppb.Graph(lambda s: ppb.run_static(op, outs, s))
    for _ in range(3):
        graph.launch()
        graph.stream.synchronize()
        recs = [dict(out=o["out"], out2=o["out2"], loss=o["loss"]) for o in outs]
        _compare(op, _host(recs), ref)
        for o in outs:  # eager work between the replays
            o["out"].fill_(-3.0)
            o["out2"].fill_(-3.0)
        torch.cuda.synchronize()
    graph.close()


def test_static_batch_rejects_more_than_its_table():
    v = _volumes(13, 1, 8, 9, "pull")[0]
    with pytest.raises(pp._lib.PPError, match="at most 256"):
        ppb.run_static("pull", [v] * (ppb.STATIC_MAX_VOLUMES + 1))
