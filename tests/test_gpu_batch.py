"""Batched uneven volumes (one persistent launch) and HIP graph replay (needs a GPU).

Checks the batch against the per-volume loop of single Tier B calls and against the
FP64 oracle, and a captured graph against eager execution on every replay.
"""
import numpy as np
import pytest

import oracle
from helpers import S, identity, rel_err, report, same_nan_mask, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull import batch as ppb  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-5


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


def _volumes(seed, n=10, lo=12, hi=40):
    rng = np.random.default_rng(seed)
    vols = []
    for i in range(n):
        d = tuple(int(v) for v in rng.integers(lo, hi, size=3))
        C = int(rng.integers(1, 5))
        deg = int(rng.choice([0, 1, 2, 3, 4])) if i % 3 == 0 else int(rng.choice([1, 3]))
        f = np.stack([smooth_field(rng, d, 1.5, 1.0)[0] for _ in range(C)]).astype(np.float32)
        phi = identity(d) + smooth_field(rng, d, 3.0, 4.0)
        if i % 4 == 1:  # a few samples far outside
            flat = phi.reshape(3, -1)
            idx = rng.choice(flat.shape[1], size=flat.shape[1] // 30, replace=False)
            flat[0, idx] = rng.uniform(-20, d[2] + 20, size=idx.size).astype(np.float32)
        aux = rng.standard_normal((C,) + d).astype(np.float32)
        bnd = tuple(int(v) for v in rng.integers(0, 3, size=3))
        vols.append(dict(f=f, phi=phi, aux=aux, deg=deg, bnd=bnd, ext=int(i % 5 != 2), d=d, C=C))
    return vols


def _dev(v, **extra):
    return dict(f=T(v["f"]), phi=T(v["phi"]), aux=T(v["aux"]), deg=v["deg"], bnd=v["bnd"], ext=v["ext"], **extra)


def _check(got, ref, what):
    assert same_nan_mask(got, ref), f"{what}: NaN mask differs"
    e = rel_err(got, ref)
    report(f"batch {what}", e)
    assert e < TOL, f"{what}: rel err {e:.3g}"


def test_batch_pull_and_pullback_match_oracle():
    vols = _volumes(1)
    plan = ppb.BatchPlan("pull", [_dev(v) for v in vols])
    outs = plan.run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        _check(H(o["out"]), oracle.pull(v["f"][None], v["phi"][None], s, prec="f64")[0], f"pull vol {i}")
    plan = ppb.BatchPlan("pull_backward", [_dev(v) for v in vols])
    outs = plan.run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        ref = oracle.backward_phi(v["f"][None], v["phi"][None], v["aux"][None], s)[0]
        _check(H(o["out"]), ref, f"pullback vol {i}")


def test_batch_push_and_push_grad_match_oracle():
    vols = _volumes(2)
    rng = np.random.default_rng(0)
    dev = []
    for v in vols:
        v["dim"] = tuple(int(x) for x in np.array(v["d"]) + rng.integers(-3, 4, size=3))
        v["g3"] = rng.standard_normal((v["C"], 3) + v["d"]).astype(np.float32)
        dev.append(dict(f=T(v["aux"]), phi=T(v["phi"]), deg=v["deg"], bnd=v["bnd"], ext=v["ext"], dim=v["dim"]))
    outs = ppb.BatchPlan("push", dev).run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        _check(H(o["out"]), oracle.push(v["aux"][None], v["phi"][None], v["dim"], s, prec="f64")[0], f"push vol {i}")
    dev = [dict(f=T(v["g3"]), phi=T(v["phi"]), deg=v["deg"], bnd=v["bnd"], ext=v["ext"], dim=v["dim"]) for v in vols]
    outs = ppb.BatchPlan("push_grad", dev).run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        ref = oracle.push_grad(v["g3"][None], v["phi"][None], v["dim"], s, prec="f64")[0]
        _check(H(o["out"]), ref, f"push_grad vol {i}")


def test_batch_objective_matches_per_volume_calls():
    vols = _volumes(3)
    plan = ppb.BatchPlan("objective", [_dev(v) for v in vols])
    outs = plan.run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = pp.Settings(v["deg"], v["bnd"], v["ext"])
        loss, dphi, df = pp.objective_grad(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
        _check(H(o["out"]), H(dphi[0]), f"objective dphi vol {i}")
        _check(H(o["out2"]), H(df[0]), f"objective df vol {i}")
        assert abs(float(o["loss"]) - float(loss)) <= TOL * abs(float(loss)), i


def test_graph_replays_match_eager():
    vols = _volumes(4, n=6)
    plan = ppb.BatchPlan("objective", [_dev(v) for v in vols])
    eager = plan.run()
    torch.cuda.synchronize()
    ref = [(H(o["out"]).copy(), H(o["out2"]).copy(), float(o["loss"])) for o in eager]
    g = ppb.Graph(lambda s: plan.run(s))
    for rep in range(4):
        for o in plan.vols:  # scribble over the outputs between replays
            o["out"].fill_(123.0)
            o["out2"].fill_(-7.0)
            o["loss"].fill_(5.0)
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        for i, (o, r) in enumerate(zip(plan.vols, ref)):
            _check(H(o["out"]), r[0], f"replay {rep} dphi vol {i}")
            _check(H(o["out2"]), r[1], f"replay {rep} df vol {i}")
            assert abs(float(o["loss"]) - r[2]) <= TOL * abs(r[2])
    g.close()
    plan.close()


def test_graph_of_tier_b_calls():
    """A captured pull + push + pullback sequence replays like the eager one."""
    import ctypes as ct

    from pushpull import _lib, ops

    rng = np.random.default_rng(8)
    d = (40, 48, 36)
    f = T(smooth_field(rng, d, 2.0, 1.0)[0])
    phi = T(identity(d) + smooth_field(rng, d, 4.0, 5.0))
    delta = T(rng.standard_normal(d).astype(np.float32))
    outs = [torch.empty(d, device=DEV), torch.empty(d, device=DEV), torch.empty((3,) + d, device=DEV)]
    desc = ops._desc(pp.Settings(3, 1, 1), d, d, 1, 1)
    L = _lib.lib()
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731

    def seq(stream):
        s = ct.c_void_p(stream.cuda_stream)
        _lib.check(L.pp_pull(ct.byref(desc), P(outs[0]), P(phi), P(f), s), "pull")
        _lib.check(L.pp_push(ct.byref(desc), P(outs[1]), P(phi), P(delta), 0, s), "push")
        _lib.check(L.pp_pull_backward(ct.byref(desc), P(outs[2]), P(phi), P(f), P(delta), s), "bwd")

    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        seq(st)
    st.synchronize()
    ref = [H(o).copy() for o in outs]
    g = ppb.Graph(seq, stream=st)
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        st.synchronize()
        for o, r in zip(outs, ref):
            _check(H(o), r, f"graph replay {rep}")
    g.close()


def test_batch_pull_vjp_matches_per_volume_calls():
    """One persistent launch per degree of the pull VJP over uneven volumes (degrees 0-4,
    mixed boundaries, some samples far outside) against per-volume pp_pull_vjp calls, and
    graph replays against the eager run."""
    vols = _volumes(5)
    plan = ppb.BatchPlan("pull_vjp", [_dev(v) for v in vols])
    outs = plan.run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = pp.Settings(v["deg"], v["bnd"], v["ext"])
        df, dphi = pp.pull_vjp(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
        _check(H(o["out"]), H(dphi[0]), f"vjp dphi vol {i}")
        _check(H(o["out2"]), H(df[0]), f"vjp df vol {i}")
    first = [(H(o["out"]).copy(), H(o["out2"]).copy()) for o in outs]
    graph = ppb.Graph(lambda st: plan.run(st))
    for _ in range(3):
        graph.launch()
        graph.stream.synchronize()
        for i, ((a, b), o) in enumerate(zip(first, outs)):
            _check(H(o["out"]), a, f"vjp replay dphi vol {i}")
            _check(H(o["out2"]), b, f"vjp replay df vol {i}")
    graph.close()
    plan.close()


ANISO_DEGS = [(1, 3, 0), 2, (3, 1, 2), (0, 0, 2), 3, (4, 2, 1), (2, 2, 1), (1, 3, 0), 0, (3, 1, 2)]


def _aniso_volumes(seed):
    vols = _volumes(seed, n=len(ANISO_DEGS))
    for v, deg in zip(vols, ANISO_DEGS):
        v["deg"] = deg
    return vols


def test_batch_aniso_degrees_match_oracle():
    """Volumes with a degree per axis beside isotropic ones in one plan: grouped by their
    largest degree and by whether the degrees differ (each group one launch, the groups on
    streams of their own), pull, the pullback, push and push_grad against the FP64 oracle."""
    vols = _aniso_volumes(6)
    outs = ppb.BatchPlan("pull", [_dev(v) for v in vols]).run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        _check(H(o["out"]), oracle.pull(v["f"][None], v["phi"][None], s, prec="f64")[0], f"aniso pull vol {i}")
    outs = ppb.BatchPlan("pull_backward", [_dev(v) for v in vols]).run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = S(v["deg"], v["bnd"], v["ext"])
        ref = oracle.backward_phi(v["f"][None], v["phi"][None], v["aux"][None], s)[0]
        _check(H(o["out"]), ref, f"aniso pullback vol {i}")
    rng = np.random.default_rng(7)
    dev, dev3 = [], []
    for v in vols:
        v["dim"] = tuple(int(x) for x in np.array(v["d"]) + rng.integers(-3, 4, size=3))
        v["g3"] = rng.standard_normal((v["C"], 3) + v["d"]).astype(np.float32)
        common = dict(phi=T(v["phi"]), deg=v["deg"], bnd=v["bnd"], ext=v["ext"], dim=v["dim"])
        dev.append(dict(f=T(v["aux"]), **common))
        dev3.append(dict(f=T(v["g3"]), **common))
    outs = ppb.BatchPlan("push", dev).run()
    outs3 = ppb.BatchPlan("push_grad", dev3).run()
    torch.cuda.synchronize()
    for i, (v, o, o3) in enumerate(zip(vols, outs, outs3)):
        s = S(v["deg"], v["bnd"], v["ext"])
        _check(H(o["out"]), oracle.push(v["aux"][None], v["phi"][None], v["dim"], s, prec="f64")[0],
               f"aniso push vol {i}")
        ref = oracle.push_grad(v["g3"][None], v["phi"][None], v["dim"], s, prec="f64")[0]
        _check(H(o3["out"]), ref, f"aniso push_grad vol {i}")


@pytest.mark.parametrize("op", ["objective", "pull_vjp"])
def test_batch_aniso_degrees_match_per_volume_calls(op):
    """The fused objective and the pull VJP over volumes with a degree per axis against
    per-volume calls with the same Settings, eagerly, through pp_batch_run_static, and on
    graph replays."""
    vols = _aniso_volumes(7)
    plan = ppb.BatchPlan(op, [_dev(v) for v in vols])
    outs = plan.run()
    torch.cuda.synchronize()
    for i, (v, o) in enumerate(zip(vols, outs)):
        s = pp.Settings(v["deg"], v["bnd"], v["ext"])
        if op == "objective":
            loss, dphi, df = pp.objective_grad(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
            assert abs(float(o["loss"]) - float(loss)) <= TOL * abs(float(loss)), i
        else:
            df, dphi = pp.pull_vjp(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
        _check(H(o["out"]), H(dphi[0]), f"aniso {op} dphi vol {i}")
        _check(H(o["out2"]), H(df[0]), f"aniso {op} df vol {i}")
    first = [(H(o["out"]).copy(), H(o["out2"]).copy()) for o in outs]
    static = ppb.run_static(op, [_dev(v) for v in vols])
    torch.cuda.synchronize()
    for i, ((a, b), o) in enumerate(zip(first, static)):
        _check(H(o["out"]), a, f"aniso {op} static dphi vol {i}")
        _check(H(o["out2"]), b, f"aniso {op} static df vol {i}")
    graph = ppb.Graph(lambda st: plan.run(st))
    for rep in range(3):
        for o in plan.vols:
            o["out"].fill_(123.0)
            o["out2"].fill_(-7.0)
        torch.cuda.synchronize()
        graph.launch()
        graph.stream.synchronize()
        for i, ((a, b), o) in enumerate(zip(first, outs)):
            _check(H(o["out"]), a, f"aniso {op} replay {rep} dphi vol {i}")
            _check(H(o["out2"]), b, f"aniso {op} replay {rep} df vol {i}")
    graph.close()
    plan.close()


CHAN_BND = [[(2, 1, 1), (1, 2, 1), (1, 1, 2)], [(1, 2, 0), (2, 1, 0)], [(2, 2, 1), (2, 2, 1), (1, 1, 1), (2, 1, 2)]]
CHAN_DEGS = [1, 3, (1, 3, 0), 2, 0, (3, 1, 2)]


def _chan_bnd_volumes(seed):
    """Volumes whose channels take boundary triples of their own (the sliding boundaries of a
    displacement field among them), a few samples far outside the field of view."""
    rng = np.random.default_rng(seed)
    vols = []
    for i, deg in enumerate(CHAN_DEGS):
        bnd = CHAN_BND[i % len(CHAN_BND)]
        C = len(bnd)
        d = tuple(int(v) for v in rng.integers(16, 34, size=3))
        f = np.stack([smooth_field(rng, d, 1.5, 1.0)[0] for _ in range(C)]).astype(np.float32)
        phi = identity(d) + smooth_field(rng, d, 3.0, 4.0)
        flat = phi.reshape(3, -1)
        idx = rng.choice(flat.shape[1], size=flat.shape[1] // 25, replace=False)
        flat[1, idx] = rng.uniform(-15, d[1] + 15, size=idx.size).astype(np.float32)
        aux = rng.standard_normal((C,) + d).astype(np.float32)
        vols.append(dict(f=f, phi=phi, aux=aux, deg=deg, bnd=bnd, ext=int(i % 2 == 0), d=d, C=C))
    return vols


def test_batch_channel_boundaries_match_oracle():
    """Per-channel boundary triples in one plan (one fold per axis, per-channel signs):
    pull, the pullback, push and push_grad against the FP64 oracle."""
    vols = _chan_bnd_volumes(11)
    outs = ppb.BatchPlan("pull", [_dev(v) for v in vols]).run()
    outs_b = ppb.BatchPlan("pull_backward", [_dev(v) for v in vols]).run()
    rng = np.random.default_rng(12)
    dev, dev3 = [], []
    for v in vols:
        v["dim"] = tuple(int(x) for x in np.array(v["d"]) + rng.integers(-3, 4, size=3))
        v["g3"] = rng.standard_normal((v["C"], 3) + v["d"]).astype(np.float32)
        common = dict(phi=T(v["phi"]), deg=v["deg"], bnd=v["bnd"], ext=v["ext"], dim=v["dim"])
        dev.append(dict(f=T(v["aux"]), **common))
        dev3.append(dict(f=T(v["g3"]), **common))
    outs_p = ppb.BatchPlan("push", dev).run()
    outs_g = ppb.BatchPlan("push_grad", dev3).run()
    torch.cuda.synchronize()
    for i, v in enumerate(vols):
        s = S(v["deg"], v["bnd"], v["ext"])
        f, phi, aux = v["f"][None], v["phi"][None], v["aux"][None]
        _check(H(outs[i]["out"]), oracle.pull(f, phi, s, prec="f64")[0], f"chan-bnd pull vol {i}")
        _check(H(outs_b[i]["out"]), oracle.backward_phi(f, phi, aux, s)[0], f"chan-bnd pullback vol {i}")
        _check(H(outs_p[i]["out"]), oracle.push(aux, phi, v["dim"], s, prec="f64")[0], f"chan-bnd push vol {i}")
        ref = oracle.push_grad(v["g3"][None], phi, v["dim"], s, prec="f64")[0]
        _check(H(outs_g[i]["out"]), ref, f"chan-bnd push_grad vol {i}")


@pytest.mark.parametrize("op", ["objective", "pull_vjp"])
def test_batch_channel_boundaries_match_per_volume_calls(op):
    """The objective and the pull VJP over volumes with per-channel boundary triples against
    per-volume calls (which launch once per channel), eagerly and through
    pp_batch_run_static."""
    vols = _chan_bnd_volumes(13)
    outs = ppb.BatchPlan(op, [_dev(v) for v in vols]).run()
    static = ppb.run_static(op, [_dev(v) for v in vols])
    torch.cuda.synchronize()
    for i, (v, o, st) in enumerate(zip(vols, outs, static)):
        s = pp.Settings(v["deg"], v["bnd"], v["ext"])
        if op == "objective":
            loss, dphi, df = pp.objective_grad(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
            assert abs(float(o["loss"]) - float(loss)) <= TOL * abs(float(loss)), i
        else:
            df, dphi = pp.pull_vjp(T(v["f"][None]), T(v["phi"][None]), T(v["aux"][None]), s)
        for got, what in ((o, "plan"), (st, "static")):
            _check(H(got["out"]), H(dphi[0]), f"chan-bnd {op} {what} dphi vol {i}")
            _check(H(got["out2"]), H(df[0]), f"chan-bnd {op} {what} df vol {i}")


@pytest.mark.parametrize("vol_set", ["isotropic", "per-axis"])
def test_batch_objective_fused_step_equals_axpy(vol_set):
    """The objective's fused descent step (phi_next = phi + step * dL/dphi) against the
    objective followed by pp_axpy, bit for bit: into a separate buffer, in place, and over
    three graph-replayed iterations of a matching loop updating phi in place."""
    import ctypes as ct

    from pushpull import _lib

    vols = _volumes(21) if vol_set == "isotropic" else _aniso_volumes(22)
    step = -0.03
    L = _lib.lib()

    def same(a, b):  # identical values, NaN where the other has NaN (invalid samples, ext=0)
        na, nb = torch.isnan(a), torch.isnan(b)
        return torch.equal(na, nb) and torch.equal(a[~na], b[~nb])

    def axpy(y, x, a):
        _lib.check(L.pp_axpy(ct.c_void_p(y.data_ptr()), ct.c_void_p(x.data_ptr()), a, y.numel(),
                             ct.c_void_p(torch.cuda.current_stream().cuda_stream)), "axpy")

    # reference: the objective, then phi + step * dphi through pp_axpy
    ref_vols = [_dev(v) for v in vols]
    ref_outs = ppb.BatchPlan("objective", ref_vols).run()
    expect = [d["phi"].clone() for d in ref_vols]
    for e, o in zip(expect, ref_outs):
        axpy(e, o["out"], step)
    # into a separate buffer
    dev = [_dev(v) for v in vols]
    nxt = [torch.full_like(d["phi"], 7.0) for d in dev]
    for d, n in zip(dev, nxt):
        d.update(phi_next=n, step=step)
    outs = ppb.BatchPlan("objective", dev).run()
    torch.cuda.synchronize()
    for i, (n, e, o, r) in enumerate(zip(nxt, expect, outs, ref_outs)):
        assert same(o["out"], r["out"]), i
        assert same(n, e), f"vol {i}: fused step differs from pp_axpy"
    # in place, replayed: three iterations against objective + axpy iterations
    dev = [_dev(v) for v in vols]
    for d in dev:
        d.update(phi_next=d["phi"], step=step)
    plan = ppb.BatchPlan("objective", dev)
    ref_plan = ppb.BatchPlan("objective", ref_vols)
    ref_phi = [d["phi"] for d in ref_vols]
    for d, r in zip(ref_vols, dev):  # both start from the same deformation
        d["phi"].copy_(r["phi"])
    graph = ppb.Graph(lambda s: plan.run(s))
    for it in range(3):
        graph.launch()
        graph.stream.synchronize()
        r_outs = ref_plan.run()
        for p, o in zip(ref_phi, r_outs):
            axpy(p, o["out"], step)
        torch.cuda.synchronize()
        for i, (d, p) in enumerate(zip(dev, ref_phi)):
            assert same(d["phi"], p), f"iteration {it} vol {i}: in-place step differs"
    graph.close()
    plan.close()
    ref_plan.close()
