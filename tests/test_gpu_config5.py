"""The matching-loop workload at its configured size (needs a GPU): 8 subjects at 256^3, 4
channels, degree 3 (bench.py's seeded make_c5_inputs), 50 iterations of the fused objective
and gradient (pull, residual, push, phi pullback) for all subjects as one batched launch,
with phi <- phi - alpha dL/dphi, each iteration replayed from one HIP graph. At iterations 1
and 50 the FP64 oracle re-evaluates dL/dphi and dL/df of two subjects from the GPU's own
iteration inputs (phi as it stood before that iteration): dL/dphi on an 8-slice slab of the
deformation, dL/df on an 8-slice slab of the image pushed from every sample that reaches it;
norm-wise relative error <= 1e-5.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from helpers import S, report

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull import _lib  # noqa: E402
from pushpull import batch as ppb  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-5
ITERS, ALPHA, DEG = 50, 0.02, 3
CHECK = (0, 5)   # subjects re-evaluated
Z0, NZ = 124, 8  # slabs


def H(t):
    return t.detach().cpu().numpy()


def _rel(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def c5():
    import ctypes as ct

    B, C = 8, 4
    f, g, phi = bench.make_c5_inputs(torch, pp, seed=5, B=B, C=C)
    dphi, df = torch.empty_like(phi), torch.empty_like(f)
    loss = torch.zeros(B, dtype=torch.float64, device=f.device)
    vols = [dict(f=f[b], phi=phi[b], aux=g[b], deg=DEG, bnd=(1, 1, 1), ext=1, out=dphi[b], out2=df[b],
                 loss=loss[b]) for b in range(B)]
    plan = ppb.BatchPlan("objective", vols)
    L = _lib.lib()

    def iteration(s):
        plan.run(s)
        _lib.check(L.pp_axpy(ct.c_void_p(phi.data_ptr()), ct.c_void_p(dphi.data_ptr()), -ALPHA, phi.numel(),
                             ct.c_void_p(s.cuda_stream)), "axpy")

    graph = ppb.Graph(iteration)
    st = graph.stream
    snaps, losses = {}, []
    with torch.cuda.stream(st):
        for it in range(1, ITERS + 1):
            keep = it in (1, ITERS)
            if keep:  # the GPU's own inputs of this iteration
                st.synchronize()
                phi_in = {b: H(phi[b]) for b in CHECK}
            graph.launch()
            if keep:
                st.synchronize()
                snaps[it] = {b: dict(phi=phi_in[b], dphi=H(dphi[b]), df=H(df[b])) for b in CHECK}
                losses.append(float(loss.sum()))
        st.synchronize()
    graph.close()
    plan.close()
    host = {b: dict(f=H(f[b]), g=H(g[b])) for b in CHECK}
    return snaps, host, losses


def _oracle(args):
    it, b, snap, h = args
    s = S((DEG,) * 3, (1, 1, 1), 1)
    f, g, phi = h["f"], h["g"], snap["phi"]
    ph = np.ascontiguousarray(phi[:, Z0:Z0 + NZ])[None]
    r = oracle.pull(f[None], ph, s, "f64")[0] - g[:, Z0:Z0 + NZ]
    dphi = oracle.backward_phi(f[None], ph, r[None], s)[0]
    zc = phi[2] - 1.0
    sel = (zc >= Z0 - DEG - 2) & (zc <= Z0 + NZ + DEG + 1)
    psel = np.ascontiguousarray(phi[:, sel])[None, :, None, None, :]
    rsel = oracle.pull(f[None], psel, s, "f64")[0] - g[:, sel][:, None, None, :]
    df = oracle.push(rsel[None], psel, f.shape[1:], s, "f64")[0][:, Z0:Z0 + NZ]
    return (it, b, _rel(snap["dphi"][:, Z0:Z0 + NZ], dphi), _rel(snap["df"][:, Z0:Z0 + NZ], df))


def test_config5_matching_loop_against_fp64_oracle(c5):
    snaps, host, losses = c5
    jobs = [(it, b, snaps[it][b], host[b]) for it in snaps for b in CHECK]
    with ThreadPoolExecutor(len(jobs)) as ex:
        res = list(ex.map(_oracle, jobs))
    for it, b, e_phi, e_f in res:
        report(f"config5 iteration {it} subject {b} dphi vs FP64", e_phi)
        report(f"config5 iteration {it} subject {b} df vs FP64", e_f)
        assert e_phi <= TOL and e_f <= TOL, (it, b, e_phi, e_f)
    # the deformation update lowers the loss
    assert losses[-1] < losses[0], losses
