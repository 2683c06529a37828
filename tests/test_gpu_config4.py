"""The batching workload at its configured size (needs a GPU): 24 volumes of uneven sizes
drawn from 96^3 up to 224 x 256 x 192, 1-4 channels, degree 1 or 3 at random, each with its
own smooth deformation (max 5 voxels) -- bench.py's seeded generator (make_c4_volumes). The
fused objective L = 0.5 ||pull(f, phi) - g||^2 with both gradients runs for all volumes as
one batched launch per degree, captured in a HIP graph and replayed, and is checked
  * replay against replay (dL/dphi is gathered per sample: bit-identical; dL/df and the
    loss are summed with atomics: to rounding),
  * against the per-volume loop of pp_objective_grad (the same tile bodies: dL/dphi
    bit-identical),
  * against the FP64 oracle, per volume: dL/dphi on an 8-slice slab of the deformation,
    dL/df on an 8-slice slab of the image (pushed from every sample that reaches it) and
    the loss over the whole volume, norm-wise relative error <= 1e-5.
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
from pushpull import batch as ppb  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = 1e-5
NZ = 8  # slab thickness


def H(t):
    return t.detach().cpu().numpy()


def _rel(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.isfinite(ref)
    return float(np.linalg.norm(got[m] - ref[m]) / np.linalg.norm(ref[m]))


@pytest.fixture(scope="module")
def c4():
    vols, descs = bench.make_c4_volumes(torch, pp, seed=4)
    plan = ppb.BatchPlan("objective", vols)
    graph = ppb.Graph(lambda s: plan.run(s))
    runs = []
    for _ in range(3):
        graph.launch()
        graph.stream.synchronize()
        runs.append([(H(v["out"]), H(v["out2"]), float(v["loss"])) for v in plan.vols])
    graph.close()
    plan.close()
    loop = []
    for v in vols:
        s = pp.Settings(v["deg"], 1, 1)
        loss, dphi, df = pp.objective_grad(v["f"], v["phi"], v["aux"], s)
        torch.cuda.synchronize()
        loop.append((H(dphi), H(df), float(loss)))
    host = [dict(f=H(v["f"]), g=H(v["aux"]), phi=H(v["phi"]), deg=v["deg"]) for v in vols]
    return runs, loop, host


def test_config4_graph_replays_agree(c4):
    runs, _, _ = c4
    for r in runs[1:]:
        for (a_phi, a_f, a_l), (b_phi, b_f, b_l) in zip(runs[0], r):
            assert np.array_equal(a_phi, b_phi, equal_nan=True)
            assert _rel(b_f, a_f) < 1e-6
            assert abs(a_l - b_l) <= 1e-9 * abs(a_l)


def test_config4_batched_matches_per_volume_loop(c4):
    runs, loop, _ = c4
    for i, ((a_phi, a_f, a_l), (b_phi, b_f, b_l)) in enumerate(zip(runs[0], loop)):
        assert np.array_equal(a_phi, b_phi, equal_nan=True), i
        assert _rel(a_f, b_f) < 1e-6, i
        assert abs(a_l - b_l) <= 1e-9 * abs(b_l), i


def _oracle_volume(args):
    """FP64 references of one volume: (dL/dphi slab, dL/df slab, loss) and the slab bounds."""
    h, (z0, z1) = args
    f, g, phi, deg = h["f"], h["g"], h["phi"], h["deg"]
    s = S((deg,) * 3, (1, 1, 1), 1)
    C, Z = f.shape[0], f.shape[1]
    # dL/dphi at the slab of samples
    ph = np.ascontiguousarray(phi[:, z0:z1])[None]
    r = oracle.pull(f[None], ph, s, "f64")[0] - g[:, z0:z1]
    dphi = oracle.backward_phi(f[None], ph, r[None], s)[0]
    # dL/df on the image slab [z0, z1): every sample whose footprint may reach it
    zc = phi[2] - 1.0  # 0-based z coordinate
    sel = (zc >= z0 - deg - 2) & (zc <= z1 + deg + 1)
    psel = np.ascontiguousarray(phi[:, sel])[None, :, None, None, :]
    rsel = oracle.pull(f[None], psel, s, "f64")[0] - g[:, sel][:, None, None, :]
    df = oracle.push(rsel[None], psel, f.shape[1:], s, "f64")[0][:, z0:z1]
    # the loss over the whole volume
    rall = oracle.pull(f[None], phi[None], s, "f64")[0] - g
    loss = 0.5 * float((rall[np.isfinite(rall)] ** 2).sum())
    return dphi, df, loss


def test_config4_against_fp64_oracle(c4):
    runs, _, host = c4
    jobs = []
    for h in host:
        Z = h["f"].shape[1]
        z0 = Z // 2 - NZ // 2
        jobs.append((h, (z0, z0 + NZ)))
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        refs = list(ex.map(_oracle_volume, jobs))
    worst = {"dphi": 0.0, "df": 0.0, "loss": 0.0}
    for (got_phi, got_f, got_l), (rphi, rf, rl), (_, (z0, z1)) in zip(runs[0], refs, jobs):
        e = {"dphi": _rel(got_phi[:, z0:z1], rphi), "df": _rel(got_f[:, z0:z1], rf),
             "loss": abs(got_l - rl) / abs(rl)}
        for k, v in e.items():
            worst[k] = max(worst[k], v)
            assert v <= TOL, (k, v)
    for k, v in worst.items():
        report(f"config4 24 volumes {k} vs FP64", v)
