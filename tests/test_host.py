"""Host-side checks that need no GPU: Settings/id normalisation, the C ABI's exports,
descriptor layout and argument validation (which returns before any HIP call)."""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

import pushpull as pp
from pushpull import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- Settings (src/pushpull_general.jl:15-60) ------------------------------------
def test_settings_defaults():
    s = pp.Settings()
    assert s.deg == (1, 1, 1) and s.bnd == [(1, 1, 1)] and s.ext == 1


def test_settings_clamps_degree_and_ext():
    s = pp.Settings((7, -2, 3), 0, 5)
    assert s.deg == (4, 0, 3) and s.ext == 1
    assert pp.Settings(2, 1, -3).ext == 0
    assert pp.Settings(9).deg == (4, 4, 4)


def test_settings_boundary_forms():
    assert pp.Settings(1, 2).bnd == [(2, 2, 2)]
    assert pp.Settings(1, (0, 1, 2)).bnd == [(0, 1, 2)]
    assert pp.Settings(1, [(0, 1, 2), (2, 2, 1)]).bnd == [(0, 1, 2), (2, 2, 1)]
    m = np.array([[0, 1], [1, 2], [2, 0]])  # 3 x C matrix, one column per channel
    assert pp.Settings(1, m).bnd == [(0, 1, 2), (1, 2, 0)]


@pytest.mark.parametrize("bad", [3, -1, (0, 1, 3), [(0, 1, 2), (0, 5, 1)], np.array([[0], [1], [4]])])
def test_settings_rejects_bad_boundary(bad):
    with pytest.raises(AssertionError):
        pp.Settings(1, bad)


def test_identity_grid_is_one_based():
    phi = pp.id(dim=(4, 3, 2))  # Julia (X, Y, Z)
    assert phi.shape == (3, 2, 3, 4)
    assert phi[0, 0, 0, :].tolist() == [1, 2, 3, 4]
    assert phi[1, 0, :, 0].tolist() == [1, 2, 3]
    assert phi[2, :, 0, 0].tolist() == [1, 2]


def test_identity_grid_affine():
    A = np.array([[2.0, 0, 0, 1], [0, 1, 0, -1], [0, 0.5, 1, 0]])
    phi = pp.id(dim=(3, 2, 2), affine=A)
    # x = 2*i + 1 at (i,j,k) = (3,1,1); z = 0.5*j + k at j=2, k=2
    assert phi[0, 0, 0, 2] == 7.0
    assert phi[2, 1, 1, 0] == 3.0


# ---- the C ABI --------------------------------------------------------------------
def test_library_exports_every_declared_function():
    names = _lib.declared_functions()
    assert {"pull", "pullg", "pullh", "push", "pushg", "pp_pull", "pp_push", "pp_last_error"} <= set(names)
    L = ct.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_desc_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "pushpull_hip.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(pp_desc), offsetof(pp_desc, nchan),'
                   ' offsetof(pp_desc, nbnd), offsetof(pp_desc, bnd)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    D = _lib.PPDesc
    assert [int(v) for v in out] == [ct.sizeof(D), D.nchan.offset, D.nbnd.offset, D.bnd.offset]


def test_volume_layout_matches_header(tmp_path):
    """pp_volume (a batch's volumes, the degree a triple per volume) as the ctypes mirror
    lays it out: every field's offset and size, and the struct's size."""
    from pushpull.batch import PPVolume

    fields = [f[0] for f in PPVolume._fields_]
    body = "".join(f'printf("%zu %zu\\n", offsetof(pp_volume, {n}), sizeof(((pp_volume*)0)->{n}));' for n in fields)
    src = tmp_path / "vol.c"
    src.write_text('#include "pushpull_hip.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   f'int main(void){{{body} printf("%zu 0\\n", sizeof(pp_volume)); return 0;}}\n')
    exe = tmp_path / "vol"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(len(fields) + 1)]
    want = [(getattr(PPVolume, n).offset, getattr(PPVolume, n).size) for n in fields] + [(ct.sizeof(PPVolume), 0)]
    assert got == want
    assert getattr(PPVolume, "deg").size == 12
    src.write_text('#include "pushpull_hip.h"\n#include <stdio.h>\nint main(void){printf("%d\\n", PP_BATCH_CHAN_BND); return 0;}\n')
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    from pushpull.batch import BATCH_CHAN_BND
    assert int(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout) == BATCH_CHAN_BND


def _desc(**kw):
    d = _lib.PPDesc()
    for a in range(3):
        d.d0[a] = 8
        d.d1[a] = 8
        d.deg[a] = 1
    d.nchan, d.nbatch, d.ext, d.nbnd = 1, 1, 1, 1
    for k, v in kw.items():
        if k in ("d0", "d1", "deg"):
            for a in range(3):
                getattr(d, k)[a] = v[a]
        elif k == "bnd":
            for i, b in enumerate(v):
                d.bnd[i] = b
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,code", [
    (dict(deg=(1, 5, 1)), -1),
    (dict(deg=(-1, 1, 1)), -1),
    (dict(d0=(0, 8, 8)), -1),
    (dict(d1=(8, 8, 0)), -1),
    (dict(nchan=0), -1),
    (dict(nbatch=0), -1),
    (dict(nchan=3, nbnd=2), -1),
    (dict(bnd=(0, 1, 3)), -1),
    (dict(d0=(1 << 30, 8, 8)), -4),
    (dict(d1=(4096, 4096, 256)), -4),
])
def test_descriptor_validation(kw, code):
    L = _lib.lib()
    d = _desc(**kw)
    p = ct.c_void_p(16)  # never dereferenced: validation fails first
    assert L.pp_pull(ct.byref(d), p, p, p, None) == code
    assert L.pp_last_status() == code
    assert len(_lib.last_error()) > 0


def test_null_arrays_rejected():
    L = _lib.lib()
    d = _desc()
    assert L.pp_pull(ct.byref(d), None, None, None, None) == -1
    assert "null" in _lib.last_error()


def test_tier_a_rejects_bad_dp_without_touching_outputs_twice():
    L = _lib.lib()
    out = np.zeros(4, np.float32)
    phi = np.ones(12, np.float32)
    f0 = np.ones(8, np.float32)
    d0 = np.array([2, 2, 2], np.uint64)
    dp = np.array([6, 2, 2], np.uint64)  # dp outside 1..5
    bnd = np.array([1, 1, 1], np.int32)
    P = lambda a: ct.c_void_p(a.ctypes.data)  # noqa: E731
    L.pull(P(out), P(phi), P(f0), P(d0), 4, P(bnd), P(dp), 1)
    assert L.pp_last_status() == -1
    assert np.all(np.isnan(out))


def test_shape_assertions_mirror_julia():
    s = pp.Settings()
    f = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(AssertionError):
        pp.pull(f, np.zeros((2, 4, 4, 4), np.float32), s)  # phi needs 3 channels
    with pytest.raises(AssertionError):
        pp.pull(np.zeros((4, 4), np.float32), np.zeros((3, 4, 4, 4), np.float32), s)
    with pytest.raises(AssertionError):
        pp.push(np.zeros((5, 4, 4), np.float32), np.zeros((3, 4, 4, 4), np.float32), (4, 4, 4), s)
    with pytest.raises(AssertionError):
        pp.pull(np.zeros((2, 1, 4, 4, 4), np.float32), np.zeros((3, 3, 4, 4, 4), np.float32), s)
    with pytest.raises(AssertionError):
        pp.pull(np.zeros((3, 4, 4, 4), np.float32), np.zeros((3, 4, 4, 4), np.float32),
                pp.Settings(1, [(1, 1, 1), (0, 0, 0)]))
    with pytest.raises(TypeError):
        pp.pull(f.astype(np.float64), np.zeros((3, 4, 4, 4), np.float32), s)


def test_tune_table_file_format(tmp_path):
    L = _lib.lib()
    p = tmp_path / "t.txt"
    p.write_text("# op dp log2 variant\n0 4 24 2\n3 2 20 0\n")
    assert L.pp_tune_clear() == 0
    assert L.pp_tune_load(str(p).encode()) == 0
    q = tmp_path / "u.txt"
    assert L.pp_tune_save(str(q).encode()) == 0
    rows = sorted(l.split() for l in q.read_text().splitlines() if not l.startswith("#"))
    assert rows == [["0", "4", "24", "2", "0"], ["3", "2", "20", "0", "0"]]
    # five columns: the box budget of a tiled choice
    p.write_text("7 4 24 1 10184\n")
    assert L.pp_tune_load(str(p).encode()) == 0 and L.pp_tune_save(str(q).encode()) == 0
    assert [l.split() for l in q.read_text().splitlines() if l.startswith("7")] == [["7", "4", "24", "1", "10184"]]
    p.write_text("7 4 24 1 100\n")  # a budget below 256 floats
    assert L.pp_tune_load(str(p).encode()) == -1
    bad = tmp_path / "bad.txt"
    bad.write_text("0 4 24 9\n")
    assert L.pp_tune_load(str(bad).encode()) == -1 and "malformed" in _lib.last_error()
    assert L.pp_tune_clear() == 0


def test_batch_plan_validation_without_gpu():
    import ctypes as ct
    from pushpull import batch as ppb

    L = ppb._sigs()
    h = ct.c_void_p()
    assert L.pp_batch_plan_create(0, None, 0, ct.byref(h)) == -1
    v = (ppb.PPVolume * 1)()
    v[0].nchan = 1
    for q in range(3):
        v[0].d0[q] = v[0].d1[q] = 4
        v[0].deg[q] = 1
    v[0].deg[1] = 7  # one axis out of range
    v[0].f = v[0].phi = v[0].out = 16
    assert L.pp_batch_plan_create(0, ct.cast(v, ct.c_void_p), 1, ct.byref(h)) == -1
    assert "degree 7 outside 0..4 on axis 1" in _lib.last_error()
    assert L.pp_batch_plan_create(9, ct.cast(v, ct.c_void_p), 1, ct.byref(h)) == -1
    v[0].deg[1] = 1
    v[0].phi_next = 16  # a fused descent step on an op other than the objective
    assert L.pp_batch_plan_create(0, ct.cast(v, ct.c_void_p), 1, ct.byref(h)) == -1
    assert "phi_next (a fused descent step) is only for PP_OP_OBJECTIVE" in _lib.last_error()


@pytest.mark.parametrize("nbnd,codes,msg", [
    (3, [(1, 1, 1)] * 3, "nbnd must be 0, 1 or nchan (2), got 3"),
    (2, [(1, 1, 1), (1, 3, 1)], "boundary code 3"),
    (2, [(0, 1, 1), (1, 1, 1)], "along axis 0 differ other than between 1 and 2"),
    (2, [(1, 1, 2), (1, 0, 2)], "along axis 1 differ other than between 1 and 2"),
])
def test_batch_channel_boundaries_validated_without_gpu(nbnd, codes, msg):
    """A volume's per-channel boundary triples: their count, codes, and the rule that the
    channels' codes along an axis agree or differ only between reflect and antisymmetric
    (checked before any device call)."""
    from pushpull import batch as ppb

    L = ppb._sigs()
    h = ct.c_void_p()
    v = (ppb.PPVolume * 1)()
    v[0].nchan, v[0].nbnd = 2, nbnd
    for q in range(3):
        v[0].d0[q] = v[0].d1[q] = 4
        v[0].deg[q] = 1
        for c, t in enumerate(codes):
            v[0].chan_bnd[3 * c + q] = t[q]
    v[0].f = v[0].phi = v[0].out = 16
    assert L.pp_batch_plan_create(0, ct.cast(v, ct.c_void_p), 1, ct.byref(h)) == -1
    assert msg in _lib.last_error(), _lib.last_error()


def _cpu_vol(op, C=2, d1=(5, 6, 7), d0=(4, 5, 6), **over):
    import torch

    v = dict(phi=torch.zeros((3,) + d1), deg=3, bnd=(1, 1, 1), ext=1)
    if op == "push":
        v.update(f=torch.zeros((C,) + d1), dim=d0)
    elif op == "push_grad":
        v.update(f=torch.zeros((C, 3) + d1), dim=d0)
    else:
        v.update(f=torch.zeros((C,) + d0), aux=torch.zeros((C,) + d1))
    v.update(over)
    return v


@pytest.mark.parametrize("op,over,exc,msg", [
    ("pull", dict(phi=None), ValueError, "required"),
    ("pull", dict(deg=5), ValueError, "degree"),
    ("pull", dict(deg=(1, 2)), ValueError, "three degrees"),
    ("pull", dict(deg=(1, 5, 0)), ValueError, "three degrees"),
    ("pull", dict(bnd=(1, 3, 1)), ValueError, "bnd"),
    ("pull", dict(bnd=[(1, 1, 1)] * 3), ValueError, "3 triples for 2 channels"),
    ("pull", dict(phi_next="phi"), ValueError, "only for the objective"),
    ("objective", dict(phi_next="small"), ValueError, "`phi_next` must be"),
    ("pull", dict(out=np.float32), TypeError, "torch tensor"),
    ("push", dict(dim=(4, 5)), ValueError, "dim"),
    ("push", dict(f="grid"), ValueError, "phi's grid"),
    ("push_grad", dict(f="no3"), ValueError, "(C, 3, Z1, Y1, X1)"),
    ("objective", dict(aux=None), ValueError, "`aux` is required"),
    ("objective", dict(aux="grid"), ValueError, "`aux` must be"),
    ("pull", dict(out="grid"), ValueError, "`out` must be"),
    ("push", dict(out="grid"), ValueError, "`out` must be"),
    ("pull_vjp", dict(out2="grid"), ValueError, "`out2` must be"),
    ("objective", dict(loss="shape"), ValueError, "`loss` must hold one value"),
    ("objective", dict(loss="f32"), TypeError, "torch.float64"),
    ("pull", dict(f="f64"), TypeError, "torch.float32"),
    ("pull", dict(), TypeError, "device tensors"),
])
def test_batch_volume_validation(op, over, exc, msg):
    """Every per-volume tensor is checked before a pointer reaches the library: shapes,
    then dtypes, then devices (the last check is the one a CPU tensor fails here)."""
    import torch
    from pushpull import batch as ppb

    d1, d0, C = (5, 6, 7), (4, 5, 6), 2
    subs = {
        "grid": lambda k: torch.zeros((C,) + tuple(x + 1 for x in d1)),
        "no3": lambda k: torch.zeros((C, 2) + d1),
        "shape": lambda k: torch.zeros(2, dtype=torch.float64),
        "f32": lambda k: torch.zeros((), dtype=torch.float32),
        "f64": lambda k: torch.zeros((C,) + d0, dtype=torch.float64),
        "phi": lambda k: torch.zeros((3,) + d1),
        "small": lambda k: torch.zeros((3,) + d0),
    }
    over = {k: (subs[v](k) if isinstance(v, str) and v in subs else v) for k, v in over.items()}
    if "out" in over and over["out"] is np.float32:
        over["out"] = np.zeros((C,) + d1, np.float32)
    v = _cpu_vol(op, C, d1, d0, **over)
    if v.get("phi") is None:
        v.pop("phi")
    if v.get("aux") is None:
        v.pop("aux", None)
    with pytest.raises(exc, match=msg.replace("(", r"\(").replace(")", r"\)")):
        ppb._volume_records(op, [v])


@pytest.mark.parametrize("fn", ["pull", "pull_grad", "push"])
def test_autograd_refuses_cpu_tensors_that_require_grad(fn):
    import torch

    d = (4, 5, 6)
    phi = torch.from_numpy(pp.id(dim=d[::-1])).requires_grad_(True)
    f = torch.zeros((1,) + d)
    s = pp.Settings(1, 1, 1)
    with pytest.raises(TypeError, match="device tensors"):
        if fn == "push":
            pp.push(f, phi, d, s)
        else:
            getattr(pp, fn)(f, phi, s)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits at once with a message (no device
    touched, so this runs without a GPU)."""
    import json
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
