"""Write tests/golden/pushpull_golden.npz from the reference C compiled from source.

Run in a container that has the PushPull.jl tree:

    make -C oracle ref            # -> oracle/_ref/libpushpull_ref.so from artifacts/C/pushpull.c
    python tests/golden/make_golden.py

Every case stores its inputs (f0, phi, f1, g1 and the settings) and the outputs of the
reference's ``pull``, ``pullg``, ``pullh``, ``push`` and ``pushg`` for one slice. The
reference ``pushg`` output is kept to pin its inverted finiteness test
(artifacts/C/pushpull.c:144): it is all zeros for finite input.

Inputs are seeded (numpy PCG64); the reference's own test data (test/runtests.jl:44-50:
8x7x1 grids, 2 channels, phi = id + randn, Settings((2,2,2),(0,1,2),1) and
((3,3,3),(0,1,2),1)) is unseeded, so those cases are regenerated here with a seed.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402


class _S:
    def __init__(self, deg, bnd, ext):
        self.deg = tuple(deg)
        self.bnd = [tuple(bnd)]
        self.ext = ext


def _identity(d1):
    Z, Y, X = d1
    phi = np.empty((3, Z, Y, X), np.float32)
    phi[0] = np.arange(1, X + 1)[None, None, :]
    phi[1] = np.arange(1, Y + 1)[None, :, None]
    phi[2] = np.arange(1, Z + 1)[:, None, None]
    return phi


def _specials(phi, d0, rng):
    """Overwrite some samples with the edge cases the kernels must reproduce."""
    flat = phi.reshape(3, -1)
    n = flat.shape[1]
    Z0, Y0, X0 = d0
    dims = (X0, Y0, Z0)
    vals = []
    for a in range(3):
        d = dims[a]
        # VOXOK thresholds (x = phi-1 compared with -0.01 and d-0.99, in FP32)
        vals += [(a, np.float32(1.0) - np.float32(0.01)), (a, np.nextafter(np.float32(0.99), np.float32(0))),
                 (a, np.float32(d) + np.float32(0.01)), (a, np.nextafter(np.float32(d) + np.float32(0.01), np.float32(1e9)))]
        # integer and half-integer knots
        vals += [(a, np.float32(2.0)), (a, np.float32(2.5)), (a, np.float32(d)), (a, np.float32(0.5))]
        # far outside: fold several periods; beyond 2^20 (large-coordinate path); int64 overflow
        vals += [(a, np.float32(-3.7 * d)), (a, np.float32(7.2 * d + 0.3)), (a, np.float32(3e6 + 0.25)),
                 (a, np.float32(-5e7)), (a, np.float32(1e19)), (a, np.float32(-2e19)), (a, np.float32(-0.49)),
                 (a, np.float32(-1.51))]
        # non-finite
        vals += [(a, np.float32(np.nan)), (a, np.float32(np.inf)), (a, np.float32(-np.inf))]
    idx = rng.choice(n, size=min(n, len(vals)), replace=False)
    for k, (a, v) in zip(idx, vals):
        flat[a, k] = v
    return phi


def _case(rng, d0, d1, deg, bnd, ext, sigma=1.5, specials=True, smooth=False):
    f0 = rng.standard_normal(d0).astype(np.float32)
    phi = _identity(d1)
    if smooth:
        u = rng.standard_normal((3, 1, 1, 1)) * 2.0
        Z, Y, X = d1
        zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        for a in range(3):
            phi[a] += (u[a] * np.sin(0.3 * xx + a) * np.cos(0.2 * yy - a) + 0.5 * np.sin(0.25 * zz)).astype(np.float32)
    else:
        phi += (sigma * rng.standard_normal(phi.shape)).astype(np.float32)
    if specials:
        phi = _specials(phi, d0, rng)
    f1 = rng.standard_normal(d1).astype(np.float32)
    f1.reshape(-1)[rng.choice(f1.size, size=max(1, f1.size // 40), replace=False)] = np.nan
    g1 = rng.standard_normal((3,) + tuple(d1)).astype(np.float32)
    # Samples with a coordinate beyond 2^20 get huge weights (up to ~1e19 once the knot
    # saturates at INT64_MIN) whose scattered contributions cancel only in the
    # reference's sequential order; any unordered (atomic) scatter differs. Those samples
    # push the value 0: their knots, folds and addresses are still exercised.
    x = phi.reshape(3, -1) - np.float32(1.0)
    extreme = np.any(np.isfinite(x) & (np.abs(x) >= 2.0 ** 20), axis=0)
    f1.reshape(-1)[extreme] = 0.0
    g1.reshape(3, -1)[:, extreme] = 0.0
    return dict(f0=f0, phi=phi, f1=f1, g1=g1, deg=np.array(deg, np.int32), bnd=np.array(bnd, np.int32),
                ext=np.array(ext, np.int32))


def _outputs(c):
    s = _S(c["deg"], c["bnd"], int(c["ext"]))
    d0 = c["f0"].shape
    out = {}
    out["pull"] = oracle.pull(c["f0"], c["phi"], s, prec="ref")[0, 0]
    out["pullg"] = oracle.pull_grad(c["f0"], c["phi"], s, prec="ref")[0, 0]
    out["pullh"] = oracle.pull_hess(c["f0"], c["phi"], s, prec="ref")[0, 0]
    out["push"] = oracle.push(c["f1"], c["phi"], d0, s, prec="ref")[0, 0]
    out["pushg_ref"] = oracle.push_grad(c["g1"], c["phi"], d0, s, prec="ref")[0, 0]
    return out


def cases():
    rng = np.random.default_rng(20240601)
    cs = []
    # 1. every isotropic degree x uniform boundary x ext
    for deg in range(5):
        for b in range(3):
            for ext in (0, 1):
                cs.append(_case(rng, (5, 6, 7), (4, 5, 6), (deg,) * 3, (b,) * 3, ext))
    # 2. all 27 per-axis boundary combinations (degree 2 and 3)
    for b0 in range(3):
        for b1 in range(3):
            for b2 in range(3):
                cs.append(_case(rng, (4, 5, 6), (3, 4, 5), (2, 3, 2) if (b0 + b1 + b2) % 2 else (3, 3, 3),
                                (b0, b1, b2), 1))
    # 3. anisotropic degrees
    for deg in [(1, 2, 3), (0, 4, 2), (3, 1, 0), (4, 4, 1), (2, 0, 4)]:
        cs.append(_case(rng, (6, 5, 7), (5, 5, 5), deg, (1, 0, 2), 1))
    # 4. tiny source grids (several folds per tap): d in {1, 2, 3}
    for d0 in [(1, 1, 1), (2, 1, 3), (3, 2, 1), (1, 3, 2)]:
        for deg in (1, 3, 4):
            for b in range(3):
                cs.append(_case(rng, d0, (3, 3, 3), (deg,) * 3, (b, (b + 1) % 3, (b + 2) % 3), 1, sigma=3.0))
    # 5. the reference test's geometry: 8x7x1 grid (array order (1, 7, 8)), phi = id + randn
    for deg in (2, 3):
        cs.append(_case(rng, (1, 7, 8), (1, 7, 8), (deg,) * 3, (0, 1, 2), 1, sigma=1.0, specials=False))
    # 6. larger smooth cases
    for deg in (1, 2, 3):
        cs.append(_case(rng, (16, 16, 16), (16, 16, 16), (deg,) * 3, (1, 1, 1), 0, smooth=True, specials=False))
    cs.append(_case(rng, (12, 14, 16), (10, 12, 18), (3, 3, 3), (0, 0, 0), 1, smooth=True, specials=False))
    return cs


def affine_cases():
    """Affine pull/push (pushpull.c:176-244): rotations, shears, scalings and shifts of
    small grids, every degree, mixed boundaries, ext 0/1."""
    rng = np.random.default_rng(20240602)
    cs = []
    for i in range(20):
        d0 = tuple(int(v) for v in rng.integers(3, 10, size=3))
        d1 = tuple(int(v) for v in rng.integers(3, 10, size=3))
        deg = int(i % 5)
        th = rng.uniform(-0.6, 0.6)
        R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        M = R @ np.diag(rng.uniform(0.6, 1.6, size=3)) + rng.uniform(-0.2, 0.2, size=(3, 3))
        Aff = np.zeros((4, 4))
        Aff[:3, :3] = M
        Aff[:3, 3] = rng.uniform(-4, 4, size=3)
        Aff[3, 3] = 1
        A0 = oracle.adjust_affine(Aff.astype(np.float32))
        bnd = tuple(int(v) for v in rng.integers(0, 3, size=3))
        ext = int(i % 2)
        s = _S((deg,) * 3, bnd, ext)
        f0 = rng.standard_normal(d0).astype(np.float32)
        f1 = rng.standard_normal(d1).astype(np.float32)
        f1.reshape(-1)[rng.choice(f1.size, size=2, replace=False)] = np.nan
        cs.append(dict(aff=Aff.astype(np.float32), aff0=A0, f0=f0, f1=f1, d1=np.array(d1, np.int64),
                       deg=np.array((deg,) * 3, np.int32), bnd=np.array(bnd, np.int32), ext=np.array(ext, np.int32),
                       pull_affine=oracle.affine_pull(f0, A0, d1, s, prec="ref")[0, 0],
                       push_affine=oracle.affine_push(f1, A0, d0, s, prec="ref")[0, 0]))
    return cs


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libpushpull_ref.so is missing: run `make -C oracle ref` first")
    arrays = {}
    manifest = []
    for i, c in enumerate(cases()):
        out = _outputs(c)
        for k, v in {**c, **out}.items():
            arrays[f"c{i:03d}_{k}"] = np.asarray(v)
        manifest.append(dict(id=i, d0=list(c["f0"].shape), d1=list(c["phi"].shape[1:]), deg=c["deg"].tolist(),
                             bnd=c["bnd"].tolist(), ext=int(c["ext"])))
    for j, c in enumerate(affine_cases()):
        for k, v in c.items():
            arrays[f"a{j:03d}_{k}"] = np.asarray(v)
    path = os.path.join(HERE, "pushpull_golden.npz")
    np.savez_compressed(path, **arrays)
    with open(os.path.join(HERE, "pushpull_golden.json"), "w") as fh:
        json.dump(dict(source="reference artifacts/C/pushpull.c compiled by oracle/Makefile (gcc -O3)",
                       layout="C-order (Z,Y,X); phi (3,Z,Y,X) 1-based; pullg (3,...); pullh (9,...)",
                       cases=manifest), fh, indent=1)
    print(f"wrote {len(manifest)} cases -> {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
