"""Benchmark: pull + push + backward at 256^3 FP32, degrees 1-3, on one MI355X.

One step = for each degree d in (1, 2, 3), on a 256^3 single-channel volume with a
smooth deformation (identity + displacement of max 6 voxels, reflect boundary, ext=1):
    f1   = pull(f, phi)                 (12 + 8C bytes/voxel algorithmic)
    df   = push(delta, phi)             (12 + 8C)
    dphi = sum_c delta_c grad pull(f)_c (24 + 8C)   -- the phi pullback
i.e. 48 + 24C = 72 bytes/voxel of compulsory HBM traffic per degree. Inputs are resident
in HBM before timing. value = 3 * 256^3 voxels / step time, in Gvoxel/s.

push(delta) and the pullback are the two cotangents of pull's rrule (src/pushpull_pb.jl:11-17):
by default (--schedule vjp) they come from one launch (pp_pull_vjp, which stages each tile of
f once and reads phi and delta once for both); the six independent launches of a step (pull
and the VJP per degree) run on three HIP streams balanced by their measured times, joined at
the end of each step. --schedule overlap runs the three
separate entry points (pull and push on the launch stream, the pullback beside them);
--schedule serial the nine separate launches in order. The per-op breakdown (separate
entry points, and the VJP) comes from serial steps run before the warm-up, which also keep
the GPU busy up to the timed steps.

The CPU baseline is the reference's own C (artifacts/C/pushpull.c, compiled from source
into oracle/_ref by oracle/Makefile, single-threaded as shipped) timed on this host on a
z-slab of the same workload; when that build is absent the project's FP32 restatement
of it (oracle/_build) is timed instead.
"""
from __future__ import annotations

import argparse
import subprocess
import ctypes as ct
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pushpull.jl_amd"))

N = 256
DEGS = (1, 2, 3)
HBM_PEAK_GBS = 8000.0
# per voxel at C = 1; the VJP (push and pullback from one launch) reads phi, f and delta
# once and writes dphi and df: 12 + 4 + 4 + 12 + 4
BYTES = {"pull": 20, "push": 20, "backward": 32, "vjp": 36}
TILED_NAMES = {"pull": "k_gather_tile<pull>", "push": "k_push_tile", "backward": "k_gather_tile<pullback>",
               "vjp": "k_gather_tile<vjp>"}
SAMPLE_NAMES = {"pull": "k_pull", "push": "k_push", "backward": "k_pull_bwd", "vjp": "k_pull_vjp"}


def kernel_name(op, deg, tiling, tuned=None):
    """Kernel that runs `op` at `deg` under the library's PP_OPT_TILING choice (and the
    autotuner's, when it chose for this op and degree)."""
    t = (tuned or {}).get(f"{TUNE_OPS.get(op, op)} deg{deg}")
    if tiling == 1 and t is not None:
        per_sample = t["kernel"] == "per-sample"
    else:
        per_sample = tiling == 0 or (tiling == 1 and deg == 1 and op in ("pull", "backward"))
    return (SAMPLE_NAMES if per_sample else TILED_NAMES)[op]


# bench op names -> autotuner op names
TUNE_OPS = {"pull": "pull", "push": "push", "backward": "pull_backward", "vjp": "pull_vjp"}


def autotune(torch, pp, f, phi, delta, schedule):
    """Let the library's autotuner (pp_tune) choose, for every op and degree the timed step
    launches, among the per-sample kernel and the tiled kernels at each tile shape and LDS
    box budget, timing each candidate on these very buffers. Returns the choices."""
    from pushpull import tune as pptune

    names = ("pull", "vjp") if schedule == "vjp" else ("pull", "push", "backward")
    out = {}
    for deg in DEGS:
        sett = pp.Settings(deg, 1, 1)
        for name in names:
            op = TUNE_OPS[name]
            if op == "push":
                pptune.tune(op, delta, phi, sett, dim=(N, N, N), reps=5)
            elif op in ("pull_backward", "pull_vjp"):
                pptune.tune(op, f, phi, sett, aux=delta, reps=5)
            else:
                pptune.tune(op, f, phi, sett, reps=5)
            cands = pptune.results()
            best = min(cands, key=lambda r: r["ms"])
            out[f"{op} deg{deg}"] = {"kernel": best["name"], "box_budget": best["cap"] or "default",
                                     "ms": round(best["ms"], 4), "candidates": len(cands)}
    return out
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_256cubed.txt")


def _traffic(kernel, deg):
    """Calibrated HBM bytes per launch of `kernel` at `deg` on this workload, from the
    committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs of this script),
    or None when it has no line for the kernel."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                p = line.split()
                if len(p) >= 5 and not line.startswith("#") and p[0] == kernel and p[1] == f"deg{deg}":
                    return int(p[4])
    except OSError:
        pass
    return None


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _smooth_noise(torch, shape, sigma, gen, dev):
    """Gaussian-smoothed standard-normal noise (separable, +-4 sigma, reflect), as a sum
    of shifted slices per axis (elementwise kernels only, no per-shape tuning)."""
    x = torch.randn(shape, generator=gen, device=dev, dtype=torch.float32)
    r = int(4 * sigma + 0.5)
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k = (k / k.sum()).astype(np.float32)
    v = x.reshape((-1,) + tuple(shape[-3:]))
    for ax in (1, 2, 3):
        n = v.shape[ax]
        idx = np.arange(-r, n + r)
        idx = np.where(idx < 0, -idx - 1, idx)             # half-sample reflection
        idx = np.where(idx >= n, 2 * n - 1 - idx, idx) % n  # (n > r is assumed)
        p = v.index_select(ax, torch.from_numpy(idx).to(dev))
        out = torch.zeros_like(v)
        for t in range(2 * r + 1):
            out.add_(p.narrow(ax, t, n), alpha=float(k[t]))
        v = out
    return v.reshape(shape).contiguous()


def make_inputs(torch, dev, seed=1):
    import pushpull as pp

    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    f = _smooth_noise(torch, d, 2.0, gen, dev)
    f /= f.std()
    u = _smooth_noise(torch, (3,) + d, 4.0, gen, dev)
    u *= 6.0 / u.abs().max()
    phi = pp.id(dim=d, like=None)
    phi = torch.from_numpy(phi).to(dev) + u
    delta = _smooth_noise(torch, d, 2.0, gen, dev)
    delta /= delta.std()
    return f.contiguous(), phi.contiguous(), delta.contiguous()


def host_cpu():
    """The host's processor as lscpu names it, and the threads this process may use. (Under a
    hypervisor the model name may be generic, e.g. "AMD EPYC"; family, model number and
    stepping still identify the generation: family 26 is AMD's Zen 5.)"""
    info = {"os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "CPU family": "family",
                "Model": "model_number", "Stepping": "stepping", "Hypervisor vendor": "hypervisor",
                "L3 cache": "l3"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    if "model" not in info:
        try:
            with open("/proc/cpuinfo") as fh:
                info["model"] = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "unknown")
        except OSError:
            info["model"] = "unknown"
    return info


def cpu_threads():
    """Worker threads for the all-cores leg: every CPU this process may run on, within the
    OMP_NUM_THREADS share when the environment sets one (a GPU host's share per GPU)."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(share))) if share.isdigit() and int(share) > 0 else n


def cpu_baseline(f, phi, delta, budget_s=40.0):
    """The reference's CPU path on z-slabs of the workload: on every host thread this
    process may use (slices are independent, as the reference's per-slice calls are), and
    single-threaded for the per-core figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    kind = "reference" if oracle.ref_available() else "port"
    prec = "ref" if kind == "reference" else "f32"
    if kind == "port" and not os.path.exists(oracle.ORACLE_SO):
        return None

    class _S:
        def __init__(self, deg):
            self.deg, self.bnd, self.ext = (deg,) * 3, [(1, 1, 1)], 1

    fz = f.cpu().numpy()
    dz_all = delta.cpu().numpy()
    pz_all = phi.cpu().numpy()
    nz = 8
    total_vox = 0
    t_total = 0.0
    while True:
        for deg in DEGS:
            s = _S(deg)
            z0 = 120
            ph = np.ascontiguousarray(pz_all[:, z0:z0 + nz])
            dl = np.ascontiguousarray(dz_all[z0:z0 + nz])
            t0 = time.perf_counter()
            oracle.pull(fz, ph, s, prec=prec)
            oracle.push(dl, ph, fz.shape, s, prec=prec)
            g = oracle.pull_grad(fz, ph, s, prec=prec)
            _ = (g[0, 0] * dl[None]).astype(np.float32)  # channel contraction of the Julia rrule
            t_total += time.perf_counter() - t0
            total_vox += ph[0].size
        if t_total > budget_s / 4 or nz >= N:
            break
        nz = min(N, nz * 2)
    # all cores: the reference is single-threaded per slice and slices are independent, so
    # one slab per worker thread (the C calls release the interpreter lock)
    from concurrent.futures import ThreadPoolExecutor

    workers = cpu_threads()
    nzw = max(1, nz // 2)

    def job(i):
        z0 = (i * nzw) % (N - nzw)
        ph = np.ascontiguousarray(pz_all[:, z0:z0 + nzw])
        dl = np.ascontiguousarray(dz_all[z0:z0 + nzw])
        for deg in DEGS:
            s = _S(deg)
            oracle.pull(fz, ph, s, prec=prec)
            oracle.push(dl, ph, fz.shape, s, prec=prec)
            g = oracle.pull_grad(fz, ph, s, prec=prec)
            _ = (g[0, 0] * dl[None]).astype(np.float32)
        return ph[0].size * len(DEGS)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        vox_all = sum(ex.map(job, range(workers)))
    t_all = time.perf_counter() - t0
    return {"value": vox_all / t_all / 1e9, "unit": "Gvoxel/s", "cores": workers, "kind": kind,
            "sample": f"pull+push+pull_grad(+contraction) at degrees 1-3 on one 256x256 x{nzw}-slice slab of "
                      f"the same inputs per thread, {workers} threads, {t_all:.1f} s wall",
            "single_thread": {"value": total_vox / t_total / 1e9, "unit": "Gvoxel/s", "cores": 1,
                              "sample": f"256x256 x{nz}-slice slabs, degrees 1-3, {t_total:.1f} s of CPU work"},
            "cpu": host_cpu()}


def _time_ms(torch, fn, reps, stream):
    """Mean ms of fn() over reps calls, HIP events on `stream`."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _warm(torch, fn, target_ms=100.0):
    """Run fn back to back for about target_ms of GPU time before it is timed: an idle GPU's
    clocks need some 30 ms of load to come back, and a short timing right after host-side
    work (input generation, a collector pass) would run partly at the lower clocks."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    fn()
    e1.record()
    e1.synchronize()
    for _ in range(min(2000, int(target_ms / max(e0.elapsed_time(e1), 0.01)) + 1)):
        fn()


def _time_ms_median(torch, fn, reps):
    """Median ms of fn() over reps calls, one pair of HIP events around each call on the
    current stream: a host stall between launches (which shows as idle GPU time inside
    that call) moves one sample, not the result."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    gc.collect()
    gc.disable()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    ev[-1][1].synchronize()
    gc.enable()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


def make_c4_volumes(torch, pp, seed=4, n=24):
    """The uneven batch: n volumes of sizes X 96-224, Y 96-256, Z 96-192 (seeded), 1-4
    channels, degree 1 or 3, each with its own smooth deformation (max 5 voxels); image f
    and target g are smoothed noise. Returns (volumes as dicts for BatchPlan, descriptors)."""
    from pushpull import ops

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    vols, descs = [], []
    for _ in range(n):
        X, Y, Z = int(rng.integers(96, 225)), int(rng.integers(96, 257)), int(rng.integers(96, 193))
        C, deg = int(rng.integers(1, 5)), int(rng.choice([1, 3]))
        d = (Z, Y, X)
        f = _smooth_noise(torch, (C,) + d, 2.0, gen, dev)
        g = _smooth_noise(torch, (C,) + d, 2.0, gen, dev)
        u = _smooth_noise(torch, (3,) + d, 4.0, gen, dev)
        u *= 5.0 / u.abs().max()
        phi = (torch.from_numpy(pp.id(dim=(X, Y, Z))).to(dev) + u).contiguous()
        vols.append(dict(f=f, phi=phi, aux=g, deg=deg, bnd=(1, 1, 1), ext=1, C=C, d=d))
        descs.append(ops._desc(pp.Settings(deg, 1, 1), d, d, C, 1))
    return vols, descs


def make_c5_inputs(torch, pp, seed=5, B=8, C=4):
    """The matching loop's subjects: B subjects at 256^3 with C channels (f, g smoothed
    noise), phi = identity + smooth displacement of max 3 voxels."""
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    f = _smooth_noise(torch, (B, C) + d, 2.0, gen, dev)
    g = _smooth_noise(torch, (B, C) + d, 2.0, gen, dev)
    u = _smooth_noise(torch, (B, 3) + d, 4.0, gen, dev)
    u *= 3.0 / u.abs().max()
    phi = (torch.from_numpy(pp.id(dim=d)).to(dev)[None] + u).contiguous()
    del u
    return f, g, phi


def bench_batched(torch, pp, reps=5, seed=4):
    """24 volumes of uneven size (X 96-224, Y 96-256, Z 96-192), 1-4 channels, degree 1
    or 3, each with its own smooth deformation (max 5 voxels): the fused objective (pull,
    residual, push, pullback) for all volumes as one persistent launch replayed from a HIP
    graph, against a per-volume loop of the same fused call, against the per-volume loop of
    separate pull / residual / push / pullback calls, and against the reference's structure:
    one call per (volume, channel) slice with the phi gradient from pull_grad."""
    from pushpull import _lib, ops
    from pushpull import batch as ppb

    dev = torch.device("cuda:0")
    vols, descs = make_c4_volumes(torch, pp, seed)
    chan_vox = sum(v["C"] * int(np.prod(v["d"])) for v in vols)
    _progress(f"batched: {len(vols)} volumes, {chan_vox / 1e6:.0f} M channel-voxels")
    plan = ppb.BatchPlan("objective", vols)
    graph = ppb.Graph(lambda s: plan.run(s))
    st = graph.stream
    L = _lib.lib()
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    sp = ct.c_void_p(st.cuda_stream)
    lo = [dict(dphi=torch.empty_like(v["phi"]), df=torch.empty_like(v["f"]),
               loss=torch.zeros((), dtype=torch.float64, device=dev), f1=torch.empty_like(v["aux"]),
               r=torch.empty_like(v["aux"])) for v in vols]

    def loop_fused():
        for v, dd, o in zip(vols, descs, lo):
            _lib.check(L.pp_objective_grad(ct.byref(dd), P(o["dphi"]), P(o["df"]), P(o["loss"]), P(v["phi"]),
                                           P(v["f"]), P(v["aux"]), 0, sp), "objective")

    def loop_unfused():
        for v, dd, o in zip(vols, descs, lo):
            _lib.check(L.pp_pull(ct.byref(dd), P(o["f1"]), P(v["phi"]), P(v["f"]), sp), "pull")
            torch.sub(o["f1"], v["aux"], out=o["r"])
            _lib.check(L.pp_push(ct.byref(dd), P(o["df"]), P(v["phi"]), P(o["r"]), 0, sp), "push")
            _lib.check(L.pp_pull_backward(ct.byref(dd), P(o["dphi"]), P(v["phi"]), P(v["f"]), P(o["r"]), sp), "bwd")

    # the reference's structure (src/pushpull_cpu.jl:42-51, src/pushpull_pb.jl:11-17): one
    # call per (volume, channel) slice, the phi gradient from a materialised pull_grad
    # contracted with the residual
    d1s = [ops._desc(pp.Settings(v["deg"], 1, 1), v["d"], v["d"], 1, 1) for v in vols]
    gbuf = [torch.empty((3,) + v["d"], device=dev) for v in vols]

    def loop_slices():
        for v, dd, o, gb in zip(vols, d1s, lo, gbuf):
            o["dphi"].zero_()
            for c in range(v["C"]):
                f_c, r_c, f1_c = v["f"][c], o["r"][c], o["f1"][c]
                _lib.check(L.pp_pull(ct.byref(dd), P(f1_c), P(v["phi"]), P(f_c), sp), "pull")
                torch.sub(f1_c, v["aux"][c], out=r_c)
                _lib.check(L.pp_push(ct.byref(dd), P(o["df"][c]), P(v["phi"]), P(r_c), 0, sp), "push")
                _lib.check(L.pp_pull_grad(ct.byref(dd), P(gb), P(v["phi"]), P(f_c), sp), "pull_grad")
                o["dphi"].addcmul_(gb, r_c.unsqueeze(0))

    with torch.cuda.stream(st):
        # some 100 ms of back-to-back work first, so that the first variant timed does not
        # run at the lower clocks of an idle GPU
        for _ in range(8):
            graph.launch()
            loop_fused()
        st.synchronize()
        t_graph = _time_ms(torch, graph.launch, reps, st)
        t_fused = _time_ms(torch, loop_fused, reps, st)
        t_unfused = _time_ms(torch, loop_unfused, reps, st)
        t_slices = _time_ms(torch, loop_slices, reps, st)
        graph.launch()
        loop_fused()
        st.synchronize()
    worst = 0.0
    for v, o in zip(plan.vols, lo):
        for a, b in ((v["out"], o["dphi"]), (v["out2"], o["df"])):
            worst = max(worst, float((a.double() - b.double()).norm() / b.double().norm()))
        worst = max(worst, abs(float(v["loss"]) - float(o["loss"])) / abs(float(o["loss"])))
    graph.close()
    plan.close()
    gv = lambda ms: round(chan_vox / (ms * 1e-3) / 1e9, 3)  # noqa: E731
    return {"volumes": 24, "channel_voxels": chan_vox,
            "batched_graph_ms": round(t_graph, 3), "loop_fused_ms": round(t_fused, 3),
            "loop_unfused_ms": round(t_unfused, 3), "loop_reference_slices_ms": round(t_slices, 3),
            "Gvoxel_s": {"batched_graph": gv(t_graph), "loop_fused": gv(t_fused), "loop_unfused": gv(t_unfused),
                         "loop_reference_slices": gv(t_slices)},
            "speedup_vs_loop_reference_slices": round(t_slices / t_graph, 2),
            "speedup_vs_loop_unfused": round(t_unfused / t_graph, 2),
            "speedup_vs_loop_fused": round(t_fused / t_graph, 2),
            "max_rel_diff_batched_vs_loop": worst}


def bench_exp(torch, pp, reps=5, seed=6, nsq=8):
    """Exp (scaling and squaring, src/compositions.jl:41-49) of a 256^3 velocity (smoothed
    noise, sigma 4, max 6 voxels): nsq compositions, each a 3-channel trilinear pull with
    sliding boundaries whose deformation Id + u is formed in the kernel. Algorithmic bytes
    per squaring: read u (12 B/voxel) and write the new u (12 B/voxel)."""
    import ctypes as ct

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    v = _smooth_noise(torch, (3,) + d, 4.0, gen, dev)
    v *= 6.0 / v.abs().max()
    v = v.contiguous()
    out, work = torch.empty_like(v), torch.empty_like(v)
    L = pp._lib.lib()
    dims = (ct.c_longlong * 3)(N, N, N)
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run():
        pp._lib.check(L.pp_exp(P(out), P(v), P(work), dims, 1, nsq, st), "pp_exp")

    _warm(torch, run)
    ms = _time_ms(torch, run, reps, None)
    n = N ** 3
    per = ms / nsq
    gbs = 24 * n / (per * 1e-3) / 1e9
    return {"volume": [N, N, N], "squarings": nsq, "ms": round(ms, 4), "ms_per_squaring": round(per, 4),
            "Gvoxel_s": round(n * nsq / (ms * 1e-3) / 1e9, 3), "GB_s_per_squaring": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_invert_def(torch, pp, reps=5, seed=7):
    """invert_def's pushes (src/multigrid.jl:266-275) at 256^3: the identity (3 channels,
    sliding boundaries) and the count of ones through one deformation, ext=0, degree 1 --
    as one pp_push_count launch against the reference's two separate pushes (identity
    materialised, then ones). Algorithmic bytes: phi 12 + g 12 + h 4 = 28 B/voxel (the
    identity and the ones are not read)."""
    import ctypes as ct

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    u = _smooth_noise(torch, (3,) + d, 4.0, gen, dev)
    u *= 4.0 / u.abs().max()
    phi = (torch.from_numpy(pp.id(dim=d)).to(dev) + u).contiguous()
    Id = torch.from_numpy(pp.id(dim=d)).to(dev).contiguous()
    ones = torch.ones(d, device=dev)
    g, h = torch.empty_like(phi), torch.empty(d, device=dev)
    L = pp._lib.lib()
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    sl = pp.Settings(1, pp.multigrid.SLIDING, 0)
    dc = pp.multigrid._count_desc(sl, d, d, 3, 1, (1, 1, 1))
    d3 = pp.ops._desc(sl, d, d, 3, 1)
    d1 = pp.ops._desc(pp.Settings(1, 1, 0), d, d, 1, 1)

    def fused():
        pp._lib.check(L.pp_push_count(ct.byref(dc), P(g), P(h), P(phi), None, 0, st), "pp_push_count")

    def separate():
        pp._lib.check(L.pp_push(ct.byref(d3), P(g), P(phi), P(Id), 0, st), "pp_push")
        pp._lib.check(L.pp_push(ct.byref(d1), P(h), P(phi), P(ones), 0, st), "pp_push")

    fused()
    separate()
    _warm(torch, fused)
    ms_f = _time_ms(torch, fused, reps, None)
    ms_s = _time_ms(torch, separate, reps, None)
    n = N ** 3
    gbs = 28 * n / (ms_f * 1e-3) / 1e9
    return {"volume": [N, N, N], "fused_ms": round(ms_f, 4), "separate_pushes_ms": round(ms_s, 4),
            "speedup": round(ms_s / ms_f, 2), "Gvoxel_s": round(n / (ms_f * 1e-3) / 1e9, 3),
            "GB_s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_affine(torch, pp, reps=5, seed=9):
    """affine_pull / affine_push (src/pushpull_cpu.jl:268-362) at 256^3 x 1 channel, degrees
    1-3, reflect boundary, ext=1, through a 10-degree rotation about z with 0.95 scaling and
    a shift (the sample coordinates come from the matrix; no deformation is read).
    Algorithmic bytes: pull reads f0 and writes f1, push reads f1 and writes f0: 8 B/voxel."""
    import ctypes as ct

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    f0 = _smooth_noise(torch, d, 3.0, gen, dev).contiguous()
    f1 = torch.empty_like(f0)
    th = np.deg2rad(10.0)
    c, sn = 0.95 * np.cos(th), 0.95 * np.sin(th)
    Aff = np.array([[c, -sn, 0, 12.0], [sn, c, 0, -9.0], [0, 0, 0.95, 6.0], [0, 0, 0, 1]], np.float32)
    a12 = pp.affine._aff12(Aff)
    L = pp._lib.lib()
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = N ** 3
    out = {"volume": [N, N, N], "affine": "0.95 x rotation of 10 degrees about z, shift (12, -9, 6)"}
    for deg in DEGS:
        dsc = pp.ops._desc(pp.Settings(deg, 1, 1), d, d, 1, 1)
        ha = pp.ops._hostptr(a12)

        def pull():
            pp._lib.check(L.pp_affine_pull(ct.byref(dsc), P(f1), P(f0), ha, st), "pp_affine_pull")

        def push():
            pp._lib.check(L.pp_affine_push(ct.byref(dsc), P(f0), P(f1), ha, 1, st), "pp_affine_push")

        _warm(torch, pull, 50.0)
        push()
        ms_pull = _time_ms(torch, pull, reps, None)
        ms_push = _time_ms(torch, push, reps, None)
        gp, gq = 8 * n / (ms_pull * 1e-3) / 1e9, 8 * n / (ms_push * 1e-3) / 1e9
        out[f"deg{deg}"] = {"pull_ms": round(ms_pull, 4), "pull_GB_s": round(gp, 1),
                            "pull_hbm_frac": round(gp / HBM_PEAK_GBS, 4), "push_ms": round(ms_push, 4),
                            "push_GB_s": round(gq, 1), "push_hbm_frac": round(gq / HBM_PEAK_GBS, 4),
                            "pull_Gvoxel_s": round(n / (ms_pull * 1e-3) / 1e9, 2)}
    return out


REG_OPS = {"membrane": [1e-3, 1.0, 0.0, 0.0], "bending+divergence": [1e-3, 0.1, 0.01, 0.05]}


def bench_regulariser(torch, pp, reps=5, seed=8, ops=tuple(REG_OPS)):
    """The sparse regulariser at 256^3 x 3 channels: vel2mom (u = L v) and one Gauss-Seidel
    sweep of (L + H) v = g with a full symmetric H (6 channels), for the multigrid default
    (membrane, 3^3 stencil, src/multigrid.jl:175-181) and bending + divergence (5^3).
    Algorithmic bytes: vel2mom reads v and writes u, 24 B/voxel; a sweep reads v, g, H and
    writes v, 60 B/voxel. A sweep is one launch per colour of each region, as the
    reference's sequential order requires (dp^3 colours). Times are medians over calls."""
    from pushpull.regulariser import DEFAULT_BND

    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    d = (N, N, N)
    n = N ** 3
    v = torch.randn((3,) + d, device=dev, generator=gen)
    g = torch.randn((3,) + d, device=dev, generator=gen)
    h = torch.rand((6,) + d, device=dev, generator=gen) + 1.0
    h[3:] *= 0.1
    u = torch.empty_like(v)
    out = {"volume": [N, N, N], "channels": 3, "hessian_channels": 6}
    for name in ops:
        k = pp.sparsify(pp.registration_operator(np.float32([1, 1, 1]), np.float32(REG_OPS[name])), d)
        pp.vel2mom_(u, v, k)  # uploads the kernel and warms up
        ms_v = _time_ms_median(torch, lambda: pp.vel2mom_(u, v, k), reps)
        w = v.clone()
        pp.relax_(g, h, k, DEFAULT_BND, 1, w)
        ms_r = _time_ms_median(torch, lambda: pp.relax_(g, h, k, DEFAULT_BND, 1, w), reps)
        gv, gr = 24 * n / (ms_v * 1e-3) / 1e9, 60 * n / (ms_r * 1e-3) / 1e9
        out[name] = {"stencil": list(k.stride), "entries": int(k.length.sum()),
                     "vel2mom_ms": round(ms_v, 4), "vel2mom_GB_s": round(gv, 1), "vel2mom_hbm_frac": round(gv / HBM_PEAK_GBS, 4),
                     "relax_sweep_ms": round(ms_r, 4), "relax_GB_s": round(gr, 1),
                     "relax_hbm_frac": round(gr / HBM_PEAK_GBS, 4), "relax_colours": int(np.prod(k.stride))}
    return out


def cpu_regulariser_baseline(res, budget_s=8.0):
    """The reference's artifacts/C/sparse_operator.c (oracle/_ref, single thread) on a
    cubic sample: vel2mom and one relax sweep per operator, scaled to voxels/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_sparse as osp
    import pushpull as pp

    if not osp.ref_available():
        return None
    from pushpull.regulariser import DEFAULT_BND

    rng = np.random.default_rng(3)
    out = {"kind": "reference", "cores": 1}
    for name, lam in REG_OPS.items():
        m = 48
        d = (m, m, m)
        k = pp.sparsify(pp.registration_operator(np.float32([1, 1, 1]), np.float32(lam)), d)
        v = rng.standard_normal((3,) + d).astype(np.float32)
        h = (1.0 + rng.random((6,) + d)).astype(np.float32)
        t0 = time.perf_counter()
        osp.ref_vel2mom(v, k, DEFAULT_BND)
        t1 = time.perf_counter()
        osp.ref_relax(v, v, h, k, DEFAULT_BND, 1)
        t2 = time.perf_counter()
        vps_v, vps_r = m ** 3 / (t1 - t0), m ** 3 / (t2 - t1)
        out[name] = {"sample": f"{m}^3 x 3 channels", "vel2mom_Mvoxel_s": round(vps_v / 1e6, 3),
                     "relax_Mvoxel_s": round(vps_r / 1e6, 3)}
        gpu = res.get("regulariser_256", {}).get(name)
        if gpu:
            out[name]["gpu_vel2mom_speedup"] = round(N ** 3 / (gpu["vel2mom_ms"] * 1e-3) / vps_v, 1)
            out[name]["gpu_relax_speedup"] = round(N ** 3 / (gpu["relax_sweep_ms"] * 1e-3) / vps_r, 1)
    return out


def bench_matching(torch, pp, iters=50, alpha=0.02, seed=5, check=True):
    """8 subjects at 256^3, 4 channels, degree 3: `iters` iterations of objective and
    gradient (pull, residual, push, pullback) with phi <- phi - alpha * dL/dphi, all
    subjects batched and each iteration replayed from a HIP graph; the descent step is fused
    into the gradient pass (each volume's phi_next is its phi), one launch per iteration
    beside the zero fill of dL/df. Host copies of subject 0's inputs and dL/dphi on a
    256x256x8 slab at iterations 1 and `iters` are returned under "_fp64_pending" for the
    CPU leg, where the FP64 oracle re-evaluates dL/dphi from the GPU's own inputs
    (fp64_checks)."""
    from pushpull import batch as ppb

    dev = torch.device("cuda:0")
    B, C = 8, 4
    f, g, phi = make_c5_inputs(torch, pp, seed, B, C)
    dphi = torch.empty_like(phi)
    df = torch.empty_like(f)
    loss = torch.zeros(B, dtype=torch.float64, device=dev)
    vols = [dict(f=f[b], phi=phi[b], aux=g[b], deg=3, bnd=(1, 1, 1), ext=1, out=dphi[b], out2=df[b], loss=loss[b],
                 phi_next=phi[b], step=-alpha) for b in range(B)]
    plan = ppb.BatchPlan("objective", vols)

    def iteration(s):
        plan.run(s)

    graph = ppb.Graph(iteration)
    st = graph.stream
    _progress("matching loop: captured")
    pending = []
    z0 = 124

    def stash(tag, phi_in):
        pending.append(dict(kind="objective_dphi", tag=tag, deg=3, f0=f[0].cpu().numpy(),
                            phi=phi_in[:, z0:z0 + 8].cpu().numpy(), g=g[0, :, z0:z0 + 8].cpu().numpy(),
                            got=dphi[0, :, z0:z0 + 8].cpu().numpy()))

    losses = []
    with torch.cuda.stream(st):
        phi_first = phi[0].clone() if check else None
        graph.launch()  # iteration 1
        st.synchronize()
        losses.append(float(loss.sum()))
        if check:
            stash("iteration_1", phi_first)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(iters - 2):
            graph.launch()
        e1.record(st)
        e1.synchronize()
        t_ms = e0.elapsed_time(e1) / max(1, iters - 2)
        phi_last = phi[0].clone() if check else None
        graph.launch()  # iteration `iters`
        st.synchronize()
        losses.append(float(loss.sum()))
        if check:
            stash(f"iteration_{iters}", phi_last)
    graph.close()
    plan.close()
    chan_vox = B * C * N ** 3
    return {"subjects": B, "channels": C, "iterations": iters, "ms_per_iteration": round(t_ms, 3),
            "Gvoxel_s": round(chan_vox / (t_ms * 1e-3) / 1e9, 3),
            "hbm_frac": round((24 + 12 * C) * B * N ** 3 / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "loss_first_last": [losses[0], losses[-1]], "_fp64_pending": pending}


def lds_atomic_peak(torch, L, blocks=4096, iters=64, reps=5):
    """Calibrated rate of 64-bit integer LDS atomic adds (adds/s over the whole GPU) under the
    scatters' own tap pattern (pp_lds_atomic_calibrate), the on-chip bound of push and the
    VJP; HIP events on the current stream."""
    from pushpull import _lib

    out = torch.empty(blocks * 256, dtype=torch.int64, device="cuda:0")
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    run = lambda: _lib.check(L.pp_lds_atomic_calibrate(ct.c_void_p(out.data_ptr()), blocks, iters, st),  # noqa: E731
                             "lds calibration")
    run()
    ms = _time_ms(torch, run, reps, None)
    return blocks * 256 * iters * 64 / (ms * 1e-3)


def gpu_slab_outputs(torch, pp, L, f, phi, delta, z0=124, nz=8):
    """GPU results of the headline ops on a 256x256x`nz` slab of the deformation (pull,
    push, pullback and the VJP, degrees 1-3), as host copies for the FP64 check of the CPU
    leg (fp64_checks)."""
    from pushpull import _lib, ops

    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    ph = phi[:, z0:z0 + nz].contiguous()
    dl = delta[z0:z0 + nz].contiguous()
    out = []
    for deg in DEGS:
        d = ct.byref(ops._desc(pp.Settings(deg, 1, 1), (N, N, N), (nz, N, N), 1, 1))
        f1, dphi, df = torch.empty_like(dl), torch.empty_like(ph), torch.empty_like(f)
        vdphi, vdf = torch.empty_like(ph), torch.empty_like(f)
        _lib.check(L.pp_pull(d, P(f1), P(ph), P(f), st), "pull")
        _lib.check(L.pp_push(d, P(df), P(ph), P(dl), 0, st), "push")
        _lib.check(L.pp_pull_backward(d, P(dphi), P(ph), P(f), P(dl), st), "backward")
        _lib.check(L.pp_pull_vjp(d, P(vdphi), P(vdf), P(ph), P(f), P(dl), 0, st), "vjp")
        torch.cuda.synchronize()
        out.append(dict(kind="headline", deg=deg, pull=f1.cpu().numpy(), push=df.cpu().numpy(),
                        backward=dphi.cpu().numpy(), vjp_df=vdf.cpu().numpy(), vjp_dphi=vdphi.cpu().numpy()))
    return dict(f=f.cpu().numpy(), phi=ph.cpu().numpy(), delta=dl.cpu().numpy(), rows=out,
                sample=f"256x256x{nz} slab of the deformation (z {z0}..{z0 + nz - 1}), all 256^3 of the image")


def fp64_checks(headline, pending):
    """CPU leg: norm-wise relative error of the GPU results against the FP64 oracle, on the
    slabs the GPU legs kept (the oracle is the checker here, never the thing measured)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def rel(got, ref):
        ref = np.asarray(ref, np.float64)
        got = np.asarray(got, np.float64).reshape(ref.shape)  # (B, C) leading dims of the oracle
        m = np.isfinite(ref)
        return float(np.linalg.norm(got[m] - ref[m]) / np.linalg.norm(ref[m]))

    class _S:
        def __init__(self, deg):
            self.deg, self.bnd, self.ext = (deg,) * 3, [(1, 1, 1)], 1

    out = {}
    if headline is not None:
        f, ph, dl = headline["f"], headline["phi"], headline["delta"]
        errs = {}
        for r in headline["rows"]:
            s = _S(r["deg"])
            ref_pull = oracle.pull(f, ph, s, prec="f64")
            ref_push = oracle.push(dl, ph, f.shape, s, prec="f64")
            ref_bwd = oracle.backward_phi(f, ph, dl, s)
            errs[f"deg{r['deg']}"] = {"pull": rel(r["pull"], ref_pull), "push": rel(r["push"], ref_push),
                                      "backward": rel(r["backward"], ref_bwd),
                                      "vjp_df": rel(r["vjp_df"], ref_push), "vjp_dphi": rel(r["vjp_dphi"], ref_bwd)}
        out["headline"] = {"sample": headline["sample"], "rel_err": errs}
    for p in pending:
        s = _S(p["deg"])
        r = oracle.pull(p["f0"][None], p["phi"][None], s, prec="f64")[0] - p["g"].astype(np.float64)
        ref = oracle.backward_phi(p["f0"][None], p["phi"][None], r[None], s)[0]
        out.setdefault("matching_loop_dphi", {})[p["tag"]] = rel(p["got"], ref)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tiling", type=int, default=1,
                    help="1: automatic kernel choice, 2: LDS-tiled kernels, 0: one work-item per sample")
    ap.add_argument("--shape", type=int, default=0, help="tile shape index (0: 16x8x8, 1: 32x8x4, 2: 16x16x4)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-uneven and matching-loop legs")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner; use the library's built-in kernel choice")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run k_fill and k_axpy over 64 Mi floats (reference kernels for HBM counter scaling)")
    ap.add_argument("--schedule", choices=["vjp", "overlap", "serial"], default="vjp",
                    help="vjp: pull on the launch stream, push and pullback from one pp_pull_vjp launch "
                         "on a second stream; overlap: the pullback runs on a second stream beside pull and "
                         "push; serial: the nine separate launches in order on one stream")
    ap.add_argument("--leg", choices=["regulariser", "affine", "exp", "invert_def", "batched", "matching"], default=None,
                    help="run only this secondary leg and print its JSON (for profiling)")
    ap.add_argument("--reg-op", choices=["all", *REG_OPS], default="all",
                    help="operators of the regulariser leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch

    import pushpull as pp

    from pushpull import _lib, ops

    if args.leg == "regulariser":
        ops_sel = list(REG_OPS) if args.reg_op == "all" else [args.reg_op]
        print(json.dumps({"regulariser_256": bench_regulariser(torch, pp, ops=ops_sel)}))
        return
    if args.leg is not None:
        leg = {"affine": bench_affine, "exp": bench_exp, "invert_def": bench_invert_def, "batched": bench_batched,
               "matching": lambda torch, pp: {k: v for k, v in bench_matching(torch, pp, iters=12, check=False).items()
                                              if not k.startswith("_")}}
        print(json.dumps({args.leg: leg[args.leg](torch, pp)}))
        return

    dev = torch.device("cuda:0")
    _progress("making inputs")
    f, phi, delta = make_inputs(torch, dev)
    n1 = N ** 3
    f1 = torch.empty_like(f)
    df = torch.empty_like(f)
    dphi = torch.empty_like(phi)
    # outputs of the launches that run beside those above in the default schedule
    f12, f13, df2, df3 = (torch.empty_like(f) for _ in range(4))
    dphi2, dphi3 = torch.empty_like(phi), torch.empty_like(phi)
    L = _lib.lib()
    _lib.set_option(_lib.PP_OPT_TILING, args.tiling)
    _lib.set_option(_lib.PP_OPT_TILE_SHAPE, args.shape)
    if args.copy_calibration:
        # counter calibration: y += a*x over 64 Mi floats reads 512 MiB and writes 256 MiB
        # (k_axpy), and the zero fill of 64 Mi floats writes 256 MiB (k_fill)
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        cs = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
        for _ in range(3):
            _lib.check(L.pp_memset_zero(ct.c_void_p(dst.data_ptr()), dst.numel(), cs), "fill")
            _lib.check(L.pp_axpy(ct.c_void_p(dst.data_ptr()), ct.c_void_p(src.data_ptr()), 0.5, src.numel(), cs),
                       "axpy")
        torch.cuda.synchronize()
        del src, dst
    descs = {deg: ops._desc(pp.Settings(deg, 1, 1), (N, N, N), (N, N, N), 1, 1) for deg in DEGS}
    P = lambda t: ct.c_void_p(t.data_ptr())  # noqa: E731
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)

    side = torch.cuda.Stream()  # default priority
    side_h = ct.c_void_p(side.cuda_stream)
    side2 = torch.cuda.Stream()  # default priority
    side2_h = ct.c_void_p(side2.cuda_stream)
    main_s = torch.cuda.current_stream()

    def step_overlap():
        # pull and push on the launch stream, the pullback on a second stream: the three ops
        # of a degree are independent (they write f1, df, dphi), and the gathers' VALU work
        # overlaps push's LDS-bound work on the same CUs; joined at the end of the step
        e = torch.cuda.Event()
        e.record(main_s)
        side.wait_event(e)
        for deg in DEGS:
            d = ct.byref(descs[deg])
            _lib.check(L.pp_pull(d, P(f1), P(phi), P(f), stream), "pull")
            _lib.check(L.pp_push(d, P(df), P(phi), P(delta), 0, stream), "push")
            _lib.check(L.pp_pull_backward(d, P(dphi), P(phi), P(f), P(delta), side_h), "backward")
        e2 = torch.cuda.Event()
        e2.record(side)
        main_s.wait_event(e2)

    def step_vjp():
        # the six independent launches of a step (pull and the VJP -- push(delta) and the
        # pullback from one launch -- per degree) on three streams balanced by their measured
        # times: VJP 3 | VJP 2, pull 1 | VJP 1, pull 2, pull 3; joined at the end of the step
        e = torch.cuda.Event()
        e.record(main_s)
        side.wait_event(e)
        side2.wait_event(e)
        d1, d2, d3 = (ct.byref(descs[deg]) for deg in DEGS)
        _lib.check(L.pp_pull_vjp(d3, P(dphi), P(df), P(phi), P(f), P(delta), 0, stream), "vjp")
        _lib.check(L.pp_pull_vjp(d2, P(dphi2), P(df2), P(phi), P(f), P(delta), 0, side_h), "vjp")
        _lib.check(L.pp_pull_vjp(d1, P(dphi3), P(df3), P(phi), P(f), P(delta), 0, side2_h), "vjp")
        _lib.check(L.pp_pull(d1, P(f1), P(phi), P(f), side_h), "pull")
        _lib.check(L.pp_pull(d2, P(f12), P(phi), P(f), side2_h), "pull")
        _lib.check(L.pp_pull(d3, P(f13), P(phi), P(f), side2_h), "pull")
        for st in (side, side2):
            e2 = torch.cuda.Event()
            e2.record(st)
            main_s.wait_event(e2)

    def step(events=None):
        for deg in DEGS:
            d = ct.byref(descs[deg])
            marks = []
            if events is not None:
                marks.append(events.record())
            _lib.check(L.pp_pull(d, P(f1), P(phi), P(f), stream), "pull")
            if events is not None:
                marks.append(events.record())
            _lib.check(L.pp_push(d, P(df), P(phi), P(delta), 0, stream), "push")
            if events is not None:
                marks.append(events.record())
            _lib.check(L.pp_pull_backward(d, P(dphi), P(phi), P(f), P(delta), stream), "backward")
            if events is not None:
                marks.append(events.record())
                events.add(deg, marks)

    class Ev:
        def __init__(self):
            self.rows = []

        def record(self):
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            return e

        def add(self, deg, m):
            self.rows.append((deg, m))

    timed_step = {"vjp": step_vjp, "overlap": step_overlap, "serial": step}[args.schedule]
    tuned = None
    if not args.no_tune and args.tiling == 1:
        _progress("autotuning")
        tuned = autotune(torch, pp, f, phi, delta, args.schedule)
    ev = Ev()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    vjp_ms = {}
    for deg in DEGS:  # the VJP per degree, zero fill of df included (as push's row)
        dd = ct.byref(descs[deg])
        vjp_ms[deg] = _time_ms_median(
            torch, lambda: _lib.check(L.pp_pull_vjp(dd, P(dphi), P(df), P(phi), P(f), P(delta), 0, stream), "vjp"),
            args.steps)
    # The GPU's clocks fall while it idles and take some 30 ms of load to come back: the host
    # work before the timed steps (the collector's pass above all) is done first, and the GPU
    # then stays busy -- the serial steps of the per-op breakdown, the warm-up -- up to the
    # timed steps (20 steps timed right after a collector pass and 5 warm-up steps ran from
    # 1.9 down to 1.52 ms a step, 1.66 ms on average).
    gc.collect()
    gc.disable()  # no collector pause from here to the end of the timed launches
    if args.schedule != "serial":  # per-op times come from separate serial steps
        _progress("serial steps for the per-op breakdown")
        for _ in range(args.steps):
            step(ev)
    _progress("warm-up")
    for _ in range(args.warmup):
        timed_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        if args.schedule == "serial":
            step(ev)
        else:
            timed_step()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gc.enable()
    ms_step = start.elapsed_time(end) / args.steps
    ms_serial = float(np.sum([m[0].elapsed_time(m[3]) for _, m in ev.rows])) / args.steps
    timed_ops = ("pull", "vjp") if args.schedule == "vjp" else ("pull", "push", "backward")
    per = {deg: {"pull": [], "push": [], "backward": []} for deg in DEGS}
    for deg, m in ev.rows:
        per[deg]["pull"].append(m[0].elapsed_time(m[1]))
        per[deg]["push"].append(m[1].elapsed_time(m[2]))
        per[deg]["backward"].append(m[2].elapsed_time(m[3]))
    detail = {}
    dominant = None
    for deg in DEGS:
        row = {}
        tri = 0.0
        for op in ("pull", "push", "backward", "vjp"):
            ms = vjp_ms[deg] if op == "vjp" else float(np.mean(per[deg][op]))
            if op != "vjp":
                tri += ms
            gbs = BYTES[op] * n1 / (ms * 1e-3) / 1e9
            row[op] = {"ms": round(ms, 4), "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
            if op in timed_ops and (dominant is None or ms > dominant[2]):
                dominant = (deg, op, ms, gbs)
        row["triple_ms"] = round(tri, 4)
        row["Gvoxel_s"] = round(n1 / (tri * 1e-3) / 1e9, 2)
        row["triple_frac"] = round(72 * n1 / (tri * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        pv = float(np.mean(per[deg]["pull"])) + vjp_ms[deg]
        row["pull_plus_vjp_ms"] = round(pv, 4)
        row["pull_plus_vjp_Gvoxel_s"] = round(n1 / (pv * 1e-3) / 1e9, 2)
        detail[f"deg{deg}"] = row

    value = len(DEGS) * n1 / (ms_step * 1e-3) / 1e9
    deg, op, ms_op, _ = dominant
    # The dominant op's kernel alone (push and the VJP without their zero fill), timed with events on the
    # stream it is launched on, as rocprofv3 --kernel-trace reports the same kernel.
    d = ct.byref(descs[deg])
    kcall = {
        "pull": lambda: L.pp_pull(d, P(f1), P(phi), P(f), stream),
        "push": lambda: L.pp_push(d, P(df), P(phi), P(delta), 1, stream),
        "backward": lambda: L.pp_pull_backward(d, P(dphi), P(phi), P(f), P(delta), stream),
        "vjp": lambda: L.pp_pull_vjp(d, P(dphi), P(df), P(phi), P(f), P(delta), 1, stream),
    }[op]
    ms = _time_ms(torch, lambda: _lib.check(kcall(), op), 20, None)
    gbs = BYTES[op] * n1 / (ms * 1e-3) / 1e9
    res = {
        "metric": "Gvoxel/s pull+push+backward at 256^3 FP32 deg1-3",
        "value": round(value, 3),
        "unit": "Gvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (smoothed Gaussian noise image and upstream gradient; identity + smooth displacement, max 6 voxels)",
        "config": {"workload": "pull+push+backward, 256^3 x 1 channel, degrees 1,2,3 (one triple each per step), "
                               "reflect boundary, ext=1", "volume": [N, N, N], "channels": 1, "degrees": list(DEGS),
                   "kernels": {0: "per-sample", 1: "automatic", 2: "tiled"}[args.tiling] +
                              (" (autotuned during warm-up)" if tuned else ""),
                   "tile_shape": args.shape, "autotuned": tuned,
                   "schedule": {"vjp": "per degree, pull and one pp_pull_vjp launch (push and pullback); the six "
                                       "launches on three streams: VJP 3 | VJP 2, pull 1 | VJP 1, pull 2, pull 3; "
                                       "joined per step",
                                "overlap": "pull and push on the launch stream, pullback on a second stream, "
                                           "joined per step",
                                "serial": "nine launches in order on one stream"}[args.schedule]},
        "roofline": {"bound": "hbm", "kernel": f"{kernel_name(op, deg, args.tiling, tuned)} deg{deg}",
                     "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": _traffic(kernel_name(op, deg, args.tiling, tuned), deg), "bytes_per_launch": BYTES[op] * n1,
                     "avg_ms": round(ms, 4), "op_ms_with_fill": round(ms_op, 4)},
        "serial_ms_per_step": round(ms_serial, 4),
        "roofline_on_chip": None,
        "per_degree": detail,
        "wall_s": round(wall, 3),
    }
    if op in ("push", "vjp"):
        # the scatters' on-chip bound: one 64-bit LDS atomic add per tap of every sample
        peak = lds_atomic_peak(torch, L)
        adds = (deg + 1) ** 3 * n1
        res["roofline_on_chip"] = {
            "bound": "lds_atomics", "kernel": res["roofline"]["kernel"], "adds_per_launch": adds,
            "achieved": round(adds / (ms * 1e-3) / 1e12, 4), "peak": round(peak / 1e12, 4), "unit": "Tadd/s",
            "frac": round(adds / (ms * 1e-3) / peak, 4),
            "calibration": "pp_lds_atomic_calibrate: 4096 x 256 lanes x 64 cubic samples, ds_add_u64 per tap"}
    # GPU results on a slab, checked against the FP64 oracle in the CPU leg below
    headline_slabs = None if args.no_cpu else gpu_slab_outputs(torch, pp, L, f, phi, delta)
    pending = []
    if not args.no_extras:
        del f1, df, dphi, f12, f13, df2, df3, dphi2, dphi3
        _progress("batched uneven volumes")
        res["batched_uneven"] = bench_batched(torch, pp)
        torch.cuda.empty_cache()
        _progress("matching loop")
        res["matching_loop"] = bench_matching(torch, pp)
        pending = res["matching_loop"].pop("_fp64_pending", [])
        torch.cuda.empty_cache()
        _progress("Exp (scaling and squaring)")
        res["exp_256"] = bench_exp(torch, pp)
        torch.cuda.empty_cache()
        _progress("invert_def pushes (push with count)")
        res["invert_def_push_count_256"] = bench_invert_def(torch, pp)
        torch.cuda.empty_cache()
        _progress("affine pull / push")
        res["affine_256"] = bench_affine(torch, pp)
        torch.cuda.empty_cache()
        _progress("sparse regulariser (vel2mom, relax)")
        res["regulariser_256"] = bench_regulariser(torch, pp)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        _progress("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(f, phi, delta)
        _progress("FP64 oracle checks")
        chk = fp64_checks(headline_slabs, pending)
        res["err_vs_fp64"] = chk["headline"]
        if "matching_loop_dphi" in chk:
            res["matching_loop"]["rel_err_vs_fp64_dphi_slab"] = chk["matching_loop_dphi"]
        if not args.no_extras:
            _progress("cpu reference of the regulariser")
            res["regulariser_cpu_reference"] = cpu_regulariser_baseline(res)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
