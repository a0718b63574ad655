#!/usr/bin/env python3
"""Benchmark: acoustic3d time steps on one MI355X, with roofline and CPU-baseline figures.

Headline workload (acoustic3d, FP32 storage and arithmetic): interior 512^3, dh 10 m, dt 0.8 ms,
fd_order 8, nbc 32 (padded 584^3 = 199,176,704 cells), seeded layered 3D model (12 jittered
interfaces, vp U[1800,4500], Gaussian perturbations of sigma 30 cells and +-5 %, Gardner rho),
12 Hz Ricker at the top centre (256, 256, 1), 256 x 256 surface receivers (every second cell,
depth 1) recording vx, vy, vz.  A step is one full time step (velocity phase, pressure phase,
receiver gather) replayed from the captured hipGraph; the workload has 1500 steps.

value = padded cell-updates per second = N_pad * steps / t, with t the hipEvent time of exactly
`steps` steps (inputs resident in HBM).  roofline: the dominant launch of a step (the largest
per-launch time, each launch timed alone with hipEvents on the plan's stream) at its algorithmic
bytes (the reference's per-phase reads and writes: velocity 40 B, pressure 24 B per cell, FP32)
over its average launch time.  traffic: HBM bytes per launch of that kernel from two rocprofv3
--pmc runs of this script (FETCH_SIZE, then WRITE_SIZE) in child processes, each scaled by what
the same counter reads on a streaming copy of known size in the same run.  cpu_baseline: the C
oracle (-O3, OpenMP on the host cores) stepping the same 512^3 workload for a bounded number of
steps (the setup cancels between two runs of different length).
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "fomo.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "cell-updates/s & % HBM roofline, FD8 2D/3D; learned-ABC+FP16 vs HABC at equal error"
HBM_PEAK_GBS = 8000.0
SEED = 20260101
# the 3D workload
N3, DH3, DT3, NT3, F03, NBC3 = 512, 10.0, 8e-4, 1500, 12.0, 32
# the 2D elastic workload (an extra line in the output)
NX, NZ, DH, DT, NT, F0, NBC, ORDER = 2000, 800, 5.0, 4e-4, 4000, 20.0, 50, 8
BYTES_VEL, BYTES_STR = 36, 44
COPY_BYTES = 1 << 30


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_threads():
    t = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    return min(t, os.cpu_count() or t)


def cpu_rate(run, cells, n_lo, n_hi, budget_s=8.0, n_max=3000):
    """Host cell-updates/s of `run(n)` (seconds for an n-step oracle run).  Runs of n_lo and n_hi
    steps estimate the time per step; then a run of n_lo + n steps, n sized to about budget_s of
    stepping, is timed against the n_lo run, so that setup cancels.  Returns (rate, steps timed,
    seconds of stepping timed)."""
    t_a = run(n_lo)
    t_b = run(n_hi)
    per = max((t_b - t_a) / (n_hi - n_lo), 1e-9)
    n = int(max(n_hi - n_lo, min(n_max, budget_s / per)))
    t_c = run(n_lo + n)
    per = max((t_c - t_a) / n, 1e-9)
    return cells / per, n, per * n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def model3d():
    from fomo.models import layered_3d
    log(f"building the seeded {N3}^3 layered model")
    return layered_3d(N3, N3, N3, seed=SEED)


def geometry3d():
    import numpy as np
    rx = np.arange(1, N3 + 1, 2)
    gx, gy = np.meshgrid(rx, rx, indexing="ij")
    return dict(sx=[N3 // 2], sy=[N3 // 2], sz=[1], rx=gx.ravel(), ry=gy.ravel(), rz=np.ones(gx.size, int))


def plan3d(vp, rho, nt, storage="f32", use_graph=True, tile_cfg=0):
    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    w = fomo.ricker_wavelet(F03, DT3, nt)
    return Plan(L.ACOUSTIC3D, vp, rho, DH3, DT3, nt, wavelet=w[None], nbc=NBC3, fd_order=8, storage=storage,
                use_graph=use_graph, tile_cfg=tile_cfg, **geometry3d())


def cpu_baseline3d(vp, rho, budget_s):
    """Oracle FP32 acoustic3d on the host cores over the full 512^3 workload: a 1-step run and a
    (1 + n)-step run, n sized to about budget_s of stepping; their difference is n steps."""
    import oracle
    oracle.lib("f32")
    th = host_threads()
    geo = geometry3d()

    def run(k):
        w = oracle.ricker(F03, DT3, k)[None]
        t = time.perf_counter()
        oracle.acoustic3d(vp, rho, DH3, DT3, k, wavelet=w, nbc=NBC3, fd_order=8, nthreads=th, **geo)
        return time.perf_counter() - t

    t1 = run(1)
    t2 = run(3)
    per = max((t2 - t1) / 2.0, 1e-3)
    n = int(max(2, min(60, budget_s / per)))
    t3 = run(1 + n)
    per = (t3 - t1) / n
    n_pad = (N3 + 2 * (NBC3 + 4)) ** 3
    return {"value": n_pad / per, "unit": "cell-updates/s", "cores": th, "kind": "port", "cpu": cpu_model(),
            "sample": f"oracle acoustic3d FP32 (C, -O3, OpenMP {th} threads on '{cpu_model()}'), the same "
                      f"{N3}^3 workload and receivers, {n} steps timed as a (1+{n})-step run minus a 1-step run "
                      f"({n * per:.1f} s of stepping)"}


# ---------------------------------------------------------------------------------------------
# HBM traffic per launch: rocprofv3 --pmc in child processes

def traffic_child(args):
    """Run under rocprofv3 --pmc: the calibration copy first, then the 3D plan's launches, each
    launched twice through the per-launch timer (the last 2 x launches dispatches)."""
    import ctypes as C

    import numpy as np

    import fomo._lib as L
    ms = C.c_double(0.0)
    L.check(L.lib().fomo_stream_copy(COPY_BYTES, 2, C.byref(ms)))
    vp = np.load(os.path.join(args.traffic_child, "vp.npy"), mmap_mode="r")
    rho = np.load(os.path.join(args.traffic_child, "rho.npy"), mmap_mode="r")
    plan = plan3d(np.asfortranarray(vp), np.asfortranarray(rho), 4, storage=args.traffic_storage, use_graph=False,
                  tile_cfg=args.tile)
    plan.advance(2)
    plan.sync()
    info = plan.launches()
    plan.launch_times(1)
    plan.close()
    print(json.dumps({"launches": [x["name"] for x in info]}), flush=True)


def tune_child(args):
    """Create the headline plan once, without a profiler, so that the 3D autotuner times its layouts
    and records the choice in the tuning file this process shares with its parent and siblings."""
    import numpy as np
    vp = np.load(os.path.join(args.tune_child, "vp.npy"), mmap_mode="r")
    rho = np.load(os.path.join(args.tune_child, "rho.npy"), mmap_mode="r")
    plan = plan3d(np.asfortranarray(vp), np.asfortranarray(rho), 4, tile_cfg=args.tile)
    print(json.dumps({"tuning": plan.tuning()}), flush=True)
    plan.close()


def _dispatches(outdir):
    """{dispatch id: (kernel name, summed counter value)} from rocprofv3's counter_collection.csv."""
    rows = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                d = int(r["Dispatch_Id"])
                name, v = rows.get(d, (r["Kernel_Name"], 0.0))
                rows[d] = (name, v + float(r["Counter_Value"]))
    return rows


def measure_traffic(vp, rho, storage="f32", timeout_s=300, tile=0):
    """HBM bytes per launch of each launch of a step: FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
    from two rocprofv3 --pmc runs of traffic_child, scaled by the readings on the streaming copy
    (COPY_BYTES read and written per dispatch).  Returns None when rocprofv3 is unavailable or a
    run fails.  Must run before this process touches the GPU (the children use it)."""
    import numpy as np
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        log("rocprofv3 not found: traffic not measured")
        return None
    work = tempfile.mkdtemp(prefix="fomo_traffic_", dir="/tmp")
    try:
        np.save(os.path.join(work, "vp.npy"), np.ascontiguousarray(vp.ravel(order="F")).reshape(vp.shape, order="F"))
        np.save(os.path.join(work, "rho.npy"), np.ascontiguousarray(rho.ravel(order="F")).reshape(rho.shape, order="F"))
        env = dict(os.environ, TMPDIR="/tmp")
        # the layout is chosen once, outside the profiler; the counter runs read it from the file
        log("traffic: tuning the layout (child process)")
        try:
            cp = subprocess.run([sys.executable, os.path.abspath(__file__), "--tune-child", work, "--tile",
                                 str(tile)], cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout_s)
        except subprocess.TimeoutExpired:
            log("traffic: tuning run timed out")
            return None
        if cp.returncode != 0:
            log(f"traffic: tuning run failed ({cp.returncode}): {cp.stderr[-400:]}")
            return None
        res = {}
        names = None
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(work, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--", sys.executable,
                   os.path.abspath(__file__), "--traffic-child", work, "--traffic-storage", storage, "--tile", str(tile)]
            log(f"traffic: rocprofv3 --pmc {counter} (child process)")
            try:
                cp = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout_s)
            except subprocess.TimeoutExpired:
                log(f"traffic: {counter} run timed out")
                return None
            lines = [ln for ln in cp.stdout.splitlines() if ln.startswith("{")]
            if cp.returncode != 0 or not lines:
                log(f"traffic: {counter} run failed ({cp.returncode}): {cp.stderr[-400:]}")
                return None
            names = json.loads(lines[-1])["launches"]
            disp = _dispatches(out)
            ids = sorted(disp)
            copy = [disp[d][1] for d in ids if "stream_copy" in disp[d][0]]
            if len(copy) < 2 or len(ids) < 2 * len(names):
                log(f"traffic: {counter}: unexpected dispatch list ({len(ids)} dispatches)")
                return None
            tail = ids[-2 * len(names):]
            per = []
            for k in range(len(names)):
                (n0, v0), (n1, v1) = disp[tail[2 * k]], disp[tail[2 * k + 1]]
                if n0 != n1:
                    log(f"traffic: dispatch pairing mismatch ({n0} / {n1})")
                    return None
                per.append((n0, 0.5 * (v0 + v1)))
            copy_kib = sum(copy) / len(copy)
            res[counter] = {"scale": COPY_BYTES / (copy_kib * 1024.0), "copy_kib": copy_kib, "per_launch": per}
        out = []
        for k, name in enumerate(names):
            kern, fk = res["FETCH_SIZE"]["per_launch"][k]
            _, wk = res["WRITE_SIZE"]["per_launch"][k]
            rd = fk * 1024.0 * res["FETCH_SIZE"]["scale"]
            wr = wk * 1024.0 * res["WRITE_SIZE"]["scale"]
            out.append({"launch": name, "kernel": kern, "fetch_kib": fk, "write_kib": wk, "read_bytes": rd,
                        "written_bytes": wr, "bytes": rd + wr})
        return {"launches": out, "fetch_scale": res["FETCH_SIZE"]["scale"], "write_scale": res["WRITE_SIZE"]["scale"],
                "copy_fetch_kib": res["FETCH_SIZE"]["copy_kib"], "copy_write_kib": res["WRITE_SIZE"]["copy_kib"]}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def elastic2d_line(args):
    """The elastic2d 2000x800 workload (dh 5 m, dt 0.4 ms, FD8, nbc 50, seeded layered model with a
    vacuum top row, 20 Hz vertical force at (1000, 5), 2000 receivers at z = 5), timed over its
    4000 steps: a secondary figure beside the 3D headline (its 153 MB per step stays in the
    256 MB Infinity Cache, so it is not an HBM figure)."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.models import layered_2d
    vp, vs, rho = layered_2d(NX, NZ, seed=SEED)
    rx = np.arange(1, NX + 1)
    w = fomo.ricker_wavelet(F0, DT, NT)
    plan = Plan(L.ELASTIC2D, vp, rho, DH, DT, NT, vs=vs, sx=[1000], sz=[5], rx=rx, rz=np.full(NX, 5),
                wavelet=w[None], nbc=NBC, fd_order=ORDER, source_type=L.SRC_FZ, use_graph=True)
    n_pad = plan.padded_cells
    plan.advance(200)
    plan.sync()
    plan.reset()
    ms = plan.timed_advance(NT)
    li, lt = plan.launches(), plan.launch_times(50)
    layout = plan.tuning()
    plan.close()
    # the autotuner picks the two sweeps or the fused interior step (frame launches beside it)
    return {"workload": f"elastic2d {NX}x{NZ}, fd8, nbc {NBC}, {NT} steps", "padded_cells": n_pad,
            "cell_updates_per_s": n_pad * NT / (ms * 1e-3), "ms_per_step": ms / NT, "layout": layout,
            "launch_ms": {x["name"]: t for x, t in zip(li, lt)},
            "step_gbs": (BYTES_VEL + BYTES_STR) * n_pad * NT / (ms * 1e-3) / 1e9}


def probe(args):
    """HBM-bound 2D probe (not the headline line): homogeneous 8192 x 8192 interior, fd8, nbc 50,
    acoustic2d, elastic2d, scalar2d (one sweep per step: read u^n, u^(n-1), c2, write u^(n+1) =
    16 B/cell) and coupled2d (velocity half-step: P, S, dP/dt, dS/dt and eight medium arrays in,
    dP/dt, dS/dt out = 56 B/cell; position half-step: 24 B/cell; its velocity launch also takes the
    interior's position half-steps, so only the step figure credits them); prints per-kernel times
    and algorithmic GB/s."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    n = args.probe_n
    vp = np.full((n, n), 3000.0, np.float32)
    vs = np.full((n, n), 1700.0, np.float32)
    rho = np.full((n, n), 2200.0, np.float32)
    nt = args.steps
    w = fomo.ricker_wavelet(15.0, 1e-3, nt)
    res = {}
    storages = args.probe_storage2d.split(",")
    # (equation, name, FP32 bytes per cell of the two phases, of which material bytes)
    eqs = [(L.ACOUSTIC2D, "acoustic2d", 28, 20, 8, 4), (L.ELASTIC2D, "elastic2d", BYTES_VEL, BYTES_STR, 8, 12),
           (L.SCALAR2D, "scalar2d", 16, 0, 4, 0), (L.COUPLED2D, "coupled2d", 56, 24, 32, 0)]
    only = set(args.probe_eq2d.split(",")) if args.probe_eq2d else None
    for eq, name, bv, bs, mv, ms_ in eqs:
        if only and name not in only:
            continue
        for st in storages:
            if st == "mat16" and eq == L.SCALAR2D:
                continue  # the scalar sweep keeps a wavefield among its material inputs
            if st != "f32" and eq == L.COUPLED2D:
                continue  # the coupled equations run with Float32 storage
            if st == "mat16":  # Float16 materials halve only the material bytes
                b_v, b_s, eb = bv - mv // 2, bs - ms_ // 2, 1.0
            else:
                b_v, b_s = bv, bs
                eb = 1.0 if st == "f32" else 0.5  # 16-bit storage halves every array's bytes
            plan = Plan(eq, vp, rho, 10.0, 1e-3, nt, vs=vs if eq in (L.ELASTIC2D, L.COUPLED2D) else None,
                        sx=[n // 2], sz=[n // 2], rx=[1, n], rz=[1, n], wavelet=w[None], nbc=50, fd_order=8,
                        storage=st)
            cells = plan.padded_cells
            plan.advance(10)
            plan.sync()
            ms = plan.timed_advance(nt)
            tv, ts, tr = plan.kernel_times(max(10, nt // 4))
            li, lt = plan.launches(), plan.launch_times(max(10, nt // 4))
            plan.close()
            key = name if st == "f32" else f"{name}_{st}"
            res[key] = {"padded_cells": cells, "ms_per_step": ms / nt, "gcell_per_s": cells * nt / ms / 1e6,
                        "step_gbs": eb * (b_v + b_s) * cells * nt / ms / 1e6,
                        "vel_ms": tv, "vel_gbs": eb * b_v * cells / tv / 1e6, "prs_ms": ts,
                        "prs_gbs": eb * b_s * cells / ts / 1e6 if b_s else 0.0, "gather_ms": tr,
                        "launches": [{"name": x["name"], "ms": t, "gbs": x["cells"] * x["bytes_per_cell"] / t / 1e6}
                                     for x, t in zip(li, lt)]}
            log(f"{key}: {res[key]}")
    print(json.dumps({"probe": f"{n}x{n} interior, fd8, nbc 50", **res}), flush=True)


def learned_line(nt=300):
    """The learned thin boundary beside plain HABC, for the default line: step times on the learned
    workload's 2400x1200 grid (dh 10 m, dt 0.8 ms, FD8, seeded layered model; the first `nt` of
    its 3000 steps) for the learned boundary at nbc 16 (FP32 and FP16 storage) and plain HABC at
    nbc 16, 32 and 100, and the reflection of each in the reference's absorption experiment
    (test/verify_patch.jl:76-91, 301x201, 1500 steps).  `bench.py --learned` adds the full width
    ladder and the CPU baseline."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo import learned
    from fomo.api import Plan
    from fomo.models import layered_2d
    nx, nz, dt = 2400, 1200, 8e-4
    vp, _, rho = layered_2d(nx, nz, seed=SEED, vacuum_top=False)
    w = fomo.ricker_wavelet(15.0, dt, nt)
    rx = np.arange(1, nx + 1, 2)
    weights = learned.load_weights()
    rows = (("learned_nbc16", 16, weights, "f32"), ("learned_nbc16_f16", 16, weights, "f16"),
            ("habc_nbc16", 16, None, "f32"), ("habc_nbc32", 32, None, "f32"), ("habc_nbc100", 100, None, "f32"))
    res = {}
    for name, nbc, wts, storage in rows:
        p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, dt, nt, sx=[nx // 2], sz=[50], rx=rx, rz=np.full(len(rx), 10),
                 wavelet=w[None], nbc=nbc, fd_order=8, storage=storage)
        try:
            if wts is not None:
                learned.enable(p, wts)
            p.advance(min(20, nt))
            p.reset()
            ms = p.timed_advance(nt)
        finally:
            p.close()
        res[name] = {"ms_per_step": ms / nt}
    vp2, rho2, sx, sz, rx2, rz2 = learned.patch_geometry()
    refl = learned.reflection_residual(vp2, rho2, sx=sx, sz=sz, rx=rx2, rz=rz2, nt=1500, L=270, judge="learned_nbc16",
                                       boundaries=[(name, nbc, wts, storage) for name, nbc, wts, storage in rows])
    for k, v in refl.items():
        if k in res:
            if v.get("storage_overflow"):
                res[k]["reflection_storage_overflow"] = True
            else:
                res[k]["reflection_db"] = v["reflection_db"]
    t = res["learned_nbc16"]["ms_per_step"]
    return {"workload": f"acoustic2d 2400x1200, fd8, dt 0.8 ms, first {nt} of 3000 steps; reflection from the "
                        "reference's absorption experiment (301x201, 1500 steps)",
            **res,
            "learned_speedup_vs_habc_nbc100": res["habc_nbc100"]["ms_per_step"] / t,
            "learned_speedup_vs_habc_nbc32": res["habc_nbc32"]["ms_per_step"] / t,
            "learned_f16_speedup_vs_habc_nbc32": res["habc_nbc32"]["ms_per_step"] / res["learned_nbc16_f16"]["ms_per_step"]}


def learned_bench(args):
    """Learned thin boundary: acoustic2d 2400x1200, dh 10 m, dt 0.8 ms, FD8, seeded layered model;
    per-step time of plain HABC at nbc 100 and of the learned boundary at nbc 16 (MLP inference
    included), and the reflection of both in the reference's absorption experiment
    (test/verify_patch.jl:76-91)."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo import learned
    from fomo.api import Plan
    from fomo.models import layered_2d
    nx, nz, dt, nt = 2400, 1200, 8e-4, args.steps
    vp, _, rho = layered_2d(nx, nz, seed=SEED, vacuum_top=False)
    w = fomo.ricker_wavelet(15.0, dt, max(nt, 2))
    rx = np.arange(1, nx + 1, 2)
    weights = learned.load_weights()
    res = {}
    ladder = (20, 24, 28, 32, 40)  # plain HABC widths between 16 and 100, for the equal-reflection comparison
    rows = (("habc_nbc100", 100, False, "f32"), ("learned_nbc16", 16, True, "f32"), ("habc_nbc16", 16, False, "f32"),
            ("learned_nbc16_f16", 16, True, "f16"), ("habc_nbc100_f16", 100, False, "f16")) + tuple(
                (f"habc_nbc{b}", b, False, "f32") for b in ladder)
    for name, nbc, use, storage in rows:
        p = Plan(L.ACOUSTIC2D, vp, rho, 10.0, dt, nt, sx=[nx // 2], sz=[50], rx=rx, rz=np.full(len(rx), 10),
                 wavelet=w[None, :nt], nbc=nbc, fd_order=8, storage=storage)
        if use:
            learned.enable(p, weights)
        p.advance(min(20, nt))
        p.reset()
        ms, overflowed = p.timed_advance_status(nt)
        res[name] = {"padded_cells": p.padded_cells, "ms_per_step": ms / nt}
        if overflowed:  # 16-bit storage saturated: the time stands, the wavefields are invalid
            res[name]["storage_overflow"] = True
        p.close()
        log(f"{name}: {res[name]}")
    for name in ("learned_nbc16", "learned_nbc16_f16"):
        res[name + "_speedup_vs_habc_nbc100"] = res["habc_nbc100"]["ms_per_step"] / res[name]["ms_per_step"]
    vp2, rho2, sx, sz, rx2, rz2 = learned.patch_geometry()
    refl = learned.reflection_residual(vp2, rho2, sx=sx, sz=sz, rx=rx2, rz=rz2, nt=1500, L=270, judge="learned_nbc16",
                                       boundaries=[("habc_nbc100", 100, None), ("learned_nbc16", 16, weights),
                                                   ("habc_nbc16", 16, None), ("learned_nbc16_f16", 16, weights, "f16"),
                                                   ("habc_nbc100_f16", 100, None, "f16")]
                                       + [(f"habc_nbc{b}", b, None) for b in ladder])
    for k, v in refl.items():
        if v.get("storage_overflow"):
            res[k]["reflection_storage_overflow"] = True
            continue
        res[k]["reflection_db"] = v["reflection_db"]
        if k == "learned_nbc16_f16" and "vs_learned_nbc16" in v:
            res[k]["trace_rel_l2_vs_learned_nbc16"] = v["vs_learned_nbc16"]
    # equal reflection: the narrowest plain HABC at least as absorbing as the learned boundary
    target = res["learned_nbc16"]["reflection_db"]
    match = next((f"habc_nbc{b}" for b in ladder + (100,)
                  if res[f"habc_nbc{b}"].get("reflection_db", float("inf")) <= target), None)
    if match:
        res["equal_reflection"] = {"habc": match, "habc_db": res[match]["reflection_db"], "learned_db": target,
                                   "learned_speedup": res[match]["ms_per_step"] / res["learned_nbc16"]["ms_per_step"],
                                   "learned_f16_speedup": res[match]["ms_per_step"] /
                                   res["learned_nbc16_f16"]["ms_per_step"]}
    if not args.no_cpu:
        import oracle
        oracle.lib("f32")
        th = host_threads()
        for name, nbc in (("habc_nbc100", 100), ("habc_nbc16", 16)):
            def run(k, nbc=nbc):
                wc = fomo.ricker_wavelet(15.0, dt, k)
                t = time.perf_counter()
                oracle.acoustic2d(vp, rho, 10.0, dt, k, sx=[nx // 2], sz=[50], rx=rx, rz=np.full(len(rx), 10),
                                  wavelet=wc[None], nbc=nbc, fd_order=8, nthreads=th)
                return time.perf_counter() - t
            log(f"cpu baseline: oracle acoustic2d FP32 nbc {nbc} on {th} threads")
            rate, ns, secs = cpu_rate(run, res[name]["padded_cells"], 4, 24, budget_s=5.0)
            res[name]["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": th, "kind": "port",
                                         "ms_per_step": res[name]["padded_cells"] / rate * 1e3,
                                         "sample": f"oracle acoustic2d FP32 (C, -O3, OpenMP {th} threads), same grid, "
                                                   f"{ns} steps timed against a 4-step run ({secs:.1f} s)"}
    print(json.dumps({"learned": f"acoustic2d {nx}x{nz}, fd8, dt 0.8 ms, {nt} steps; reflection from the "
                                 "reference's absorption experiment (301x201, 1500 steps)", **res}), flush=True)


def batch_bench(args):
    """Uneven multi-shot batch: a 4000x1000 seeded layered model (the 2000x800 model's rules), dh 10
    m, dt 0.8 ms, fd8, nbc 40, 64 shots (acoustic, or elastic with --batch-eq elastic2d) in one
    windowed plan.  Window widths are drawn
    uniformly from 256..1536 and window starts uniformly over the model (SplitMix64, seeded); each
    shot has its source at the window centre 5 cells deep and receivers on every second cell at
    depth 5.  Timed with the autotuned strip length and with each fixed one, so the line shows
    whether the autotuner picked the fastest.  Times --steps steps (the workload is 3000)."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.models import SplitMix64, layered_2d
    nx, nz, dt, nbc = 4000, 1000, 8e-4, 40
    nt = min(args.steps, 3000)
    log("building the seeded 4000x1000 layered model")
    vp, vs, rho = layered_2d(nx, nz, seed=SEED)
    elastic = args.batch_eq == "elastic2d"
    eq2 = L.ELASTIC2D if elastic else L.ACOUSTIC2D
    bv, bp = (BYTES_VEL, BYTES_STR) if elastic else (28, 20)
    w = fomo.ricker_wavelet(15.0, dt, nt)
    rng = SplitMix64(SEED + 3)
    shots = []
    for _ in range(64):
        wd = 256 + int(rng.uniform(0.0, 1281.0 - 1e-9))
        x0 = 1 + int(rng.uniform(0.0, float(nx - wd + 1) - 1e-9))
        rxl = np.arange(1, wd + 1, 2)
        shots.append(dict(x0=x0, nx=wd, sx=[wd // 2], sz=[5], rx=rxl, rz=np.full(len(rxl), 5), wavelet=w[None]))
    res = {}
    fused_family = not os.environ.get("FOMO_2D_UNFUSED")
    for cfg in (0, 1, 2, 3):
        log(f"64-shot plan, tile_cfg {cfg} (0: autotuned): setup, graph capture, {nt} timed steps")
        plan = Plan(eq2, vp, rho, 10.0, dt, nt, vs=vs if elastic else None, shots=shots, nbc=nbc, fd_order=8,
                    tile_cfg=cfg)
        cells = plan.padded_cells
        plan.advance(min(20, nt))
        plan.sync()
        plan.reset()
        ms = plan.timed_advance(nt)
        plan.reset()
        tv, tp, tr = plan.kernel_times(min(50, nt))
        plan.close()
        res[cfg] = {"ms_per_step": ms / nt, "cell_updates_per_s": cells * nt / (ms * 1e-3),
                    "step_gbs": (bv + bp) * cells * nt / ms / 1e6, "vel_ms": tv, "vel_gbs": bv * cells / tv / 1e6,
                    "prs_ms": tp, "prs_gbs": bp * cells / tp / 1e6, "step_rest_ms": tr}
    out = {"batch": f"64 uneven windows (widths 256..1536) of a 4000x1000 model, {args.batch_eq} fd8, nbc 40, "
                    f"dt 0.8 ms, "
                    f"{nt} steps", "padded_cells": cells, "autotuned": res[0],
           # a forced tile shape of an acoustic2d or elastic2d plan is one of the fused interior's strip lengths
           "fixed_strip_rows": {str(r): res[c] for c, r in zip((1, 2, 3), (24, 32, 48) if fused_family else (16, 32, 64))},
           "widths": [s["nx"] for s in shots]}
    if not args.no_cpu:
        import oracle
        oracle.lib("f32")
        th = host_threads()
        sample = shots[:8]

        def run(k):
            wc = fomo.ricker_wavelet(15.0, dt, k)
            t = time.perf_counter()
            for s in sample:
                x0, wd = s["x0"], s["nx"]
                win = slice(x0 - 1, x0 - 1 + wd)
                kw = dict(sx=s["sx"], sz=s["sz"], rx=s["rx"], rz=s["rz"], wavelet=wc[None], nbc=nbc, fd_order=8,
                          nthreads=th)
                if elastic:
                    oracle.elastic2d(vp[win], vs[win], rho[win], 10.0, dt, k, **kw)
                else:
                    oracle.acoustic2d(vp[win], rho[win], 10.0, dt, k, **kw)
            return time.perf_counter() - t
        cells8 = sum((s["nx"] + 2 * (nbc + 4)) * (nz + 2 * (nbc + 4)) for s in sample)
        log(f"cpu baseline: oracle {args.batch_eq} FP32 on {th} threads over the first 8 windows")
        rate, ns, secs = cpu_rate(run, cells8, 10, 60)
        out["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": th, "kind": "port",
                               "sample": f"oracle {args.batch_eq} FP32 (C, -O3, OpenMP {th} threads) on the first 8 windows "
                                         f"run one after another, {ns} steps timed against 10-step runs ({secs:.1f} s)"}
    print(json.dumps(out), flush=True)


def stream_copy_gbs(nbytes=2 << 30, iters=20):
    """Practical HBM ceiling: float4 streaming copy of `nbytes` (read + write counted)."""
    import ctypes as C

    import fomo._lib as L
    ms = C.c_double(0.0)
    L.check(L.lib().fomo_stream_copy(int(nbytes), int(iters), C.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def probe3d(args):
    """acoustic3d (or elastic3d with --probe-eq elastic3d) at the 3D workload's size (512^3
    interior, nbc 32, fd8; 584^3 padded) for each storage precision: per-kernel times and
    algorithmic GB/s (acoustic 64 B/cell FP32 = 40 velocity + 24 pressure; elastic 140 B/cell =
    60 velocity + 80 stress; 16-bit storage halves them)."""
    import numpy as np

    import fomo
    import fomo._lib as L
    from fomo.api import Plan
    from fomo.models import layered_3d
    n = args.probe_n3
    log(f"building the seeded {n}^3 layered model")
    vp, rho = layered_3d(n, n, n, seed=SEED)
    elastic = args.probe_eq == "elastic3d"
    vs = (vp / np.float32(1.73)).astype(np.float32) if elastic else None
    eq3 = L.ELASTIC3D if elastic else L.ACOUSTIC3D
    bv, bs = (15, 20) if elastic else (10, 6)  # 4-byte words per cell of the two sweeps
    nt = args.steps
    w = fomo.ricker_wavelet(12.0, 8e-4, nt)
    rx = np.arange(1, n + 1, 2)
    gx, gy = np.meshgrid(rx, rx, indexing="ij")
    res = {}
    codes = {"f32": L.STORE_F32, "f16": L.STORE_F16, "bf16": L.STORE_BF16}
    for name in args.probe_storage.split(","):
        storage = codes[name]
        eb = 4 if storage == L.STORE_F32 else 2
        plan = Plan(eq3, vp, rho, 10.0, 8e-4, nt, vs=vs, sx=[n // 2], sy=[n // 2], sz=[1], rx=gx.ravel(),
                    ry=gy.ravel(), rz=np.ones(gx.size, int), wavelet=w[None], nbc=32, fd_order=8, storage=storage,
                    tile_cfg=args.probe_tile)
        cells = plan.padded_cells
        plan.advance(4)
        plan.sync()
        ms = plan.timed_advance(nt)
        tv, tp, tr = plan.kernel_times(max(4, nt // 4))
        li, lt = plan.launches(), plan.launch_times(max(4, nt // 4))
        layout = plan.tuning()
        plan.close()
        res[name] = {"padded_cells": cells, "ms_per_step": ms / nt, "gcell_per_s": cells * nt / ms / 1e6,
                     "step_gbs": (bv + bs) * eb * cells * nt / ms / 1e6, "vel_ms": tv,
                     "vel_gbs": bv * eb * cells / tv / 1e6, "prs_ms": tp, "prs_gbs": bs * eb * cells / tp / 1e6,
                     "gather_ms": tr, "layout": layout,
                     "launches": [{"name": x["name"], "ms": t, "gbs": x["cells"] * x["bytes_per_cell"] / t / 1e6}
                                  for x, t in zip(li, lt)]}
        log(f"{name}: {res[name]}")
    gbs, cms = stream_copy_gbs()
    res["stream_copy"] = {"bytes": 2 << 30, "ms": cms, "gbs": gbs}
    log(f"stream copy: {gbs:.0f} GB/s")
    if not args.no_cpu:
        import oracle
        oracle.lib("f32")
        th = host_threads()

        def run(k):
            wc = fomo.ricker_wavelet(12.0, 8e-4, k)
            kw = dict(sx=[n // 2], sy=[n // 2], sz=[1], rx=gx.ravel(), ry=gy.ravel(), rz=np.ones(gx.size, int),
                      wavelet=wc[None], nbc=32, fd_order=8, nthreads=th)
            t = time.perf_counter()
            if elastic:
                oracle.elastic3d(vp, vs, rho, 10.0, 8e-4, k, **kw)
            else:
                oracle.acoustic3d(vp, rho, 10.0, 8e-4, k, **kw)
            return time.perf_counter() - t
        log(f"cpu baseline: oracle {args.probe_eq} FP32 on {th} threads")
        rate, ns, secs = cpu_rate(run, res[next(iter(res))]["padded_cells"], 1, 3, budget_s=10.0)
        res["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": th, "kind": "port",
                               "sample": f"oracle {args.probe_eq} FP32 (C, -O3, OpenMP {th} threads), same {n}^3 "
                                         f"workload, {ns} steps timed against a 1-step run ({secs:.1f} s of stepping)"}
    print(json.dumps({"probe3d": f"{args.probe_eq} {n}^3 interior, fd8, nbc 32", **res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=NT3)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--launch-reps", type=int, default=10, help="back-to-back launches per per-launch timing")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic children")
    ap.add_argument("--no-extras", action="store_true", help="skip the FP16/BF16 and elastic2d extra lines")
    ap.add_argument("--tile", type=int, default=0, help="3D z-segment choice of the headline plan (0: default)")
    ap.add_argument("--traffic-child", default="", help=argparse.SUPPRESS)
    ap.add_argument("--tune-child", default="", help=argparse.SUPPRESS)
    ap.add_argument("--traffic-storage", default="f32", help=argparse.SUPPRESS)
    ap.add_argument("--probe", action="store_true", help="HBM-bound 2D probe instead of the headline line")
    ap.add_argument("--probe-n", type=int, default=8192)
    ap.add_argument("--probe3d", action="store_true", help="acoustic3d 512^3 probe (f32/f16/bf16)")
    ap.add_argument("--probe-n3", type=int, default=512)
    ap.add_argument("--probe-storage", default="f32,f16,bf16", help="storages of the 3D probe")
    ap.add_argument("--probe-storage2d", default="f32", help="storages of the 2D probe (f32,f16,bf16,mat16)")
    ap.add_argument("--probe-eq2d", default="", help="equations of the 2D probe (default all: acoustic2d,elastic2d,"
                    "scalar2d,coupled2d)")
    ap.add_argument("--probe-eq", default="acoustic3d", choices=["acoustic3d", "elastic3d"], help="3D probe equation")
    ap.add_argument("--batch-eq", default="acoustic2d", choices=["acoustic2d", "elastic2d"], help="batch equation")
    ap.add_argument("--stream-copy", action="store_true", help="only the HBM streaming-copy calibration")
    ap.add_argument("--learned", action="store_true", help="learned nbc-16 boundary vs plain HABC nbc 100 (acoustic2d 2400x1200)")
    ap.add_argument("--probe-tile", type=int, default=0, help="3D z-segment choice (0 default, k: 16<<(k-1) planes)")
    ap.add_argument("--copy-mb", default="", help="with --stream-copy: comma-separated copy sizes in MB")
    ap.add_argument("--stream-probe", action="store_true",
                    help="the sweeps' array counts without stencils (5r2w, 4r1w, 7r2w, 8r3w, 3r1w) at --probe-mb MiB "
                         "per array")
    ap.add_argument("--probe-mb", default="16,256",
                    help="with --stream-probe: comma-separated MiB per array (7.3 is one 2108x908 FP32 field)")
    ap.add_argument("--batch", action="store_true",
                    help="uneven 64-shot windowed acoustic batch on a 4000x1000 model (timed over --steps)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (shots and slabs would shard "
              "across GPUs).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.traffic_child:
        traffic_child(args)
        return
    if args.tune_child:
        tune_child(args)
        return
    if args.probe:
        probe(args)
        return
    if args.probe3d:
        probe3d(args)
        return
    if args.learned:
        learned_bench(args)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.stream_probe:
        import ctypes as C

        import fomo._lib as L
        out = []
        for mb in (float(v) for v in args.probe_mb.split(",")):
            for nr, nw, what in ((5, 2, "acoustic velocity"), (4, 1, "acoustic pressure"), (7, 2, "elastic velocity"),
                                 (8, 3, "elastic stress"), (3, 1, "scalar")):
                ms = C.c_double(0.0)
                nb = int(mb * (1 << 20)) // 16 * 16
                L.check(L.lib().fomo_stream_probe(nr, nw, nb, max(10, min(200, int(640 / mb))), C.byref(ms)))
                out.append({"pattern": f"{nr} read + {nw} written ({what})", "mb_per_array": mb, "ms": ms.value,
                            "gbs": (nr + nw) * nb / (ms.value * 1e-3) / 1e9})
                log(f"{out[-1]}")
        print(json.dumps({"stream_probe": out}), flush=True)
        return
    if args.stream_copy:
        if args.copy_mb:
            res = []
            for mb in (float(x) for x in args.copy_mb.split(",")):
                nb = int(mb * 1e6) // 16 * 16
                gbs, ms = stream_copy_gbs(nb, 200 if mb < 300 else 20)
                res.append({"bytes": nb, "ms": ms, "gbs": gbs})
            print(json.dumps({"stream_copy": res}), flush=True)
            return
        gbs, ms = stream_copy_gbs(1 << 30, 4)
        print(json.dumps({"stream_copy": {"bytes": 1 << 30, "ms": ms, "gbs": gbs}}), flush=True)
        return

    headline(args)


def headline(args):
    import numpy as np

    from fomo.models import model_digest

    vp, rho = model3d()
    # This process and its children share one tuning file, so the counter runs measure the layout
    # the timed run uses (a fresh file per bench run unless FOMO_TUNE_CACHE names one).
    if not os.environ.get("FOMO_TUNE_CACHE"):
        os.environ["FOMO_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="fomo_tune_", dir="/tmp"), "tune.txt")
    traffic = None
    if not args.no_traffic:
        traffic = measure_traffic(vp, rho, tile=args.tile)  # before this process touches the GPU
    log("creating the plan (setup, upload, graph capture)")
    plan = plan3d(vp, rho, NT3, tile_cfg=args.tile)
    layout = plan.tuning()
    log(f"layout: {layout}")
    n_pad = plan.padded_cells
    pnx, pny, pnz = plan.padded[0]
    log(f"warmup {args.warmup} steps")
    plan.advance(args.warmup)
    plan.sync()
    plan.reset()
    plan.sync()
    log(f"timing {args.steps} steps")
    ms = plan.timed_advance(args.steps)
    seis = plan.seis(3)
    finite = bool(all(np.all(np.isfinite(x)) for x in seis))
    peak_vz = float(np.abs(seis[2]).max())
    log(f"{args.steps} steps in {ms:.2f} ms; seismograms finite: {finite}")
    # per-launch times on a live wavefield (the launches are idempotent: the state does not move)
    info = plan.launches()
    lt = plan.launch_times(args.launch_reps)
    plan.close()
    launches = []
    for k, (li, t) in enumerate(zip(info, lt)):
        gbs = li["bytes_per_cell"] * li["cells"] / (t * 1e-3) / 1e9
        launches.append({**li, "ms": t, "algorithmic_gbs": gbs})
        log(f"launch {k} {li['name']}: {t:.4f} ms, {gbs:.0f} GB/s algorithmic")
    dom = max(range(len(lt)), key=lambda k: lt[k])
    d = launches[dom]
    alg_bytes = d["bytes_per_cell"] * d["cells"]
    achieved = alg_bytes / (d["ms"] * 1e-3) / 1e9
    bytes_step = 64 * n_pad
    value = n_pad * args.steps / (ms * 1e-3)
    roof = {"bound": "hbm", "kernel": d["name"], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes_per_launch": alg_bytes,
            "bytes_per_cell": d["bytes_per_cell"], "cells_per_launch": d["cells"], "launch_ms": d["ms"]}
    if traffic:
        t = traffic["launches"][dom]
        roof["traffic"] = t["bytes"]
        roof["traffic_detail"] = {"read_bytes": t["read_bytes"], "written_bytes": t["written_bytes"],
                                  "fetch_kib": t["fetch_kib"], "write_kib": t["write_kib"],
                                  "fetch_scale": traffic["fetch_scale"], "write_scale": traffic["write_scale"],
                                  "copy_fetch_kib": traffic["copy_fetch_kib"],
                                  "copy_write_kib": traffic["copy_write_kib"], "copy_bytes": COPY_BYTES,
                                  "source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in child runs of this "
                                            "script, scaled by the same counters on a 1 GiB streaming copy",
                                  "traffic_over_algorithmic": t["bytes"] / alg_bytes}
        for l, tr in zip(launches, traffic["launches"]):
            l["traffic_bytes"] = tr["bytes"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"acoustic3d {N3}^3 interior, padded {pnx}x{pny}x{pnz}, fd8, nbc {NBC3}, dh {DH3} m, "
                        f"dt {DT3 * 1e3} ms, Ricker {F03} Hz at the top centre, 256x256 surface receivers "
                        f"(vx, vy, vz), FP32 storage",
            "padded_cells": n_pad,
            "interior_cell_updates_per_s": N3 ** 3 * args.steps / (ms * 1e-3),
            "model": f"seeded layered 3D with Gaussian perturbations (SplitMix64 seed {SEED}, digest "
                     f"{model_digest(vp, rho)})",
            "step_algorithmic_gbs": bytes_step / (ms * 1e-3 / args.steps) / 1e9,
            "step_frac_of_peak": bytes_step / (ms * 1e-3 / args.steps) / 1e9 / HBM_PEAK_GBS,
            "launches": launches,
            "layout": layout,
            "outputs_finite": finite,
            "peak_abs_vz": peak_vz,
        },
        "roofline": roof,
    }
    if not args.no_extras:
        ex = {}
        for st in ("f16", "bf16"):
            log(f"extra: the same workload with {st} storage")
            p16 = plan3d(vp, rho, NT3, storage=st)
            p16.advance(16)
            p16.sync()
            p16.reset()
            m16 = p16.timed_advance(args.steps)
            p16.close()
            ex[f"acoustic3d_{st}"] = {"ms_per_step": m16 / args.steps, "cell_updates_per_s": n_pad * args.steps /
                                      (m16 * 1e-3), "speedup_vs_f32": ms / m16}
        log("extra: elastic2d 2000x800")
        ex["elastic2d_2000x800"] = elastic2d_line(args)
        log("extra: learned thin boundary against plain HABC")
        ex["learned_2400x1200"] = learned_line()
        out["extras"] = ex
    if not args.no_cpu:
        log("timing the CPU baseline (oracle acoustic3d at 512^3)")
        out["cpu_baseline"] = cpu_baseline3d(vp, rho, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
