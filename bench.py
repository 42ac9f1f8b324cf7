"""Throughput of the grid-kriging hot path on one MI355X.

Workload: ordinary kriging of a 100x100x50 SugarboxGrid (500,000 cells, 5%
seeded hard data, exponential model ranges (20,20,4), sill 1, nugget 0.05,
radiuses (20,20,4), max_neighbours 24) -- synthetic data from
geo_bsd.synthetic.ok_c2.  One step = one complete pass of the hot path
(mask padding, cell hash, neighbour search, count bucketing, covariance
assembly, Cholesky, solve, combine, variance) over the grid, with inputs
resident in HBM, in the cached-plan steady state: the parameter set's search
template and covariance tables are built on the host by the first call and
reused by the steps.  The line's `cold_call_mixed` (a parameter set not seen
before) and `end_to_end_*` (geo_bsd.ordinary_kriging on host arrays, the
reference's own entry) figures report what the steps leave out.  Before the
steps, hpglx_autotune picks the launch settings (`autotune`).

The metric is kriged cells/s of the FP32-Cholesky + FP64-refinement solver
("mixed"); the plain FP64 solver is timed the same way and reported beside
it, together with both solvers' largest estimate error against the FP64 CPU
oracle.  Prints one JSON line.

`--workload` selects one of the other workloads instead (same line layout,
its own metric): `mik` (median IK, 5 thresholds, 200x200x100), `wells` (OK
from 40 wells, 256x256x64, cluster path), `ik` (indicator kriging over 4
categories, 200x200x100), `lvm` (kriging with a locally varying mean,
200x200x40), `io` (INC property file of 4M values: GPU parse and print) and
`variogram` (experimental semivariogram of the 200x200x100 median-IK hard
data, 40,000 cells, and of 20,000 scattered points) and `sgs` (sequential
Gaussian simulation of the 100x100x50 grid, 475,000 cells; one realization,
and `--realizations` of them advancing together).
Each is timed beside the reference's own CPU build where it is present (for
`sgs`, beside the CPU oracle: the reference's simulation cannot be built).
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hpgl-reborn_amd"))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6    # MI355X vector FP64
PEAK_FP32_TFLOPS = 157.3   # MI355X vector FP32 (packed)

# HBM bytes per launch of the main solve kernel on this workload, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a copy kernel of
# known size (--calibrate-copy; profiles/hbm_traffic_ok_100x100x50.md):
# FETCH_SIZE x 1.9998 + WRITE_SIZE x 1.0000, KiB x 1024.
TRAFFIC_BYTES = {
    "k_solve_sorted_lds<24,4,mixed,OK>": 72.86e6,
    "k_solve_sorted_lds<24,4,fp64,OK>": 72.80e6,
}
# Algorithmic bytes of one solve launch: per kriged cell its work-list entry
# and count (8 B) and 96 B of offsets in, 13 B estimate/mask/variance out;
# the value array and the octant table once.
ALGO_BYTES = 475000 * 117 + 500000 * 4 + 121 * 1024


def flops_ok(n):
    """Algorithmic flops of one OK cell with n neighbours: covariance evaluations
    (25 flop each) + Cholesky + two solves + tail."""
    import numpy as np
    n = n.astype(np.float64)
    e = n * (n + 1) / 2 + n + 1
    return 25 * e + n ** 3 / 3 + n ** 2 / 2 + n / 6 + 4 * n ** 2 + 10 * n


def flops_mik(n, nind):
    """Algorithmic flops of one median-IK cell with n neighbours and nind
    thresholds: covariance evaluations + Cholesky + one solve + tail."""
    import numpy as np
    n = n.astype(np.float64)
    e = n * (n + 1) / 2 + n + 1
    return 25 * e + n ** 3 / 3 + n ** 2 / 2 + n / 6 + 2 * n ** 2 + 5 * n + 3 * nind * n


def main_kernel_roofline(flops, ms, peak, kernel):
    achieved = flops / (ms / 1e3) / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "flops_per_launch": flops, "launch_ms": ms}


def cpu_baseline(wl):
    """The reference's own OpenMP CPU path (oracle/_ref, compiled from its
    sources) on the same workload, or the oracle when that build is absent."""
    import numpy as np
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, 16))
    cov = O.make_cov(wl.cov["type"], wl.cov["ranges"], wl.cov["angles"], wl.cov["sill"], wl.cov["nugget"])
    kriged = int((wl.mask == 0).sum())
    times = []
    if O.ref_available():
        kind = "reference"
        for _ in range(3):
            t0 = time.perf_counter()
            O.ref_ordinary_kriging(cov, wl.radiuses, wl.k, wl.shape, wl.values, wl.mask, threads=threads)
            times.append(time.perf_counter() - t0)
    else:
        kind = "port"
        os.environ["OMP_NUM_THREADS"] = str(threads)
        for _ in range(3):
            t0 = time.perf_counter()
            O.krige(0, cov, wl.radiuses, wl.k, wl.shape, wl.mask, values=wl.values, want_neighbours=False)
            times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {"value": kriged / t, "unit": "kriged cells/s", "cores": threads, "kind": kind,
            "host_cores_total": os.cpu_count(),
            "sample": f"full 100x100x50 OK grid ({kriged} kriged cells), median of 3 runs, {threads} OpenMP threads "
                      f"(the GPU box's 16-core share of a {os.cpu_count()}-core host)"}


def host_call_figures(wl, model):
    """End-to-end figures beside the steady-state step: the reference's own
    entry (geo_bsd.ordinary_kriging on host arrays: host-to-device copies,
    kernels, copy back) and a device call with a parameter set not seen
    before (host plan: search template sort and covariance tables, their
    upload, then the kernels)."""
    import numpy as np
    import torch
    import geo_bsd
    from geo_bsd import CovarianceModel, device as D
    nx, ny, nz = wl.shape
    prop = geo_bsd.ContProperty(wl.values.reshape((nx, ny, nz), order="F"), wl.mask.reshape((nx, ny, nz), order="F"))
    grid = geo_bsd.SugarboxGrid(nx, ny, nz)
    kriged = int((wl.mask == 0).sum())
    out = {}
    # the entry reports its progress through the output handler; keep stdout
    # for the bench line
    geo_bsd.set_output_handler(lambda msg, param: 0, None)
    for prec in ("mixed", "fp64"):
        D.set_default_precision(prec)
        geo_bsd.ordinary_kriging(prop, grid, wl.radiuses, wl.k, model)  # plan built once
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            geo_bsd.ordinary_kriging(prop, grid, wl.radiuses, wl.k, model)
            times.append(time.perf_counter() - t0)
        t = statistics.median(times)
        out[f"end_to_end_{prec}"] = {"ms": t * 1e3, "kriged_cells_per_s": kriged / t,
                                     "call": "geo_bsd.ordinary_kriging on host arrays, median of 5"}
    D.set_default_precision("fp64")
    geo_bsd.set_output_handler(None, None)
    dev = torch.device("cuda:0")
    v = torch.from_numpy(wl.values).to(dev)
    m = torch.from_numpy(wl.mask).to(dev)
    cold = []
    for i in range(3):
        cov = dict(wl.cov, nugget=wl.cov["nugget"] * (1.0 + 1e-6 * (i + 1)))  # a parameter set not seen before
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        D.krige(0, wl.shape, wl.radiuses, wl.k, CovarianceModel(**cov), m, values=v, precision="mixed")
        torch.cuda.synchronize()
        cold.append(time.perf_counter() - t0)
    out["cold_call_mixed"] = {"ms": statistics.median(cold) * 1e3,
                              "call": "device buffers, new parameter set (template sort, tables, upload, kernels), "
                                      "median of 3"}
    return out


def loo_bench(torch, check_oracle):
    """Leave-one-out kriging of 20,000 scattered points for 64 spherical
    parameter sets (k = 24): k-NN, forward, adjoint backward, 200 Adam steps."""
    import numpy as np
    from geo_bsd import loo as L, synthetic as S
    wl = S.loo_c5()
    dev = torch.device("cuda:0")
    xyz = torch.from_numpy(wl.xyz).to(dev)
    v = torch.from_numpy(wl.values).to(dev)
    prm = torch.from_numpy(wl.params).to(dev)
    nbr = L.knn(xyz, wl.k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nbr = L.knn(xyz, wl.k)
    torch.cuda.synchronize()
    knn_ms = (time.perf_counter() - t0) * 1e3
    fwd, bwd = [], []
    for it in range(6):
        p = prm.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = L.loo_loss(p, xyz, v, nbr)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= 2:
            fwd.append(t1 - t0)
            bwd.append(t2 - t1)
    # the optimizer's own torch kernels load on their first use (about 1 s,
    # once per process): two untimed steps first
    L.fit(prm, xyz, v, nbr, steps=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = L.fit(prm, xyz, v, nbr, steps=200)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    h = hist.cpu().numpy()
    nsys = wl.params.shape[0] * wl.xyz.shape[0]
    out = {"workload": f"{wl.xyz.shape[0]} scattered points x {wl.params.shape[0]} spherical parameter sets, k=24",
           "knn_ms": knn_ms, "forward_ms": 1e3 * statistics.median(fwd), "backward_ms": 1e3 * statistics.median(bwd),
           "systems_per_s_forward": nsys / statistics.median(fwd),
           "adam_200_steps_s": fit_s, "loss_best_start": [float(h[0].min()), float(h[-1].min())],
           "loss_mean": [float(h[0].mean()), float(h[-1].mean())], "dtype": "f64"}
    if check_oracle:
        from oracle import loo_oracle as O
        sub = np.arange(0, wl.xyz.shape[0], 100)
        with torch.no_grad():
            est = L.loo_estimates(prm[:2], xyz, v, nbr).cpu().numpy()[:, sub]
        # the oracle on a subsample of points, with their neighbour lists
        nb = nbr.cpu().numpy()
        xyz_np, v_np = wl.xyz, wl.values
        ref = np.empty((2, len(sub)))
        for j, i in enumerate(sub):
            pts = xyz_np[nb[i]].astype(np.float64)
            for p in range(2):
                k = nb.shape[1]
                A = O.cov("spherical", wl.params[p], pts[:, None, :] - pts[None, :, :])
                b = O.cov("spherical", wl.params[p], pts - xyz_np[i].astype(np.float64))
                K = np.ones((k + 1, k + 1))
                K[:k, :k] = A
                K[k, k] = 0
                y = np.linalg.solve(K, np.concatenate([b, [1.0]]))
                ref[p, j] = y[:k] @ v_np[nb[i]].astype(np.float64)
        out["max_abs_err_vs_oracle_sample"] = float(np.abs(est - ref).max())
    return out


def _ref_threads():
    return max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 16))


def _timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def _cpu(kind, fn, unit_count, unit, sample):
    from oracle import oracle as O
    if not O.ref_available():
        return None
    t0 = time.perf_counter()
    fn()
    t = time.perf_counter() - t0
    return {"value": unit_count / t, "unit": unit, "cores": _ref_threads(), "kind": kind, "sample": sample,
            "seconds": t}


def side_workload(name, args):
    """One JSON line for a workload other than the headline one."""
    import numpy as np
    import torch
    from geo_bsd import CovarianceModel, device as D, synthetic as S
    from oracle import oracle as O
    dev = torch.device("cuda:0")
    sync = torch.cuda.synchronize
    th = _ref_threads()
    line = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic (geo_bsd.synthetic)"}
    per = {}
    if name in ("mik", "wells", "lvm"):
        wl = {"mik": S.mik_c3, "wells": S.ok_wells_c4, "lvm": lambda: S.lvm_c1(4.0)}[name]()
        model = CovarianceModel(**wl.cov)
        m = torch.from_numpy(wl.mask).to(dev)
        cov = O.make_cov(wl.cov["type"], wl.cov["ranges"], wl.cov["angles"], wl.cov["sill"], wl.cov["nugget"])
        if name == "mik":
            z = torch.from_numpy(wl.values).to(dev)
            call = lambda prec: D.indicator_cdf(z, m, wl.shape, wl.thresholds, wl.levels, wl.radiuses, wl.k, model,
                                                precision=prec, variance=False)
            tune = lambda prec: D.autotune(2, wl.shape, wl.radiuses, wl.k, model, m, zfield=z, thresholds=wl.thresholds,
                                           marginal=wl.levels, precision=prec, order_relations=True, variance=False,
                                           reps=2)
            units = wl.mask.size
            ind = [(wl.values <= t).astype(np.uint8) for t in wl.thresholds]
            ref = lambda: [O.ref_median_ik(cov, wl.radiuses, wl.k, wl.shape, i, wl.mask, (1 - p, p), threads=th)
                           for i, p in zip(ind, wl.levels)]
            sample = f"{len(ind)} reference median_ik runs, one per threshold indicator, full grid"
        elif name == "wells":
            v = torch.from_numpy(wl.values).to(dev)
            call = lambda prec: D.krige(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec,
                                        variance=False)
            tune = lambda prec: D.autotune(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec,
                                           variance=False, reps=2)
            units = int((wl.mask == 0).sum())
            ref = lambda: O.ref_ordinary_kriging(cov, wl.radiuses, wl.k, wl.shape, wl.values, wl.mask, threads=th)
            sample = "reference ordinary_kriging, full grid"
        else:
            v = torch.from_numpy(wl.values).to(dev)
            lm = torch.from_numpy(wl.lvm_mean).to(dev)
            call = lambda prec: D.lvm_kriging(v, m, lm, wl.shape, wl.radiuses, wl.k, model, precision=prec,
                                              variance=False)
            tune = lambda prec: D.autotune(1, wl.shape, wl.radiuses, wl.k, model, m, values=v, lvm_mean=lm,
                                           precision=prec, variance=False, reps=2)
            units = int((wl.mask == 0).sum())
            ref = lambda: O.ref_lvm_kriging(cov, wl.radiuses, wl.k, wl.shape, wl.values, wl.mask, wl.lvm_mean,
                                            threads=th)
            sample = "reference lvm_kriging, full grid"
        # neighbour counts and search paths (untimed) for the main solve
        # kernel's algorithmic flops: with the LDS table (median IK, LVM) the
        # template-path cells of every bucket; without it (wells) every cell
        # of the 25-32 bucket, which the 4-lane 32-neighbour kernel solves
        nbk = dict(zfield=z, thresholds=wl.thresholds, marginal=wl.levels) if name == "mik" else dict(values=v)
        nb = D.krige(2 if name == "mik" else (1 if name == "lvm" else 0), wl.shape, wl.radiuses, wl.k, model, m,
                     neighbours=True, variance=False, **nbk)
        sync()
        counts, cl = nb["count"].cpu().numpy(), nb["cluster_path"].cpu().numpy()
        cnt = np.maximum(counts, 0)
        if name == "wells":
            main = cnt >= 25
            kernel = "k_solve_sorted<32,4,{},OK>"
        else:
            main = (counts > 0) & ~cl
            kernel = "k_solve_sorted_lds<{},4,{{}},{}>".format(16 if name == "mik" else 24,
                                                                "MIK" if name == "mik" else "SK")
        flops = float((flops_mik(cnt[main], len(wl.thresholds)) if name == "mik" else flops_ok(cnt[main])).sum())
        kern_ms, tuning = {}, {}
        for prec in ("fp64", "mixed"):
            if not args.no_autotune:
                _, tuning[prec] = tune(prec)
            per[prec] = _timed(lambda: call(prec), args.steps, args.warmup, sync)
            D.set_profiling(True)
            _timed(lambda: call(prec), args.steps, 0, sync)
            D.set_profiling(False)
            kern_ms[prec] = D.profile_summary()
        best = min(per, key=per.get)
        line.update(metric=f"kriged cells/s on 1 MI355X ({name})", value=units / per[best], unit="kriged cells/s",
                    ms_per_step=per[best] * 1e3, dtype="f64" if best == "fp64" else "f32+f64",
                    config={"workload": wl.name, "cells": int(wl.mask.size), "kriged_cells": units,
                            "best_precision": best},
                    fp64_ms=per["fp64"] * 1e3, mixed_ms=per["mixed"] * 1e3,
                    kernels_ms={p: {"search": k["search"], "solve_main": k["solve_main"]} for p, k in kern_ms.items()},
                    roofline=main_kernel_roofline(flops, kern_ms[best]["solve_main"],
                                                  PEAK_FP64_TFLOPS if best == "fp64" else PEAK_FP32_TFLOPS,
                                                  kernel.format("fp64" if best == "fp64" else "mixed")),
                    main_kernel_cells=int(main.sum()))
        if tuning:
            line["autotune"] = tuning
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = _cpu("reference", ref, units, "kriged cells/s", sample)
    elif name == "ik":
        wl = S.ik_c3()
        cats = [dict(cov_model=CovarianceModel(**c["cov"]), radiuses=c["radiuses"], max_neighbours=c["k"])
                for c in wl.categories]
        d = torch.from_numpy(wl.values).to(dev)
        m = torch.from_numpy(wl.mask).to(dev)
        for prec in ("fp64", "mixed"):
            per[prec] = _timed(lambda: D.indicator_kriging(d, m, wl.shape, cats, wl.marginal, precision=prec),
                               args.steps, args.warmup, sync)
        best = min(per, key=per.get)
        units = wl.mask.size
        line.update(metric="cells/s on 1 MI355X (indicator kriging, 4 categories)", value=units / per[best],
                    unit="cells/s", ms_per_step=per[best] * 1e3, dtype="f64" if best == "fp64" else "f32+f64",
                    config={"workload": wl.name, "cells": units, "categories": 4, "distinct_models": 3,
                            "best_precision": best},
                    fp64_ms=per["fp64"] * 1e3, mixed_ms=per["mixed"] * 1e3)
        if not args.no_cpu_baseline:
            ocats = [dict(cov=O.make_cov(c["cov"]["type"], c["cov"]["ranges"], c["cov"]["angles"], c["cov"]["sill"],
                                         c["cov"]["nugget"]), radiuses=c["radiuses"], max_nb=c["k"], marginal=p)
                     for c, p in zip(wl.categories, wl.marginal)]
            line["cpu_baseline"] = _cpu("reference", lambda: O.ref_indicator_kriging(wl.shape, wl.values, wl.mask,
                                                                                      ocats, threads=th),
                                        units, "cells/s", "reference indicator_kriging, full grid")
    elif name == "io":
        import tempfile
        import geo_bsd
        n = 4_000_000
        rng = np.random.default_rng(1)
        f = (0.2 + 0.05 * rng.standard_normal(n)).astype(np.float32)
        mk = (rng.random(n) > 0.01).astype(np.uint8)
        tmp = tempfile.mkdtemp()
        path = os.path.join(tmp, "prop.inc")
        df, dm = torch.from_numpy(f).to(dev), torch.from_numpy(mk).to(dev)
        per["write"] = _timed(lambda: D.write_inc(path, df, dm, -99.0, "PROP"), args.steps, args.warmup, sync)
        size = os.path.getsize(path)
        per["read"] = _timed(lambda: D.read_inc_float(path, -99.0, n), args.steps, args.warmup, sync)
        per["geo_read"] = _timed(lambda: geo_bsd.read_inc_file_float(path, -99, n), 3, 1, sync)
        per["geo_write"] = _timed(lambda: geo_bsd.write_property(geo_bsd.ContProperty(f, mk), path, "PROP", -99.0),
                                  3, 1, sync)
        line.update(metric="property values/s read from an INC file into HBM on 1 MI355X",
                    value=n / per["read"], unit="values/s", ms_per_step=per["read"] * 1e3, dtype="u8->f32",
                    config={"workload": f"INC file of {n} float values (%E text, {size / 1e6:.1f} MB), file in the "
                                        "page cache", "values": n, "bytes": size},
                    device_write_ms=per["write"] * 1e3, geo_read_ms=per["geo_read"] * 1e3,
                    geo_write_ms=per["geo_write"] * 1e3)
        if not args.no_cpu_baseline and O.ref_available():
            ref_path = os.path.join(tmp, "ref.inc")
            wr = _cpu("reference", lambda: O.ref_write(ref_path, "PROP", mk, floats=f, undefined=-99.0), n,
                      "values/s", "reference property_writer_t (fprintf %E), 1 thread")
            rd = _cpu("reference", lambda: O.ref_read_inc(ref_path, n), n, "values/s",
                      "reference read_inc_file_float (fscanf/sscanf), 1 thread")
            rd["cores"] = wr["cores"] = 1
            line["cpu_baseline"] = rd
            line["cpu_baseline_write"] = wr
            line["files_identical"] = open(ref_path, "rb").read() == open(path, "rb").read()
    elif name == "variogram":
        from geo_bsd import cvariogram as CV
        wl = S.mik_c3()
        f = np.where(wl.mask == 1, wl.values, np.float32(0)).astype(np.float32)
        templ = CV.VariogramSearchTemplate(2.0, 2.0, 1.0, 20, 0.0, CV.Ellipsoid(40.0, 3.0, 1.5, 30.0, 0.0, 0.0))
        df, dm = torch.from_numpy(f).to(dev), torch.from_numpy(wl.mask).to(dev)
        per["grid"] = _timed(lambda: CV.calc_variograms_device(templ, df, dm, wl.shape), args.steps, args.warmup,
                             sync)
        lw = S.loo_c5(nsets=1)
        npts = lw.xyz.shape[0]
        pts = torch.from_numpy(np.concatenate([lw.xyz, lw.values[:, None]], 1).astype(np.float32)).to(dev)
        per["points"] = _timed(lambda: CV.calc_variograms_points_device(templ, pts), max(1, args.steps // 5), 1, sync)
        ninf = int(wl.mask.sum())
        g = CV.calc_variograms_device(templ, df, dm, wl.shape).cpu().numpy()
        line.update(metric="hard-data cells/s through the grid experimental variogram on 1 MI355X",
                    value=ninf / per["grid"], unit="informed cells/s", dtype="f64", ms_per_step=per["grid"] * 1e3,
                    config={"workload": f"semivariogram of {ninf} hard data on a 200x200x100 grid, 20 lags of 2 "
                                        "cells, azimuth 30, tunnel half-widths 3 x 1.5; and of 20,000 scattered "
                                        "points", "informed_cells": ninf},
                    points_ms=per["points"] * 1e3, point_pairs_per_s=npts * npts / per["points"],
                    lags=[float(x) for x in g])
        if not args.no_cpu_baseline and O.ref_available():
            e = templ.templ.ellipsoid
            t = dict(dirs=np.array([list(e.direction1.data), list(e.direction2.data), list(e.direction3.data)]),
                     R2=3.0, R3=1.5, lag_width=2.0, lag_sep=2.0, tol=1.0, num_lags=20, first=0.0)
            t0 = time.perf_counter()
            rg = O.ref_vario_grid(t, f, wl.mask, wl.shape)
            tg = time.perf_counter() - t0
            line["cpu_baseline"] = {"value": ninf / tg, "unit": "informed cells/s", "cores": 1, "kind": "reference",
                                    "sample": "reference calc_variograms on the full grid, 1 thread (as it runs)",
                                    "seconds": tg}
            line["max_rel_diff_vs_reference"] = float(np.max(np.abs(g - rg) / np.maximum(np.abs(rg), 1e-30)))
            sub = 3000
            rp = _cpu("reference", lambda: O.ref_vario_points(t, lw.xyz[:sub, 0], lw.xyz[:sub, 1], lw.xyz[:sub, 2],
                                                              lw.values[:sub]),
                      sub * sub, "pairs/s", f"reference calc_variograms_from_point_set on {sub} points, 1 thread")
            rp["cores"] = 1
            line["cpu_baseline_points"] = rp
    elif name == "sgs":
        # Sequential Gaussian simulation of the 100x100x50 grid (5 % hard data
        # as normal scores), simple kriging, k = 24, radiuses (20, 20, 4).
        wl = S.ok_c2()
        inf = wl.mask == 1
        z = ((wl.values - wl.values[inf].mean()) / wl.values[inf].std()).astype(np.float32)
        v0 = np.where(inf, z, np.float32(0)).astype(np.float32)
        model = CovarianceModel(**wl.cov)
        seed = 20240611
        dv0, dm0 = torch.from_numpy(v0).to(dev), torch.from_numpy(wl.mask).to(dev)
        dv, dm = dv0.clone(), dm0.clone()

        def run():
            dv.copy_(dv0)
            dm.copy_(dm0)
            D.sgs_simulation(dv, dm, wl.shape, wl.radiuses, wl.k, model, seed, "sk")

        t = _timed(run, args.steps, args.warmup, sync)
        nsim = int((~inf).sum())
        line.update(metric="simulated cells/s of sequential Gaussian simulation on 1 MI355X", value=nsim / t,
                    unit="simulated cells/s", dtype="f64", ms_per_step=t * 1e3,
                    config={"workload": "SGS of a 100x100x50 grid with 5% hard data (normal scores), simple "
                                        "kriging, exponential (20,20,4), nugget 0.05, radiuses (20,20,4), "
                                        "max_neighbours 24, the reference's path and draws for seed 20240611",
                            "simulated_cells": nsim})
        # several realizations advancing together (one launch per level
        # index over all of them): seeds seed .. seed + R - 1
        R = args.realizations
        if R > 1:
            seeds = [seed + i for i in range(R)]

            def run_batch():
                return D.sgs_realizations(dv0, dm0, wl.shape, wl.radiuses, wl.k, model, seeds, "sk")

            tb = _timed(run_batch, max(2, args.steps // 4), 1, sync)
            bv, _ = run_batch()
            dv.copy_(dv0)
            dm.copy_(dm0)
            D.sgs_simulation(dv, dm, wl.shape, wl.radiuses, wl.k, model, seeds[-1], "sk")
            sync()
            line["realizations"] = {"count": R, "value": R * nsim / tb, "unit": "simulated cells/s",
                                    "ms_per_call": tb * 1e3,
                                    "last_equals_single_call": bool(torch.equal(bv[-1], dv))}
        # agreement with the oracle's one-by-one walk, and its speed, on a
        # 100x100x10 slab (the reference's own simulation cannot be built:
        # its path generator needs a prime table it does not ship)
        if not args.no_cpu_baseline:
            sub = (100, 100, 10)
            ns = int(np.prod(sub))
            sv, smask = v0[:ns].copy(), wl.mask[:ns].copy()
            cov = O.make_cov(wl.cov["type"], wl.cov["ranges"], wl.cov["angles"], wl.cov["sill"], wl.cov["nugget"])
            t0 = time.perf_counter()
            ov, om, order = O.sgs(1, cov, wl.radiuses, wl.k, sub, sv, smask, seed)
            tc = time.perf_counter() - t0
            gv, gm = torch.from_numpy(sv).to(dev), torch.from_numpy(smask).to(dev)
            D.sgs_simulation(gv, gm, sub, wl.radiuses, wl.k, model, seed, "sk")
            sync()
            line["max_abs_diff_vs_oracle_slab"] = float(np.abs(gv.cpu().numpy() - ov).max())
            line["cpu_baseline"] = {"value": order.size / tc, "unit": "simulated cells/s", "cores": 1,
                                    "kind": "port",
                                    "sample": f"the oracle's sequential walk over a 100x100x10 slab ({order.size} "
                                              "cells), 1 thread", "seconds": tc}
    else:
        raise SystemExit(f"unknown workload {name}")
    return line


def oracle_est64(wl):
    from oracle import oracle as O
    cov = O.make_cov(wl.cov["type"], wl.cov["ranges"], wl.cov["angles"], wl.cov["sill"], wl.cov["nugget"])
    return O.krige(0, cov, wl.radiuses, wl.k, wl.shape, wl.mask, values=wl.values, want_neighbours=False)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-loo", action="store_true", help="skip the leave-one-out kriging figures")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch settings (profiled runs: no tuning trials among the kernels)")
    ap.add_argument("--workload", default="ok", choices=["ok", "mik", "wells", "ik", "lvm", "io", "variogram", "sgs"],
                    help="ok (default): the headline OK 100x100x50 workload; others: see the module docstring")
    ap.add_argument("--realizations", type=int, default=32,
                    help="sgs workload: also time this many realizations run together (1 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the run, one device-to-device copy of MIB MiB (a kernel of known traffic for "
                         "scaling memory counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported (the kriging path runs on a "
              "single GPU; see DESIGN.md for how it would shard).", file=sys.stderr)
        return 2
    if args.workload != "ok":
        print(json.dumps(side_workload(args.workload, args)), flush=True)
        return 0

    import numpy as np
    import torch
    from geo_bsd import CovarianceModel, device as D, synthetic as S

    wl = S.ok_c2()
    model = CovarianceModel(**wl.cov)
    dev = torch.device("cuda:0")
    v = torch.from_numpy(wl.values).to(dev)
    m = torch.from_numpy(wl.mask).to(dev)
    kriged = int((wl.mask == 0).sum())
    stream = torch.cuda.current_stream(dev)

    # Neighbour counts and search paths for the algorithmic flop count (untimed).
    nb = D.krige(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, neighbours=True, variance=False)
    torch.cuda.synchronize()
    counts = nb["count"].cpu().numpy()
    cluster = nb["cluster_path"].cpu().numpy()
    flops_all = float(flops_ok(counts[counts > 0]).sum())
    # the main solve kernel takes the template-path cells; cluster-path cells
    # go to a second, list-driven launch
    main_cells = (counts > 0) & ~cluster
    flops_main = float(flops_ok(counts[main_cells]).sum())

    results = {}
    tuning = {}
    for prec in ("fp64", "mixed"):
        # launch settings for this request, found by timing (kept for the steps)
        if not args.no_autotune:
            _, tuning[prec] = D.autotune(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec,
                                         variance=True, reps=3)
        for _ in range(args.warmup):
            D.krige(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec, variance=True)
        torch.cuda.synchronize()
        # Timed region: the dispatches of every step's search kernel and main
        # solve kernel record start/stop HIP events on the launch stream.
        D.set_profiling(True)
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            D.krige(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec, variance=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        D.set_profiling(False)
        prof = D.profile_summary()
        assert prof["calls"] == args.steps, prof
        ms = ev0.elapsed_time(ev1) / args.steps
        stats = D.last_call_stats()
        est = D.krige(0, wl.shape, wl.radiuses, wl.k, model, m, values=v, precision=prec, est64=True,
                      status=True)
        torch.cuda.synchronize()
        results[prec] = dict(ms=ms, wall_ms=wall * 1e3 / args.steps, prof=prof, est64=est["est64"].cpu().numpy(),
                             fallbacks=stats["fp64_fallbacks"], cluster_cells=stats["cluster_path_cells"])

    err = {}
    if not args.no_oracle:
        o = oracle_est64(wl)
        kr = o["status"] == 0
        for prec in results:
            err[prec] = float(np.abs(results[prec]["est64"][kr] - o["est64"][kr]).max())

    def phase(r):
        p = r["prof"]
        return {"ms_per_step": r["ms"], "host_wall_ms_per_step": r["wall_ms"], "search_kernel_ms": p["search"],
                "solve_kernel_ms": p["solve_main"]}

    def roof(r, peak, kernel):
        achieved = flops_main / (r["prof"]["solve_main"] / 1e3) / 1e12
        return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": TRAFFIC_BYTES.get(kernel), "algorithmic_bytes": ALGO_BYTES,
                "flops_per_launch": flops_main, "launch_ms": r["prof"]["solve_main"],
                "hbm_gbs": (TRAFFIC_BYTES.get(kernel) or 0) / (r["prof"]["solve_main"] / 1e3) / 1e9}

    mixed = results["mixed"]
    fp64 = results["fp64"]
    k_mixed = "k_solve_sorted_lds<24,4,mixed,OK>"
    k_fp64 = "k_solve_sorted_lds<24,4,fp64,OK>"
    line = {
        "metric": "kriged cells/s on 1 MI355X: FP32+refine vs FP64, at <=1e-6 estimate error",
        "value": kriged / (mixed["ms"] / 1e3),
        "unit": "kriged cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded moving-average field, 5% hard data)",
        "config": {"workload": "ordinary kriging 100x100x50 (500k cells, 475k kriged), exponential (20,20,4), "
                               "nugget 0.05, radiuses (20,20,4), max_neighbours 24", "cells": 500000,
                   "kriged_cells": kriged,
                   "solver": "FP32 Cholesky + 1 FP64 residual refinement step, FP64 re-solve on small pivots"},
        "mixed": dict(phase(mixed), max_abs_err_vs_oracle=err.get("mixed"), fp64_fallbacks=mixed["fallbacks"],
                      cluster_path_cells=mixed["cluster_cells"]),
        "fp64": dict(phase(fp64), value=kriged / (fp64["ms"] / 1e3), max_abs_err_vs_oracle=err.get("fp64"),
                     roofline=roof(fp64, PEAK_FP64_TFLOPS, k_fp64)),
        "speedup_mixed_vs_fp64": fp64["ms"] / mixed["ms"],
        "autotune": tuning,
        "steady_state": "device buffers resident; the parameter set's plan (template, tables) is built on the "
                        "first call and reused by every step, as it is by the cold_call figure's later calls",
        "roofline": roof(mixed, PEAK_FP32_TFLOPS, k_mixed),
        "flops_per_step": flops_all,
    }
    line.update(host_call_figures(wl, model))
    if not args.no_loo:
        line["loo"] = loo_bench(torch, not args.no_oracle)
    if args.calibrate_copy > 0:
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["calibration_copy_bytes"] = src.numel() * 4
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
