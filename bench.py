#!/usr/bin/env python3
"""Benchmark of the MI355X SpMV hot path (FP64 drop-in Morpheus::multiply).

Workload (one "step"): one FP64 SpMV y = A x in each of the six formats on
its benchmark matrix, matrices and vectors resident in HBM:
  * CSR, COO, HYB on Z15: 1M rows, power-law(1.5) row lengths, mean 16,
    capped at 20k, uniform random columns, N(0,1) values;
  * DIA, HDC, ELL on S7: 3D 7-point stencil on a 128^3 grid.
ELL on Z15 is inadmissible (the reference converter throws
FormatConversionException; the library reproduces that), so ELL runs on S7.
The formats are produced on the GPU by the library's own Convert from CSR.

value = total FLOPs of the step (2 * nnz per SpMV) / step time.  Every SpMV
is bracketed by HIP events on the stream it is launched on; per-format
averages give the roofline figures (algorithmic bytes / kernel time).
The CPU baseline is the OpenMP restatement of the reference multiply timed
on the host cores on the same six (format, matrix) pairs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "SpMV GFLOP/s & % HBM roofline per format; batched SpMM FP32/BF16 vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=2.0,
                    help="CPU baseline time budget per (format, matrix) pair")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched SpMM section (512 matrices x 8 RHS)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-batch-plain", action="store_true",
                    help="skip the plain FP64 forms beside the batch (per-matrix SpMM / SpMV "
                         "loops and their hipGraphs): counter passes")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (PMC traffic calibration)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default launch shapes (no autotuning)")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the FP64 products on S27 (3D 27-point, 160^3) and Z12 "
                         "(2M rows, power-law(1.2), mean 30), the large parity configs")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the format x precision sweep (FP64 / FP32 / BF16 on S7, Z15, B9)")
    ap.add_argument("--only", default="",
                    help="run one section alone and print its JSON (profiling): batch, "
                         "large (FP64 formats on S27 and Z12), large-S27, large-Z12, p2, "
                         "spmm (FP64 CSR SpMM on Z12, k = 1..16)")
    ap.add_argument("--formats", default="",
                    help="comma-separated subset of the step's formats (profiling one "
                         "format at a time; the JSON line then covers that subset)")
    return ap.parse_args()


def product_kernels(fmt, tuning):
    """Kernels one FP64 product launches for `fmt` on the bench matrices, in
    launch order, with the launch shape `tuning` picked: Z15's long rows take
    the x-tiled pass; row-sorted COO runs the row-block plan unless a
    segmented shape (coo_items) was picked; Z15's HYB (ELL slots about 10 %
    filled) runs over its merged CSR copy; S7's HDC has an empty CSR part;
    S7's DIA stages its in-plane diagonals' x entries in LDS (the default
    launch shape) or takes the plain kernel (a rows_per_thread shape)."""
    rpt = 2 if tuning.get("rows_per_thread") == 2 else 1
    dia = (["dia_staged_kernel<double>"] if not tuning.get("rows_per_thread")
           else [f"dia_kernel<double, 1, {rpt}>"])
    items = tuning.get("coo_items") or 0
    xt = ["xtile_kernel<double, false>", "xtile_combine_kernel<double, 0, false>"]
    rows = ["rowblock_spmv_kernel<double, 0, false>", *xt]
    return {
        "CSR": rows,
        "COO": rows if items == 0 else [
            "coo_zero_kernel<double>", *xt, f"coo_segmented_kernel<double, {items}, false>",
            f"coo_fixup_kernel<double, {items}, false>"],
        "HYB": rows,
        "HDC": dia,
        "DIA": dia,
        "ELL": [f"ell_kernel<double, 1, {rpt}>"],
    }[fmt]


def build_workload(only=None, tune=True):
    import torch
    from morpheus_amd import capi, device
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import (COO_FORMAT, DIA_FORMAT, ELL_FORMAT, HDC_FORMAT,
                                      HYB_FORMAT)

    t0 = time.time()
    z15 = gen.Z15()
    s7 = gen.S7()
    log(f"generated Z15 (nnz {z15.nnz}) and S7 (nnz {s7.nnz}) in {time.time() - t0:.1f}s")
    dz = device.to_device(z15, "f64")
    ds = device.to_device(s7, "f64")
    mats = {
        "CSR": (dz, z15),
        "COO": (device.convert(dz, COO_FORMAT), z15),
        "HYB": (device.convert(dz, HYB_FORMAT), z15),
        "DIA": (device.convert(ds, DIA_FORMAT), s7),
        "HDC": (device.convert(ds, HDC_FORMAT), s7),
        "ELL": (device.convert(ds, ELL_FORMAT), s7),
    }
    try:
        device.convert(dz, ELL_FORMAT)
        ell_z15 = "converted"
    except capi.FormatConversionException:
        ell_z15 = "FormatConversionException (as the reference)"
    torch.cuda.synchronize()
    # interleave matrices so consecutive products never share matrix data
    order = ["CSR", "DIA", "COO", "ELL", "HYB", "HDC"]
    if only:
        order = [n for n in order if n in only]
    vecs = {}
    tuned = {}
    for name in order:
        D, host = mats[name]
        x = torch.empty(D.ncols, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0)
        y = torch.empty(D.nrows, dtype=torch.float64, device="cuda")
        D.analyse()
        # the launch shape of each handle is tuned once, by timed trials, as
        # part of its analysis (mph_matrix_autotune)
        if tune:
            device.autotune_launch(D, 1, 10)
            tuned[name] = {k: v for k, v in D.get_tuning().items() if v}
        vecs[name] = (x, y)
    torch.cuda.synchronize()
    return mats, vecs, order, ell_z15, (z15, s7), tuned


def run_gpu(mats, vecs, order, steps, warmup):
    import torch
    stream = torch.cuda.current_stream()

    def step(evs=None):
        for name in order:
            D, _ = mats[name]
            x, y = vecs[name]
            if evs is not None:
                evs[name][0].record(stream)
            D.multiply(x, y, True, stream)
            if evs is not None:
                evs[name][1].record(stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    per = {n: [] for n in order}
    evlist = [{n: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for n in order} for _ in range(steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for k in range(steps):
        step(evlist[k])
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = t_start.elapsed_time(t_end)
    for k in range(steps):
        for n in order:
            per[n].append(evlist[k][n][0].elapsed_time(evlist[k][n][1]))
    return total_ms, wall, per


def cold_and_graph(mats, vecs, order, reps):
    """Per format, two more figures for the same product: the median over
    `reps` products each started after the 256 MB MALL is flushed (a 1 GiB
    read probe pass between products, outside the events: clean lines, so the
    product does not also pay for write-backs), and the mean of `reps`
    back-to-back products replayed from one hipGraph (warm caches, no host
    launch gaps)."""
    import numpy as np
    import torch
    stream = torch.cuda.current_stream()
    scratch = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    sink = torch.zeros(1, dtype=torch.float64, device="cuda")
    cold, graph = {}, {}
    for n in order:
        D, _ = mats[n]
        x, y = vecs[n]
        evs = []
        for r in range(reps):
            probe_read(scratch, scratch.numel() * 8, sink, stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            D.multiply(x, y, True, stream)
            e1.record(stream)
            evs.append((e0, e1))
        torch.cuda.synchronize()
        cold[n] = float(np.median([a.elapsed_time(b) for a, b in evs]))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(reps):
                D.multiply(x, y, True)
        g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        g.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        graph[n] = e0.elapsed_time(e1) / reps
        del g
    del scratch, sink
    return cold, graph


def run_large(reps=10, autotune=True, matrices=("S27", "Z12")):
    """FP64 products on the two large configurations (S27: 3D 27-point
    stencil on 160^3, 4.1M rows, 109M nnz; Z12: 2M rows, power-law(1.2) row
    lengths, mean 30, cap 200k), each format the reference admits: time
    (median of `reps` back-to-back products), fraction of 8 TB/s on the
    algorithmic and on the plan's bytes and of a same-size read stream, the
    FP64 per-row bound against the oracle, and the cost of switching: the
    on-GPU Convert from CSR and the analysis of the new handle (wall clock,
    synchronised, allocations included)."""
    import numpy as np
    import torch
    import oracle  # the oracle: FP64 reference products and row scales
    from morpheus_amd import capi, device
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import FORMAT_NAMES
    stream = torch.cuda.current_stream()
    out = {}
    for mname, make in (("S27", gen.S27), ("Z12", gen.Z12)):
        if mname not in matrices:
            continue
        t0 = time.time()
        A = make()
        log(f"large: generated {mname} (nnz {A.nnz}) in {time.time() - t0:.1f}s")
        x = gen.random_vector(A.ncols, 31)
        ref = oracle.spmv(A, x)
        bound = 4.0 * np.maximum(np.diff(A.row_offsets), 1) * 2.0 ** -53 * oracle.absrow(A, x)
        xt = torch.from_numpy(x).cuda()
        y = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
        base = device.to_device(A, "f64")
        out[mname] = {"nnz": int(A.nnz)}
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        base.analyse()
        torch.cuda.synchronize()
        csr_analyse_ms = (time.perf_counter() - w0) * 1e3
        for f, fname in enumerate(FORMAT_NAMES):
            convert_ms = 0.0
            try:
                if f == base.format:
                    D = base
                else:
                    torch.cuda.synchronize()
                    w0 = time.perf_counter()
                    D = device.convert(base, f)
                    torch.cuda.synchronize()
                    convert_ms = (time.perf_counter() - w0) * 1e3
            except capi.FormatConversionException:
                out[mname][fname] = "FormatConversionException"
                continue
            if D is base:
                analyse_ms = csr_analyse_ms
            else:
                w0 = time.perf_counter()
                D.analyse()
                torch.cuda.synchronize()
                analyse_ms = (time.perf_counter() - w0) * 1e3
            D.multiply(xt, y, True, stream)
            torch.cuda.synchronize()
            ok = bool(np.all(np.abs(y.cpu().numpy() - ref) <= bound))
            ts = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                D.multiply(xt, y, True, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            ms = float(np.median(ts))
            nbytes = D.algorithmic_bytes()
            pbytes = D.plan_bytes()
            rs = read_stream_flushed(nbytes, 5)
            out[mname][fname] = {
                "ms": round(ms, 4),
                "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2),
                "bytes_alg": int(nbytes),
                "hbm_frac": round(data_gbs(nbytes, ms) / HBM_PEAK_GBS, 4),
                "bytes_plan": int(pbytes),
                "hbm_frac_plan": round(data_gbs(pbytes, ms) / HBM_PEAK_GBS, 4),
                "read_stream_same_size_gbs": round(rs, 1),
                "frac_of_read_stream": round(data_gbs(nbytes, ms) / rs, 4),
                "convert_from_csr_ms": round(convert_ms, 3),
                "analyse_ms": round(analyse_ms, 3),
                "within_fp64_bound": ok,
            }
            log(f"large {mname} {fname}: {ms * 1e3:.0f} us, "
                f"{out[mname][fname]['hbm_frac']:.2f} of 8 TB/s "
                f"({out[mname][fname]['hbm_frac_plan']:.2f} on the plan's bytes), "
                f"{out[mname][fname]['frac_of_read_stream']:.2f} of a read stream, bound {ok}; "
                f"convert {convert_ms:.1f} ms, analyse {analyse_ms:.1f} ms")
            if D is not base:
                D.close()
        base.close()
        if autotune:
            out[mname]["autotuned"] = run_large_autotuned(A, mname)
        del A
    return out


def run_large_autotuned(A, mname, ks=(1, 4, 16), reps=5):
    """The DynamicMatrix switching run on one large matrix: per value type
    (FP64, FP32, BF16; int32 indices) the autotuner times every admissible
    format at k = 4 right-hand sides and picks one; the picked format's
    product with k in `ks` (row-major X / Y, U[-1,1]) is then timed (median
    of `reps` back to back) and its first column checked against the FP64
    oracle: the per-row FP64 bound, or the max-row relative error
    max_i |y_i - y_ref_i| / sum_j |a_ij x_j| for FP32 / BF16."""
    import numpy as np
    import torch
    import oracle  # the oracle: FP64 reference products and row scales
    from morpheus_amd import device
    from morpheus_amd.formats import FORMAT_NAMES
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(41)
    X = {k: rng.uniform(-1.0, 1.0, (A.ncols, k)) for k in ks}
    col0 = {k: np.ascontiguousarray(X[k][:, 0]) for k in ks}
    ref = {k: oracle.spmv_compensated(A, col0[k]) for k in ks}
    scale = {k: oracle.absrow(A, col0[k]) for k in ks}
    nrow = np.diff(A.row_offsets)
    res = {}
    for dtype in ("f64", "f32", "bf16"):
        t0 = time.time()
        D = device.to_device(A, dtype)
        best, trials = device.autotune_format(D, k=4, reps=3)
        M = D if best == D.format else device.convert(D, best)
        M.analyse()
        rec = {"format": FORMAT_NAMES[best],
               "trial_ms_k4": {FORMAT_NAMES[f]: round(t, 4) for f, t in trials.items()}}
        vdt = device.vector_dtype(dtype)
        for k in ks:
            Xt = torch.from_numpy(X[k] if k > 1 else col0[k]).cuda().to(vdt)
            Y = torch.empty((A.nrows, k) if k > 1 else (A.nrows,), dtype=vdt, device="cuda")
            run = (lambda: M.multiply(Xt, Y, True, stream)) if k == 1 else \
                (lambda: M.multiply_multi(Xt, Y, True, stream))
            run()
            torch.cuda.synchronize()
            y0 = (Y if k == 1 else Y[:, 0]).double().cpu().numpy()
            ts = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                run()
                e1.record(stream)
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            ms = float(np.median(ts))
            nbytes = M.algorithmic_bytes(k)
            mask = scale[k] > 0
            err = float(np.max(np.abs(y0[mask] - ref[k][mask]) / scale[k][mask])) if mask.any() else 0.0
            r = {"ms": round(ms, 4),
                 "gflops": round(2.0 * A.nnz * k / (ms * 1e-3) / 1e9, 2),
                 "matrhs_per_s": round(k / (ms * 1e-3), 1),
                 "bytes_alg": int(nbytes),
                 "hbm_frac": round(data_gbs(nbytes, ms) / HBM_PEAK_GBS, 4),
                 "max_rel_err_col0": err}
            if dtype == "f64":
                r["within_fp64_bound_col0"] = bool(np.all(
                    np.abs(y0 - ref[k]) <= 4.0 * np.maximum(nrow, 1) * 2.0 ** -53 * scale[k]))
            rec[f"k{k}"] = r
            del Xt, Y
        if M is not D:
            M.close()
        D.close()
        res[dtype] = rec
        log(f"large {mname} autotuned {dtype}: {rec['format']}; " + ", ".join(
            f"k={k} {rec[f'k{k}']['ms'] * 1e3:.0f} us ({rec[f'k{k}']['hbm_frac']:.2f} of 8 TB/s, "
            f"err {rec[f'k{k}']['max_rel_err_col0']:.1e})" for k in ks)
            + f" [{time.time() - t0:.1f}s]")
    for dtype in ("f32", "bf16"):
        res[dtype]["matrhs_per_s_vs_f64"] = {
            f"k{k}": round(res[dtype][f"k{k}"]["matrhs_per_s"] / res["f64"][f"k{k}"]["matrhs_per_s"], 3)
            for k in ks}
    res["tolerance"] = {"f32": 1e-5, "bf16": 1e-2}
    return res


def run_spmm_z12(ks=(1, 2, 4, 8, 16), reps=10):
    """FP64 CSR SpMM on Z12 (2M rows, power-law(1.2), 57.5M nnz; its long
    rows x-tiled) for k right-hand sides (row-major X / Y, U[-1, 1]): median
    of `reps` products back to back, the algorithmic bytes and fraction of
    8 TB/s, and column 0 against the per-row FP64 bound of the oracle.  With
    the value copy the x-tiled rows take the k-wide x-tiled pass (groups of
    4 / 2 columns) up to k = 8, the position chunks beyond."""
    import numpy as np
    import torch
    import oracle  # the oracle: FP64 reference products and row scales
    from morpheus_amd import device
    from morpheus_amd import generators as gen
    A = gen.Z12()
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(43)
    D = device.to_device(A, "f64")
    D.analyse()
    nrow = np.diff(A.row_offsets)
    res = {"matrix": "Z12", "nnz": int(A.nnz)}
    for k in ks:
        X = rng.uniform(-1.0, 1.0, (A.ncols, k))
        x0 = np.ascontiguousarray(X[:, 0])
        Xt = torch.from_numpy(X if k > 1 else x0).cuda()
        Y = torch.empty((A.nrows, k) if k > 1 else (A.nrows,), dtype=torch.float64, device="cuda")
        run = (lambda: D.multiply(Xt, Y, True, stream)) if k == 1 else \
            (lambda: D.multiply_multi(Xt, Y, True, stream))
        run()
        torch.cuda.synchronize()
        y0 = (Y if k == 1 else Y[:, 0]).cpu().numpy()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = float(np.median(ts))
        nbytes = D.algorithmic_bytes(k)
        ref = oracle.spmv(A, x0)
        bound = 4.0 * np.maximum(nrow, 1) * 2.0 ** -53 * oracle.absrow(A, x0)
        res[f"k{k}"] = {"ms": round(ms, 4), "bytes_alg": int(nbytes),
                        "hbm_frac": round(data_gbs(nbytes, ms) / HBM_PEAK_GBS, 4),
                        "within_fp64_bound_col0": bool(np.all(np.abs(y0 - ref) <= bound))}
        log(f"spmm Z12 k={k}: {ms * 1e3:.0f} us ({res[f'k{k}']['hbm_frac']:.2f} of 8 TB/s)")
        del Xt, Y
    D.close()
    return res


def probe_read(buf, nbytes, sink, stream):
    """One pass of the library's read probe (16-byte non-temporal loads,
    nothing written) over the first `nbytes` of `buf`."""
    from morpheus_amd import capi, device
    capi.call("mph_probe_read", buf.data_ptr(), int(nbytes), sink.data_ptr(),
              device.stream_handle(stream))


def read_stream_flushed(nbytes, reps=20):
    """The device's read rate at a given size, as the flushed products see
    it: one pass of the library's read probe (a hand-written kernel, 16-byte
    non-temporal loads, nothing written) over `nbytes`, median of `reps` runs
    each after the same 1 GiB MALL-flushing read sweep, outside the events.
    GB/s."""
    import numpy as np
    import torch
    stream = torch.cuda.current_stream()
    scratch = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    data = torch.ones(max(2, (nbytes + 15) // 8), dtype=torch.float64, device="cuda")
    sink = torch.zeros(1, dtype=torch.float64, device="cuda")
    probe_read(data, nbytes, sink, stream)
    evs = []
    for _ in range(reps):
        probe_read(scratch, scratch.numel() * 8, sink, stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        probe_read(data, nbytes, sink, stream)
        e1.record(stream)
        evs.append((e0, e1))
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
    del scratch, data
    return data_gbs(nbytes, ms)


def gather_ceiling(z15, reps=20):
    """The x-gather floor of Z15's short-row pass: the library's gather probe
    over the column indices of the rows of at most 512 entries (3.3M items;
    the longer rows take the x-tiled pass) into an 8 MB x, index loads
    coalesced, eight gathers in flight per thread, no row bookkeeping;
    median of `reps` runs back to back (steady state).  Returns us and the
    item count."""
    import numpy as np
    import torch
    from morpheus_amd import capi, device
    stream = torch.cuda.current_stream()
    L = np.diff(z15.row_offsets)
    keep = np.repeat(L <= 512, L)
    ci = torch.from_numpy(np.ascontiguousarray(z15.column_indices[keep], dtype=np.int32)).cuda()
    x = torch.empty(z15.ncols, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0)
    n = ci.numel()
    out = torch.empty((n + 7) // 8, dtype=torch.float64, device="cuda")
    run = lambda: capi.call("mph_probe_gather", ci.data_ptr(), n, x.data_ptr(), out.data_ptr(),
                            device.stream_handle(stream))
    run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run()
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts)) * 1e3, int(n)


def data_gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def run_cg(mats, iters):
    """The consumer side: `iters` fused CG iterations on S7 (the SPD 3-D
    Laplacian) with the DIA and ELL handles of the step, captured once per
    handle in a hipGraph of 10 iterations and replayed.  Bytes per iteration:
    the SpMV's bytes_alg plus 12 vector passes of N * 8 bytes (p.q: p, q;
    update: x, r, p, q read, x, r written; direction: r, p read, p written;
    the partial sums are a few KB)."""
    import numpy as np
    import torch
    from morpheus_amd import device
    out = {}
    for name in ("DIA", "ELL"):
        if name not in mats:
            continue
        D, host = mats[name]
        n = host.nrows
        b = torch.empty(n, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg = device.CG(D)
        cg.start(b, x)
        cg.iterate(x, 2)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cg.iterate(x, 10)
        g.replay()
        torch.cuda.synchronize()
        reps = max(1, iters // 10)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / (10 * reps)
        rr = cg.residual()
        nbytes = D.algorithmic_bytes() + 12 * n * 8
        out[name] = {"matrix": "S7", "iterations": 10 * reps, "ms_per_iteration": round(ms, 5),
                     "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "bytes_per_iteration": int(nbytes),
                     "relative_residual": float(np.sqrt(rr) / np.sqrt(float(torch.dot(b, b))))}
        log(f"CG on S7 ({name}): {ms * 1e3:.1f} us per iteration, "
            f"{out[name]['gbs']:.0f} GB/s, relative residual {out[name]['relative_residual']:.1e} "
            f"after {out[name]['iterations'] + 2} iterations")
        del g
        cg.close()
    return out or None


def run_cpu(mats, order, budget):
    import numpy as np
    import oracle  # the oracle's OpenMP restatement: cpu_baseline leg only

    res = {}
    for name in order:
        D, _ = mats[name]
        Ah = D.to_host()
        fn, _, _ = oracle.omp_spmv_fn(Ah)
        for _ in range(3):
            fn()
        times = []
        t_end = time.perf_counter() + budget
        while time.perf_counter() < t_end or len(times) < 5:
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
            if len(times) >= 200:
                break
        res[name] = (float(np.median(times)), len(times))
        log(f"cpu {name}: median {res[name][0] * 1e3:.2f} ms over {len(times)} reps")
    return res, oracle.omp_threads()


def run_batch(steps, plain=True):
    """Batched SpMM: 512 matrices (log-uniform 1k-64k rows; 2D 5-point
    stencils, power-law(1.5) rows with mean 12, banded with bandwidth 5-33),
    k = 8 row-major right-hand sides U[-1, 1]; formats chosen per matrix by
    the autotuner; one launch per product, replayed from a hipGraph.

    Per value type: time, algorithmic bytes (each member's stored arrays in
    its chosen format read once, X read once, Y written once; and the same
    at CSR-equivalent storage), GB/s and fraction of 8 TB/s, and the error of
    column 0 of every member against the oracle (FP64: the per-row bound
    4 n_row u sum_j |a_ij x_j|; FP32 / BF16: max_i |y_i - y_ref_i| /
    sum_j |a_ij x_j| against the compensated FP64 oracle).  The primary
    ratios are FP32 / BF16 against the same batched launch in FP64; the plain
    FP64 forms (one SpMM per matrix, one SpMV per matrix and column, eager
    and graph-replayed) are reported beside them."""
    import numpy as np
    import torch
    import oracle  # the oracle: FP64 reference products and row scales
    from morpheus_amd import device
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import FORMAT_NAMES

    K = 8
    t0 = time.time()
    mats = [A for _, A in gen.batch_set(512)]
    nnz = sum(A.nnz for A in mats)
    rows = sum(A.nrows for A in mats)
    log(f"batch: generated {len(mats)} matrices, {nnz} nnz in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(88)
    Xh = [rng.uniform(-1.0, 1.0, (A.ncols, K)) for A in mats]
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    # oracle column 0 of every member, and its row scales
    t0 = time.time()
    x0 = [np.ascontiguousarray(x[:, 0]) for x in Xh]
    ref64 = [oracle.spmv(A, x) for A, x in zip(mats, x0)]
    refc = [oracle.spmv_compensated(A, x) for A, x in zip(mats, x0)]
    scale = [oracle.absrow(A, x) for A, x in zip(mats, x0)]
    bound = [4.0 * np.maximum(np.diff(A.row_offsets), 1) * 2.0 ** -53 * sc
             for A, sc in zip(mats, scale)]
    log(f"batch: oracle column 0 of {len(mats)} members in {time.time() - t0:.1f}s")

    t0 = time.time()
    D64 = [device.to_device(A, "f64") for A in mats]
    X64 = [torch.from_numpy(x).cuda() for x in Xh]
    for D in D64:
        D.analyse()
    # format choice per matrix, by timed trials of the FP32 product with k = 8
    Dcsr32 = [device.to_device(A, "f32") for A in mats]
    choice = [device.autotune_format(D, k=K, reps=3)[0] for D in Dcsr32]
    hist = {FORMAT_NAMES[f]: choice.count(f) for f in sorted(set(choice))}
    log(f"batch: autotuned formats {hist} in {time.time() - t0:.1f}s")

    res = {"matrices": len(mats), "k": K, "rows": int(rows), "nnz": int(nnz),
           "formats_chosen": hist}
    for dtype in ("f64", "f32", "bf16"):
        base = D64 if dtype == "f64" else (Dcsr32 if dtype == "f32" else
                                            [device.to_device(A, "bf16") for A in mats])
        F = [device.convert(D, f) if f != D.format else D for D, f in zip(base, choice)]
        vdt = device.vector_dtype(dtype)
        X = [x.to(vdt) for x in X64]
        Y = [torch.empty((D.nrows, K), dtype=vdt, device="cuda") for D in F]
        B = device.Batch(F, K, X, Y)
        ntiles, ncomp = B.info()
        ms = timed(lambda: B.run_graph(True, stream), steps)
        nbytes = sum(D.algorithmic_bytes(K) for D in F)
        csr_bytes = sum(D.algorithmic_bytes(K) for D in base)
        gbs = data_gbs(nbytes, ms)
        rec = {"ms": round(ms, 4),
               "matrhs_per_s": round(len(mats) * K / (ms * 1e-3), 1),
               "gflops": round(2.0 * nnz * K / (ms * 1e-3) / 1e9, 1),
               "bytes_alg": int(nbytes), "gbs": round(gbs, 1),
               "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
               "bytes_csr_equiv": int(csr_bytes),
               "hbm_frac_csr_equiv": round(data_gbs(csr_bytes, ms) / HBM_PEAK_GBS, 4),
               "tiles": int(ntiles), "dot2_rows": int(ncomp)}
        worst, outside = 0.0, 0
        for A, y, r64, rc, sc, bd in zip(mats, Y, ref64, refc, scale, bound):
            y0 = y[:, 0].double().cpu().numpy()
            if dtype == "f64":
                outside += int(np.count_nonzero(~(np.abs(y0 - r64) <= bd)))
            m = sc > 0
            if m.any():
                worst = max(worst, float(np.max(np.abs(y0[m] - rc[m]) / sc[m])))
        rec["max_rel_err_col0"] = worst
        if dtype == "f64":
            rec["rows_outside_fp64_bound_col0"] = outside
        res[dtype] = rec
        B.close()
        for D, Db in zip(F, base):
            if D is not Db:
                D.close()
        if dtype != "f64":
            for D in base:
                D.close()
        log(f"batch {dtype}: {ms * 1e3:.1f} us per product of {len(mats)} matrices x {K} RHS, "
            f"{gbs:.0f} GB/s ({rec['hbm_frac']:.3f} of 8 TB/s), max rel err {worst:.1e}")
    for dtype in ("f32", "bf16"):
        res[f"speedup_{dtype}_vs_batched_f64"] = round(
            res[dtype]["matrhs_per_s"] / res["f64"]["matrhs_per_s"], 3)
    res["tolerance"] = {"f32": 1e-5, "bf16": 1e-2}
    if not plain:
        for D in D64:
            D.close()
        return res
    # plain FP64 GPU forms: one SpMM per matrix, and one SpMV per matrix and RHS
    Y64 = [torch.empty((D.nrows, K), dtype=torch.float64, device="cuda") for D in D64]
    plain = {}
    ms = timed(lambda: [D.multiply_multi(x, y, True, stream) for D, x, y in zip(D64, X64, Y64)], 3)
    plain["per_matrix_spmm_ms"] = round(ms, 4)
    xcols = [[x[:, j].contiguous() for j in range(K)] for x in X64]
    ycols = [[torch.empty(D.nrows, dtype=torch.float64, device="cuda") for _ in range(K)]
             for D in D64]
    ms = timed(lambda: [D.multiply(xc[j], yc[j], True, stream)
                        for D, xc, yc in zip(D64, xcols, ycols) for j in range(K)], 2)
    plain["per_column_spmv_ms"] = round(ms, 4)
    # the same launches captured in hipGraphs (no host overhead per launch)
    g1, g2 = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g1):
        for D, x, y in zip(D64, X64, Y64):
            D.multiply_multi(x, y, True)
    with torch.cuda.graph(g2):
        for D, xc, yc in zip(D64, xcols, ycols):
            for j in range(K):
                D.multiply(xc[j], yc[j], True)
    plain["per_matrix_spmm_graph_ms"] = round(timed(g1.replay, 5), 4)
    plain["per_column_spmv_graph_ms"] = round(timed(g2.replay, 5), 4)
    del g1, g2
    for D in D64:
        D.close()
    best = min(plain.values())
    plain["fastest_matrhs_per_s"] = round(len(mats) * K / (best * 1e-3), 1)
    res["plain_f64"] = plain
    for dtype in ("f32", "bf16"):
        res[f"speedup_{dtype}_vs_fastest_plain_f64"] = round(
            res[dtype]["matrhs_per_s"] / plain["fastest_matrhs_per_s"], 3)
    log(f"batch: FP32 {res['speedup_f32_vs_batched_f64']:.2f}x, BF16 "
        f"{res['speedup_bf16_vs_batched_f64']:.2f}x the batched FP64 launch; plain FP64 forms "
        + ", ".join(f"{k} {v:.2f}" for k, v in plain.items() if k.endswith("_ms")) + " ms")
    return res


def run_sweep(z15, s7, steps):
    """Format x precision sweep: S7, Z15 and B9 (4M rows, bandwidth 9, 2 %
    off-band) stored in FP64, FP32 and BF16 (int32 indices; FP32 accumulation
    with TwoSum on rows longer than 1k for the reduced types), every format
    the reference admits for the matrix; per (matrix, format, value type) the
    product time, its share of the 8 TB/s roofline on the algorithmic bytes,
    and the max-row relative error max_i |y_i - y_ref_i| / sum_j |a_ij x_j|
    against the compensated FP64 oracle on the original FP64 matrix and x
    (FP64 also reports whether every row is within 4 nnz_row u sum_j
    |a_ij x_j|)."""
    import numpy as np
    import torch
    import oracle  # the oracle: FP64 reference products and row scales
    from morpheus_amd import capi, device
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import FORMAT_NAMES

    t0 = time.time()
    b9 = gen.B9()
    log(f"sweep: generated B9 (nnz {b9.nnz}) in {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()
    out = {}
    for mname, A in (("S7", s7), ("Z15", z15), ("B9", b9)):
        x = gen.random_vector(A.ncols, 21)
        yref = oracle.spmv_compensated(A, x)
        scale = oracle.absrow(A, x)
        mask = scale > 0
        bound64 = 4.0 * np.diff(A.row_offsets) * 2.0 ** -53 * scale
        out[mname] = {}
        for dtype in ("f64", "f32", "bf16"):
            base = device.to_device(A, dtype)
            vdt = device.vector_dtype(dtype)
            xt = torch.from_numpy(x).cuda().to(vdt)
            y = torch.empty(A.nrows, dtype=vdt, device="cuda")
            for f, fname in enumerate(FORMAT_NAMES):
                try:
                    D = base if f == base.format else device.convert(base, f)
                except capi.FormatConversionException:
                    out[mname].setdefault(fname, {})[dtype] = "FormatConversionException"
                    continue
                D.analyse()
                D.multiply(xt, y, True, stream)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(steps):
                    D.multiply(xt, y, True, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / steps
                yh = y.double().cpu().numpy()
                err = float(np.max(np.abs(yh[mask] - yref[mask]) / scale[mask])) if mask.any() else 0.0
                gbs = D.algorithmic_bytes() / (ms * 1e-3) / 1e9
                rec = {
                    "ms": round(ms, 5),
                    "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2),
                    "gbs": round(gbs, 1),
                    "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                    "max_rel_err": err,
                }
                if dtype == "f64":
                    rec["within_fp64_bound"] = bool(np.all(np.abs(yh - yref) <= bound64))
                out[mname].setdefault(fname, {})[dtype] = rec
                if D is not base:
                    D.close()
            base.close()
        log("sweep " + mname + ": " + ", ".join(
            f"{f} " + "/".join(f"{v['max_rel_err']:.1e}" if isinstance(v, dict) else "n/a"
                               for v in d.values()) for f, d in out[mname].items()))
    out["tolerance"] = {"f32": 1e-5, "bf16": 1e-2}
    out["max_rel_err"] = {dt: max(v[dt]["max_rel_err"] for m in ("S7", "Z15", "B9")
                                  for v in out[m].values() if isinstance(v.get(dt), dict))
                          for dt in ("f32", "bf16")}
    out["fp64_all_within_bound"] = all(v["f64"]["within_fp64_bound"] for m in ("S7", "Z15", "B9")
                                       for v in out[m].values() if isinstance(v.get("f64"), dict))
    return out


def run_one_shot(z15, reps=20):
    """The one-shot C-ABI multiply (mph_csr_multiply: a transient handle,
    its analysis, the product, a stream synchronisation and the release, per
    call) on Z15 in FP64, wall clock per call, median of `reps`, beside the
    handle's product."""
    import numpy as np
    import torch
    from morpheus_amd import capi, device
    ro = torch.from_numpy(z15.row_offsets.astype(np.int32)).cuda()
    ci = torch.from_numpy(z15.column_indices.astype(np.int32)).cuda()
    v = torch.from_numpy(z15.values).cuda()
    x = torch.empty(z15.ncols, dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0)
    y = torch.empty(z15.nrows, dtype=torch.float64, device="cuda")
    h = device.stream_handle(torch.cuda.current_stream())

    def call():
        capi.call("mph_csr_multiply", z15.nrows, z15.ncols, z15.nnz, ro.data_ptr(),
                  ci.data_ptr(), v.data_ptr(), 0, x.data_ptr(), y.data_ptr(), 1, h)

    call()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        w0 = time.perf_counter()
        call()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - w0) * 1e3)
    ms = float(np.median(ts))
    log(f"one-shot mph_csr_multiply on Z15: {ms:.3f} ms per call (analysis, product, sync, release)")
    return {"ms_per_call": round(ms, 4), "reps": reps,
            "note": "transient analysis + product + stream sync + release, wall clock"}


def run_no_copies(mats, vecs, order, reps=20):
    """The products of the Morpheus drop-in containers: the same matrices
    through handles with MPH_POLICY_NO_VALUE_COPIES (the x-tiled pass reads
    the values in place, no HYB merged copy: every product reads the caller's
    value arrays), for the
    formats whose default plan keeps a copy on these matrices; mean of `reps`
    back-to-back products, and the plan's bytes."""
    import numpy as np
    import torch
    from morpheus_amd import capi, device
    stream = torch.cuda.current_stream()
    out = {}
    for n in order:
        D, _ = mats[n]
        if n not in ("CSR", "COO", "HYB"):
            continue
        M = device.DeviceMatrix(D.format, D.nrows, D.ncols, D.nnz, D.dtype, D.arrays, D.meta)
        M.set_policy(capi.POLICY_NO_VALUE_COPIES)
        M.analyse()
        x, y = vecs[n]
        M.multiply(x, y, True, stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            M.multiply(x, y, True, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        pb = M.plan_bytes()
        out[n] = {"ms": round(ms, 5), "bytes_plan": int(pb),
                  "hbm_frac_alg": round(data_gbs(D.algorithmic_bytes(), ms) / HBM_PEAK_GBS, 4),
                  "hbm_frac_plan": round(data_gbs(pb, ms) / HBM_PEAK_GBS, 4)}
        log(f"{n} without value copies (drop-in policy): {ms * 1e3:.1f} us, plan bytes {pb / 1e6:.0f} MB")
        M.close()
    return out


def run_p2(reps=200):
    """The CPU-reference configuration: the 2D 5-point Poisson matrix on a
    256x256 grid (65,536 rows, 326,656 nnz, values 4 / -1) in FP64, x ~
    U[-1, 1].  The OpenMP restatement of the reference CSR multiply on the
    host cores and the Serial restatement of Convert CSR -> COO -> DIA -> CSR
    (checked bit-exact), timed on the host; beside them the same product and
    round trip through the library on the GPU (checked against the oracle:
    bit-exact round trip, the FP64 per-row bound)."""
    import numpy as np
    import torch
    import oracle  # the oracle: CPU-reference leg and checks
    from morpheus_amd import device
    from morpheus_amd import generators as gen
    from morpheus_amd.formats import COO_FORMAT, CSR_FORMAT, DIA_FORMAT
    A = gen.P2()
    x = gen.random_vector(A.ncols, 7)
    ref = oracle.spmv(A, x)
    out = {"matrix": "P2 (2D 5-point Poisson, 256x256)", "nrows": int(A.nrows), "nnz": int(A.nnz)}
    fn, _, _ = oracle.omp_spmv_fn(A)
    for _ in range(5):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out["cpu_omp_spmv"] = {"ms": round(t * 1e3, 4), "gflops": round(2.0 * A.nnz / t / 1e9, 3),
                           "cores": oracle.omp_threads(), "reps": reps}
    t0 = time.perf_counter()
    back = oracle.convert(oracle.convert(oracle.to_coo(A), DIA_FORMAT), CSR_FORMAT)
    rt = (time.perf_counter() - t0) * 1e3
    same = bool(np.array_equal(back.row_offsets, A.row_offsets)
                and np.array_equal(back.column_indices, A.column_indices)
                and np.array_equal(back.values.view(np.uint64), A.values.view(np.uint64)))
    out["cpu_serial_roundtrip_csr_coo_dia_csr"] = {"ms": round(rt, 3), "bit_exact": same}
    # the GPU side of the same configuration
    D = device.to_device(A, "f64")
    xt = torch.from_numpy(x).cuda()
    y = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    D.analyse()
    D.multiply(xt, y)
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    bound = 4.0 * np.maximum(np.diff(A.row_offsets), 1) * 2.0 ** -53 * oracle.absrow(A, x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        D.multiply(xt, y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    C = device.convert(D, COO_FORMAT)
    Dd = device.convert(C, DIA_FORMAT)
    back = device.convert(Dd, CSR_FORMAT)
    torch.cuda.synchronize()
    grt = (time.perf_counter() - w0) * 1e3
    bh = back.to_host()
    gsame = bool(np.array_equal(bh.row_offsets, A.row_offsets)
                 and np.array_equal(bh.column_indices, A.column_indices)
                 and np.array_equal(bh.values.view(np.uint64), A.values.view(np.uint64)))
    for M in (back, Dd, C, D):
        M.close()
    out["gpu_csr_spmv"] = {"ms": round(ms, 5), "gflops": round(2.0 * A.nnz / (ms * 1e-3) / 1e9, 2),
                           "within_fp64_bound": bool(np.all(np.abs(yh - ref) <= bound))}
    out["gpu_roundtrip_csr_coo_dia_csr"] = {"ms": round(grt, 3), "bit_exact": gsame}
    log(f"P2: CPU OpenMP {out['cpu_omp_spmv']['gflops']:.2f} GFLOP/s on "
        f"{out['cpu_omp_spmv']['cores']} threads, Serial round trip {rt:.1f} ms (bit-exact {same}); "
        f"GPU {out['gpu_csr_spmv']['gflops']:.1f} GFLOP/s, round trip {grt:.1f} ms (bit-exact {gsame})")
    return out


def calibrate_copy():
    """An elementwise kernel reading 1 GiB and writing 1 GiB, three times
    (the known-size reference for the FETCH_SIZE / WRITE_SIZE counters; a
    plain copy_ may be served by a DMA engine, invisible to kernel
    counters)."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the SpMV hot path runs on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from morpheus_amd import capi

    capi.call("mph_initialize", 0)
    if args.only:
        sec = args.only
        if sec == "batch":
            res = run_batch(args.batch_steps, not args.no_batch_plain)
        elif sec.startswith("large"):
            res = run_large(autotune=False,
                            matrices=("S27", "Z12") if sec == "large" else (sec.split("-")[1],))
        elif sec == "p2":
            res = run_p2()
        elif sec == "spmm":
            res = run_spmm_z12()
        else:
            sys.exit(f"unknown section {sec}")
        if args.calibrate_copy:
            calibrate_copy()
        print(json.dumps({"section": sec, "result": res}), flush=True)
        return
    only = [f.strip().upper() for f in args.formats.split(",") if f.strip()]
    mats, vecs, order, ell_z15, (z15, s7), tuned = build_workload(only, not args.no_tune)
    total_ms, wall, per = run_gpu(mats, vecs, order, args.steps, args.warmup)
    ms_per_step = total_ms / args.steps
    flops_step = sum(2.0 * mats[n][1].nnz for n in order)
    value = flops_step / (ms_per_step * 1e-3) / 1e9

    formats = {}
    for n in order:
        D, host = mats[n]
        t = float(np.mean(per[n]))
        bytes_alg = D.algorithmic_bytes()
        gbs = bytes_alg / (t * 1e-3) / 1e9
        bytes_plan = D.plan_bytes()
        formats[n] = {
            "matrix": "Z15" if host is z15 else "S7",
            "nnz": int(host.nnz),
            "ms": round(t, 5),
            "ms_median": round(float(np.median(per[n])), 5),
            "gflops": round(2.0 * host.nnz / (t * 1e-3) / 1e9, 2),
            "bytes_alg": int(bytes_alg),
            "gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "bytes_plan": int(bytes_plan),
            "hbm_frac_plan": round(data_gbs(bytes_plan, t) / HBM_PEAK_GBS, 4),
        }
        log(f"{n} on {formats[n]['matrix']}: {t * 1e3:.1f} us  {formats[n]['gflops']} GFLOP/s  "
            f"{gbs:.0f} GB/s  ({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s)")
    cold, graph = cold_and_graph(mats, vecs, order, 20)
    for n in order:
        bytes_alg = formats[n]["bytes_alg"]
        formats[n]["ms_mall_flushed"] = round(cold[n], 5)
        formats[n]["hbm_frac_mall_flushed"] = round(bytes_alg / (cold[n] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        formats[n]["ms_graph_warm"] = round(graph[n], 5)
        formats[n]["hbm_frac_graph_warm"] = round(bytes_alg / (graph[n] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        log(f"{n}: {cold[n] * 1e3:.1f} us after a MALL flush, {graph[n] * 1e3:.1f} us "
            f"graph-replayed back to back")
    for n in order:
        formats[n]["kernels"] = product_kernels(n, tuned.get(n, {}))
    # the dominant kernels: those of the slowest product; of the formats that
    # run exactly those kernels (CSR, row-sorted COO and merged HYB share
    # them on Z15) the one with the fewest algorithmic bytes is reported, so
    # that bytes a format counts but its plan never reads (HYB's ELL padding)
    # do not raise the fraction
    dom = max(order, key=lambda n: formats[n]["ms"])
    dom = min((n for n in order if formats[n]["kernels"] == formats[dom]["kernels"]),
              key=lambda n: formats[n]["bytes_alg"])
    roof = {
        "bound": "hbm",
        "kernel": " + ".join(formats[dom]["kernels"]),
        "format": dom,
        "matrix": formats[dom]["matrix"],
        "achieved": formats[dom]["gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(formats[dom]["gbs"] / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_alg": formats[dom]["bytes_alg"],
        "bytes_plan": formats[dom]["bytes_plan"],
    }
    # HBM traffic per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes on this workload (scaled by a copy-kernel calibration), committed
    # under profiles/
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "spmv_f64_six_formats_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tf = json.load(f).get("formats", {})
        if dom in tf:
            roof["traffic"] = int(tf[dom]["fetch_bytes"] + tf[dom]["write_bytes"])
            roof["traffic_unit"] = ("bytes per product (rocprofv3 FETCH_SIZE + WRITE_SIZE over "
                                    "the format's kernels)")

    # practical streaming ceiling on this device: an elementwise kernel that
    # reads 1 GiB and writes 1 GiB (context for the roofline fraction)
    sa = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    sb = torch.empty_like(sa)
    for _ in range(2):
        torch.mul(sa, 1.0, out=sb)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        torch.mul(sa, 1.0, out=sb)
    e1.record()
    torch.cuda.synchronize()
    stream_gbs = 5 * 2 * sa.numel() * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del sa, sb
    log(f"elementwise 1 GiB read + 1 GiB write: {stream_gbs:.0f} GB/s")
    # per format: the read rate of a plain one-pass sum over the product's
    # algorithmic bytes, MALL flushed as for ms_mall_flushed: how far each
    # product is from what a pure read stream of its size reaches here
    for n, rec in formats.items():
        rs = read_stream_flushed(rec["bytes_alg"])
        rec["read_stream_same_size_gbs"] = round(rs, 1)
        rec["frac_of_read_stream_mall_flushed"] = round(
            data_gbs(rec["bytes_alg"], rec["ms_mall_flushed"]) / rs, 4)
        log(f"{n}: a {rec['bytes_alg'] / 1e6:.0f} MB read stream runs at {rs:.0f} GB/s; "
            f"the product (MALL flushed) at {rec['frac_of_read_stream_mall_flushed']:.2f} of it")

    # ceilings: the read probe over 1 GiB (MALL flushed), and the x-gather
    # floor of Z15's short rows
    ceilings = {"read_1gib_gbs": round(read_stream_flushed(1 << 30, 10), 1)}
    if "CSR" in order:
        g_us, g_items = gather_ceiling(z15)
        ceilings["gather_z15_short_rows"] = {
            "us": round(g_us, 2), "items": g_items,
            "note": ("coalesced int32 column loads + 8-byte x gathers from an 8 MB x, "
                     "8 in flight per thread, no row bookkeeping")}
        log(f"ceilings: 1 GiB read {ceilings['read_1gib_gbs']:.0f} GB/s; Z15 short-row gather "
            f"floor {g_us:.1f} us over {g_items} items")
        formats["CSR"]["one_shot"] = run_one_shot(z15)
    for n, rec in run_no_copies(mats, vecs, order).items():
        formats[n]["no_value_copies"] = rec

    cg = run_cg(mats, 100)
    p2 = run_p2()
    batch = None if args.no_batch else run_batch(args.batch_steps, not args.no_batch_plain)
    sweep = None if (args.no_sweep or only) else run_sweep(z15, s7, 20)
    large = None if (args.no_large or only) else run_large()

    cpu = None
    if not args.no_cpu:
        res, threads = run_cpu(mats, order, args.cpu_seconds)
        t_cpu = sum(res[n][0] for n in order)
        cpu = {
            "value": round(flops_step / t_cpu / 1e9, 3),
            "unit": "GFLOP/s",
            "cores": threads,
            "kind": "port",
            "sample": ("OpenMP restatement of the reference multiply (static row / nnz split) "
                       "on the same six (format, matrix) FP64 SpMVs, median of "
                       + ", ".join(f"{n}:{res[n][1]}" for n in order)
                       + f" reps; {cpu_model()}"),
        }

    if args.calibrate_copy:
        calibrate_copy()

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generators: Z15 seed 15, S7 closed-form stencil, x ~ U[-1,1])",
        "config": {
            "workload": ("FP64 drop-in multiply, one SpMV per format per step: CSR/COO/HYB on "
                         "Z15 (1M rows, power-law(1.5) row lengths, mean 16, cap 20k), "
                         "DIA/HDC/ELL on S7 (3D 7-point stencil 128^3)"),
            "index": "int32",
            "ell_on_z15": ell_z15,
            "stream_copy_gbs": round(stream_gbs, 1),
            "tuning": tuned,
            "wall_s": round(wall, 4),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "formats": formats,
        "ceilings": ceilings,
        "p2": p2,
        "batched_spmm": batch,
        "cg_iteration": cg,
        "precision_sweep": sweep,
    }
    if large is not None:
        line["large_configs"] = large
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
