"""Sparse gemv benchmark on one MI355X.

Headline: plain CSR SpMV y = A x, FP32 values, int32 indices, A 16M x 16M
with row lengths ~ Zipf(1.5) clipped to [1, 4096], sorted distinct columns
(stratified uniform), vals N(0,1), x U[-1,1] -- inputs generated on the GPU
and resident in HBM before timing. One step = one drx_spmv over the whole
matrix. value = algorithmic GB/s (the reference's byte formula,
examples/sp/gemv_benchmark.cpp:135-138) over the timed steps.

Also reported in the same JSON line (`forms`): the 2-D Laplacian FP64 form
(1e6 rows, four rotating copies so the 256 MB MALL cannot hold them), the
banded forward / adjoint / dvals / fused-residual forms (4M rows, band +-16),
the low-precision storage forms (3-D Laplacian 256^3 and the Zipf matrix in
BF16 / FP16 vs FP32 / FP64) with their error against the FP64 form, the
batched form, and the reference CPU benchmark's own input (GemvEq_DR: band
+-1000, 50000 rows, 8 right-hand sides, FP64) on the GPU and through the
restated CPU path.

cpu_baseline: the reference's CPU path (mp-bench GemvEq_DR without --sycl:
csr_eq shares, one single-threaded rank per core, sequential local loop,
Gatherv + boundary-row assembly) as restated in oracle/, run on a bounded
sample of the headline matrix (its first rows), on rank 0 only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "distributed-ranges_amd"))

import drx  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFS = {"f32": 157.3, "f64": 78.6}  # dense matrix-core peaks of one MI355X
SZ = {torch.float64: 8, torch.float32: 4, torch.bfloat16: 2, torch.float16: 2,
      torch.int32: 4, torch.int64: 8}


def spmv_bytes(m, n, nnz, sv, si, sx, sy, k=1):
    """B = nnz*(sv+si) + si*(m+1) + k*(sx*n + sy*m)."""
    return nnz * (sv + si) + si * (m + 1) + k * (sx * n + sy * m)


def time_launches(fn, steps, warmup, stream=None):
    """(wall ms per step over `steps` back-to-back launches, mean event ms per
    launch). Events are recorded on the stream the kernels run on."""
    stream = stream or torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    ev_ms = sum(a.elapsed_time(b) for a, b in ev) / steps
    return wall_ms, ev_ms


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Headline
# ---------------------------------------------------------------------------
def headline(steps, warmup):
    n = 16_000_000
    t0 = time.perf_counter()
    A = drx.gen.zipf_csr(n, n, s=1.5, lmax=4096, seed=0, vdtype=torch.float32,
                         itype=torch.int32)
    x = drx.gen.uniform(n, torch.float32, -1.0, 1.0, seed=1)
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    log("zipf matrix: %d rows, %d nnz (mean %.2f/row), generated in %.1f s"
        % (n, A.nnz, A.nnz / n, time.perf_counter() - t0))
    B = spmv_bytes(n, n, A.nnz, 4, 4, 4, 4)
    # the same CSR arrays through a plain plan (no panel layout), for reference
    plain = drx.CsrMatrix(A.rowptr, A.colind, A.vals, A.shape)
    _, plain_ms = time_launches(lambda: plain.matvec(x, y=y), max(3, steps // 4), 2)
    plain.close()
    t0 = time.perf_counter()
    npan, width, nseg = A.set_panels(0)
    log("panel layout: %d panels of %d columns, %d row segments, built in %.1f s"
        % (npan, width, nseg, time.perf_counter() - t0))
    wall_ms, ev_ms = time_launches(lambda: A.matvec(x, y=y), steps, warmup)
    log("zipf fp32: %.3f ms/step wall, %.3f ms/step (events); plain CSR %.3f ms"
        % (wall_ms, ev_ms, plain_ms))
    layout = {"panels": npan, "panel_cols": width, "row_segments": nseg,
              "plain_csr_ms_per_step": round(plain_ms, 4),
              "plain_csr_GBps": round(B / plain_ms / 1e6, 2)}
    # The bound that matters here: random 4-byte gathers inside one panel's x
    # slice. drx_gather does only that (streamed indices, gathered values
    # written out) for one panel's share of the entries.
    per = A.nnz // npan
    g = torch.Generator(device="cuda").manual_seed(3)
    idx = torch.randint(0, width, (per,), dtype=torch.int32, device="cuda", generator=g)
    out = torch.empty(per, dtype=torch.float32, device="cuda")
    xs = x[:width].contiguous()
    _, g_ms = time_launches(lambda: drx.gen.gather(idx, xs, out=out), max(5, steps // 2), 2)
    del idx, out
    torch.cuda.empty_cache()
    gather = {"G_gathers_per_s": round(per / g_ms / 1e6, 1),
              "ms_per_step": round(g_ms * npan, 4),
              "kernel_frac_of_gather_ceiling": round(g_ms * npan / ev_ms, 4),
              "probe": "drx_gather: %d random int32 indices into %d floats (one panel's x "
                       "slice), x %d panels" % (per, width, npan)}
    return A, x, y, B, wall_ms, ev_ms, layout, gather


def form_panel_sweep(A, x, steps, widths=(262144, 393216, 524288, 786432, 1048576)):
    """The headline product through column panels of each width (columns of x
    per panel): time per step and the row segments the layout makes. Narrow
    panels cost row-segment work; a panel whose x slice and streamed entries
    outgrow an XCD's 4 MB L2 misses on its gathers."""
    n = A.shape[0]
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    B = spmv_bytes(n, A.shape[1], A.nnz, 4, 4, 4, 4)
    out = []
    for w in widths:
        M = drx.CsrMatrix(A.rowptr, A.colind, A.vals, A.shape)
        npan, width, nseg = M.set_panels(w)
        _, ms = time_launches(lambda: M.matvec(x, y=y), steps, 2)
        out.append({"panel_cols": width, "MB_of_x": round(width * 4 / 2**20, 3),
                    "panels": npan, "row_segments": nseg, "ms_per_step": round(ms, 4),
                    "GBps": round(B / ms / 1e6, 1)})
        log("panel sweep: %s" % out[-1])
        M.close()
        del M
        torch.cuda.empty_cache()
    return out


def host_cpu_info():
    """(ranks to use, description): one single-threaded rank per core of the
    CPU share this process may run on, capped at the 16 cores a one-GPU job
    is granted on the GPU host (more threads than the share would time
    oversubscription, not the reference's loop)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    P = max(1, min(16, affinity))
    return P, {"nproc": os.cpu_count(), "affinity_cores": affinity, "cpu_model": model,
               "ranks": P,
               "launcher": "P pinned std::thread ranks in one process, shared-memory "
                           "Gatherv (the reference's MPI path cannot be built: no "
                           "range-v3, mdspan or MPI-4 _c calls in this image)",
               "rank_cap": "16 = the CPU share of a one-GPU job on the GPU host"}


def cpu_baseline(A, x, y_gpu, target_s=10.0):
    """Restated GemvEq_DR CPU path on the first rows of the headline matrix."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import _oracle  # the CPU restatement (test infrastructure)

    orc = _oracle.Oracle()
    n = A.shape[1]
    rp_all = A.rowptr
    # sample: first R rows, about 1/8 of the matrix
    R = A.shape[0] // 8
    nnz_s = int(rp_all[R].item())
    rp = rp_all[:R + 1].cpu().numpy()
    ci = A.colind[:nnz_s].cpu().numpy()
    v = A.vals[:nnz_s].cpu().numpy()
    xs = x.cpu().numpy()
    P, host = host_cpu_info()
    # one untimed probe to size the repetitions
    _, t1, _ = orc.gemveq_cpu_bench(rp, ci, v, xs, (R, n), P, warmup=0, reps=1)
    reps = int(max(3, min(50, target_s / max(t1, 1e-3))))
    res, med, _ = orc.gemveq_cpu_bench(rp, ci, v, xs, (R, n), P, warmup=1, reps=reps)
    B_s = spmv_bytes(R, n, nnz_s, 4, 4, 4, 4)
    yg = y_gpu[:R].double().cpu().numpy()
    yc = res.astype(np.float64)
    rel = float(np.linalg.norm(yg - yc) / max(np.linalg.norm(yc), 1e-300))
    return {
        "value": round(B_s / med / 1e9, 3), "unit": "GB/s", "cores": P, "kind": "port",
        "sample": ("first %d rows (%d nnz, 1/8) of the 16M-row Zipf FP32 matrix, full x; "
                   "median of %d GemvEq_DR calls (csr_eq shares over %d single-threaded "
                   "std::thread ranks, Gatherv + boundary-row assembly), %.4f s/call"
                   % (R, nnz_s, reps, P, med)),
        "ms_per_call": round(med * 1e3, 3),
        "gpu_vs_reference_cpu_rel_err_norm2": rel,
        "host": host,
    }


# ---------------------------------------------------------------------------
# Forms
# ---------------------------------------------------------------------------
def rel_err(y, y64):
    y = y.double()
    return float((y - y64).norm() / y64.norm())


def form_laplacian2d(steps):
    """2-D 5-point Laplacian 1000 x 1000, FP64, int32: 80 MB per copy, four
    copies rotated so every launch streams from HBM."""
    mats = [drx.gen.laplacian(2, 1000, vdtype=torch.float64) for _ in range(4)]
    xs = [drx.gen.uniform(1_000_000, torch.float64, -1, 1, seed=20261202 + i) for i in range(4)]
    ys = [torch.empty(1_000_000, dtype=torch.float64, device="cuda") for _ in range(4)]
    it = [0]

    def rot():
        i = it[0] % 4
        it[0] += 1
        mats[i].matvec(xs[i], y=ys[i])

    B = spmv_bytes(1_000_000, 1_000_000, mats[0].nnz, 8, 4, 8, 8)
    _, ev_ms = time_launches(rot, steps * 4, 8)
    _, hot_ms = time_launches(lambda: mats[0].matvec(xs[0], y=ys[0]), steps, 5)
    return {"rows": 1_000_000, "nnz": mats[0].nnz, "dtype": "f64",
            "GBps": round(B / ev_ms / 1e6, 1), "frac": round(B / ev_ms / 1e6 / HBM_PEAK_GBS, 4),
            "us_per_launch": round(ev_ms * 1e3, 2),
            "GBps_hot_cache": round(B / hot_ms / 1e6, 1)}


def form_band_adjoint(steps):
    """Band +-16 (33 diagonals), 4M rows, FP32: forward, adjoint (CSC with
    permuted vals), dvals and the fused forward+residual+dvals pass."""
    m = 4_000_000
    A = drx.gen.band(m, 16, 16, vdtype=torch.float32, seed=31)
    x = drx.gen.uniform(m, torch.float32, -1, 1, seed=32)
    t = drx.gen.uniform(m, torch.float32, -1, 1, seed=33)
    y = torch.empty(m, dtype=torch.float32, device="cuda")
    dx = torch.empty(m, dtype=torch.float32, device="cuda")
    dv = torch.empty(A.nnz, dtype=torch.float32, device="cuda")
    r = torch.empty(m, dtype=torch.float32, device="cuda")
    A.transpose()
    nnz = A.nnz
    fwd_B = spmv_bytes(m, m, nnz, 4, 4, 4, 4)
    adj_B = spmv_bytes(m, m, nnz, 4, 4, 4, 4) + 4 * nnz  # + permutation
    _, a_perm_ms = time_launches(lambda: A.rmatvec(y, dx=dx), steps, 3)
    A.set_transposed_values()  # CSC-order copy: the adjoint streams it
    dv_B = 4 * nnz + 4 * (m + 1) + 4 * nnz + 4 * m + 4 * m
    fu_B = fwd_B + 4 * m + 4 * m + 4 * nnz  # + t read, r write, dvals write
    _, f_ms = time_launches(lambda: A.matvec(x, y=y), steps, 3)
    _, a_ms = time_launches(lambda: A.rmatvec(y, dx=dx), steps, 3)
    _, d_ms = time_launches(lambda: A.dvals(x, y, out=dv), steps, 3)
    _, u_ms = time_launches(lambda: A.fused_residual(x, t, y=y, r=r, dvals=dv), steps, 3)
    A64 = drx.CsrMatrix(A.rowptr, A.colind, A.vals.double(), A.shape)
    y64 = A64.matvec(x.double())
    e_fwd = rel_err(A.matvec(x), y64)
    e_adj = rel_err(A.rmatvec(y64.float()), A64.rmatvec(y64))
    out = {"rows": m, "nnz": nnz, "dtype": "f32",
           "adjoint_through_permutation": {"GBps": round(adj_B / a_perm_ms / 1e6, 1),
                                           "us_per_launch": round(a_perm_ms * 1e3, 1)}}
    adj_B = spmv_bytes(m, m, nnz, 4, 4, 4, 4)  # stored CSC-order values: no permutation read
    for name, B, ms, err in (("forward", fwd_B, f_ms, e_fwd), ("adjoint", adj_B, a_ms, e_adj),
                             ("dvals", dv_B, d_ms, None), ("fused_fwd_residual_dvals", fu_B, u_ms, None)):
        out[name] = {"GBps": round(B / ms / 1e6, 1), "frac": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
                     "us_per_launch": round(ms * 1e3, 1)}
        if err is not None:
            out[name]["rel_err_vs_fp64"] = err
    sep = f_ms + d_ms
    out["fused_speedup_vs_separate_fwd_plus_dvals"] = round(sep / u_ms, 3)
    return out


def form_lowprec(steps, zipf=None):
    """3-D 7-point Laplacian 256^3 (16.7M rows, 117M nnz) with vals and x
    stored as FP64 / FP32 / BF16 / FP16 (FP32 accumulation for the 16-bit
    storage), plus the Zipf matrix in BF16; error vs the FP64 form."""
    g = 256
    m = g ** 3
    A64 = drx.gen.laplacian(3, g, vdtype=torch.float64)
    x64 = drx.gen.uniform(m, torch.float64, -1, 1, seed=41)
    y64 = A64.matvec(x64)
    out = {"rows": m, "nnz": A64.nnz}
    base_ms = {}
    for dt, name in ((torch.float64, "f64"), (torch.float32, "f32"),
                     (torch.bfloat16, "bf16"), (torch.float16, "f16")):
        vals = A64.vals.to(dt)
        x = x64.to(dt)
        ydt = torch.float64 if dt == torch.float64 else torch.float32
        y = torch.empty(m, dtype=ydt, device="cuda")
        _, ms = time_launches(lambda: A64.matvec(x, y=y, vals=vals), steps, 3)
        B = spmv_bytes(m, m, A64.nnz, SZ[dt], 4, SZ[dt], SZ[ydt])
        base_ms[name] = ms
        out[name] = {"GBps": round(B / ms / 1e6, 1), "frac": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
                     "us_per_launch": round(ms * 1e3, 1),
                     "rel_err_vs_fp64": rel_err(A64.matvec(x, vals=vals), y64)}
    for name in ("bf16", "f16"):
        out[name]["speedup_vs_f32"] = round(base_ms["f32"] / base_ms[name], 3)
        out[name]["speedup_vs_f64"] = round(base_ms["f64"] / base_ms[name], 3)
    del A64
    if zipf is not None:
        A, x = zipf  # FP32, with its panel layout
        vb = A.vals.to(torch.bfloat16)
        xb = x.to(torch.bfloat16)
        Ab = drx.CsrMatrix(A.rowptr, A.colind, vb, A.shape)
        Ab.set_panels(0)
        y = torch.empty(A.shape[0], dtype=torch.float32, device="cuda")
        _, ms = time_launches(lambda: Ab.matvec(xb, y=y), steps, 3)
        _, ms32 = time_launches(lambda: A.matvec(x, y=y), steps, 3)
        B = spmv_bytes(A.shape[0], A.shape[1], A.nnz, 2, 4, 2, 4)
        y32 = A.matvec(x)
        yb = Ab.matvec(xb)
        Ab.close()
        del Ab, vb
        A64z = drx.CsrMatrix(A.rowptr, A.colind, A.vals.double(), A.shape)
        yz64 = A64z.matvec(x.double())
        A64z.close()
        del A64z
        out["zipf_bf16"] = {"GBps": round(B / ms / 1e6, 1), "frac": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
                            "us_per_launch": round(ms * 1e3, 1), "layout": "column panels",
                            "speedup_vs_f32": round(ms32 / ms, 3),
                            "rel_err_vs_fp64": rel_err(yb, yz64),
                            "f32_rel_err_vs_fp64": rel_err(y32, yz64)}
    return out


def form_batched(steps, nmat=512, ks=(1, 4, 16)):
    """512 independent square CSRs, rows U[1k, 200k], nnz/row U[3, 64], FP32,
    k right-hand sides each: forward (y = A x) and adjoint (dx = A^T dy) of
    every matrix in one persistent, autotuned, hipGraph-captured launch,
    against the plain GPU form (one drx_spmv + one drx_spmv_adjoint call per
    matrix)."""
    rng = np.random.default_rng(0)
    rows = rng.integers(1000, 200001, nmat)
    mats = [drx.gen.uniform_rows_csr(int(m), int(m), 3, 64, seed=7000 + i)
            for i, m in enumerate(rows)]
    for A in mats:
        A.transpose()
        A.set_transposed_values()
    torch.cuda.synchronize()
    nnz = sum(A.nnz for A in mats)
    mtot = int(rows.sum())
    out = {"matrices": nmat, "rows": mtot, "nnz": nnz, "dtype": "f32"}
    for k in ks:
        def vec(n, s):
            return drx.gen.uniform(k * n, torch.float32, -1, 1, seed=s).reshape(k, n)
        xs = [vec(A.shape[1], 1 + i) for i, A in enumerate(mats)]
        dys = [vec(A.shape[0], 100000 + i) for i, A in enumerate(mats)]
        ys = [torch.empty_like(x) for x in xs]
        dxs = [torch.empty_like(d) for d in dys]
        b = drx.Batch(forward=list(zip(mats, xs, ys)), adjoint=list(zip(mats, dys, dxs)), k=k)
        tune = b.autotune(iters=2)
        b.graph()
        _, g_ms = time_launches(lambda: b.graph_launch(), steps, 2)

        def plain():
            for A, x, y, d, dx in zip(mats, xs, ys, dys, dxs):
                A.matvec(x, y=y)
                A.rmatvec(d, dx=dx)
        _, p_ms = time_launches(plain, max(3, steps // 4), 1)
        B = (spmv_bytes(mtot, mtot, nnz, 4, 4, 4, 4, k=k) +
             spmv_bytes(mtot, mtot, nnz, 4, 4, 4, 4, k=k))
        # error of the first matrices against their FP64 forms
        b.graph_launch()
        errs = []
        for A, x, y, d, dx in list(zip(mats, xs, ys, dys, dxs))[:8]:
            A64 = drx.CsrMatrix(A.rowptr, A.colind, A.vals.double(), A.shape)
            errs.append(rel_err(y, A64.matvec(x.double())))
            errs.append(rel_err(dx, A64.rmatvec(d.double())))
            A64.close()
        out["k%d" % k] = {"GBps": round(B / g_ms / 1e6, 1), "frac": round(B / g_ms / 1e6 / HBM_PEAK_GBS, 4),
                          "ms_fwd_plus_adjoint": round(g_ms, 3),
                          "plain_gpu_form_ms": round(p_ms, 3),
                          "speedup_vs_plain_gpu_form": round(p_ms / g_ms, 3),
                          "tuning": tune, "max_rel_err_vs_fp64_first8": max(errs)}
        b.close()
        del xs, dys, ys, dxs
        torch.cuda.empty_cache()
    return out


def form_gemveq_band(steps, cpu_seconds=10.0, with_cpu=True):
    """The reference's own CPU benchmark input (benchmarks/gbench/mp/gemv.cpp:
    GemvEq_DR): generate_band_csr<double, long>(50000, 1000, 1000), all-ones
    values, width-8 slim matrix b[i + j*n] = i*j + 1, FP64 / int64. The GPU
    form (k = 8 right-hand sides) and the restated GemvEq_DR CPU path time
    the same matrix and right-hand sides."""
    n, band, k = 50_000, 1000, 8
    A = drx.gen.band(n, band, band, vdtype=torch.float64, itype=torch.int64, seed=1)
    A.vals.fill_(1.0)
    i = torch.arange(n, dtype=torch.float64, device="cuda")
    x = torch.stack([i * j + 1 for j in range(k)])  # (k, n): column j of the slim matrix
    y = torch.empty((k, n), dtype=torch.float64, device="cuda")
    B = spmv_bytes(n, n, A.nnz, 8, 8, 8, 8, k=k)
    _, ms = time_launches(lambda: A.matvec(x, y=y), steps, 3)
    with drx.tuned(mrhs=0):
        _, ms_cols = time_launches(lambda: A.matvec(x, y=y), steps, 3)
    y = A.matvec(x)
    torch.cuda.synchronize()
    out = {"rows": n, "nnz": A.nnz, "k": k, "dtype": "f64/i64",
           "GBps": round(B / ms / 1e6, 1), "frac": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
           "us_per_call": round(ms * 1e3, 1),
           "GBps_column_gathers": round(B / ms_cols / 1e6, 1)}
    # exact answer: row r sums columns [r-1000, r+1000] of b
    lo = torch.clamp(i - band, min=0)
    hi = torch.clamp(i + band, max=n - 1)
    cnt = hi - lo + 1
    ssum = (lo + hi) * cnt / 2
    exact = torch.stack([j * ssum + cnt for j in range(k)])
    out["max_rel_err_vs_exact"] = float(((y - exact).abs() / exact.abs()).max())
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the CPU restatement (test infrastructure)
        orc = _oracle.Oracle()
        rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
        xs = x.cpu().numpy()
        P, _ = host_cpu_info()
        _, t1, _ = orc.gemveq_cpu_bench(rp, ci, v, xs, (n, n), P, warmup=0, reps=1)
        reps = int(max(3, min(30, cpu_seconds / max(t1, 1e-3))))
        res, med, _ = orc.gemveq_cpu_bench(rp, ci, v, xs, (n, n), P, warmup=1, reps=reps)
        out["cpu_gemveq"] = {"GBps": round(B / med / 1e9, 3), "ms_per_call": round(med * 1e3, 2),
                             "ranks": P, "reps": reps,
                             "gpu_speedup": round(med * 1e3 / ms, 1),
                             "max_abs_diff_gpu_vs_cpu": float(np.abs(res - y.cpu().numpy()).max())}
    A.close()
    return out


def form_dense_gemm(steps, n=16000):
    """The reference's dense benchmark shape (benchmarks/gbench/sp/gemm.cpp:
    Gemm_DR, m = n = k = 16000, float): one tile product c += a b on the
    matrix cores (drx_gemm), FP32 and FP64, beside torch.mm (hipBLASLt) on the
    same operands. Inputs U[0, 100) as Gemm_Reference fills them; 16 rows of
    the product are checked against float64 within (k + 2) u |a||b|."""
    out = {"m": n, "n": n, "k": n, "flop_per_call": 2 * n ** 3}
    peaks = {torch.float32: MFMA_PEAK_TFS["f32"], torch.float64: MFMA_PEAK_TFS["f64"]}
    for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        g = torch.Generator(device="cuda").manual_seed(3)
        a = torch.rand(n, n, device="cuda", dtype=dt, generator=g) * 100
        b = torch.rand(n, n, device="cuda", dtype=dt, generator=g) * 100
        c = torch.zeros(n, n, device="cuda", dtype=dt)
        reps = max(3, min(steps, 5))
        _, ms = time_launches(lambda: drx.gemm(a, b, c), reps, 1)
        _, ms_lib = time_launches(lambda: torch.mm(a, b, out=c), reps, 1)
        c.zero_()
        drx.gemm(a, b, c)
        rows = torch.arange(0, n, n // 16, device="cuda")[:16]
        ref = a[rows].double() @ b.double()
        absref = a[rows].double().abs() @ b.double().abs()
        u = 2.0 ** -24 if dt == torch.float32 else 2.0 ** -53
        err = float(((c[rows].double() - ref).abs() / absref).max())
        tfs = 2 * n ** 3 / ms / 1e9
        out[name] = {"TFLOPs": round(tfs, 1), "frac": round(tfs / peaks[dt], 4),
                     "peak_TFLOPs": peaks[dt], "ms_per_call": round(ms, 3),
                     "torch_mm_TFLOPs": round(2 * n ** 3 / ms_lib / 1e9, 1),
                     "vs_torch_mm": round(ms_lib / ms, 3),
                     "rel_err_sample_rows": err, "bound_(k+2)u": (n + 2) * u}
        del a, b, c, ref, absref
        torch.cuda.empty_cache()
    return out


def stream_ceiling(steps=10):
    """HBM read ceiling of this box: drx_stream_read (16-byte non-temporal
    loads, 8 in flight per lane) over a 4 GiB buffer."""
    buf = torch.empty(1 << 30, dtype=torch.float32, device="cuda")
    buf.fill_(1.0)
    digest = torch.empty(1 << 20, dtype=torch.int32, device="cuda")
    _, ms = time_launches(lambda: drx.gen.stream_read(buf, digest), steps, 2)
    nbytes = buf.numel() * 4
    del buf, digest
    torch.cuda.empty_cache()
    return {"GBps": round(nbytes / ms / 1e6, 1), "frac_of_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4),
            "probe": "drx_stream_read over 4 GiB (16-byte non-temporal loads), mean of %d launches"
                     % steps}


def copy_ceiling(steps=10):
    """Read + write ceiling of this box: the runtime's device-to-device copy
    (torch copy_) of 1 GiB, counting the bytes read and the bytes written --
    the traffic mix of dvals and the fused pass (about as many bytes written
    as read)."""
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    src.fill_(1.0)
    dst = torch.empty_like(src)
    _, ms = time_launches(lambda: dst.copy_(src), steps, 2)
    nbytes = 2 * src.numel() * 4
    del src, dst
    torch.cuda.empty_cache()
    return {"GBps": round(nbytes / ms / 1e6, 1), "frac_of_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4),
            "probe": "torch copy_ of 1 GiB device to device (1 GiB read + 1 GiB written), mean of "
                     "%d launches" % steps}


def form_mp_host(reps=20):
    """The reference benchmark's call shape, gemv(0, res, m, b) on a csr_eq
    distributed_sparse_matrix (benchmarks/gbench/mp/gemv.cpp:41-64), timed
    through the C++ host interface on the headline matrix
    (distributed-ranges_amd/bin/mp_gemv_bench, one rank)."""
    import subprocess
    exe = os.path.join(REPO, "distributed-ranges_amd", "bin", "mp_gemv_bench")
    if not os.path.exists(exe):
        return {"error": "mp_gemv_bench not built"}
    p = subprocess.run([exe, "16000000", str(reps)], capture_output=True, text=True, timeout=600)
    line = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if p.returncode != 0 or not line:
        return {"error": "rc %d: %s" % (p.returncode, (p.stderr or p.stdout)[-400:])}
    out = json.loads(line[-1])
    out["call"] = "dr::mp::gemv(0, res, m, b): share product + partial to host + gather + fence"
    return out


# ---------------------------------------------------------------------------
def read_traffic():
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel scaled)."""
    path = os.path.join(REPO, "profiles", "pmc_spmv_zipf16m_f32.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return int(float(line.split(":", 1)[1].strip()))
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--forms", default="all",
                    help="all | none | comma list (panel_sweep runs only when listed)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--xstage", type=int, default=None,
                    help="x windows in LDS (the library's xstage knob: 0 never, 1 default, "
                         "2 always); for profiling the forms with and without")
    ap.add_argument("--vec", type=int, default=None,
                    help="consecutive entries per lane and load (the library's vec knob: 0 "
                         "default, 1, 2, 4); for profiling the forms with each")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 2 GiB device copy 5 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        a = torch.ones(1 << 29, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": a.numel() * 4,
                          "copy_bytes_written_per_launch": b.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the kernel on one GPU's "
              "share only (see the sharding notes); --gpus %d is not supported."
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    torch.cuda.set_device(0)
    if args.xstage is not None:
        drx.tune_set("xstage", args.xstage)
    if args.vec is not None:
        drx.tune_set("vec", args.vec)
    A, x, y, B, wall_ms, ev_ms, layout, gather = headline(args.steps, args.warmup)
    value = B / (wall_ms * 1e-3) / 1e9
    achieved = B / (ev_ms * 1e-3) / 1e9
    result = {
        "metric": "SpMV GB/s (% of 8 TB/s) fwd/adjoint, batched & bf16 vs plain, at stated rel error",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain CSR SpMV y=A*x, FP32 vals, int32 indices, 16M x 16M, "
                               "nnz/row Zipf(s=1.5) clipped to [1,4096], sorted distinct "
                               "stratified-uniform columns, vals N(0,1), x U[-1,1]",
                   "rows": A.shape[0], "cols": A.shape[1], "nnz": A.nnz,
                   "mean_nnz_per_row": round(A.nnz / A.shape[0], 3),
                   "algorithmic_bytes_per_step": B, "parallelism": "single GPU",
                   "layout": "CSR + column-panel plan (2 MB x slices, L2-resident gathers)",
                   **layout},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(),
                     "kernel": "drx::k_spmv_fwd<256,2048,kRowMap,f32,i32> (one launch per "
                               "column panel); achieved = bytes / summed time of one step's "
                               "launches",
                     "ms_per_step_kernels": round(ev_ms, 4),
                     "gather_ceiling": gather},
    }
    log("stream ceiling ...")
    result["roofline"]["stream_ceiling"] = stream_ceiling()
    result["roofline"]["copy_ceiling"] = copy_ceiling()
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(A, x, y, args.cpu_seconds)
    forms = {}
    want = set(["laplacian2d_f64", "band_adjoint", "lowprec", "batched", "gemveq_band",
                "dense_gemm", "mp_host"]) \
        if args.forms == "all" else \
        set() if args.forms == "none" else set(args.forms.split(","))
    steps_f = max(5, min(args.steps, 20))
    if "lowprec" in want:
        log("form lowprec ...")
        forms["lowprec"] = form_lowprec(steps_f, zipf=(A, x))
    if "panel_sweep" in want:
        log("form panel_sweep ...")
        forms["panel_sweep"] = form_panel_sweep(A, x, steps_f)
    del A
    torch.cuda.empty_cache()
    if "laplacian2d_f64" in want:
        log("form laplacian2d_f64 ...")
        forms["laplacian2d_f64"] = form_laplacian2d(steps_f)
    if "band_adjoint" in want:
        log("form band_adjoint ...")
        forms["band_adjoint"] = form_band_adjoint(steps_f)
    if "batched" in want:
        log("form batched ...")
        forms["batched"] = form_batched(steps_f)
    if "gemveq_band" in want:
        log("form gemveq_band ...")
        forms["gemveq_band"] = form_gemveq_band(steps_f, args.cpu_seconds, not args.no_cpu)
    if "dense_gemm" in want:
        log("form dense_gemm ...")
        forms["dense_gemm"] = form_dense_gemm(steps_f)
    if "mp_host" in want:
        log("form mp_host ...")
        torch.cuda.empty_cache()
        forms["mp_host"] = form_mp_host(steps_f)
    if forms:
        result["forms"] = forms
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
