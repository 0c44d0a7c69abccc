"""Benchmark: spherical-harmonic round trips (spectral -> grid -> spectral) per second at
T255 (lmax=257, mmax=256) on the octahedral Gaussian grid nlat_half=192, 16 layers x 4
fields (K = 64 columns), split-precision MFMA Legendre vs the plain FP32 path.

One step = one round trip of all 64 columns (sht_synthesis + sht_analysis) with inputs
resident in HBM.  Prints ONE JSON line.  --gpus above 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP32_PEAK = 157.3e12       # flop/s, MI355X FP32 (matrix)
F16_PEAK = 2.5e15          # flop/s, MI355X dense FP16/BF16 MFMA
KERNELS = ["legendre_synthesis", "fft_synthesis", "fft_analysis", "legendre_analysis"]


def geometry(sw, lmax, mmax, n):
    grid = sw.OctahedralGaussianGrid(n)
    nlon = np.array([grid.nlon_ring(j) for j in range(n)])
    mt = np.minimum(nlon // 2, mmax - 1)
    # M = sum_j sum_{m <= mt_j} (lmax - m): complex x real MACs per column and direction
    M = int(sum(sum(lmax - m for m in range(int(t) + 1)) for t in mt))
    nfreq_used = int((mt + 1).sum())
    nharm = sw.Spectrum(lmax, mmax).nharm
    return grid, nlon, mt, M, nfreq_used, nharm


MFMA_PRODUCTS = {"fp32": 1, "fp16x3": 3, "bf16x3": 6}  # MFMA products per algorithmic product


def kernel_cost(name, K, M, nharm, n, npoints, nlon, precision):
    """Algorithmic (bytes, flops, peak flop/s) of one launch of each stage kernel: the
    round trip's compulsory bytes B_rt = 2 nharm K 8 + 2 npoints K 4 + 2 nharm n 4 split by
    stage (Legendre synthesis: spectral input + FP32 Λ; FFT synthesis: grid output; FFT
    analysis: grid input; Legendre analysis: FP32 Λ + spectral output).  The Fourier rows
    handed between the stages are intermediates and not counted."""
    if name.startswith("legendre"):
        peak = FP32_PEAK if precision == "fp32" else F16_PEAK / MFMA_PRODUCTS[precision]
        return nharm * K * 8 + nharm * n * 4, 4.0 * K * M, peak
    fft_flops = float(sum(5.0 * N * math.log2(N) for N in nlon)) * K  # N and S rings in one complex FFT
    return npoints * K * 4, fft_flops, FP32_PEAK


def stage_ms(sw, S, coeffs, field, back, reps):
    """Mean launch duration (ms) of each stage kernel: HIP events around `reps` back-to-back
    launches on the launch stream (sht_time_stage), after a round trip that fills the
    stream's Fourier scratch."""
    L = sw.lib()
    st = S._stream()
    sw.transform_(field, coeffs, S)
    sw.transform_(back, field, S)
    out = {}
    for i, k in enumerate(KERNELS):
        ms = C.c_double()
        c = coeffs if i == 0 else back  # the Legendre analysis writes `back`
        check = L.sht_time_stage(S._plan, i, c.data.data_ptr(), S.spectrum.nharm, field.data.data_ptr(),
                                 S.grid.npoints, c.nlayers, reps, st, C.byref(ms))
        assert check == 0, L.sht_last_error()
        out[k] = ms.value
    return out


def run_mode(sw, torch, precision, spectrum, grid, coeffs_host, steps, warmup):
    K = coeffs_host.shape[0]
    S = sw.SpectralTransform(spectrum, grid, nlayers=K, precision=precision)
    tuning = S.autotune(K)  # plan-time choice of FFT policy and launch orders (not timed)
    coeffs = sw.LowerTriangularArray(torch.from_numpy(coeffs_host).to(torch.complex64).cuda(), spectrum)
    field = sw.Field(torch.empty((K, grid.npoints), device="cuda"), grid)
    back = sw.LowerTriangularArray(torch.empty((K, spectrum.nharm), dtype=torch.complex64, device="cuda"), spectrum)
    for _ in range(warmup):
        sw.transform_(field, coeffs, S)
        sw.transform_(back, field, S)
    torch.cuda.synchronize()
    # the timed region: `steps` round trips, synchronised on both sides
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sw.transform_(field, coeffs, S)
        sw.transform_(back, field, S)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = back.data.cpu().numpy().astype(np.complex128)
    kms = stage_ms(sw, S, coeffs, field, back, max(steps, 20))
    S.close()
    return {"s_per_step": dt / steps, "kernel_ms": kms, "roundtrip": out, "tuning": tuning}


def time_loop(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_c2(sw, torch, steps, warmup):
    """T127 (lmax=129, mmax=128) octahedral nlat_half=96, 8 layers x 4 fields: scalar round trip
    (K=32) plus the vector path UV_from_vordiv (vor, div: K=8) -> U, V -> grid u, v with
    unscale_coslat (K=16), as the model's transform step does it."""
    sp, grid = sw.Spectrum(129, 128), sw.OctahedralGaussianGrid(96)
    S = sw.SpectralTransform(sp, grid, nlayers=32, precision="fp32")
    c = sw.LowerTriangularArray(torch.from_numpy(sw.seeded_spectrum(sp, 32, 5)).to(torch.complex64).cuda(), sp)
    f = sw.Field(torch.empty((32, grid.npoints), device="cuda"), grid)
    b = sw.LowerTriangularArray(torch.empty_like(c.data), sp)
    uv = sw.LowerTriangularArray(torch.empty((16, sp.nharm), dtype=torch.complex64, device="cuda"), sp)
    U = sw.LowerTriangularArray(uv.data[:8], sp)
    V = sw.LowerTriangularArray(uv.data[8:], sp)
    vor = sw.LowerTriangularArray(c.data[:8], sp)
    div = sw.LowerTriangularArray(c.data[8:16], sp)
    fuv = sw.Field(torch.empty((16, grid.npoints), device="cuda"), grid)

    def step():
        sw.transform_(f, c, S)
        sw.transform_(b, f, S)
        sw.UV_from_vordiv_(U, V, vor, div, S)
        sw.transform_(fuv, uv, S, unscale_coslat=True)
    dt = time_loop(torch, step, steps, warmup)
    S.close()
    return {"ms": dt * 1e3, "per_s": 1.0 / dt,
            "workload": "T127 octahedral nlat_half=96: K=32 round trip + UV_from_vordiv (K=8) + U,V synthesis "
                        "with unscale_coslat (K=16), FP32"}


def bench_dense(sw, torch, steps, warmup):
    """Round trips of the dense matrix transform (one FP32 GEMM per direction) against the
    regular transform at small truncations, where the reference uses the dense one on GPUs:
    T31 (nlat_half 24, K=8) and T63 (nlat_half 48, K=32), FP32; the regular transform also
    with its calls replayed from cached graphs (gpu_graphs)."""
    out = {}
    for name, n, lmax, mmax, K in (("T31_K8", 24, 33, 32, 8), ("T63_K32", 48, 65, 64, 32)):
        sp, grid = sw.Spectrum(lmax, mmax), sw.OctahedralGaussianGrid(n)
        c = sw.LowerTriangularArray(torch.from_numpy(sw.seeded_spectrum(sp, K, 9)).to(torch.complex64).cuda(), sp)
        f = sw.Field(torch.empty((K, grid.npoints), device="cuda"), grid)
        b = sw.LowerTriangularArray(torch.empty_like(c.data), sp)
        row = {}
        for tag, cls, graphs in (("dense_ms", sw.MatrixSpectralTransform, False),
                                 ("sht_ms", sw.SpectralTransform, False), ("sht_graphs_ms", sw.SpectralTransform, True)):
            S = cls(sp, grid, nlayers=K, precision="fp32")
            if graphs:
                S.set_graphs(True)

            def step():
                sw.transform_(f, c, S)
                sw.transform_(b, f, S)
            row[tag] = time_loop(torch, step, steps, warmup) * 1e3
            S.close()
        out[name] = row
    out["workload"] = "round trip (spectral -> grid -> spectral), octahedral grids, FP32"
    return out


def bench_fused_laplace(sw, torch, steps, warmup):
    """T255 K=64 (FP16x3): the grid -> spectral transform fused with ∇² added into the output
    with flipped sign (the Bernoulli potential into the divergence tendency) against the same
    as transform_ then laplace_; and ∇⁻² fused into the spectral -> grid transform (the
    streamfunction from vorticity) against laplace_ then transform_."""
    sp, grid, K = sw.Spectrum(257, 256), sw.OctahedralGaussianGrid(192), 64
    S = sw.SpectralTransform(sp, grid, nlayers=K, precision="fp16x3")
    S.autotune(K)
    c = sw.LowerTriangularArray(torch.from_numpy(sw.seeded_spectrum(sp, K, 13)).to(torch.complex64).cuda(), sp)
    f = sw.transform(c, S)
    out = sw.LowerTriangularArray(torch.zeros_like(c.data), sp)
    tmp = sw.LowerTriangularArray(torch.zeros_like(c.data), sp)
    g = sw.Field(torch.empty_like(f.data), grid)
    res = {
        "analysis_laplace_fused_ms": time_loop(torch, lambda: sw.transform_laplace_(out, f, S, flipsign=True, add=True),
                                               steps, warmup) * 1e3,
        "analysis_then_laplace_ms": time_loop(torch, lambda: (sw.transform_(tmp, f, S),
                                                              sw.laplace_(out, tmp, S, flipsign=True, add=True)),
                                              steps, warmup) * 1e3,
        "inverse_laplace_synthesis_fused_ms": time_loop(torch, lambda: sw.transform_laplace_(g, c, S, inverse=True),
                                                        steps, warmup) * 1e3,
        "inverse_laplace_then_synthesis_ms": time_loop(torch, lambda: (sw.laplace_(tmp, c, S, inverse=True),
                                                                       sw.transform_(g, tmp, S)),
                                                       steps, warmup) * 1e3,
        "workload": "T255 octahedral nlat_half=192, K=64, FP16x3",
    }
    S.close()
    return res


def bench_pullbacks(sw, torch, steps, warmup):
    """T255 K=64 (FP16x3): the adjoint transforms (spec2grid_pullback!, grid2spec_pullback!)
    accumulating into their outputs as the reference's Enzyme rules do, beside the forward
    transforms they mirror (ms per call)."""
    sp, grid, K = sw.Spectrum(257, 256), sw.OctahedralGaussianGrid(192), 64
    S = sw.SpectralTransform(sp, grid, nlayers=K, precision="fp16x3")
    S.autotune(K)
    c = sw.LowerTriangularArray(torch.from_numpy(sw.seeded_spectrum(sp, K, 17)).to(torch.complex64).cuda(), sp)
    f = sw.transform(c, S)
    cb = sw.LowerTriangularArray(torch.zeros_like(c.data), sp)
    fb = sw.Field(torch.zeros_like(f.data), grid)
    res = {
        "synthesis_ms": time_loop(torch, lambda: sw.transform_(fb, c, S), steps, warmup) * 1e3,
        "analysis_ms": time_loop(torch, lambda: sw.transform_(cb, f, S), steps, warmup) * 1e3,
        "spec2grid_pullback_ms": time_loop(torch, lambda: sw.spec2grid_pullback_(cb, f, S), steps, warmup) * 1e3,
        "grid2spec_pullback_ms": time_loop(torch, lambda: sw.grid2spec_pullback_(fb, c, S), steps, warmup) * 1e3,
        "workload": "T255 octahedral nlat_half=192, K=64, FP16x3, pullbacks accumulating",
    }
    S.close()
    return res


def bench_c4(sw, torch, steps, warmup):
    """64 independent round trips, truncations from {T31,T63,T127,T255} (octahedral), layers
    U{1..16} x fields U{1..4}, seed 1234, FP32: one batched launch per stage from a hipGraph,
    against the same 64 round trips issued one transform call at a time."""
    sizes = {31: (24, 33, 32), 63: (48, 65, 64), 127: (96, 129, 128), 255: (192, 257, 256)}
    plans = {T: sw.SpectralTransform(sw.Spectrum(l, m), sw.OctahedralGaussianGrid(n), nlayers=64, precision="fp32")
             for T, (n, l, m) in sizes.items()}
    tuned = {f"T{T}": S.autotune(16) for T, S in plans.items()}  # per size class, plan time
    rng = np.random.default_rng(1234)
    items = []
    for i in range(64):
        T = int(rng.choice([31, 63, 127, 255]))
        K = int(rng.integers(1, 17)) * int(rng.integers(1, 5))
        S = plans[T]
        c = sw.LowerTriangularArray(torch.from_numpy(sw.seeded_spectrum(S.spectrum, K, 1000 + i))
                                    .to(torch.complex64).cuda(), S.spectrum)
        mid = sw.Field(torch.empty((K, S.grid.npoints), device="cuda"), S.grid)
        out = sw.LowerTriangularArray(torch.empty_like(c.data), S.spectrum)
        items.append((S, c, out, mid))
    B = sw.SpectralBatch(items)
    dt_batch = time_loop(torch, B.launch, steps, warmup)

    def sequential():
        for S, c, out, mid in items:
            sw.transform_(mid, c, S)
            sw.transform_(out, mid, S)
    dt_seq = time_loop(torch, sequential, max(steps // 4, 5), max(warmup // 4, 2))
    # compulsory bytes of every round trip (spectral in and out, grid out and in, Λ both ways)
    b_rt = sum(2 * S.spectrum.nharm * c.nlayers * 8 + 2 * S.grid.npoints * c.nlayers * 4
               + 2 * S.spectrum.nharm * S.grid.nlat_half * 4 for S, c, _, _ in items)
    res = {"batch_ms": dt_batch * 1e3, "transforms_per_s": 64.0 / dt_batch, "sequential_ms": dt_seq * 1e3,
           "speedup_vs_sequential": dt_seq / dt_batch, "uses_graph": B.uses_graph,
           "total_columns": int(sum(it[1].nlayers for it in items)), "compulsory_bytes": b_rt,
           "roofline_frac": b_rt / dt_batch / HBM_PEAK, "autotuned": tuned}
    B.close()
    for S in plans.values():
        S.close()
    return res


def bench_c5(sw, seeds=200, heldout=8, steps=3000):
    """Dry primitive-equation step at T31 L8 with the learned column correction: data from
    T127 L8 runs (zonal jets, N(0, 1 K) noise), MLP training on the engine's kernels, and the
    A/B scores of the corrected coarse model against the plain one on held-out runs."""
    import torch
    ln, dy = sw.learned, sw.dycore
    t0 = time.perf_counter()
    x, y = ln.generate(list(range(seeds)), nsamples=2)
    xv, yv = ln.generate(list(range(10000, 10000 + heldout)), nsamples=2)
    t_gen = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net, nx, ny, losses = ln.fit(x, y, steps=steps, batch=4096)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    held = ln.ab_score(net, nx, ny, xv, yv)
    # how the correction couples into the model: chosen on validation runs, scored on others
    variant, kw, val = ln.choose_corrector(net, nx, ny, list(range(30000, 30008)))
    fc = ln.forecast_score(net, nx, ny, list(range(20000, 20008)), variants={variant: kw})[variant]
    # step cost of the coarse model with and without the correction (8 members)
    m = dy.DryPrimitiveModel(31, 8, nmembers=8)
    m.set_state(m.jet_state(list(range(8))))
    m.run(4)
    corr = ln.corrector(net, nx, ny, **kw)
    ms = {}
    for tag, c in (("plain", None), ("corrected", corr)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.run(50, correction=c)
        torch.cuda.synchronize()
        ms[tag] = (time.perf_counter() - t0) / 50 * 1e3
    m.close()
    net.close()
    return {"workload": f"T31 L8 coarse step (8 members) with a 32-128-128-32 column MLP trained on "
                        f"{seeds} T127 L8 runs x 2 samples ({x.shape[0]} columns), {steps} Adam steps of 4096",
            "data_gen_s": t_gen, "train_s": t_fit, "train_columns_per_s": steps * 4096 / t_fit,
            "final_loss": losses[-1], "heldout_mse_ratio": held, "forecast_1h_error_ratio": fc,
            "coupling": variant, "coupling_validation_ratios": val,
            "step_ms": ms}


def profile_table(name):
    """Rows `kernel precision value...` of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(REPO, "profiles", name)
    rows = {}
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) >= 3 and not line.startswith("#"):
                for prec in f[1].split("+"):
                    rows[(f[0], prec)] = f[2:]
    return rows


def measured_traffic(kernel, precision):
    """HBM bytes per launch (fetch + write) of `kernel` from the committed rocprofv3
    counter summary (FETCH_SIZE / WRITE_SIZE, scaled by a plain copy), or None."""
    r = profile_table("counters_t255_k64.txt").get((kernel, precision))
    try:
        return int(r[0]) + int(r[1]) if r else None
    except ValueError:  # a pass without FETCH_SIZE / WRITE_SIZE
        return None


def profiled_ms(kernel, precision):
    """Mean GPU duration (ms) of back-to-back launches of `kernel` in the committed rocprofv3
    kernel-trace summary (the launches kernel_ms times)."""
    r = profile_table("kernel_times_t255_k64.txt").get((kernel, precision))
    return float(r[0]) * 1e-3 if r else None


MODE_NAMES = {"0": "fp32", "1": "fp16x3", "2": "bf16x3"}
XCDS = 8  # MI355X


def summarize_trace(path):
    """Per-(stage kernel, precision) launch statistics of a rocprofv3 kernel trace
    (kernel_trace.csv) of a `--no-extra --tune-cache` run, over the most frequent launch
    shape: lines `kernel precision b2b_mean_us mean_us median_us launches b2b_launches
    workgroups vgpr lds_bytes`.  b2b: the launches that follow a launch of the same kernel
    (sht_time_stage's back-to-back launches, which bench.py's kernel_ms times with HIP
    events); mean / median over every launch (inside round trips too)."""
    import csv
    groups = {}
    prev = None
    for kernel, mode, r in _stage_dispatches(list(csv.DictReader(open(path))), "Start_Timestamp"):
        wg = int(r["Grid_Size_X"]) // max(int(r["Workgroup_Size_X"]), 1)
        g = groups.setdefault((kernel, mode), {}).setdefault((wg, r["VGPR_Count"], r["LDS_Block_Size"]), [])
        g.append(((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3, prev == (kernel, mode)))
        prev = (kernel, mode)
    lines = ["# kernel precision b2b_mean_us mean_us median_us launches b2b_launches workgroups vgpr lds_bytes"]
    for (kernel, mode), byshape in sorted(groups.items()):
        (wg, vgpr, lds), d = max(byshape.items(), key=lambda kv: len(kv[1]))
        allt = [t for t, _ in d]
        b2b = [t for t, rep in d if rep] or allt
        lines.append(f"{kernel} {mode} {np.mean(b2b):.2f} {np.mean(allt):.2f} {np.median(allt):.2f} {len(allt)} "
                     f"{len(b2b)} {wg} {vgpr} {lds}")
    return "\n".join(lines)


def _stage_dispatches(rows, key_start):
    """(kernel, precision, launch shape, row) of the stage-kernel launches in `rows`, in launch
    order: the Legendre kernels carry the precision mode as their first template argument;
    an FFT launch belongs to the mode of the Legendre launch before it."""
    import re
    pat = re.compile(r"sht::(\w+)_kernel<([^>]*)>")
    mode = None
    for r in sorted(rows, key=lambda r: int(r[key_start])):
        mt = pat.search(r["Kernel_Name"])
        if not mt:
            continue
        name, targs = mt.group(1), [a.strip() for a in mt.group(2).split(",")]
        if name.startswith("legendre"):
            mode = MODE_NAMES.get(targs[0], targs[0])
        if mode is None:
            continue
        kernel = {"dft_synthesis": "fft_synthesis", "dft_analysis": "fft_analysis"}.get(name, name)
        yield kernel, mode, r


def summarize_counters(paths):
    """Per-(stage kernel, precision) means per launch of rocprofv3 --pmc counter collections
    (counter_collection.csv of `--no-extra --tune-cache --copy-calibration` runs), over the most
    frequent launch shape, with the derived figures: HBM bytes (FETCH_SIZE and WRITE_SIZE in
    KiB, each scaled by what it reads for the largest plain device copy of the run, whose bytes
    are known), MFMA busy share (SQ_VALU_MFMA_BUSY_CYCLES, summed over the 1024 SIMDs, over
    1024 x the kernel's cycles; GRBM_GUI_ACTIVE sums the 8 XCDs' cycles: it reads 22.7 to 23.4
    per ns of kernel time on MI355X, 8 x 2.4 GHz plus the dispatch),
    MFMA flops (SQ_INSTS_VALU_MFMA_MOPS_* x 512), the share of LDS-active cycles stalled by
    bank conflicts (SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE), and where the waves' time goes,
    as shares of SQ_WAVE_CYCLES: issuing any instruction, VALU, LDS or scalar instructions
    (SQ_ACTIVE_INST_ANY / _VALU / _LDS / _SCA), waiting for a dependency (s_waitcnt,
    SQ_WAIT_ANY) and waiting for issue (SQ_WAIT_INST_ANY)."""
    import csv
    per = {}   # dispatch -> {"row": row, counters...}
    scale = {}
    for path in paths:
        disp = {}
        for r in csv.DictReader(open(path)):
            d = disp.setdefault(r["Dispatch_Id"], {"row": r})
            d[r["Counter_Name"]] = float(r["Counter_Value"])
        copies = [d for d in disp.values() if d["row"]["Kernel_Name"].startswith("__amd_rocclr_copyBuffer")]
        if copies:  # the calibration copy: 256 MiB read and written (the largest copy of the run)
            big = max(copies, key=lambda d: int(d["row"]["Grid_Size"]))
            for c in ("FETCH_SIZE", "WRITE_SIZE"):
                if c in big and big[c] > 0:
                    scale[c] = (256 << 20) / (big[c] * 1024)
        for k, d in disp.items():
            per.setdefault(k, {"row": d["row"]}).update({c: v for c, v in d.items() if c != "row"})
    rows = [dict(d["row"], _c=d) for d in per.values()]
    groups = {}
    for kernel, mode, r in _stage_dispatches(rows, "Dispatch_Id"):
        shape = (int(r["Grid_Size"]) // max(int(r["Workgroup_Size"]), 1), r["VGPR_Count"], r["LDS_Block_Size"])
        groups.setdefault((kernel, mode), {}).setdefault(shape, []).append(r["_c"])
    lines = [f"# copy calibration: bytes = FETCH_SIZE x 1024 x {scale.get('FETCH_SIZE', float('nan')):.4f}, "
             f"WRITE_SIZE x 1024 x {scale.get('WRITE_SIZE', float('nan')):.4f}",
             "# kernel precision fetch_bytes write_bytes launches workgroups mfma_busy_pct mfma_gflop "
             "lds_conflict_pct_of_active issue_any valu lds salu wait_dep wait_issue"]
    for (kernel, mode), byshape in sorted(groups.items()):
        (wg, vgpr, lds), ds = max(byshape.items(), key=lambda kv: len(kv[1]))

        def mean(c):
            v = [d[c] for d in ds if c in d]
            return float(np.mean(v)) if v else float("nan")
        fetch = mean("FETCH_SIZE") * 1024 * scale.get("FETCH_SIZE", float("nan"))
        write = mean("WRITE_SIZE") * 1024 * scale.get("WRITE_SIZE", float("nan"))
        busy = 100 * mean("SQ_VALU_MFMA_BUSY_CYCLES") / (mean("GRBM_GUI_ACTIVE") / XCDS * 1024)
        flop = 512 * sum(mean(c) for c in ("SQ_INSTS_VALU_MFMA_MOPS_F16", "SQ_INSTS_VALU_MFMA_MOPS_BF16",
                                             "SQ_INSTS_VALU_MFMA_MOPS_F32")) * 1e-9
        conf = 100 * mean("SQ_LDS_BANK_CONFLICT") / mean("SQ_LDS_IDX_ACTIVE")
        wc = mean("SQ_WAVE_CYCLES")
        shares = " ".join(f"{mean(c) / wc:.2f}" for c in ("SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS",
                                                          "SQ_ACTIVE_INST_SCA", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY"))
        lines.append(f"{kernel} {mode} {fetch:.0f} {write:.0f} {len(ds)} {wg} {busy:.1f} {flop:.3f} {conf:.1f} {shares}")
    return "\n".join(lines)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spectrum_o, grid_o, coeffs, seconds=10.0):
    """The oracle's FP32 C restatement of the reference CPU path, timed on whole round trips
    of the same workload until `seconds` have elapsed (at least one): on the host cores this
    process may use, capped at 16 (one GPU's share of the host), and on one core."""
    from oracle.cref import CRefTransform
    from oracle.sht_oracle import OracleTransform
    o = OracleTransform(spectrum_o, grid_o)
    c = CRefTransform(o)
    avail = len(os.sched_getaffinity(0))
    threads = min(16, avail)
    a = coeffs.astype(np.complex64)

    def rate(nt, secs):
        c.analysis(c.synthesis(a, np.float32, nthreads=nt), np.float32, nthreads=nt)  # warm
        n, t0 = 0, time.perf_counter()
        while True:
            c.analysis(c.synthesis(a, np.float32, nthreads=nt), np.float32, nthreads=nt)
            n += 1
            if time.perf_counter() - t0 >= secs:
                break
        return n, time.perf_counter() - t0
    n, dt = rate(threads, seconds)
    n1, dt1 = rate(1, seconds)
    return {"value": n / dt, "unit": "round-trips/s", "cores": threads, "kind": "port",
            "sample": f"{n} full T255 K={coeffs.shape[0]} FP32 round trips ({dt:.1f} s, OpenMP {threads} threads; "
                      f"{avail} host cores available, {cpu_model()})",
            "single_thread_value": n1 / dt1,
            "single_thread_sample": f"{n1} round trips in {dt1:.1f} s on one thread",
            "cpu_model": cpu_model(), "host_cores_available": avail}


def copy_bandwidth(torch):
    """Device copy / read (sum) / write (fill) rates of plain torch kernels over 64 MiB, 256 MiB and
    1 GiB buffers, HIP events around 50 back-to-back launches: the memory rates a streaming kernel
    reaches with its data in the Infinity Cache (64 MiB: 128 MiB moved by the copy) and in HBM."""
    out = {}
    for mib in (64, 256, 1024):
        n = mib << 18  # floats
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for name, fn, nbytes in (("copy", lambda: dst.copy_(src), 8 * n), ("read_sum", lambda: src.sum(), 4 * n),
                                 ("write_fill", lambda: dst.fill_(1.0), 4 * n)):
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(50):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 50
            out[f"{name}_{mib}MiB"] = {"ms": ms, "bytes": nbytes, "TB_per_s": nbytes / ms / 1e9}
        del src, dst
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="fp16x3", choices=["fp16x3", "bf16x3", "fp32"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only (for profiling)")
    ap.add_argument("--c5-seeds", type=int, default=200)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run three plain 256 MiB device copies (scale for HBM byte counters)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="print the plain device copy / read / write rates as one JSON line and exit")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: loaded first if it exists (plans then tune without timing), "
                         "written at the end")
    ap.add_argument("--summarize-trace", default=None, metavar="KERNEL_TRACE_CSV",
                    help="print the per-kernel summary of a rocprofv3 kernel trace and exit")
    ap.add_argument("--summarize-counters", nargs="+", default=None, metavar="COUNTER_COLLECTION_CSV",
                    help="print the per-kernel summary of rocprofv3 --pmc collections and exit")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace))
        return
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the transform shards over columns without communication, see the design notes",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_bandwidth:
        torch.cuda.set_device(0)
        print(json.dumps({"copy_bandwidth": copy_bandwidth(torch)}), flush=True)
        return
    import __graft_entry__
    from oracle.sht_oracle import Grid as OGrid, OracleTransform, Spectrum as OSpectrum
    sw = __graft_entry__.load_package()
    torch.cuda.set_device(0)
    if args.tune_cache and os.path.exists(args.tune_cache):
        sw.load_tuning_cache(args.tune_cache)

    lmax, mmax, n, layers, fields = 257, 256, 192, 16, 4
    K = layers * fields
    spectrum = sw.Spectrum(lmax, mmax)
    grid, nlon, mt, M, nfreq_used, nharm = geometry(sw, lmax, mmax, n)
    coeffs = sw.seeded_spectrum(spectrum, K, seed=0)
    coeffs = coeffs.astype(np.complex64).astype(np.complex128)  # the FP32 values the GPU sees

    res = {p: run_mode(sw, torch, p, spectrum, grid, coeffs, args.steps, args.warmup)
           for p in dict.fromkeys([args.precision, "fp32", "fp16x3", "bf16x3"])}
    head = res[args.precision]
    # the other GPU workloads run before any host-heavy work (the FP64 oracle, the CPU
    # baseline), so host threads do not slow their launches
    extra = {}
    if not args.no_extra:
        extra["c2_t127_vector_path"] = bench_c2(sw, torch, args.steps, args.warmup)
        extra["c4_uneven_batch64"] = bench_c4(sw, torch, args.steps, args.warmup)
        extra["dense_matrix_transform"] = bench_dense(sw, torch, args.steps, args.warmup)
        extra["fused_laplace"] = bench_fused_laplace(sw, torch, args.steps, args.warmup)
        extra["pullbacks"] = bench_pullbacks(sw, torch, args.steps, args.warmup)
        extra["c5_learned_correction"] = bench_c5(sw, seeds=args.c5_seeds)

    # accuracy against the FP64 oracle of the same algorithm (per field and degree band)
    o = OracleTransform(OSpectrum(lmax, mmax), OGrid("octahedral", n))
    ref = o.analysis(o.synthesis(coeffs))
    bands = [(0, 15), (16, 63), (64, 127), (128, lmax - 1)]
    errs = {}
    for p, r in res.items():
        worst = 0.0
        per = {}
        for f in range(fields):
            cols = slice(f * layers, (f + 1) * layers)
            for lo, hi in bands:
                sel = (o.spectrum.l_of >= lo) & (o.spectrum.l_of <= hi)
                e = float(np.linalg.norm(r["roundtrip"][cols][:, sel] - ref[cols][:, sel]) /
                          np.linalg.norm(ref[cols][:, sel]))
                per[f"f{f}_l{lo}-{hi}"] = e
                worst = max(worst, e)
        errs[p] = {"worst": worst, "per_field_band": per}

    def roofline(kernel, precision, ms):
        byts, flops, peak = kernel_cost(kernel, K, M, nharm, n, grid.npoints, nlon, precision)
        t = ms * 1e-3
        if byts / HBM_PEAK >= flops / peak:
            r = {"bound": "hbm", "achieved": byts / t / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s"}
        else:
            r = {"bound": "mfma", "achieved": flops / t / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s"}
        r["frac"] = r["achieved"] / r["peak"]
        r["bytes"] = byts
        return r

    kms = head["kernel_ms"]
    dominant = max(KERNELS, key=lambda k: kms[k])
    roof = roofline(dominant, args.precision, kms[dominant])
    roof["traffic"] = measured_traffic(dominant, args.precision)
    roof["traffic_source"] = "profiles/counters_t255_k64.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
    roof["kernel"] = dominant
    roof["kernel_ms"] = kms[dominant]
    pm = profiled_ms(dominant, args.precision)
    if pm:  # the same figure from the committed rocprofv3 kernel-trace mean
        roof["rocprof_kernel_ms"] = pm
        roof["rocprof_frac"] = roofline(dominant, args.precision, pm)["frac"]
    stage_roofline = {p: {k: {"ms": r["kernel_ms"][k], "frac": roofline(k, p, r["kernel_ms"][k])["frac"],
                              "bound": roofline(k, p, r["kernel_ms"][k])["bound"]} for k in KERNELS}
                      for p, r in res.items()}

    # whole round trip against its compulsory bytes / flops
    B_rt = 2 * nharm * K * 8 + 2 * grid.npoints * K * 4 + 2 * nharm * n * 4
    F_rt = 2 * 4.0 * K * M
    P_leg = FP32_PEAK if args.precision == "fp32" else F16_PEAK / MFMA_PRODUCTS[args.precision]
    t_roof = max(B_rt / HBM_PEAK, F_rt / P_leg)

    value = 1.0 / head["s_per_step"]
    out = {
        "metric": "SH round-trips/sec at T255 L16x4 fields, split-prec MFMA vs FP32 @1e-5 rel",
        "value": value,
        "unit": "round-trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp16x3": "f16x3-f32acc", "bf16x3": "bf16x3-f32acc", "fp32": "f32"}[args.precision],
        "data": "synthetic (seeded N(0,1)(l+1)^-1.5 spectrum)",
        "config": {"workload": "T255 spectral->grid->spectral round trip, OctahedralGaussianGrid nlat_half=192, "
                               "16 layers x 4 fields (K=64)", "lmax": lmax, "mmax": mmax, "nlat_half": n, "K": K,
                   "precision": args.precision},
        "roofline": roof,
        "round_trips_per_s": {p: 1.0 / r["s_per_step"] for p, r in res.items()},
        "fp32_round_trips_per_s": 1.0 / res["fp32"]["s_per_step"],
        "speedup_vs_fp32": res["fp32"]["s_per_step"] / head["s_per_step"],
        "kernel_ms": {p: r["kernel_ms"] for p, r in res.items()},
        "stage_roofline": stage_roofline,
        "autotuned": {p: r["tuning"] for p, r in res.items()},
        "roundtrip_roofline_frac": t_roof / head["s_per_step"],
        "max_rel_err_vs_fp64": {p: e["worst"] for p, e in errs.items()},
        "rel_err_per_field_band": {p: e["per_field_band"] for p, e in errs.items()},
    }
    out.update(extra)
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu:
        cb = cpu_baseline(OSpectrum(lmax, mmax), OGrid("octahedral", n), coeffs, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    if args.tune_cache:
        sw.save_tuning_cache(args.tune_cache)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
