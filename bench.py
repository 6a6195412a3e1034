"""Throughput of the spectral random-field generator on one MI355X.

Headline (value): batched FP32 Hermitian generator, 3D lognormal von Karman fields
(512^3, L = 32 cells, nu = 0.3, lognormal mean 1 CoV 0.3), B realisations per step with device
Philox white noise, replayed as one hipGraph, plans picked by the library's autotuner at batch
creation: Gpoints/s of generated output points.  A step = one generation of the B fields
(spectrum + noise, z/y/x passes, marginal, moments).

Also reported on the same line:
  roofline          -- the pass with the most time per step: algorithmic bytes per launch / average
                       duration (HIP events on the library stream), against 8 TB/s; roofline_passes
                       gives every pass, roofline_least_efficient the pass furthest from 8 TB/s
  default_plans     -- the same batch with the planner's fixed rules (no autotuning)
  plain_fp64        -- the same spec, one FP64 field per launch (the plain drop-in path)
  dropin_redft01    -- config 2: the reference-parity REDFT01 drop-in gen_std_gauss_FFT on 256^3
                       (FP64, lognormal), device-resident, and its allocation-free form
  config1_2d        -- 2D 256^2 Gaussian FP64 field, checked against the oracle
  config4_blend     -- 1024 x 1024 x 256 from 8 x 8 x 2 blended subdomains, FP32, one hipGraph
  config5_ensemble  -- 256 mixed-size realizations, FP32 Stockham vs bf16x3 MFMA radix-16 stages vs
                       autotuned, plus the per-axis lag covariance at lags 0-32
  cpu_baseline      -- the oracle's C restatement of the reference FFTW path (REDFT01 FP64 + lognormal
                       marginal, OpenMP over the host cores) on config 2, beside scipy.fft.dctn
                       (pocketfft) on the same array
  fp32_vs_fp64      -- field 0 of the headline against the FP64 oracle on the same noise
  deck_program      -- the deck-driven program end to end on the reference's MAX_Test deck
                       (SIZE 115: 461^3 points, 8 ranks' subdomains in one blend): generation and
                       the HDF5/XDMF output, timed separately.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gpoints/s per GPU (3D FP32 GRF), batched vs plain FP64; % HBM roofline; L2 err"

# algorithmic HBM bytes per output point of each pass (FP32, device noise; the half spectrum
# holds (n1/2+1) n2 n3 complex ~ N/2 complex = 4 B per output point)
def pass_bytes(n, prec_bytes):
    n1, n2, n3 = n
    H = (n1 // 2 + 1) * n2 * n3
    c = 2 * prec_bytes
    N = n1 * n2 * n3
    return {"h_zpass": H * c,            # writes the half spectrum (noise generated on chip)
            "h_ypass": 2 * H * c,        # reads + writes it
            "h_xpass": H * c + N * prec_bytes}  # reads it, writes the real field


def log(msg):
    print(msg, file=sys.stderr, flush=True)


COPY_PROBE_BYTES = 1 << 30


def copy_probe(stream, nbytes=COPY_PROBE_BYTES, reps=5):
    """Device-to-device copy of nbytes on the library stream, timed with HIP events.

    Returns (GB/s counting read + write bytes, average ms per copy).  The same copy kernel
    calibrates the FETCH_SIZE / WRITE_SIZE counters (known bytes per dispatch)."""
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    if hip.hipMemsetAsync(src, 1, C.c_size_t(nbytes), stream) != 0:
        raise RuntimeError("hipMemsetAsync failed in the copy probe")
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(ev0))
    hip.hipEventCreate(C.byref(ev1))
    kind_d2d = 3
    if hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), kind_d2d, stream) != 0:
        raise RuntimeError("hipMemcpyAsync failed in the copy probe")
    hip.hipEventRecord(ev0, stream)
    for _ in range(reps):
        hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), kind_d2d, stream)
    hip.hipEventRecord(ev1, stream)
    if hip.hipEventSynchronize(ev1) != 0:
        raise RuntimeError("copy probe failed")
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    avg = ms.value / reps
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipFree(src)
    hip.hipFree(dst)
    return 2 * nbytes / (avg * 1e-3) / 1e9, avg


# FP32 pass kernels of the headline workload, by the template signature rocprof reports
PASS_KERNELS = {"h_zpass": "k_h_strided<float, true", "h_ypass": "k_h_strided<float, false",
                "h_xpass": "k_h_xpass<float"}


def traffic_from_counters(fetch_csv, write_csv, out_json):
    """HBM bytes per launch of each pass from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE,
    KiB per dispatch), each scaled by what the counter reads for the copy probe's dispatch of
    COPY_PROBE_BYTES (largest copyBuffer dispatch in the same run)."""
    import csv

    def per_kernel(path, counter):
        vals = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return vals

    res = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        vals = per_kernel(path, counter)
        copies = [max(v) for k, v in vals.items() if "copyBuffer" in k]
        if not copies:
            raise SystemExit(f"no copy kernel in {path}")
        factor = COPY_PROBE_BYTES / (max(copies) * 1024.0)
        res[counter] = {"factor": factor}
        for name, pat in PASS_KERNELS.items():
            ks = [v for k, v in vals.items() if pat in k]
            if ks:
                allv = [x for v in ks for x in v]
                res[counter][name] = sum(allv) / len(allv) * 1024.0 * factor
    out = {"method": "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by "
                     "COPY_PROBE_BYTES / counter of the copy probe dispatch",
           "fetch_factor": res["FETCH_SIZE"]["factor"], "write_factor": res["WRITE_SIZE"]["factor"]}
    for name in PASS_KERNELS:
        f, w = res["FETCH_SIZE"].get(name), res["WRITE_SIZE"].get(name)
        if f is not None and w is not None:
            out[name] = {"read_bytes": f, "write_bytes": w, "bytes": f + w}
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def secondary_rows():
    """Kernel time of the other rows' entry points on one 256^3 FP64 field (the library's HIP-event
    timing of each launch; the host-array copies of these drop-in calls are not counted), with
    GB/s against each kernel's algorithmic bytes, or products per second for the lag rows."""
    from scarl_amd import generator as Gm
    from scarl_amd import main_ScaRL, normal_ScaRL, spectra_ScaRL, statistics

    n = 256
    npts = n ** 3
    rng = np.random.default_rng(5)
    field = np.asfortranarray(rng.standard_normal((n, n, n)))
    rows = {}

    def timed(name, fn, kernels, nbytes=None, products=None):
        fn()  # scratch allocation and first-touch outside the timed call
        Gm.timing(True)
        Gm.timing_reset()
        try:
            fn()
            rep = Gm.timing_report()
        finally:
            Gm.timing(False)
        ms = sum(rep[k][1] for k in kernels if k in rep)
        r = {"kernels": kernels, "kernel_ms": round(ms, 4)}
        if nbytes is not None and ms > 0:
            r["alg_bytes"] = nbytes
            r["GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 1)
            r["frac_of_8TBps"] = round(nbytes / (ms * 1e-3) / 8e12, 3)
        if products is not None and ms > 0:
            r["Gproducts_per_s"] = round(products / (ms * 1e-3) / 1e9, 1)
        rows[name] = r

    nl = 1 << 24
    timed("a4 r8vec_uniform_01 (2^24 values)", lambda: normal_ScaRL.r8vec_uniform_01(nl, 12345), ["lcg_uniform"],
          nbytes=12 * nl)
    f1 = field.copy(order="F")
    timed("a5 apply_UnityPartition_mtx (interior of 3x3x3)",
          lambda: main_ScaRL.apply_UnityPartition_mtx((n, n, n), (21, 21, 21), 1, (1, 1, 1), (3, 3, 3), f1),
          ["pou"], nbytes=16 * npts)
    f2 = field.copy(order="F")
    timed("a7 multiVariateTransformation (lognormal)",
          lambda: main_ScaRL.multiVariateTransformation(1.0, 0.3, 2, f2), ["marginal"], nbytes=16 * npts)
    f3 = field.copy(order="F")
    timed("a8 normalize_field (one rank)",
          lambda: main_ScaRL.normalize_field((0, 0, 0), (1, 1, 1), (n, n, n), (0, 0, 0), 0, 0, f3),
          ["box_sums", "normalize"], nbytes=32 * npts)  # mean pass, centred-square pass, normalise
    timed("a9 calculate_average_and_stdVar", lambda: statistics.calculate_average_and_stdVar(field.reshape(-1, 1,
                                                                                                         order="F")),
          ["sums"], nbytes=8 * npts)
    timed("a10 rebuild_Sk_FFT along x (256 periodic lags)",
          lambda: statistics.rebuild_Sk_FFT(field, (n, n, n), 1, 3, 0.0, 1.0), ["lag"], nbytes=8 * npts,
          products=n * npts)
    timed("config 5 lag_covariance (3 axes, 33 lags)", lambda: statistics.lag_covariance(field, 0.0, 33), ["lag"],
          nbytes=24 * npts, products=3 * 33 * npts)
    def build_sk():
        return spectra_ScaRL.build_Sk_mtx(1, (3.0, 3.0, 3.0), (1.0, 1.0, 1.0), (64, 64, 64), (128, 128, 128), 7, 0)

    timed("a3 build_Sk_mtx (Np 64^3, sz_c 128^3)", build_sk, ["build_sk"], nbytes=int(build_sk().nbytes))
    return rows


def deck_program():
    """main_ScaRL.main on the MAX_Test deck with #SIZE# = 115, for 8 ranks, in a scratch folder."""
    import shutil
    import tempfile

    from scarl_amd import main_ScaRL
    src = os.path.join(ROOT, "tests", "golden", "decks", "MAX_Test")
    d = tempfile.mkdtemp(prefix="scarl_deck_")
    try:
        for f in os.listdir(src):
            shutil.copyfile(os.path.join(src, f), os.path.join(d, f))
        p = os.path.join(d, "input_ScaRL.txt")
        txt = open(p).read().replace("#SIZE#", "115")
        open(p, "w").write(txt)
        log("bench: deck program, MAX_Test SIZE 115, 8 ranks")
        t0 = time.perf_counter()
        (res,) = main_ScaRL.main(d, nb_procs=8)
        total = time.perf_counter() - t0
        L = res["geometry"]["L"]
        npts = int(L[0]) * int(L[1]) * int(L[2])
        t = res["times"]
        size = sum(os.path.getsize(os.path.join(d, "SAMPLES", f)) for f in os.listdir(os.path.join(d, "SAMPLES")))
        return {"deck": "bash_ScaRL/MAX_Test/MODEL (SIZE 115), 8 ranks", "L": [int(v) for v in L],
                "topo_shape": res["topo_shape"], "points": npts, "s_total": round(total, 3),
                "s_generation": round(t["Generation"], 4), "s_write_file": round(t["Write_file"], 3),
                "output_MB": round(size / 1e6, 1), "write_MBps": round(size / 1e6 / max(t["Write_file"], 1e-9), 1)}
    except Exception as exc:  # the deck figure never costs the bench line
        return {"error": str(exc)}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _batch_bytes(shapes, prec_bytes):
    """Algorithmic bytes per step of each pass summed over fields of the given extents."""
    tot = {}
    for n in shapes:
        for k, v in pass_bytes(n, prec_bytes).items():
            tot[k] = tot.get(k, 0) + v
    return tot


def _steps(batch, steps, warmup):
    """Seconds per step over `steps` graph replays after `warmup` ones."""
    for _ in range(max(warmup, 1)):
        batch.run(use_graph=True)
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(use_graph=True)
    batch.sync()
    return (time.perf_counter() - t0) / steps


def _pass_table(batch, nruns, step_bytes):
    """Per-pass HIP-event durations of eager runs (every launch alone on the library stream) with
    GB/s and the fraction of 8 TB/s against the algorithmic bytes of each launch."""
    from scarl_amd import generator as Gm
    Gm.timing(True)
    Gm.timing_reset()
    try:
        for _ in range(nruns):
            batch.run(use_graph=False)
        rep = Gm.timing_report()
    finally:
        Gm.timing(False)
    passes = {}
    for k, (cnt, tot) in rep.items():
        avg = tot / cnt
        per_step = cnt / nruns
        passes[k] = {"avg_ms": round(avg, 5), "launches_per_step": per_step, "ms_per_step": round(tot / nruns, 5)}
        if k in step_bytes:
            per_launch = step_bytes[k] / per_step
            gbs = per_launch / (avg * 1e-3) / 1e9
            passes[k].update({"alg_bytes": per_launch, "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
    return passes


def _plans(batch):
    return [f"{r['pass']} N={r['N']} {r['kernel']} {r['radices']} T={r['T']} {r['choice']}"
            + (f" tuned_ms={r['tuned_ms']} default_ms={r['default_ms']}" if "tuned_ms" in r else "")
            + (f" step_ms={r['step_ms']} default_step_ms={r['default_step_ms']}" if "step_ms" in r else "")
            + (" cached" if r.get("cached") else "") for r in batch.plan_report()]


def _make(kind, *args, tune=False, mfma=0):
    """A batch created with the autotuner and/or the MFMA radix-16 stages (1: every pass, 2: the x
    passes) switched on."""
    from scarl_amd import _capi as A
    from scarl_amd import generator as Gm
    Gm.set_tuning(A.TUNE_AUTOTUNE, int(tune))
    Gm.set_tuning(A.TUNE_MFMA_DFT16, int(mfma))
    try:
        return getattr(Gm.Batch, kind)(*args)
    finally:
        Gm.set_tuning(A.TUNE_AUTOTUNE, 0)
        Gm.set_tuning(A.TUNE_MFMA_DFT16, 0)


def _mfma_flops(shapes, x_only=False):
    """bf16 MFMA flops of one generation with every radix-16 first stage on the matrix cores (or
    those of the x passes only): a pass of length N over P pencils issues P N / 256 groups of 16
    columns, each 12 products of 16 x 16 x 32 (2 output halves x 6 split terms), 16384 flop apiece
    (P N x 768 flop); the 256-point x pass of a 512-wide field runs both of its stages there."""
    tot = 0
    for n1, n2, n3 in shapes:
        passes = [(n1 // 2, n2 * n3, True)]
        if not x_only:
            passes += [(n3, (n1 // 2 + 1) * n2, False), (n2, (n1 // 2 + 1) * n3, False)]
        for N, pens, xp in passes:
            if N >= 32 and N % 16 == 0:
                tot += pens * N * 768 * (2 if xp and N == 256 else 1)
    return tot


def config5(args):
    """256 realizations, edges from {64, ..., 192} by seed, half Gaussian L = 12 and half
    exponential L = 20, standard-normal marginal, FP32: Stockham, MFMA and autotuned plans."""
    from scarl_amd import _capi as A
    from scarl_amd import generator as Gm
    edges = np.random.default_rng(5).choice([64, 96, 128, 160, 192], size=256)
    specs = [Gm.FieldSpec(n=(int(e),) * 3, model=1 if i < 128 else 3, corrL=(12.0,) * 3 if i < 128 else (20.0,) * 3,
                          seed=55, stream=i, margi=1, avg=0.0, std_dev=1.0) for i, e in enumerate(edges)]
    shapes = [s.n for s in specs]
    pts = int(sum(int(e) ** 3 for e in edges))
    sb = _batch_bytes(shapes, 4)
    out = {"workload": "256 realizations, cube edges drawn by seed 5 from {64,96,128,160,192}, half Gaussian "
                       "(L=12) half exponential (L=20), standard-normal marginal, FP32, one hipGraph",
           "points": pts, "alg_bytes_per_point": round(sum(sb.values()) / pts, 3)}
    flops = _mfma_flops(shapes)
    for name, tune, mfma in (("stockham", False, 0), ("mfma_bf16x3", False, 1), ("autotuned", True, 0)):
        log(f"bench: config 5 ensemble, {name}")
        b = _make("ensemble", specs, A.PREC_FP32, tune=tune, mfma=mfma)
        try:
            dt = _steps(b, args.steps, args.warmup)
            passes = _pass_table(b, 3, sb)
            r = {"value": round(pts / dt / 1e9, 3), "unit": "Gpoints/s", "ms_per_step": round(dt * 1e3, 4),
                 "step_frac_of_8TBps": round(sum(sb.values()) / dt / 1e9 / HBM_PEAK_GBS, 4), "passes": passes}
            if mfma:
                ms = sum(passes[k]["ms_per_step"] for k in ("h_zpass", "h_ypass", "h_xpass") if k in passes)
                r["mfma_flop_per_step"] = flops
                r["mfma_TFLOPs_over_pass_time"] = round(flops / (ms * 1e-3) / 1e12, 2)
                r["mfma_frac_of_bf16_dense_peak"] = round(flops / (ms * 1e-3) / 2.5e15, 4)
            if tune:
                r["plans"] = _plans(b)
            if name == "stockham":
                # per-realization moments (fused) and per-axis covariance at lags 0-32 (own pass)
                Gm.timing(True)
                Gm.timing_reset()
                try:
                    b.lag_covariance(33)
                    rep = Gm.timing_report()
                finally:
                    Gm.timing(False)
                lag_ms = rep["lag"][1]
                r["lag_cov_33_ms"] = round(lag_ms, 4)
                r["with_lag_cov_Gpts"] = round(pts / (dt + lag_ms * 1e-3) / 1e9, 3)
            out[name] = r
        finally:
            b.close()
    return out


def config4(args):
    """1024 x 1024 x 256 blended from 8 x 8 x 2 subdomains of seeded extents, overlap 16,
    exponential L = 24, lognormal mean 1 CoV 0.5, FP32, one hipGraph (default and autotuned)."""
    from scarl_amd import _capi as A
    from scarl_amd import generator as Gm
    gdim = (1024, 1024, 256)
    starts, exts = Gm.seeded_tiling(gdim, (8, 8, 2), 16, (128, 144, 160, 176, 192), 2024)
    spec = Gm.FieldSpec(n=(2, 2, 2), model=3, corrL=(24.0,) * 3, seed=31, margi=2, avg=1.0, std_dev=0.5)
    shapes = [(exts[0][i], exts[1][j], exts[2][k]) for k in range(2) for j in range(8) for i in range(8)]
    gen_pts = sum(a * b * c for a, b, c in shapes)
    gpts = gdim[0] * gdim[1] * gdim[2]
    sb = _batch_bytes(shapes, 4)
    out = {"workload": "global 1024x1024x256 from 8x8x2 subdomains (extents 128-192 by seed 2024, overlap 16, "
                       "cosine PoU), exponential L=24, lognormal mean 1 CoV 0.5, FP32, one hipGraph",
           "global_points": gpts, "generated_points": gen_pts}
    for name, tune in (("default", False), ("autotuned", True)):
        log(f"bench: config 4 blend, {name}")
        b = _make("blend", gdim, starts, exts, spec, A.PREC_FP32, tune=tune)
        try:
            dt = _steps(b, args.steps, args.warmup)
            passes = _pass_table(b, 3, sb)
            out[name] = {"value": round(gpts / dt / 1e9, 3), "unit": "Gpoints/s (global points)",
                         "generated_Gpts": round(gen_pts / dt / 1e9, 3), "ms_per_step": round(dt * 1e3, 4),
                         "step_frac_of_8TBps": round(sum(sb.values()) / dt / 1e9 / HBM_PEAK_GBS, 4),
                         "passes": passes}
            if tune:
                out[name]["plans"] = _plans(b)
        finally:
            b.close()
    return out


def config1(args):
    """2D 256 x 256, Gaussian C = exp(-(r/8)^2), sigma 1, standard-normal marginal, FP64, one field
    from Philox seed 1: throughput (launch-bound at this size) and parity with the oracle."""
    from scarl_amd import _capi as A
    from scarl_amd import generator as Gm
    from oracle import oracle as O
    spec = Gm.FieldSpec(n=(256, 256, 1), model=A.CM_GAUSSIAN, corrL=(8.0, 8.0, 1.0), seed=1, margi=A.FOM_GAUSSIAN,
                        avg=0.0, std_dev=1.0)
    b = Gm.Batch.ensemble([spec], A.PREC_FP64)
    try:
        dt = _steps(b, max(args.steps, 20), args.warmup)
        f = b.output(0)
    finally:
        b.close()
    ref = O.gen_mode_h((256, 256, 1), (1, 1, 1), A.CM_GAUSSIAN, (8.0, 8.0, 1.0), seed=1, stream=0, margi=1, avg=0.0,
                       std=1.0)
    return {"workload": "2D 256x256 Gaussian L=8, FP64, one field per graph replay", "value": round(256 ** 2 / dt / 1e9, 4),
            "unit": "Gpoints/s", "ms_per_field": round(dt * 1e3, 4),
            "rel_l2_vs_fp64_oracle": float(np.linalg.norm(f - ref) / np.linalg.norm(ref)),
            "sample_mean": float(f.mean()), "sample_var": float(f.var())}


def cpu_baseline_redft01(threads):
    """Config 2 on the host: phases, the oracle's REDFT01 restatement of gen_std_gauss_FFT
    (SRC_F/fftw_ScaRL.f90:293-381, OpenMP) and the lognormal marginal (main_ScaRL.f90:1355-1395) --
    the reference's Generation + Multivar_trans window (main_ScaRL.f90:527-583); one warm-up, then the
    median of three.  Beside it, scipy.fft.dctn(type=3) (pocketfft, `threads` workers) on the same
    octant amplitude, with numpy's exp for the marginal."""
    from oracle import oracle as O
    Np, xr, cl = (256, 256, 256), (255.0, 255.0, 255.0), (16.0, 16.0, 8.0)
    m = tuple(v // 2 for v in Np)
    rng = np.random.default_rng(7)

    def oracle_run():
        phi = np.asfortranarray(rng.random(m))
        return O.marginal(O.gen_std_gauss_fft(Np, xr, cl, 1, phi, nthreads=threads), 2, 1.0, 0.3)

    def med(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    t_or = med(oracle_run)
    lib = None
    try:
        import scipy.fft as sf
        mu, sd = O.marginal_params(2, 1.0, 0.3)

        def scipy_run():
            phi = np.asfortranarray(rng.random(m))
            a = O.mode_r_amplitude(Np, m, xr, cl, 1, phi)
            y = sf.dctn(a, type=3, workers=threads)
            with np.errstate(over="ignore"):  # the reference feeds the un-normalised field to exp
                return np.exp(y * sd + mu)

        t_sp = med(scipy_run)
        lib = {"library": "scipy.fft.dctn type 3 (pocketfft) + numpy exp", "workers": threads,
               "value": round(256 ** 3 / t_sp / 1e9, 5), "unit": "Gpoints/s", "s_per_field": round(t_sp, 4)}
    except Exception as exc:  # the comparator never costs the bench line
        lib = {"error": str(exc)}
    return {"value": round(256 ** 3 / t_or / 1e9, 5), "unit": "Gpoints/s", "cores": threads, "kind": "port",
            "sample": f"config 2: one 256^3 gen_std_gauss_FFT (REDFT01 FP64 on the octant, phases drawn on the host) "
                      f"+ lognormal marginal, oracle C restatement with OpenMP, median of 3 after a warm-up: "
                      f"{t_or:.3f} s per field",
            "library_comparator": lib}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the GPU needs a few dozen replays to reach its steady clocks (the first 20 replays after
    # creation run about 2 % slower than later ones); the warmup is untimed
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--edge", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 oracle check")
    ap.add_argument("--no-deck", action="store_true", help="skip the deck-program figure")
    ap.add_argument("--no-secondary", action="store_true", help="skip the other rows' kernel timings")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 1, 4 and 5 lines")
    ap.add_argument("--no-tune", action="store_true", help="headline with the planner's fixed rules")
    ap.add_argument("--probe-only", action="store_true", help="run the copy probe and the headline only")
    ap.add_argument("--only", choices=("config1", "config4", "config5"),
                    help="run one config line alone (for profiling) and print it")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc counter CSVs and exit")
    args = ap.parse_args()
    if args.traffic:
        traffic_from_counters(args.traffic[0], args.traffic[1], os.path.join(ROOT, "profiles", "traffic.json"))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs on several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from scarl_amd import _capi as A
    from scarl_amd import generator as Gm

    A.call("scarl_hip_init", 0)
    if args.only:
        fn = {"config1": config1, "config4": config4, "config5": config5}[args.only]
        print(json.dumps({args.only: fn(args)}), flush=True)
        return
    lib = A.lib()
    stream = C.c_void_p()
    A.call("scarl_hip_get_stream", C.byref(stream))
    copy_gbs, copy_ms = copy_probe(stream)
    log(f"bench: device copy probe {copy_gbs:.0f} GB/s")
    e = args.edge
    n = (e, e, e)
    L = (32.0, 32.0, 32.0)
    B = args.batch
    specs = [Gm.FieldSpec(n=n, model=A.CM_VON_KARMAN, corrL=L, nu=0.3, seed=2024, stream=i, margi=A.FOM_LOGNORMAL,
                          avg=1.0, std_dev=0.3) for i in range(B)]
    pts = B * e ** 3
    pb = pass_bytes(n, 4)
    sb = {k: B * v for k, v in pb.items()}

    # ---- headline: batched FP32, hipGraph, autotuned plans --------------------------------
    tune = not args.no_tune
    log(f"bench: batched fp32 {B} x {e}^3 ({'autotuned' if tune else 'default'} plans)")
    batch = _make("ensemble", specs, A.PREC_FP32, tune=tune)
    plans = _plans(batch)
    dt = _steps(batch, args.steps, args.warmup)
    value = pts / dt / 1e9
    log(f"bench: {dt * 1e3:.3f} ms/step, {value:.1f} Gpts/s")
    nruns = max(3, min(args.steps, 10))
    passes = _pass_table(batch, nruns, sb)
    log(f"bench: passes {json.dumps(passes)}")
    f32_field0 = batch.output(0)
    moments = batch.moments()
    batch.close()
    if args.probe_only:
        print(json.dumps({"copy_probe_GBps": copy_gbs, "value": value}), flush=True)
        return

    traffic_all = {}
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic_all = json.load(open(tfile))
    roof_passes = {}
    for k in pb:
        if k in passes:
            t = traffic_all.get(k)
            roof_passes[k] = {"bound": "hbm", "achieved": passes[k]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": passes[k]["frac"], "traffic": None if t is None else t["bytes"],
                              "bytes_per_launch": passes[k]["alg_bytes"], "avg_ms": passes[k]["avg_ms"],
                              "ms_per_step": passes[k]["ms_per_step"]}
    dom = max(roof_passes, key=lambda k: roof_passes[k]["ms_per_step"])
    worst = min(roof_passes, key=lambda k: roof_passes[k]["frac"])
    roofline = dict(roof_passes[dom], kernel=dom, frac_of_copy_probe=round(roof_passes[dom]["achieved"] / copy_gbs, 4))
    step_bytes = sum(sb.values())
    step_frac = step_bytes / dt / 1e9 / HBM_PEAK_GBS

    # ---- the same batch with the planner's fixed rules ------------------------------------
    default_plans = None
    if tune:
        log("bench: batched fp32, default plans")
        db = _make("ensemble", specs, A.PREC_FP32)
        try:
            dtd = _steps(db, args.steps, args.warmup)
            default_plans = {"value": round(pts / dtd / 1e9, 3), "unit": "Gpoints/s", "ms_per_step": round(dtd * 1e3, 4),
                             "plans": _plans(db)}
        finally:
            db.close()

    # ---- the same batch with the x pass as split-precision bf16 MFMA radix-16 stages ----------
    mfma = {"scheme": "radix-16 stages as [Re -Im; Im Re] bf16 MFMA products, three-term split, six products, "
                      "FP32 accumulation"}
    m0 = None
    for mode, name, kind in ((2, "x_pass", "x pass only: 256 = 16 x 16, both stages"),
                             (1, "all_passes", "every radix-16 first stage (z, y: 512 = 16 x 8 x 4; x: 256, both)")):
        log(f"bench: batched fp32, MFMA radix-16 stages ({name})")
        mb = _make("ensemble", specs, A.PREC_FP32, mfma=mode)
        try:
            dtm = _steps(mb, args.steps, args.warmup)
            mpass = _pass_table(mb, nruns, sb)
            mo = mb.output(0)
        finally:
            mb.close()
        if m0 is None:
            m0 = mo
        mflops = _mfma_flops([n] * B, x_only=mode == 2)
        mks = ("h_xpass",) if mode == 2 else ("h_zpass", "h_ypass", "h_xpass")
        mms = sum(mpass[k]["ms_per_step"] for k in mks)
        mfma[name] = {"value": round(pts / dtm / 1e9, 3), "unit": "Gpoints/s", "ms_per_step": round(dtm * 1e3, 4),
                      "stages": kind, "passes": mpass, "mfma_flop_per_step": mflops,
                      "mfma_TFLOPs_over_pass_time": round(mflops / (mms * 1e-3) / 1e12, 2),
                      "mfma_frac_of_bf16_dense_peak": round(mflops / (mms * 1e-3) / 2.5e15, 4),
                      "rel_l2_vs_fp32_stockham": float(np.linalg.norm(mo.astype(np.float64) - f32_field0) /
                                                       np.linalg.norm(f32_field0.astype(np.float64)))}

    # ---- plain FP64, one field per launch ------------------------------------------------
    log("bench: plain fp64, one field per launch")
    plain_rows = {}
    for name, tn in (("default", False), ("autotuned", True)):
        plain = _make("ensemble", specs[:1], A.PREC_FP64, tune=tn)
        try:
            plain.run()
            plain.sync()
            ks = max(2, args.steps // 2)
            t0 = time.perf_counter()
            for _ in range(ks):
                plain.run()
            plain.sync()
            dtp = (time.perf_counter() - t0) / ks
            pp = _pass_table(plain, 2, pass_bytes(n, 8))
            plain_rows[name] = {"value": round(e ** 3 / dtp / 1e9, 3), "unit": "Gpoints/s",
                                "ms_per_field": round(dtp * 1e3, 4), "passes": pp}
            if tn:
                plain_rows[name]["plans"] = _plans(plain)
        finally:
            plain.close()
    plain_value = max(r["value"] for r in plain_rows.values())

    # ---- config 2: reference-parity REDFT01 drop-in, 256^3 FP64, inputs resident ------------
    log("bench: REDFT01 drop-in 256^3 fp64")
    hip = C.CDLL("libamdhip64.so")
    dptr = C.c_void_p()
    rc = hip.hipMalloc(C.byref(dptr), C.c_size_t(256 ** 3 * 8))
    if rc != 0:
        raise RuntimeError("hipMalloc failed")
    Np = np.array([256, 256, 256], np.int32)
    xr = np.full(3, 255.0)
    cl = np.array([16.0, 16.0, 8.0])
    A.call("scarl_hip_set_marginal", 2, 1.0, 0.3)
    ks = max(2, args.steps // 2)

    def redft_rate(name):
        fn = getattr(lib, name)

        def call():
            A.check(fn(dptr, A.ptr(Np), A.ptr(xr), A.ptr(cl), 1, 7, 0, None, 1))

        call()
        A.check(lib.scarl_hip_batch_sync(None))
        t0 = time.perf_counter()
        for _ in range(ks):
            call()
        A.check(lib.scarl_hip_batch_sync(None))
        return (time.perf_counter() - t0) / ks

    dtr = redft_rate("scarl_hip_gen_std_gauss_fft_dev")
    dtn = redft_rate("scarl_hip_gen_std_gauss_fft_dev_noalloc")
    Gm.timing(True)
    Gm.timing_reset()
    A.check(lib.scarl_hip_gen_std_gauss_fft_dev(dptr, A.ptr(Np), A.ptr(xr), A.ptr(cl), 1, 7, 0, None, 1))
    rrep = Gm.timing_report()
    Gm.timing(False)
    hip.hipFree(dptr)
    # the same call with the host output array of the Fortran interface: PCIe-inclusive rate
    from scarl_amd import fftw_ScaRL
    hout = np.empty((256, 256, 256), np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(hout, (256, 256, 256), tuple(xr), tuple(cl), 1, 7, 0)
    t0 = time.perf_counter()
    for _ in range(3):
        fftw_ScaRL.gen_std_gauss_FFT(hout, (256, 256, 256), tuple(xr), tuple(cl), 1, 7, 0)
    dth = (time.perf_counter() - t0) / 3
    rbytes = {"r_xpass": 8 * 256 ** 3 * (1 / 8 + 1 / 4), "r_ypass": 8 * 256 ** 3 * (1 / 4 + 1 / 2),
              "r_zpass": 8 * 256 ** 3 * (1 / 2 + 1)}
    rpass = {k: {"ms": round(v[1] / v[0], 5), "GBps": round(rbytes[k] / (v[1] / v[0] * 1e-3) / 1e9, 1),
                 "frac": round(rbytes[k] / (v[1] / v[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
             for k, v in rrep.items() if k in rbytes}

    # ---- configs 1, 4, 5 --------------------------------------------------------------------
    c1 = c4 = c5 = None
    if not args.no_configs:
        c1 = config1(args)
        c4 = config4(args)
        c5 = config5(args)

    # ---- CPU baseline (config 2 on the host) and FP32 error on shared noise -----------------
    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"bench: CPU baseline, config 2 256^3 REDFT01, {threads} threads")
        cpu = cpu_baseline_redft01(threads)
        log(f"bench: FP64 oracle of headline field 0, {threads} threads")
        ref = O.gen_mode_h(n, (1, 1, 1), A.CM_VON_KARMAN, L, 0.3, seed=2024, stream=0, margi=A.FOM_LOGNORMAL,
                           avg=1.0, std=0.3, nthreads=threads)
        err = {"rel_l2": float(np.linalg.norm(f32_field0 - ref) / np.linalg.norm(ref)),
               "max_abs": float(np.abs(f32_field0.astype(np.float64) - ref).max()),
                      "mfma_x_pass_rel_l2": float(np.linalg.norm(m0 - ref) / np.linalg.norm(ref))}

    # ---- the other rows' kernels on one 256^3 FP64 field ------------------------------------
    secondary = None
    if not args.no_secondary:
        try:
            secondary = secondary_rows()
        except Exception as exc:  # never costs the bench line
            secondary = {"error": str(exc)}

    # ---- the deck-driven program (generation + HDF5/XDMF output) ----------------------------
    deck = None
    if not args.no_deck:
        deck = deck_program()

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: device Philox4x32-10 Hermitian white noise",
        "config": {"workload": f"batched Hermitian C2R, {B} x {e}^3 lognormal von Karman (L=32 cells, nu=0.3, "
                               f"mean 1, CoV 0.3), one hipGraph per step, "
                               f"{'autotuned' if tune else 'default'} plans",
                   "fields_per_step": B, "edge": e, "precision": "fp32 butterflies, fp64-rounded-once twiddles"},
        "roofline": roofline,
        "roofline_passes": roof_passes,
        "roofline_least_efficient": dict(roof_passes[worst], kernel=worst),
        "step_roofline_frac": round(step_frac, 4),
        "copy_probe": {"GBps": round(copy_gbs, 1), "bytes": COPY_PROBE_BYTES, "avg_ms": round(copy_ms, 4),
                       "step_frac_of_copy": round(step_bytes / dt / 1e9 / copy_gbs, 4)},
        "alg_bytes_per_point": sum(pb.values()) / e ** 3,
        "passes": passes,
        "plans": plans,
        "default_plans": default_plans,
        "plain_fp64": dict(plain_rows[max(plain_rows, key=lambda k: plain_rows[k]["value"])],
                           workload=f"one {e}^3 field per launch, FP64", variants=plain_rows),
        "batched_vs_plain": round(value / plain_value, 3),
        "dropin_redft01": {"value": round(256 ** 3 / dtr / 1e9, 3), "unit": "Gpoints/s",
                           "ms_per_field": round(dtr * 1e3, 4), "passes": rpass,
                           "noalloc_value": round(256 ** 3 / dtn / 1e9, 3),
                           "workload": "config 2: gen_std_gauss_FFT 256^3 FP64 + lognormal, device phi, resident"},
        "dropin_redft01_host_array": {"value": round(256 ** 3 / dth / 1e9, 3), "unit": "Gpoints/s",
                                      "ms_per_field": round(dth * 1e3, 4),
                                      "workload": "the same call into a host array (PCIe copy included)"},
        "config1_2d": c1,
        "config4_blend": c4,
        "config5_ensemble": c5,
        "fp32_vs_fp64": err,
        "mfma_variant": mfma,
        "moments_field0": {"mean": float(moments[0, 3]), "std": float(moments[0, 4])},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "deck_program": deck,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
