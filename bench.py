"""Throughput of the spectral random-field generator on one MI355X.

Headline (value): batched FP32 Hermitian generator, 3D lognormal von Karman fields
(512^3, L = 32 cells, nu = 0.3, lognormal mean 1 CoV 0.3), B realisations per step with device
Philox white noise, replayed as one hipGraph: Gpoints/s of generated output points.
A step = one generation of the B fields (spectrum + noise, z/y/x passes, marginal, moments).

Also reported on the same line:
  plain_fp64        -- the same spec, one FP64 field per launch (the plain drop-in path)
  dropin_redft01    -- the reference-parity REDFT01 drop-in gen_std_gauss_FFT on 256^3 (FP64)
  roofline          -- dominant kernel: algorithmic bytes per launch / average duration (HIP
                       events on the library stream), against 8 TB/s
  cpu_baseline      -- the FP64 oracle (C port, OpenMP) on one field of the same workload; that
                       field also gives the FP32-vs-FP64 relative L2 error on shared noise.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--edge", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-deck", action="store_true", help="skip the deck-program figure")
    ap.add_argument("--no-secondary", action="store_true", help="skip the other rows' kernel timings")
    ap.add_argument("--probe-only", action="store_true", help="run the copy probe and the headline only")
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

    # ---- headline: batched FP32, hipGraph ------------------------------------------------
    log(f"bench: batched fp32 {B} x {e}^3")
    batch = Gm.Batch.ensemble(specs, A.PREC_FP32)
    for _ in range(max(args.warmup, 1)):
        batch.run(use_graph=True)
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run(use_graph=True)
    batch.sync()
    dt = (time.perf_counter() - t0) / args.steps
    pts = B * e ** 3
    value = pts / dt / 1e9
    log(f"bench: {dt * 1e3:.3f} ms/step, {value:.1f} Gpts/s")

    # per-kernel durations with HIP events on the library stream (eager launches)
    # (with timing on every launch runs alone on the library stream; a pass may take several
    # launches per step, one per field group)
    Gm.timing(True)
    Gm.timing_reset()
    nruns = max(3, min(args.steps, 10))
    for _ in range(nruns):
        batch.run(use_graph=False)
    rep = Gm.timing_report()
    Gm.timing(False)
    pb = pass_bytes(n, 4)
    passes = {}
    for k, (cnt, tot) in rep.items():
        avg = tot / cnt
        per_step = cnt / nruns
        byts = pb.get(k)
        passes[k] = {"avg_ms": avg, "launches": cnt, "launches_per_step": per_step, "ms_per_step": tot / nruns}
        if byts is not None:
            per_launch = B * byts / per_step
            gbs = per_launch / (avg * 1e-3) / 1e9
            passes[k].update({"alg_bytes": per_launch, "GBps": gbs, "frac": gbs / HBM_PEAK_GBS})
    dom = max((k for k in passes if k in pb), key=lambda k: passes[k]["ms_per_step"])
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile)).get(dom)
        traffic = None if t is None else t["bytes"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": passes[dom]["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": passes[dom]["frac"], "traffic": traffic,
                "bytes_per_launch": passes[dom]["alg_bytes"], "avg_ms": passes[dom]["avg_ms"],
                "frac_of_copy_probe": passes[dom]["GBps"] / copy_gbs}
    step_bytes = sum(pb.values()) * B
    step_frac = step_bytes / dt / 1e9 / HBM_PEAK_GBS
    log(f"bench: passes {json.dumps(passes)}")

    # keep field 0 for the error checks
    f32_field0 = batch.output(0)
    moments = batch.moments()
    batch.close()
    if args.probe_only:
        print(json.dumps({"copy_probe_GBps": copy_gbs, "value": value}), flush=True)
        return

    # ---- the same batch with the x pass as split-precision bf16 MFMA radix-16 stages ----------
    log("bench: batched fp32, MFMA radix-16 x pass")
    Gm.set_tuning(A.TUNE_MFMA_DFT16, 1)
    try:
        mb = Gm.Batch.ensemble(specs, A.PREC_FP32)
    finally:
        Gm.set_tuning(A.TUNE_MFMA_DFT16, 0)
    for _ in range(max(args.warmup, 1)):
        mb.run(use_graph=True)
    mb.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mb.run(use_graph=True)
    mb.sync()
    dtm = (time.perf_counter() - t0) / args.steps
    Gm.timing(True)
    Gm.timing_reset()
    for _ in range(nruns):
        mb.run(use_graph=False)
    mrep = Gm.timing_report()
    Gm.timing(False)
    mx = mrep["h_xpass"][1] / mrep["h_xpass"][0]
    m0 = mb.output(0)
    mb.close()
    mfma = {"value": round(pts / dtm / 1e9, 3), "unit": "Gpoints/s", "ms_per_step": round(dtm * 1e3, 4),
            "h_xpass_avg_ms": mx, "h_xpass_GBps": pb["h_xpass"] * B / (mrep["h_xpass"][0] / nruns) / (mx * 1e-3) / 1e9,
            # per pencil: 2 stages x 2 output tiles (Re, Im) x 6 split products of a 16x16x32 bf16 MFMA
            "mfma_TFLOPs": (e * e) * 2 * 2 * 6 * (16 * 16 * 32 * 2) / (mx * 1e-3) / 1e12,
            "mfma_frac_of_bf16_dense_peak": (e * e) * 2 * 2 * 6 * (16 * 16 * 32 * 2) / (mx * 1e-3) / 2.5e15,
            "rel_l2_vs_fp32_stockham": float(np.linalg.norm(m0.astype(np.float64) - f32_field0) /
                                             np.linalg.norm(f32_field0.astype(np.float64))),
            "scheme": "x pass radix-16 stages as [Re -Im; Im Re] bf16 MFMA products, three-term split, "
                      "six products, FP32 accumulation"}

    # ---- plain FP64, one field per launch ------------------------------------------------
    log("bench: plain fp64, one field per launch")
    plain = Gm.Batch.ensemble(specs[:1], A.PREC_FP64)
    plain.run()
    plain.sync()
    ks = max(2, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(ks):
        plain.run()
    plain.sync()
    dtp = (time.perf_counter() - t0) / ks
    plain_value = e ** 3 / dtp / 1e9
    plain.close()

    # ---- reference-parity REDFT01 drop-in, 256^3 FP64, inputs resident --------------------
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

    def redft():
        A.check(lib.scarl_hip_gen_std_gauss_fft_dev(dptr, A.ptr(Np), A.ptr(xr), A.ptr(cl), 1, 7, 0, None, 1))

    redft()
    A.check(lib.scarl_hip_batch_sync(None))
    t0 = time.perf_counter()
    for _ in range(ks):
        redft()
    A.check(lib.scarl_hip_batch_sync(None))
    dtr = (time.perf_counter() - t0) / ks
    hip.hipFree(dptr)
    # the same call with the host output array of the Fortran interface: PCIe-inclusive rate
    from scarl_amd import fftw_ScaRL
    hout = np.empty((256, 256, 256), np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(hout, (256, 256, 256), tuple(xr), tuple(cl), 1, 7, 0)
    t0 = time.perf_counter()
    for _ in range(3):
        fftw_ScaRL.gen_std_gauss_FFT(hout, (256, 256, 256), tuple(xr), tuple(cl), 1, 7, 0)
    dth = (time.perf_counter() - t0) / 3

    # ---- CPU baseline (oracle, FP64) and FP32 error on shared noise ----------------------
    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"bench: CPU oracle, one {e}^3 field, {threads} threads")
        t0 = time.perf_counter()
        ref = O.gen_mode_h(n, (1, 1, 1), A.CM_VON_KARMAN, L, 0.3, seed=2024, stream=0, margi=A.FOM_LOGNORMAL,
                           avg=1.0, std=0.3, nthreads=threads)
        tc = time.perf_counter() - t0
        cpu = {"value": e ** 3 / tc / 1e9, "unit": "Gpoints/s", "cores": threads, "kind": "port",
               "sample": f"one {e}^3 lognormal von Karman field (same spec and Philox noise as field 0), "
                         f"FP64 oracle C port, {tc:.2f} s"}
        err = {"rel_l2": float(np.linalg.norm(f32_field0 - ref) / np.linalg.norm(ref)),
               "max_abs": float(np.abs(f32_field0.astype(np.float64) - ref).max()),
               "mfma_rel_l2": float(np.linalg.norm(m0 - ref) / np.linalg.norm(ref))}

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
                               f"mean 1, CoV 0.3), one hipGraph per step",
                   "fields_per_step": B, "edge": e, "precision": "fp32 butterflies, fp64-rounded-once twiddles"},
        "roofline": roofline,
        "step_roofline_frac": round(step_frac, 4),
        "copy_probe": {"GBps": round(copy_gbs, 1), "bytes": COPY_PROBE_BYTES, "avg_ms": round(copy_ms, 4),
                       "step_frac_of_copy": round(step_bytes / dt / 1e9 / copy_gbs, 4)},
        "alg_bytes_per_point": sum(pb.values()) / e ** 3,
        "passes": passes,
        "plain_fp64": {"value": round(plain_value, 3), "unit": "Gpoints/s", "ms_per_field": round(dtp * 1e3, 4),
                       "workload": f"one {e}^3 field per launch, FP64"},
        "batched_vs_plain": round(value / plain_value, 3),
        "dropin_redft01": {"value": round(256 ** 3 / dtr / 1e9, 3), "unit": "Gpoints/s",
                           "ms_per_field": round(dtr * 1e3, 4),
                           "workload": "gen_std_gauss_FFT 256^3 FP64 + lognormal, device phi, resident"},
        "dropin_redft01_host_array": {"value": round(256 ** 3 / dth / 1e9, 3), "unit": "Gpoints/s",
                                      "ms_per_field": round(dth * 1e3, 4),
                                      "workload": "the same call into a host array (PCIe copy included)"},
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
