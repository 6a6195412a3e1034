"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the same inputs.

Bars: FP64 paths <= 1e-12 relative L2 (reference-parity REDFT01 and Hermitian C2R);
integer LCG bit-exact; FP32 Hermitian path <= 1e-5 relative L2 against the FP64 oracle on the
same white noise.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _say(msg):
    print(msg, flush=True)


# ---- fftw_ScaRL: gen_std_gauss_FFT (REDFT01, FP64) --------------------------------------
@pytest.mark.parametrize("Np", [(24, 20, 16), (5, 8, 6), (41, 41, 41), (81, 30, 17), (2, 3, 4), (1, 6, 6),
                                (64, 64, 64), (121, 40, 9)])
def test_gen_std_gauss_fft_host_phi(hip, Np):
    from scarl_amd import fftw_ScaRL
    _say(f"gen_std_gauss_FFT Np={Np}")
    rng = np.random.default_rng(sum(Np))
    phi = np.asfortranarray(rng.random(tuple(v // 2 for v in Np)))
    xr = tuple(float(v - 1) for v in Np)
    cl = (3.0, 2.5, 1.5)
    gpu = np.empty(Np, np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(gpu, Np, xr, cl, 1, 1, 0, phi=phi)
    ref = O.gen_std_gauss_fft(Np, xr, cl, 1, phi)
    if np.linalg.norm(ref) == 0:
        assert np.all(gpu == 0)
    else:
        assert rel_l2(gpu, ref) <= FP64_TOL


def test_gen_std_gauss_fft_256_cube(hip):
    """256^3 anisotropic Gaussian (corrL 16,16,8 cells), seed-7 phases, lognormal marginal."""
    from scarl_amd import fftw_ScaRL
    _say("gen_std_gauss_FFT 256^3")
    Np = (256, 256, 256)
    phi = np.asfortranarray(np.random.default_rng(7).random((128, 128, 128)))
    xr, cl = (255.0, 255.0, 255.0), (16.0, 16.0, 8.0)
    gpu = np.empty(Np, np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(gpu, Np, xr, cl, 1, 7, 0, phi=phi)
    ref = O.gen_std_gauss_fft(Np, xr, cl, 1, phi)
    assert rel_l2(gpu, ref) <= FP64_TOL
    # the marginal fused into the z pass equals the separate transform.  The reference feeds the
    # unnormalised field (|x| ~ 1e5 here) to the lognormal, so exp overflows to inf at the same
    # points on both sides (SRC_F/main_ScaRL.f90:565-567,1388-1393).
    fused = np.empty(Np, np.float64, order="F")
    fftw_ScaRL.set_marginal(1, 1.0, 0.3)
    fftw_ScaRL.gen_std_gauss_FFT(fused, Np, xr, cl, 1, 7, 0, phi=phi, fuse_marginal=True)
    assert rel_l2(fused, O.marginal(ref, 1, 1.0, 0.3)) <= FP64_TOL
    fftw_ScaRL.set_marginal(2, 1.0, 0.3)
    fftw_ScaRL.gen_std_gauss_FFT(fused, Np, xr, cl, 1, 7, 0, phi=phi, fuse_marginal=True)
    m = O.marginal(ref, 2, 1.0, 0.3)
    np.testing.assert_array_equal(np.isinf(fused), np.isinf(m))
    np.testing.assert_array_equal(fused == 0, m == 0)
    # subnormal results (exponent below -708) carry too few significant bits to invert
    tiny = np.finfo(np.float64).tiny
    ok = np.isfinite(m) & (m >= tiny) & (fused >= tiny)
    mu, sd = O.marginal_params(2, 1.0, 0.3)
    assert rel_l2((np.log(fused[ok]) - mu) / sd, ref[ok]) <= FP64_TOL


def test_gen_std_gauss_fft_device_phi(hip):
    from scarl_amd import fftw_ScaRL
    Np = (20, 18, 12)
    gpu = np.empty(Np, np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(gpu, Np, (19.0, 17.0, 11.0), (2.0, 2.0, 2.0), 1, 5, 3)
    phi = O.phi_octant((10, 9, 6), 5 + 3, 0)
    ref = O.gen_std_gauss_fft(Np, (19.0, 17.0, 11.0), (2.0, 2.0, 2.0), 1, phi)
    assert rel_l2(gpu, ref) <= FP64_TOL


def test_gen_std_gauss_fft_mpi_one_rank(hip):
    from scarl_amd import fftw_ScaRL
    L = (16, 12, 10)
    phi = np.asfortranarray(np.random.default_rng(3).random((8, 6, 10)))
    gpu = fftw_ScaRL.gen_Std_Gauss_FFT_MPI((15.0, 11.0, 9.0), (1.0, 1.0, 1.0), 1, 1, 0, L, L, 1, phi=phi)
    ref = O.gen_std_gauss_fft_mpi(L, (15.0, 11.0, 9.0), (1.0, 1.0, 1.0), 1, phi)
    assert rel_l2(gpu, ref) <= FP64_TOL


def test_gen_std_gauss_fft_errors(hip):
    from scarl_amd import fftw_ScaRL, ScarlError
    out = np.empty((8, 8, 8), np.float64, order="F")
    with pytest.raises(ScarlError):
        fftw_ScaRL.gen_std_gauss_FFT(out, (8, 8, 8), (7.0, 7.0, 7.0), (1.0, 1.0, 1.0), 2, 1, 0)  # von Karman
    with pytest.raises(ScarlError):
        fftw_ScaRL.gen_Std_Gauss_FFT_MPI((7.0, 7.0, 7.0), (1.0, 1.0, 1.0), 1, 1, 0, (8, 8, 8), (8, 8, 8), 2)


# ---- normal_ScaRL ------------------------------------------------------------------------
@pytest.mark.parametrize("n,seed", [(1, 12345), (3, 12345), (2, 1), (1000, 777), (100_000, -123456789),
                                    (5, 2147483646), (3_000_001, 987654321)])
def test_lcg_bit_exact(hip, n, seed):
    """n = 3000001 exceeds the kernel's 4096 x 256 threads, so threads step by 16807^G."""
    from scarl_amd import normal_ScaRL
    r, s = normal_ScaRL.r8vec_uniform_01(n, seed)
    ro, so = O.r8vec_uniform_01(n, seed)
    assert s == so
    np.testing.assert_array_equal(r, ro)
    x, s = normal_ScaRL.r8vec_normal_01(n, seed)
    xo, so = O.r8vec_normal_01(n, seed)
    assert s == so
    np.testing.assert_allclose(x, xo, rtol=1e-15, atol=1e-15)


def test_lcg_known_answer(hip):
    from scarl_amd import normal_ScaRL
    r, s = normal_ScaRL.r8_uniform_01(12345)
    assert s == 207482415
    r, s = normal_ScaRL.r8vec_uniform_01(3, 12345)
    assert s == 2035175616


# ---- main_ScaRL element routines ---------------------------------------------------------
@pytest.mark.parametrize("shape", [(2, 1, 1), (2, 2, 1), (2, 2, 2), (3, 3, 3)])
def test_unity_partition(hip, shape):
    """x rows of 20 points (x weights held in registers) and of 1100 points (formed per point)."""
    from scarl_amd import main_ScaRL
    rng = np.random.default_rng(1)
    for Np, Nov in (((20, 17, 12), (5, 4, 3)), ((1100, 3, 2), (40, 2, 2))):
        for pos in [tuple(p) for p in np.ndindex(*shape)][:5]:
            f = np.asfortranarray(rng.standard_normal(Np))
            ref = O.apply_pou(f, Nov, pos, shape)
            main_ScaRL.apply_UnityPartition_mtx(Np, Nov, 1, pos, shape, f)
            np.testing.assert_allclose(f, ref, rtol=1e-15, atol=1e-15)


def test_marginal(hip):
    from scarl_amd import main_ScaRL, ScarlError
    x = np.asfortranarray(np.random.default_rng(2).standard_normal((30, 20, 10)))
    for margi, avg, sd in ((1, 2.0, 0.5), (2, 1.0, 0.3), (2, 3.0, 1.5)):
        y = main_ScaRL.multiVariateTransformation(avg, sd, margi, x.copy(order="F"))
        np.testing.assert_allclose(y, O.marginal(x, margi, avg, sd), rtol=1e-15)
    with pytest.raises(ScarlError):
        main_ScaRL.multiVariateTransformation(-1.0, 0.3, 2, x.copy(order="F"))


def test_moments_compensated(hip):
    from scarl_amd import statistics
    rng = np.random.default_rng(4)
    rf = np.asfortranarray(1e8 + rng.standard_normal((1_000_003, 3)) * 1e-3)
    ea, es, ga, gs = statistics.calculate_average_and_stdVar(rf)
    oa, os_, oga, ogs = O.moments(rf)
    np.testing.assert_allclose(ea, oa, rtol=1e-15)
    np.testing.assert_allclose(ga, oga, rtol=1e-15)
    # the reference's E[x^2]-E[x]^2 cancels: both sides see the same cancellation
    np.testing.assert_allclose(es, os_, rtol=1e-3, atol=1e-3)


# ---- special functions -------------------------------------------------------------------
def test_kv_device(hip):
    from scarl_amd import statistics
    xs = np.array([1e-3, 0.05, 0.5, 1.0, 1.9, 3.0, 6.0, 8.9, 9.1, 20.0, 40.0, 80.0])
    for nu in (0.0, 0.1, 0.3, 0.5, 1.0, 1.5, 2.7):
        dev = statistics.kv(nu, xs)
        for x, d in zip(xs, dev):
            ref = O.kv(nu, x)
            tol = 1e-7 if 2 < x <= 9 else (2e-9 if 9 < x < 20 else 1e-13)
            assert abs(d - ref) <= tol * abs(ref), (nu, x, d, ref)


def test_target_corr_device(hip):
    from scarl_amd import statistics
    rho = np.linspace(0, 4, 41)
    for model, nu in ((1, 0.5), (3, 0.5), (2, 0.3), (2, 1.5)):
        dev = statistics.target_corr(model, rho, nu)
        ref = [O.target_corr(model, 1, (r, 0, 0), (1, 1, 1), nu) for r in rho]
        np.testing.assert_allclose(dev, ref, rtol=1e-7, atol=1e-12)


# ---- Hermitian generator -----------------------------------------------------------------
HSHAPES = [(64, 48, 1), (256, 256, 1), (16, 12, 10), (32, 24, 20), (16, 9, 7), (40, 44, 36), (64, 64, 64)]


@pytest.mark.parametrize("n", HSHAPES)
@pytest.mark.parametrize("model,nu", [(1, 0.5), (2, 0.3), (3, 0.5)])
def test_hermitian_fp64_host_noise(hip, n, model, nu):
    from scarl_amd import generator
    _say(f"hermitian fp64 n={n} model={model}")
    L = (8.0, 6.0, 4.0)
    spec = generator.FieldSpec(n=n, model=model, corrL=L, nu=nu, seed=11)
    W = O.noise_hermitian(n, 11, 0)
    gpu = generator.gen_field_h(spec, 64, W)
    ref = O.gen_mode_h(n, (1, 1, 1), model, L, nu, W=W)
    assert rel_l2(gpu, ref) <= FP64_TOL


@pytest.mark.parametrize("n", [(32, 24, 20), (256, 256, 1), (48, 40, 30)])
def test_hermitian_fp64_device_noise(hip, n):
    from scarl_amd import generator
    spec = generator.FieldSpec(n=n, model=2, corrL=(6.0, 5.0, 4.0), nu=0.3, seed=3, stream=9, margi=2, avg=1.0,
                               std_dev=0.3)
    gpu = generator.gen_field_h(spec, 64)
    ref = O.gen_mode_h(n, (1, 1, 1), 2, (6.0, 5.0, 4.0), 0.3, seed=3, stream=9, margi=2, avg=1.0, std=0.3)
    assert rel_l2(gpu, ref) <= FP64_TOL


@pytest.mark.parametrize("n,model", [((64, 64, 64), 2), ((128, 96, 80), 3), ((256, 256, 1), 1), ((176, 144, 160), 3)])
def test_hermitian_fp32_vs_fp64_oracle(hip, n, model):
    from scarl_amd import generator
    _say(f"hermitian fp32 n={n}")
    L = (12.0, 10.0, 8.0)
    spec = generator.FieldSpec(n=n, model=model, corrL=L, nu=0.3, seed=5, margi=2, avg=1.0, std_dev=0.5)
    W = O.noise_hermitian(n, 5, 0)
    gpu = generator.gen_field_h(spec, 32, W.astype(np.complex64))
    ref = O.gen_mode_h(n, (1, 1, 1), model, L, 0.3, W=W.astype(np.complex64).astype(np.complex128), margi=2,
                       avg=1.0, std=0.5)
    err = rel_l2(gpu, ref)
    _say(f"  fp32 rel L2 = {err:.3e}, max abs = {np.abs(gpu - ref).max():.3e}")
    assert err <= FP32_TOL


def test_hermitian_fp32_device_noise(hip):
    from scarl_amd import generator
    n = (96, 64, 48)
    spec = generator.FieldSpec(n=n, model=2, corrL=(16.0, 16.0, 16.0), nu=0.3, seed=77)
    gpu = generator.gen_field_h(spec, 32)
    ref = O.gen_mode_h(n, (1, 1, 1), 2, (16.0, 16.0, 16.0), 0.3, seed=77)
    assert rel_l2(gpu, ref) <= FP32_TOL


@pytest.mark.parametrize("prec,n", [(32, (16, 12, 64)), (32, (15, 8, 96)), (32, (8, 8, 128)), (32, (12, 10, 160)),
                                    (32, (10, 8, 176)), (32, (8, 8, 256)), (32, (8, 6, 512)), (64, (9, 6, 256)),
                                    (64, (8, 4, 512)), (32, (512, 6, 4)), (32, (512, 5, 1))])
def test_hermitian_device_noise_generated_in_first_stage(hip, prec, n):
    """z lengths with a compile-time plan whose first radix is even: the pairs (q, q + n3/2) are
    generated straight into the first butterfly's registers. Both self-conjugate planes (kx = 0
    and, for even n1, kx = n1/2) take the per-point path inside the same tiles."""
    from scarl_amd import generator
    _say(f"hermitian device noise fp{prec} n={n}")
    L = (4.0, 3.0, 24.0)
    spec = generator.FieldSpec(n=n, model=2, corrL=L, nu=0.3, seed=19, stream=4, margi=1, avg=0.5, std_dev=2.0)
    ref = O.gen_mode_h(n, (1, 1, 1), 2, L, 0.3, seed=19, stream=4, margi=1, avg=0.5, std=2.0)
    assert rel_l2(generator.gen_field_h(spec, prec), ref) <= (FP32_TOL if prec == 32 else FP64_TOL)


@pytest.mark.parametrize("prec,n", [(32, (64, 10, 64)), (32, (256, 6, 128)), (32, (512, 4, 64)), (32, (40, 6, 96)),
                                    (32, (33, 7, 160)), (64, (64, 5, 64)), (64, (42, 5, 96)), (32, (2, 9, 64))])
def test_z_tile_kinds(hip, prec, n):
    """The generating z pass's tile kinds: rows of one or a few aligned tiles (the first tile of a
    row from kx = 1, the kx = 0 pencils in column tiles), rows of 8 or more aligned tiles (row tiles,
    the first holding kx = 0), and rows that are not a multiple of the tile (tiles over the pencils
    off kx = 0, straddling rows), even and odd n1, with device noise against the oracle."""
    from scarl_amd import generator
    _say(f"z tile kinds fp{prec} n={n}")
    L = (6.0, 3.0, 10.0)
    spec = generator.FieldSpec(n=n, model=3, corrL=L, seed=29, stream=2, margi=2, avg=1.0, std_dev=0.3)
    ref = O.gen_mode_h(n, (1, 1, 1), 3, L, seed=29, stream=2, margi=2, avg=1.0, std=0.3)
    assert rel_l2(generator.gen_field_h(spec, prec), ref) <= (FP32_TOL if prec == 32 else FP64_TOL)


@pytest.mark.parametrize("n", [(23, 12, 10), (41, 41, 41), (81, 30, 17), (5, 6, 1), (45, 33, 1)])
def test_hermitian_odd_x_extent(hip, n):
    """Odd x extents (the reference geometry's L = 1 + ceiling(...) is usually odd): rows of
    h1 = (n1+1)/2, no Nyquist column, x pass as a full-length C2C of the Hermitian extension."""
    from scarl_amd import generator
    _say(f"hermitian odd n={n}")
    L = (5.0, 4.0, 3.0)
    spec = generator.FieldSpec(n=n, model=3, corrL=L, seed=13, margi=2, avg=1.0, std_dev=0.4)
    W = O.noise_hermitian(n, 13, 0)
    ref = O.gen_mode_h(n, (1, 1, 1), 3, L, W=W, margi=2, avg=1.0, std=0.4)
    assert rel_l2(generator.gen_field_h(spec, 64, W), ref) <= FP64_TOL
    refd = O.gen_mode_h(n, (1, 1, 1), 3, L, seed=13, margi=2, avg=1.0, std=0.4)
    assert rel_l2(generator.gen_field_h(spec, 32), refd) <= FP32_TOL


# ---- batches -----------------------------------------------------------------------------
def test_ensemble_batch_mixed_sizes(hip):
    from scarl_amd import generator
    specs = []
    sizes = [64, 96, 128, 160, 192, 96]
    for i, e in enumerate(sizes):
        model = 1 if i % 2 == 0 else 3
        L = 12.0 if model == 1 else 20.0
        specs.append(generator.FieldSpec(n=(e, e, e), model=model, corrL=(L, L, L), seed=1000 + i, stream=i,
                                         margi=1, avg=0.0, std_dev=1.0))
    for prec, tol in ((64, FP64_TOL), (32, FP32_TOL)):
        with generator.Batch.ensemble(specs, prec) as b:
            b.run()
            mom = b.moments()
            for i, s in enumerate(specs):
                out = b.output(i)
                ref = O.gen_mode_h(s.shape, (1, 1, 1), s.model, s.corrL, seed=s.seed, stream=s.stream, margi=1,
                                   avg=0.0, std=1.0)
                assert rel_l2(out, ref) <= tol, (prec, i)
                x = out.astype(np.float64)
                assert abs(mom[i, 3] - x.mean()) <= 1e-12 * max(1.0, abs(x).max())
                assert abs(mom[i, 2] - x.size) == 0


def test_ensemble_graph_replay_matches_eager(hip):
    from scarl_amd import generator
    specs = [generator.FieldSpec(n=(64, 48, 40), model=2, corrL=(8.0, 8.0, 8.0), nu=0.3, seed=9, stream=i)
             for i in range(3)]
    with generator.Batch.ensemble(specs, 32) as b:
        b.run(use_graph=False)
        eager = [b.output(i) for i in range(3)]
        for rep in range(3):
            b.run(use_graph=True)
            b.sync()
            for i in range(3):
                np.testing.assert_array_equal(b.output(i), eager[i])
            b.run(use_graph=False)  # eager work between replays
            b.sync()


@pytest.mark.parametrize("schedule", [1, 2, 0])
def test_pipelined_groups_match_single_fields(hip, schedule):
    """Three 256^3 fields make two groups, run pipelined (z pass of one group overlapping the y and
    x passes of the other on a second stream), one after another with each pass's launches fanned
    out, or as the cache-fit rule picks (sequential: the groups' mean workspace, 100 MB, fits the
    Infinity Cache); every field and its moments must equal, bit for bit, the same field generated alone,
    through the hipGraph and eagerly."""
    from scarl_amd import _capi as A, generator
    specs = [generator.FieldSpec(n=(256, 256, 256), model=2, corrL=(16.0, 16.0, 16.0), nu=0.3, seed=5, stream=i,
                                 margi=2, avg=1.0, std_dev=0.3) for i in range(3)]
    alone, alone_mom = [], []
    for sp in specs:
        with generator.Batch.ensemble([sp], 32) as s:
            s.run()
            alone.append(s.output(0))
            alone_mom.append(s.moments()[0])
    generator.set_tuning(A.TUNE_SCHEDULE, schedule)
    try:
        b = generator.Batch.ensemble(specs, 32)
    finally:
        generator.set_tuning(A.TUNE_SCHEDULE, 0)
    with b:
        rep = b.plan_report()
        assert {r.get("schedule") for r in rep} == {"pipelined" if schedule == 1 else "sequential"}, rep
        for use_graph in (True, False, True):
            b.run(use_graph=use_graph)
            b.sync()
            mom = b.moments()
            for i in range(3):
                np.testing.assert_array_equal(b.output(i), alone[i])
                np.testing.assert_array_equal(mom[i], alone_mom[i])


def test_mixed_2d_3d_groups(hip):
    """Three 256^3 fields and two 2D fields in one batch (several groups, the kx = 0 planes of each
    group's 3D fields generated just before its z pass, 2D fields without one; the 2D fields' y
    launches generate, the 3D fields' do not), pipelined and one group after another: every field
    equals, bit for bit, the same field generated alone, through the hipGraph and eagerly."""
    from scarl_amd import _capi as A, generator
    shapes = [(256, 256, 256), (128, 96, 1), (256, 256, 256), (96, 160, 1), (256, 256, 256)]
    specs = [generator.FieldSpec(n=n, model=1 + i % 3, corrL=(12.0, 10.0, 8.0), nu=0.4, seed=8, stream=i, margi=2,
                                 avg=1.0, std_dev=0.3) for i, n in enumerate(shapes)]
    alone = []
    for sp in specs:
        with generator.Batch.ensemble([sp], 32) as s:
            s.run()
            alone.append(s.output(0))
    for schedule in (1, 2):
        generator.set_tuning(A.TUNE_SCHEDULE, schedule)
        try:
            b = generator.Batch.ensemble(specs, 32)
        finally:
            generator.set_tuning(A.TUNE_SCHEDULE, 0)
        with b:
            assert {r.get("schedule") for r in b.plan_report()} == {"pipelined" if schedule == 1 else "sequential"}
            for use_graph in (True, False):
                b.run(use_graph=use_graph)
                b.sync()
                for i in range(len(specs)):
                    np.testing.assert_array_equal(b.output(i), alone[i])


def _blend_reference(gdim, starts, exts, spec, dx=(1, 1, 1)):
    fields, st, ws = [], [], []
    sid = 0
    for k in range(len(starts[2])):
        for j in range(len(starts[1])):
            for i in range(len(starts[0])):
                n = (exts[0][i], exts[1][j], exts[2][k])
                f = O.gen_mode_h(n, dx, spec.model, spec.corrL, spec.nu, seed=spec.seed, stream=sid)
                pos = (i, j, k)
                w = []
                for a in range(3):
                    p = pos[a]
                    lo = starts[a][p - 1] + exts[a][p - 1] - starts[a][p] if p > 0 else 0
                    hi = starts[a][p] + exts[a][p] - starts[a][p + 1] if p + 1 < len(starts[a]) else 0
                    w.append(O.pou_weights_1d(n[a], lo, hi, lo > 0, hi > 0))
                fields.append(f)
                st.append((starts[0][i], starts[1][j], starts[2][k]))
                ws.append(w)
                sid += 1
    G = O.blend_global(fields, st, gdim, ws)
    return O.marginal(G, spec.margi, spec.avg, spec.std_dev) if spec.margi else G


@pytest.mark.parametrize("prec,tol", [(64, FP64_TOL), (32, FP32_TOL)])
def test_blend_subdomains(hip, prec, tol):
    from scarl_amd import generator
    starts = [[0, 40, 80], [0, 36], [0, 24]]
    exts = [[48, 48, 40], [44, 44], [32, 24]]
    gdim = (120, 80, 48)
    spec = generator.FieldSpec(n=(2, 2, 2), model=3, corrL=(10.0, 10.0, 10.0), seed=21, margi=2, avg=1.0,
                               std_dev=0.5)
    with generator.Batch.blend(gdim, starts, exts, spec, prec) as b:
        b.run()
        G = b.output(-1)
        mom = b.moments()
    ref = _blend_reference(gdim, starts, exts, spec)
    assert rel_l2(G, ref) <= tol
    g = G.astype(np.float64)
    assert abs(mom[0, 3] - g.mean()) <= 1e-10 * abs(g.mean())
    assert mom[0, 2] == g.size


def test_blend_wide_overlaps(hip):
    """Overlaps wider than half a subdomain, as the reference's expansion rule allows them: with
    two subdomains on an axis only Np >= Np_ovlp is required (main_ScaRL.f90:161); with three, the
    middle one may be all ramp (Np = 2 Np_ovlp, :162)."""
    from scarl_amd import generator
    starts = [[0, 10], [0, 12, 24], [0]]
    exts = [[40, 40], [24, 24, 24], [20]]
    gdim = (50, 48, 20)
    spec = generator.FieldSpec(n=(2, 2, 2), model=1, corrL=(6.0, 6.0, 6.0), seed=5, margi=1, avg=0.5, std_dev=2.0)
    with generator.Batch.blend(gdim, starts, exts, spec, 64) as b:
        b.run()
        G = b.output(-1)
    assert rel_l2(G, _blend_reference(gdim, starts, exts, spec)) <= FP64_TOL


def test_blend_x_extent_512(hip):
    """Subdomains 512 wide along x: the FP32 x pass takes the row-mirrored pair loads (each X_k
    loaded once, X_{M-k} from the mirrored lane) with the blend epilogue."""
    from scarl_amd import generator
    starts = [[0, 448], [0, 4], [0]]
    exts = [[512, 512], [8, 8], [6]]
    gdim = (960, 12, 6)
    spec = generator.FieldSpec(n=(2, 2, 2), model=3, corrL=(40.0, 3.0, 3.0), seed=8, margi=2, avg=1.0, std_dev=0.5)
    with generator.Batch.blend(gdim, starts, exts, spec, 32) as b:
        b.run()
        G = b.output(-1)
    assert rel_l2(G, _blend_reference(gdim, starts, exts, spec)) <= FP32_TOL


def test_blend_partition_of_unity_variance(hip):
    """The HIP apply_UnityPartition_mtx on fields of ones over 2x1x1, 2x2x1 and 2x2x2 topologies:
    the squared weights of every owner of a global point sum to one (the blend preserves the
    variance), as the oracle's brute force does (tests/test_oracle.py)."""
    from scarl_amd import main_ScaRL
    Np, Nov = (10, 9, 8), (4, 3, 2)
    for shape in ((2, 1, 1), (2, 2, 1), (2, 2, 2)):
        G = np.zeros([(Np[d] - Nov[d]) * shape[d] + Nov[d] for d in range(3)])
        for pos in np.ndindex(*shape):
            f = np.ones(Np, order="F")
            main_ScaRL.apply_UnityPartition_mtx(Np, Nov, 1, pos, shape, f)
            np.testing.assert_allclose(f, O.apply_pou(np.ones(Np), Nov, pos, shape), rtol=0, atol=1e-15)
            s = [(Np[d] - Nov[d]) * pos[d] for d in range(3)]
            G[s[0]:s[0] + Np[0], s[1]:s[1] + Np[1], s[2]:s[2] + Np[2]] += f ** 2
        np.testing.assert_allclose(G, 1.0, rtol=0, atol=1e-14)


# ---- spectra_ScaRL: build_Sk_mtx and the Shinozuka consumer -------------------------------
@pytest.mark.parametrize("corrMod", [1, 2])
def test_build_sk_mtx(hip, corrMod):
    from scarl_amd import spectra_ScaRL
    for Np, szc, xr, seed in (((16, 16, 16), (32, 32, 32), (3.0, 3.0, 3.0), 7), ((20, 12, 9), (64, 32, 32),
                                                                                  (4.0, 2.5, 1.7), 123)):
        sk = spectra_ScaRL.build_Sk_mtx(corrMod, xr, (1.0, 1.0, 1.0), Np, szc, seed, 0)
        ref = O.build_sk(corrMod, xr, Np, szc, seed)
        assert sk.shape == ref.shape
        np.testing.assert_allclose(sk, ref, rtol=1e-14, atol=1e-15 * np.abs(ref).max())


# (Np, xRange, corrL, seed): the y/z quadrants exactly filling a 32^3 padded grid; unequal padded
# extents 64 x 32 x 16; dx = 0.125 corrL with k_max = 8 > k_cut on a 64 x 64 x 32 grid
SHINO_GPU_CASES = [((12, 10, 9), (2.75, 1.0, 1.0), (1.0, 1.0, 1.0), 5),
                   ((20, 9, 5), (19 / 8, 1.0, 0.5), (1.0, 1.0, 1.0), 11),
                   ((20, 18, 16), (9.5, 8.5, 7.5), (4.0, 4.0, 4.0), 3)]


@pytest.mark.parametrize("case", range(len(SHINO_GPU_CASES)))
def test_shino_fft(hip, case):
    from scarl_amd import spectra_ScaRL
    Np, xr, cl, seed = SHINO_GPU_CASES[case]
    out = np.empty(Np, np.float64, order="F")
    spectra_ScaRL.gen_std_gauss_Shino_FFT(out, Np, xr, cl, (5, 5, 5), 1, seed, 0)
    ref = O.gen_std_gauss_shino_fft(Np, xr, cl, 1, seed)
    assert rel_l2(out, ref) <= FP64_TOL


def test_shino_fft_coarse_mesh_is_an_error(hip):
    from scarl_amd import spectra_ScaRL, ScarlError
    Np = (20, 18, 16)
    out = np.empty(Np, np.float64, order="F")
    # too coarse a mesh: the reference would write outside c_in; here it is an error
    with pytest.raises(ScarlError):
        spectra_ScaRL.gen_std_gauss_Shino_FFT(out, Np, (19.0, 17.0, 15.0), (4.0, 4.0, 3.0), (5, 5, 5), 1, 3, 0)


# ---- Stat_ScaRL: corrL estimator, per-axis covariance; normalize_field ------------------------
def _expected_lag_cov(n, model, L, nlag, nu=0.5):
    """Exact expectation of the mean-centred per-axis lag covariance of the periodic Hermitian
    generator: the sample mean is the k = 0 mode, so E = C_per(l e_a) - amp(0)^2 with C_per the
    unnormalised C2R of amp^2; plus the large-N variance 2 sum_r C_per(r)^2 / N of one lag
    estimate (Gaussian field) per axis."""
    amp = O.amplitude_h(n, (1, 1, 1), model, L, nu)
    Cp = O.c2r_3d(amp.astype(np.complex128) ** 2, n)
    p0 = amp[0, 0, 0] ** 2
    N = float(np.prod(n))
    exp = np.full((3, nlag), np.nan)
    for a in range(3):
        m = min(nlag, n[a])
        idx = [0, 0, 0]
        for l in range(m):
            idx[a] = l
            exp[a, l] = Cp[tuple(idx)] - p0
    return exp, 2.0 * np.sum(Cp ** 2) / N
def test_rebuild_sk_fft_matches_oracle(hip):
    from scarl_amd import statistics
    rng = np.random.default_rng(21)
    x3 = np.asfortranarray(rng.standard_normal((48, 40, 36)) * 1.5 + 0.25)
    ref = O.rebuild_sk_corrl(x3, 1.0, 0.25, 1.5)
    for d in (1, 2, 3):
        sk = statistics.rebuild_Sk_FFT(x3, x3.shape, d, 3, 0.25, 1.5)
        np.testing.assert_allclose(sk, ref[d - 1], rtol=0, atol=1e-12 * np.abs(ref[d - 1]).max())
    x2 = np.asfortranarray(rng.standard_normal((64, 50)))
    ref2 = O.rebuild_sk_corrl(x2, 1.0, 0.0, 1.0)
    for d in (1, 2):
        sk = statistics.rebuild_Sk_FFT(x2, x2.shape, d, 2, 0.0, 1.0)
        np.testing.assert_allclose(sk, ref2[d - 1], rtol=0, atol=1e-12 * np.abs(ref2[d - 1]).max())
    # rebuild_Sk keeps L/2 lags per direction; rebuild_corrL integrates them
    sktot, nhalf = statistics.rebuild_Sk(x3, x3.shape, 3, 0.25, 1.5)
    corrl = statistics.rebuild_corrL(sktot, nhalf, (0.5, 0.5, 0.5))
    for d in range(3):
        exp = O.corrl_from_sk(ref[d][:x3.shape[d] // 2], 0.5)
        assert abs(corrl[d] - exp) <= 1e-12 * max(1.0, abs(exp))


@pytest.mark.parametrize("lag_fft", [1, 0], ids=["fft", "direct"])
def test_lag_sums_both_paths(hip, lag_fft):
    """The periodic lag sums by FFT (|FFT|^2 and back) and by direct products give the oracle's
    corrL estimator to 1e-12, and the per-axis covariance (bounded lags, direct products on both
    settings) stays exact.  n = 47 has no Stockham plan, so its periodic lags fall back to the direct
    products even on the FFT path; 512 periodic lags take a 512-point FFT."""
    from scarl_amd import _capi as A, generator, statistics
    rng = np.random.default_rng(23)
    generator.set_tuning(A.TUNE_LAG_FFT, lag_fft)
    try:
        x3 = np.asfortranarray(rng.standard_normal((47, 40, 36)) * 0.7 - 0.1)
        ref = O.rebuild_sk_corrl(x3, 1.0, -0.1, 0.7)
        for d in (1, 2, 3):
            sk = statistics.rebuild_Sk_FFT(x3, x3.shape, d, 3, -0.1, 0.7)
            np.testing.assert_allclose(sk, ref[d - 1], rtol=0, atol=1e-12 * np.abs(ref[d - 1]).max())
        x1 = np.asfortranarray(rng.standard_normal((512, 24, 3)))
        ref1 = O.rebuild_sk_corrl(x1, 1.0, 0.0, 1.0)
        sk = statistics.rebuild_Sk_FFT(x1, x1.shape, 1, 3, 0.0, 1.0)
        np.testing.assert_allclose(sk, ref1[0], rtol=0, atol=1e-12 * np.abs(ref1[0]).max())
        x = np.asfortranarray(rng.standard_normal((47, 33, 192)) + 1.0)
        mu = float(x.mean())
        cov = statistics.lag_covariance(x, mu, 33)
        rc = O.lag_covariance(x, mu, 33)
        assert np.array_equal(np.isnan(cov), np.isnan(rc))
        ok = ~np.isnan(rc)
        np.testing.assert_allclose(cov[ok], rc[ok], rtol=0, atol=1e-12 * np.abs(rc[ok]).max())
    finally:
        generator.set_tuning(A.TUNE_LAG_FFT, 1)


def test_lag_covariance_matches_oracle(hip):
    from scarl_amd import statistics
    rng = np.random.default_rng(22)
    x = np.asfortranarray(rng.standard_normal((40, 33, 28)) + 2.0)
    mu = float(x.mean())
    cov = statistics.lag_covariance(x, mu, 33)
    ref = O.lag_covariance(x, mu, 33)
    assert np.array_equal(np.isnan(cov), np.isnan(ref))
    ok = ~np.isnan(ref)
    np.testing.assert_allclose(cov[ok], ref[ok], rtol=0, atol=1e-12 * np.abs(ref[ok]).max())


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_lag_cov_matches_oracle(hip, prec):
    from scarl_amd import generator
    specs = [generator.FieldSpec(n=n, model=m, corrL=(6.0, 6.0, 6.0), seed=4, stream=i)
             for i, (n, m) in enumerate((((64, 48, 40), 1), ((48, 48, 48), 3), ((32, 40, 24), 1)))]
    with generator.Batch.ensemble(specs, prec) as b:
        b.run()
        cov = b.lag_covariance(33)
        outs = [b.output(i).astype(np.float64) for i in range(3)]
        mom = b.moments()
    for i, x in enumerate(outs):
        ref = O.lag_covariance(x, mom[i, 3], 33)
        assert np.array_equal(np.isnan(cov[i]), np.isnan(ref))
        ok = ~np.isnan(ref)
        np.testing.assert_allclose(cov[i][ok], ref[ok], rtol=0, atol=1e-10 * np.abs(ref[ok]).max())


def test_device_noise_covariance_in_confidence_band(hip):
    """Device Philox noise, FP32 batch: the ensemble-mean per-axis covariance at lags 0-32 lies
    within 4 standard errors of the target correlation C(l) minus the sample-mean bias
    sum_r C(r)/N.  Standard error of one field's lag estimate ~ sqrt(2 sum_r C(r)^2 / N) (Gaussian
    field), divided by sqrt(fields x axes)."""
    from scarl_amd import generator
    for model, L, edge, nf in ((1, 12.0, 96, 12), (3, 20.0, 192, 6)):
        specs = [generator.FieldSpec(n=(edge,) * 3, model=model, corrL=(L, L, L), seed=77, stream=i)
                 for i in range(nf)]
        with generator.Batch.ensemble(specs, 32) as b:
            b.run()
            cov = b.lag_covariance(33)
            mom = b.moments()
        exp, var1 = _expected_lag_cov((edge,) * 3, model, (L, L, L), 33)
        se = np.sqrt(var1 / (nf * 3))
        est = cov.mean(axis=(0, 1))
        assert np.all(np.abs(est - exp.mean(axis=0)) <= 4 * se), (model, est - exp.mean(axis=0), se)
        # the sample mean is the k = 0 mode: sd amp(0)
        a0 = O.amplitude_h((edge,) * 3, (1, 1, 1), model, (L, L, L))[0, 0, 0]
        assert np.all(np.abs(mom[:, 3]) <= 4 * a0)


def test_normalize_field_matches_oracle(hip):
    from scarl_amd import main_ScaRL, ScarlError
    rng = np.random.default_rng(23)
    x = np.asfortranarray(rng.standard_normal((30, 22, 17)) * 4.0 - 1.0)
    ref, avg_r, sd_r = O.normalize_field(x)
    y = x.copy(order="F")
    avg, sd = main_ScaRL.normalize_field((0, 0, 0), (1, 1, 1), x.shape, (0, 0, 0), 0, 0, y)
    assert abs(avg - avg_r) <= 1e-14 * abs(avg_r) + 1e-15 and abs(sd - sd_r) <= 1e-14 * sd_r
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-13)
    with pytest.raises(ScarlError):
        main_ScaRL.normalize_field((0, 0, 0), (2, 1, 1), x.shape, (4, 0, 0), 0, 0, y)


# ---- the workloads of the benchmark configurations --------------------------------------------
def test_config1_2d_gaussian_against_numpy_irfftn(hip):
    """2D 256x256, unit spacing, Gaussian C = exp(-(r/8)^2), sigma 1, standard-normal marginal,
    Hermitian N(0,1) noise of seed 1, FP64: against numpy's irfftn of amp * W (unnormalised)."""
    from scarl_amd import generator
    n = (256, 256, 1)
    spec = generator.FieldSpec(n=n, model=1, corrL=(8.0, 8.0, 8.0), seed=1, margi=1, avg=0.0, std_dev=1.0)
    W = O.noise_hermitian(n, 1, 0)
    gpu = generator.gen_field_h(spec, 64, W)
    amp = O.amplitude_h(n, (1, 1, 1), 1, (8.0, 8.0, 8.0))
    ref = np.fft.irfftn((amp * W)[:, :, 0], s=(256, 256), axes=(1, 0)) * (256 * 256)
    assert rel_l2(gpu[:, :, 0], ref) <= FP64_TOL
    _say(f"config 1: mean {gpu.mean():.4f} variance {gpu.var():.4f}")
    assert abs(gpu.mean()) < 0.2 and abs(gpu.var() - 1.0) < 0.3  # ~ (256/8)^2 = 1024 correlation areas


def test_blend_without_overlap_equals_ensemble(hip):
    """Regression: 8x8x2 subdomains of 112-144 points tiling 1024x1024x256 without overlap. The
    workspace slab is exactly 520 x 2 MiB, and the last subdomain (144^3) runs the 144-point y pass
    (9 threads per pencil, 28 pencils: four threads of each block fall outside the plan).  Those
    threads once ran a butterfly past the pencil, reading one element beyond the slab's end, which
    here is an unmapped page.  With no overlap the blend stores each subdomain's own field, so the
    global field equals the same subdomains generated as an ensemble."""
    from scarl_amd import generator
    gdim = (1024, 1024, 256)
    starts, exts = generator.seeded_tiling(gdim, (8, 8, 2), 0, (112, 128, 144), 2024)
    base = dict(model=3, corrL=(24.0, 24.0, 24.0), seed=31, margi=2, avg=1.0, std_dev=0.5)
    _say(f"blend without overlap: extents {exts}")
    with generator.Batch.blend(gdim, starts, exts, generator.FieldSpec(n=(2, 2, 2), **base), 32) as b:
        b.run(use_graph=True)
        b.sync()
        G = b.output(-1)
    specs, st = [], []
    for k in range(2):
        for j in range(8):
            for i in range(8):
                specs.append(generator.FieldSpec(n=(exts[0][i], exts[1][j], exts[2][k]), stream=len(specs), **base))
                st.append((starts[0][i], starts[1][j], starts[2][k]))
    with generator.Batch.ensemble(specs, 32) as e:
        e.run()
        e.sync()
        worst = 0.0
        for f, (s, spec) in enumerate(zip(st, specs)):
            n = spec.n
            sub = G[s[0]:s[0] + n[0], s[1]:s[1] + n[1], s[2]:s[2] + n[2]]
            worst = max(worst, rel_l2(sub, e.output(f)))
    _say(f"  worst subdomain rel L2 {worst:.2e}")
    assert worst <= 1e-6


def test_config4_blend_1024x1024x256(hip):
    """Global 1024x1024x256 from 8x8x2 subdomains of seeded extents (128-192, multiples of 16),
    16-point overlaps, cosine PoU, exponential L = 24, lognormal mean 1 CoV 0.5, FP32, one batch.
    The blended Gaussian field equals the gather-sum of sqrt(w) times the same subdomains generated
    as an ensemble, and the lognormal field's moments sit in their bands."""
    from scarl_amd import generator
    gdim = (1024, 1024, 256)
    starts, exts = generator.seeded_tiling(gdim, (8, 8, 2), 16, (128, 144, 160, 176, 192), 2024)
    base = dict(model=3, corrL=(24.0, 24.0, 24.0), seed=31)
    with generator.Batch.blend(gdim, starts, exts, generator.FieldSpec(n=(2, 2, 2), **base), 32) as b:
        b.run(use_graph=True)
        b.sync()
        G = b.output(-1)
    specs, st, ws = [], [], []
    sid = 0
    for k in range(2):
        for j in range(8):
            for i in range(8):
                pos, n = (i, j, k), (exts[0][i], exts[1][j], exts[2][k])
                specs.append(generator.FieldSpec(n=n, stream=sid, **base))
                st.append((starts[0][i], starts[1][j], starts[2][k]))
                w = []
                for a in range(3):
                    p = pos[a]
                    lo = starts[a][p - 1] + exts[a][p - 1] - starts[a][p] if p > 0 else 0
                    hi = starts[a][p] + exts[a][p] - starts[a][p + 1] if p + 1 < len(starts[a]) else 0
                    w.append(O.pou_weights_1d(n[a], lo, hi, lo > 0, hi > 0))
                ws.append(w)
                sid += 1
    with generator.Batch.ensemble(specs, 32) as e:
        e.run()
        fields = [e.output(i).astype(np.float64) for i in range(len(specs))]
    # the subdomain fields themselves against the FP64 oracle (same device noise stream): the first
    # and the last subdomain, and one of every x extent in between
    for sid in sorted({0, len(specs) - 1} | {next(q for q, sp in enumerate(specs) if sp.n[0] == e0)
                                              for e0 in set(exts[0])}):
        sp = specs[sid]
        ref_f = O.gen_mode_h(sp.n, (1, 1, 1), 3, (24.0, 24.0, 24.0), seed=31, stream=sid, nthreads=16)
        err_f = rel_l2(fields[sid], ref_f)
        _say(f"config 4: subdomain {sid} {sp.n} vs FP64 oracle rel L2 {err_f:.2e}")
        assert err_f <= FP32_TOL
    ref = O.blend_global(fields, st, gdim, ws)
    err = rel_l2(G, ref)
    _say(f"config 4: blend vs gather-sum rel L2 {err:.2e}")
    assert err <= 1e-6
    with generator.Batch.blend(gdim, starts, exts, generator.FieldSpec(n=(2, 2, 2), margi=2, avg=1.0, std_dev=0.5,
                                                                       **base), 32) as b:
        b.run(use_graph=True)
        b.sync()
        mom = b.moments()[0]
    _say(f"config 4: lognormal mean {mom[3]:.4f} std {mom[4]:.4f}")
    assert mom[2] == 1024 * 1024 * 256
    # ~770 correlation volumes: the underlying Gaussian mean has sd sqrt(sum C / N) ~ 0.036, the
    # lognormal mean ~ 0.47 of that; bands of 4 sd
    assert abs(mom[3] - 1.0) < 0.07 and abs(mom[4] / mom[3] - 0.5) < 0.1


@pytest.mark.parametrize("mfma", [False, True], ids=["stockham", "mfma_bf16x3"])
def test_config5_ensemble_256(hip, mfma):
    """256 realizations, edges drawn by seed from {64, 96, 128, 160, 192}, half Gaussian (L = 12) and
    half exponential (L = 20), standard-normal marginal, FP32, one batch: per-axis covariance at lags
    0-32 in the confidence band of each model; two realizations of each of the smallest and largest
    edges against the FP64 oracle (field and lag covariance).  mfma: every radix-16 first stage (all
    five edges on y and z, the x half-lengths 32-96) as split-precision bf16 MFMA products."""
    from scarl_amd import _capi as A, generator
    rng = np.random.default_rng(5)
    edges = rng.choice([64, 96, 128, 160, 192], size=256)
    specs = [generator.FieldSpec(n=(int(e),) * 3, model=1 if i < 128 else 3, corrL=(12.0,) * 3 if i < 128 else (20.0,) * 3,
                                 seed=55, stream=i, margi=1, avg=0.0, std_dev=1.0) for i, e in enumerate(edges)]
    _say(f"config 5 ensemble, mfma={mfma}")
    generator.set_tuning(A.TUNE_MFMA_DFT16, int(mfma))
    try:
        b = generator.Batch.ensemble(specs, 32)
    finally:
        generator.set_tuning(A.TUNE_MFMA_DFT16, 0)
    with b:
        b.run(use_graph=True)
        b.sync()
        cov = b.lag_covariance(33)
        mom = b.moments()
        picks = []
        for e in (64, 192):
            sel = [i for i in range(256) if edges[i] == e]
            picks += [sel[0], [i for i in sel if i >= 128][0]]
        outs = {i: b.output(i) for i in picks}
    assert mom[:, 2].sum() == float(np.sum(edges.astype(np.float64) ** 3))
    stats = {}  # (edge, model) -> exact expectation and one-axis variance
    for model, L, sel in ((1, 12.0, slice(0, 128)), (3, 20.0, slice(128, 256))):
        dev, var = [], []
        for i in range(256)[sel]:
            e = int(edges[i])
            if (e, model) not in stats:
                stats[(e, model)] = _expected_lag_cov((e,) * 3, model, (L, L, L), 33)
            exp, v = stats[(e, model)]
            dev.append(cov[i].mean(axis=0) - exp.mean(axis=0))
            var.append(v / 3)
        band = 4 * np.sqrt(np.sum(var)) / len(var)
        assert np.all(np.abs(np.mean(dev, axis=0)) <= band), (model, np.mean(dev, axis=0), band)
    for i in picks:
        sp = specs[i]
        ref = O.gen_mode_h(sp.n, (1, 1, 1), sp.model, sp.corrL, 0.5, seed=55, stream=i)
        assert rel_l2(outs[i], ref) <= FP32_TOL
        rc = O.lag_covariance(outs[i].astype(np.float64), mom[i, 3], 33)
        np.testing.assert_allclose(cov[i], rc, rtol=0, atol=1e-10 * np.abs(rc).max())


# FP32 shapes whose y/z lengths and x half-lengths are multiples of 16 (32 ... 192): every first
# stage runs on the matrix cores with SCARL_TUNE_MFMA_DFT16; one 2D field (the y pass generates)
# and one with host-exported noise
MFMA_SHAPES = [((64, 96, 128), 1, None), ((96, 160, 64), 3, None), ((128, 192, 32), 2, None),
               ((160, 64, 96), 1, "host"), ((192, 128, 48), 3, None), ((96, 176, 1), 1, None)]


@pytest.mark.parametrize("case", range(len(MFMA_SHAPES)))
def test_mfma_generic_radix16_stage(hip, case):
    """The split-precision MFMA radix-16 first stage inside runtime plans of any length that is a
    multiple of 16: against the FP64 oracle (FP32 bound) and the FP32 Stockham field."""
    from scarl_amd import _capi as A, generator
    n, model, noise = MFMA_SHAPES[case]
    _say(f"mfma generic stage n={n} model={model} noise={noise or 'device'}")
    L = (10.0, 14.0, 8.0)
    spec = generator.FieldSpec(n=n, model=model, corrL=L, nu=0.3, seed=21, stream=case, margi=2, avg=1.0,
                               std_dev=0.3)
    W = O.noise_hermitian(n, 21, case) if noise else None
    ref = O.gen_mode_h(n, (1, 1, 1), model, L, 0.3, seed=21, stream=case, W=W, margi=2, avg=1.0, std=0.3)
    Wc = None if W is None else W.astype(np.complex64)
    stock = generator.gen_field_h(spec, 32, Wc)
    generator.set_tuning(A.TUNE_MFMA_DFT16, 1)
    try:
        mf = generator.gen_field_h(spec, 32, Wc)
    finally:
        generator.set_tuning(A.TUNE_MFMA_DFT16, 0)
    e_ref, e_st = rel_l2(mf, ref), rel_l2(mf, stock)
    _say(f"  mfma vs fp64 oracle {e_ref:.2e} (stockham {rel_l2(stock, ref):.2e}), vs stockham {e_st:.2e}")
    assert e_ref <= FP32_TOL and e_st <= FP32_TOL


def test_mfma_generic_blend(hip):
    """Blended subdomains (x extents 96 and 128, y 64, z 32) with the MFMA first stages: the blend
    epilogue after the MFMA x pass equals the oracle's gather-sum within the FP32 bound."""
    from scarl_amd import _capi as A, generator
    starts, exts = [[0, 80], [0, 48], [0]], [[96, 128], [64, 64], [32]]
    gdim = (208, 112, 32)
    spec = generator.FieldSpec(n=(2, 2, 2), model=3, corrL=(12.0, 9.0, 6.0), seed=4, margi=2, avg=1.0, std_dev=0.5)
    generator.set_tuning(A.TUNE_MFMA_DFT16, 1)
    try:
        b = generator.Batch.blend(gdim, starts, exts, spec, 32)
    finally:
        generator.set_tuning(A.TUNE_MFMA_DFT16, 0)
    with b:
        b.run()
        G = b.output(-1)
    assert rel_l2(G, _blend_reference(gdim, starts, exts, spec)) <= FP32_TOL


def test_autotuned_batches_match_oracle(hip):
    """Batches created with SCARL_TUNE_AUTOTUNE: every plan reports its pick and timings, the fields
    equal the untuned batch's within the FP32 bound and the FP64 oracle (FP64 at 1e-12); a second
    batch of the same shapes reuses the process cache."""
    from scarl_amd import _capi as A, generator
    shapes = [(96, 64, 128), (64, 160, 48), (128, 96, 64), (80, 48, 176)]
    specs = [generator.FieldSpec(n=n, model=1 + i % 3, corrL=(9.0, 7.0, 5.0), nu=0.7, seed=77, stream=i, margi=2,
                                 avg=1.0, std_dev=0.4) for i, n in enumerate(shapes)]
    with generator.Batch.ensemble(specs, 32) as b0:
        b0.run()
        plain = [b0.output(i) for i in range(len(specs))]
    generator.set_tuning(A.TUNE_AUTOTUNE, 1)
    try:
        bt = generator.Batch.ensemble(specs, 32)
        b64 = generator.Batch.ensemble(specs[:2], 64)
        again = generator.Batch.ensemble(specs, 32)
    finally:
        generator.set_tuning(A.TUNE_AUTOTUNE, 0)
    with bt, b64, again:
        rep = bt.plan_report()
        _say("  " + "\n  ".join(f"{r['pass']} N={r['N']} {r['kernel']} {r['radices']} T={r['T']} {r['choice']} "
                                  f"tuned={r.get('tuned_ms')} default={r.get('default_ms')}" for r in rep))
        assert rep and all("tuned_ms" in r and float(r["tuned_ms"]) <= float(r["default_ms"]) * 1.5 for r in rep)
        assert all(r.get("cached") for r in again.plan_report())
        for bb in (bt, again):
            bb.run(use_graph=True)
            bb.sync()
            for i, sp in enumerate(specs):
                ref = O.gen_mode_h(sp.n, (1, 1, 1), sp.model, sp.corrL, sp.nu, seed=77, stream=i, margi=2, avg=1.0,
                                   std=0.4)
                out = bb.output(i)
                assert rel_l2(out, ref) <= FP32_TOL and rel_l2(out, plain[i]) <= FP32_TOL
        b64.run()
        for i, sp in enumerate(specs[:2]):
            ref = O.gen_mode_h(sp.n, (1, 1, 1), sp.model, sp.corrL, sp.nu, seed=77, stream=i, margi=2, avg=1.0, std=0.4)
            assert rel_l2(b64.output(i), ref) <= FP64_TOL


def test_autotune_cache_file(hip, tmp_path):
    """SCARL_TUNE_CACHE: a process that tunes writes its picks; a later process reads them and
    tunes nothing (its plans report "cached" with the same choices)."""
    import subprocess
    import sys
    from conftest import ROOT
    cache = tmp_path / "tune.txt"
    code = (
        "from scarl_amd import _capi as A, generator as G\n"
        "A.call('scarl_hip_init', 0)\n"
        "G.set_tuning(A.TUNE_AUTOTUNE, 1)\n"
        "b = G.Batch.ensemble([G.FieldSpec(n=(64, 96, 80), model=1, corrL=(6.0, 6.0, 6.0), seed=1)], 32)\n"
        "print(';'.join(r['choice'] + ('/cached' if r.get('cached') else '') for r in b.plan_report()))\n"
        "b.close()\n")
    env = dict(os.environ, SCARL_TUNE_CACHE=str(cache))
    outs = []
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, env=env,
                           timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        outs.append(r.stdout.strip().splitlines()[-1])
    lines = cache.read_text().split("\n")
    assert len([l for l in lines if l.startswith("fp32/")]) == 3
    first, second = outs[0].split(";"), outs[1].split(";")
    assert not any(c.endswith("/cached") for c in first)
    assert [c + "/cached" for c in first] == second


# Lengths with a prime factor above 13 (the reference's geometry gives 41, 233, 241, 461): Bluestein
# over an inner 2^a 3^b 5^c length M >= 2N - 1 for N >= 32, two backward Stockham FFTs of M
BLUESTEIN_SHAPES = [(241, 233, 41), (482, 40, 24), (461, 48, 32), (233, 241, 1), (64, 461, 34)]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case", range(len(BLUESTEIN_SHAPES)))
def test_bluestein_lengths_mode_h(hip, case, prec):
    from scarl_amd import generator
    n = BLUESTEIN_SHAPES[case]
    _say(f"bluestein mode H n={n} fp{prec}")
    L = (12.0, 9.0, 5.0)
    spec = generator.FieldSpec(n=n, model=2, corrL=L, nu=0.6, seed=9, stream=case, margi=2, avg=1.0, std_dev=0.4)
    with generator.Batch.ensemble([spec], prec) as b:
        b.run()
        out = b.output(0)
        rep = b.plan_report()
    assert any(r["radices"].startswith("bluestein") for r in rep), rep
    ref = O.gen_mode_h(n, (1, 1, 1), 2, L, 0.6, seed=9, stream=case, margi=2, avg=1.0, std=0.4)
    assert rel_l2(out, ref) <= (FP64_TOL if prec == 64 else FP32_TOL)


@pytest.mark.parametrize("Np", [(241, 40, 233), (41, 461, 36)])
def test_bluestein_redft01(hip, Np):
    """gen_std_gauss_FFT (FP64 REDFT01) on prime extents against the oracle, 1e-12."""
    from scarl_amd import fftw_ScaRL
    xr = tuple(float(v - 1) for v in Np)
    cl = (8.0, 6.0, 5.0)
    phi = np.asfortranarray(np.random.default_rng(41).random(tuple(v // 2 for v in Np)))
    out = np.empty(Np, np.float64, order="F")
    fftw_ScaRL.gen_std_gauss_FFT(out, Np, xr, cl, 1, 3, 0, phi=phi)
    assert rel_l2(out, O.gen_std_gauss_fft(Np, xr, cl, 1, phi)) <= FP64_TOL


def test_bluestein_deck_size(hip):
    """The reference's MAX_Test deck at SIZE 115 gives 241^3 subdomains: one blended FP64 batch of
    2 x 2 x 2 such subdomains against the oracle's gather-sum on a reduced z extent (241 x 241 x 48)."""
    from scarl_amd import generator
    starts, exts = [[0, 220], [0, 220], [0]], [[241, 241], [241, 241], [48]]
    gdim = (461, 461, 48)
    spec = generator.FieldSpec(n=(2, 2, 2), model=1, corrL=(20.0, 20.0, 20.0), seed=5, margi=2, avg=1.0, std_dev=1.0)
    with generator.Batch.blend(gdim, starts, exts, spec, 64) as b:
        b.run()
        G = b.output(-1)
    assert rel_l2(G, _blend_reference(gdim, starts, exts, spec)) <= FP64_TOL


# Lengths planned with the composite radices 6, 9, 12 (108 = 12 x 9, 216 = 12 x 9 x 2) and the
# compile-time kernels 144 = 16 x 9 (y/z and x half), 192 = 16 x 12, 72 = 8 x 9 (x half of 144)
COMPOSITE_SHAPES = [(288, 108, 144), (144, 192, 216), (144, 144, 96)]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("case", range(len(COMPOSITE_SHAPES)))
def test_composite_radices(hip, case, prec):
    from scarl_amd import generator
    n = COMPOSITE_SHAPES[case]
    _say(f"composite radices n={n} fp{prec}")
    L = (14.0, 10.0, 7.0)
    spec = generator.FieldSpec(n=n, model=3, corrL=L, seed=12, stream=case, margi=2, avg=1.0, std_dev=0.3)
    with generator.Batch.ensemble([spec], prec) as b:
        b.run()
        out = b.output(0)
        rep = b.plan_report()
    _say("  " + "; ".join(f"{r['pass']} N={r['N']} {r['kernel']} {r['radices']}" for r in rep))
    if prec == 32:
        assert any(r["kernel"] == "fixed" and ("16x9" in r["radices"] or "16x12" in r["radices"] or
                                               "8x9" in r["radices"]) for r in rep)
    ref = O.gen_mode_h(n, (1, 1, 1), 3, L, seed=12, stream=case, margi=2, avg=1.0, std=0.3)
    assert rel_l2(out, ref) <= (FP64_TOL if prec == 64 else FP32_TOL)


# (shape, {plan key: choice}) written into a SCARL_TUNE_CACHE file: the batch then uses exactly these
# plans, so candidates the autotuner would not pick are checked too
FORCED = [((192, 96, 144), {"fp32/144/0/0/1": "fixed:19", "fp32/96/0/0/0": "fixed:22", "fp32/96/1/1/2": "fixed:11"}),
          ((192, 96, 144), {"fp32/144/0/0/1": "generic:e16:o1:t2:s1:c1", "fp32/96/0/0/0": "generic:e32:o0:t1:s1:c0",
                            "fp32/96/1/1/2": "generic:e16:o0:t1:s0:c1"}),
          ((144, 216, 80), {"fp32/80/0/0/1": "mfma:t2:s1", "fp32/216/0/0/0": "generic:e16:o0:t2:s1:c1",
                            "fp32/72/1/1/2": "fixed:21"}),
          ((144, 216, 72), {"fp64/72/0/0/1": "generic:e16:o1:t1:s1:c1", "fp64/216/0/0/0": "generic:e16:o0:t2:s1:c1",
                            "fp64/72/1/1/2": "generic:e16:o1:t2:s0:c1"}),
          # tiles cut to a multiple of 16 pencils (t0): 21 -> 16 for 144, 21 -> 16 for 192
          ((160, 144, 192), {"fp32/192/0/0/1": "generic:e16:o0:t0:s1:c1", "fp32/144/0/0/0": "generic:e16:o0:t0:s1:c1",
                             "fp32/80/1/1/2": "generic:e16:o0:t1:s1:c0"})]


@pytest.mark.parametrize("case", range(len(FORCED)))
def test_forced_plan_choices(hip, tmp_path, case):
    import subprocess
    import sys
    from conftest import ROOT
    n, picks = FORCED[case]
    prec = 64 if next(iter(picks)).startswith("fp64") else 32
    cache = tmp_path / "tune.txt"
    cache.write_text("".join(f"{k} {v}\n" for k, v in picks.items()))
    code = (
        "import numpy as np\n"
        "from scarl_amd import _capi as A, generator as G\n"
        "from oracle import oracle as O\n"
        "A.call('scarl_hip_init', 0)\n"
        "G.set_tuning(A.TUNE_AUTOTUNE, 1)\n"
        f"sp = G.FieldSpec(n={n!r}, model=1, corrL=(9.0, 8.0, 6.0), seed=3, margi=2, avg=1.0, std_dev=0.3)\n"
        f"b = G.Batch.ensemble([sp], {prec})\n"
        "b.run(); f = b.output(0)\n"
        "print(';'.join(r['choice'] + ('/cached' if r.get('cached') else '') for r in b.plan_report()))\n"
        f"ref = O.gen_mode_h({n!r}, (1, 1, 1), 1, (9.0, 8.0, 6.0), seed=3, margi=2, avg=1.0, std=0.3)\n"
        "print(float(np.linalg.norm(f - ref) / np.linalg.norm(ref)))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT,
                       env=dict(os.environ, SCARL_TUNE_CACHE=str(cache)), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    used, err = r.stdout.strip().splitlines()[-2:]
    _say(f"  forced {used} rel L2 {err}")
    assert sorted(c.split("/")[0] for c in used.split(";")) == sorted(picks.values())
    assert float(err) <= (FP64_TOL if prec == 64 else FP32_TOL)


def test_deck_geometry_blend(hip):
    """test/3_TEST sample 1 (x in [0, 10], corrL 1, overlap 1, 5 points per corrL, Gaussian,
    lognormal 1/1, seed 1) on a 2x2x2 topology: the eight ranks' subdomains (23^3, overlap 5,
    spacing 0.25) blended in one FP64 batch equal the oracle's gather-sum."""
    from scarl_amd import geometry, generator
    g = geometry.sample_geometry((0, 0, 0), (10, 10, 10), (1, 1, 1), (1, 1, 1), (5, 5, 5), (2, 2, 2))
    with geometry.blend_batch(g, (1, 1, 1), 1, 1, margiFirst=2, avg=1.0, std_dev=1.0, precision=64) as b:
        b.run()
        Gf = b.output(-1)
    gdim, starts, exts = geometry.blend_layout(g)
    spec = generator.FieldSpec(n=(2, 2, 2), model=1, corrL=(1.0, 1.0, 1.0), seed=2, margi=2, avg=1.0, std_dev=1.0)
    ref = _blend_reference(gdim, starts, exts, spec, dx=(0.25, 0.25, 0.25))
    assert Gf.shape == (41, 41, 41)
    assert rel_l2(Gf, ref) <= FP64_TOL


@pytest.mark.parametrize("n", [(512, 16, 8), (512, 20, 1)])
def test_mfma_dft16_x_pass(hip, n):
    """The FP32 x pass of a 512-wide field as split-precision bf16 MFMA radix-16 stages: against
    the FP64 oracle (FP32 bound) and against the FP32 Stockham pass, plain and blended."""
    from scarl_amd import _capi as A, generator
    _say(f"mfma x pass n={n}")
    spec = generator.FieldSpec(n=n, model=2, corrL=(24.0, 6.0, 4.0), nu=0.3, seed=3, margi=2, avg=1.0, std_dev=0.3)
    W = O.noise_hermitian(n, 3, 0).astype(np.complex64)
    ref = O.gen_mode_h(n, (1, 1, 1), 2, (24.0, 6.0, 4.0), 0.3, W=W.astype(np.complex128), margi=2, avg=1.0, std=0.3)
    stock = generator.gen_field_h(spec, 32, W)
    gdim = (1000, n[1], n[2])
    try:
        generator.set_tuning(A.TUNE_MFMA_DFT16, 1)
        mf = generator.gen_field_h(spec, 32, W)
        with generator.Batch.blend(gdim, [[0, 488], [0], [0]], [[512, 512], [n[1]], [n[2]]], spec, 32) as b:
            b.run()
            bm = b.output(-1)
    finally:
        generator.set_tuning(A.TUNE_MFMA_DFT16, 0)
    with generator.Batch.blend(gdim, [[0, 488], [0], [0]], [[512, 512], [n[1]], [n[2]]], spec, 32) as b:
        b.run()
        bs = b.output(-1)
    e_ref, e_st = rel_l2(mf, ref), rel_l2(mf, stock)
    _say(f"  mfma vs fp64 oracle {e_ref:.2e} (stockham {rel_l2(stock, ref):.2e}), vs stockham {e_st:.2e}")
    assert e_ref <= FP32_TOL and e_st <= FP32_TOL
    assert rel_l2(bm, bs) <= FP32_TOL


def test_config3_512_von_karman_fp32_vs_fp64_oracle(hip):
    """512^3 von Karman (L = 32 cells, nu = 0.3), Gaussian marginal, FP32 butterflies with
    long-double twiddles rounded once, FP64 normalisation and moments, device noise: against the
    FP64 oracle on the same Philox noise, relative L2 <= 1e-5; max-abs and moments stated."""
    from scarl_amd import generator
    n = (512, 512, 512)
    spec = generator.FieldSpec(n=n, model=2, corrL=(32.0, 32.0, 32.0), nu=0.3, seed=2024, margi=1, avg=0.0,
                               std_dev=1.0)
    with generator.Batch.ensemble([spec], 32) as b:
        b.run()
        gpu = b.output(0)
        mom = b.moments()[0]
    ref = O.gen_mode_h(n, (1, 1, 1), 2, (32.0, 32.0, 32.0), 0.3, seed=2024)
    err = rel_l2(gpu, ref)
    mx = float(np.abs(gpu.astype(np.float64) - ref).max())
    _say(f"config 3: rel L2 {err:.3e}, max abs {mx:.3e}, mean {mom[3]:.4f}, std {mom[4]:.4f}")
    assert err <= FP32_TOL
    g = gpu.astype(np.float64)
    assert abs(mom[3] - g.mean()) <= 1e-9 and abs(mom[4] - g.std()) <= 1e-9
