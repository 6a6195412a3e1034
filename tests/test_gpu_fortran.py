"""The Fortran drop-in modules (scarl_amd/fortran: fftw_ScaRL, normal_ScaRL) called from a
Fortran driver the way create_fields calls the reference (SRC_F/main_ScaRL.f90:510-521)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "scarl_amd", "fortran", "_build", "dropin_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_dropin_driver(tmp_path):
    Np = (36, 30, 22)
    xr, cl = (35.0, 29.0, 21.0), (4.0, 3.0, 2.0)
    phi = np.asfortranarray(np.random.default_rng(17).random(tuple(v // 2 for v in Np)))
    out = tmp_path / "field.bin"
    phi.ravel(order="F").tofile(str(out) + ".phi")
    stdin = f"{Np[0]} {Np[1]} {Np[2]} {xr[0]} {xr[1]} {xr[2]} {cl[0]} {cl[1]} {cl[2]} 5 2\n"
    # spectra_ScaRL / Shinozuka case: unequal padded extents (64, 32, 16) for the consumer
    NpS, xrS, clS, szc, seedS = (20, 9, 5), (4.75, 2.0, 1.0), (2.0, 2.0, 2.0), (64, 32, 16), 11
    stdin += " ".join(str(v) for v in (*NpS, *xrS, *clS, *szc, seedS)) + "\n"
    r = subprocess.run([DRIVER, str(out)], input=stdin, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1].strip() == "EXIT CODE: OK"
    N = int(np.prod(Np))
    raw = np.fromfile(str(out), dtype=np.uint8)
    k_mtx = raw[:8 * N].view(np.float64).reshape(Np, order="F")
    k_mpi = raw[8 * N:16 * N].view(np.float64).reshape(Np, order="F")
    r5 = raw[16 * N:16 * N + 40].view(np.float64)
    seed = int(raw[16 * N + 40:16 * N + 44].view(np.int32)[0])
    ref = O.gen_std_gauss_fft(Np, xr, cl, 1, phi)
    assert rel_l2(k_mtx, ref) <= 1e-12
    phi_mpi = O.phi_octant((Np[0] // 2, Np[1] // 2, Np[2]), 5 + 2, 0)
    assert rel_l2(k_mpi, O.gen_std_gauss_fft_mpi(Np, xr, cl, 1, phi_mpi)) <= 1e-12
    ro, so = O.r8vec_uniform_01(5, 12345)
    np.testing.assert_array_equal(r5, ro)
    assert seed == so
    o = 16 * N + 44
    shp = tuple(int(v) for v in raw[o:o + 12].view(np.int32))
    o += 12
    nsk = int(np.prod(shp))
    sk = raw[o:o + 16 * nsk].view(np.complex128).reshape(shp, order="F")
    o += 16 * nsk
    xrc = tuple(a / b for a, b in zip(xrS, clS))
    ref_sk = O.build_sk(1, xrc, NpS, szc, seedS)
    assert shp == ref_sk.shape
    np.testing.assert_allclose(sk, ref_sk, rtol=1e-14, atol=1e-15 * np.abs(ref_sk).max())
    shino = raw[o:o + 8 * int(np.prod(NpS))].view(np.float64).reshape(NpS, order="F")
    assert rel_l2(shino, O.gen_std_gauss_shino_fft(NpS, xrS, clS, 1, seedS)) <= 1e-12


PROGRAM = os.path.join(ROOT, "scarl_amd", "fortran", "_build", "scarl_main")
DECKS = os.path.join(ROOT, "tests", "golden", "decks")
INT_ATTRS = {"nb_procs", "nDim", "Nmc", "method", "seedStart", "corrMod", "margiFirst", "nFields", "Np", "Np_ovlp",
             "L", "pos_0", "pos_N"}


def _copy_deck(name, dest, subst=None, config_subst=None):
    import shutil
    dest.mkdir(parents=True)
    for f in os.listdir(os.path.join(DECKS, name)):
        shutil.copyfile(os.path.join(DECKS, name, f), dest / f)
    for fname, sub in (("input_ScaRL.txt", subst), ("config_ScaRL.txt", config_subst)):
        if sub:
            p = dest / fname
            txt = p.read_text()
            for k, v in sub.items():
                txt = txt.replace(k, v)
            p.write_text(txt)
    return dest


def _attrs(f, obj):
    return {n: f.read_attr(obj, n, "i" if n in INT_ATTRS else "d").tolist() for n in f.attr_names(obj)}


@pytest.mark.skipif(not os.path.exists(PROGRAM), reason="Fortran program not built")
@pytest.mark.parametrize("deck,nb,subst,csub,datasets", [
    ("3_TEST", 8, None, None, {"Sample_Test_1.h5": ["samples"]}),
    ("GENERAL_BASH", 8, {"#SIZE#": "10"}, {"1 #Delete sample in the end": "0 #Delete sample in the end"},
     {f"Sample_proc_{r:05d}.h5": [f"sample_1_p{r}"] for r in range(8)}),
    ("testFusion", 4, None, None, {"BASIC_Sample.h5": ["samples"]}),
    ("3_TEST", 4, None, {"1\n#Input file": "0\n#Input file"}, {"Sample_Test_1.h5": [f"sample_1_p{r}" for r in range(4)]}),
])
def test_fortran_program_matches_python_driver(hip, tmp_path, deck, nb, subst, csub, datasets):
    """scarl_main (Fortran: read_input, write_output, the geometry and topology restated) and
    main_ScaRL.main (Python) on the same deck write the same files: same names, same data, same
    attributes, byte-identical .xmf. Only the times differ. This also pins the Python deck parser
    against Fortran's own list-directed READ."""
    from scarl_amd import main_ScaRL, write_output as W
    print(f"case: Fortran program vs Python driver, deck {deck}, {nb} ranks", flush=True)
    fd = _copy_deck(deck, tmp_path / "f", subst, csub)
    pd = _copy_deck(deck, tmp_path / "p", subst, csub)
    r = subprocess.run([PROGRAM, str(nb), "32"], cwd=str(fd), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "EXIT CODE: OK" in r.stdout, r.stdout + r.stderr
    main_ScaRL.main(str(pd), nb_procs=nb)
    fo = sorted(os.listdir(fd / "SAMPLES"))
    assert fo == sorted(os.listdir(pd / "SAMPLES"))
    for name in fo:
        a, b = fd / "SAMPLES" / name, pd / "SAMPLES" / name
        if name.endswith(".xmf"):
            assert a.read_bytes() == b.read_bytes(), name
            continue
        with W.H5File(str(a)) as fa, W.H5File(str(b)) as fb:
            ra, rb = _attrs(fa, ""), _attrs(fb, "")
            assert ra == rb, name
            for ds in datasets.get(name, []):
                assert fa.dataset_info(ds) == fb.dataset_info(ds)
                np.testing.assert_array_equal(fa.read(ds), fb.read(ds))
                assert _attrs(fa, ds) == _attrs(fb, ds), (name, ds)
            la, ta = fa.read_times()
            lb, tb = fb.read_times()
            assert la == lb and ta.shape == tb.shape


STAT_PROGRAM = os.path.join(ROOT, "scarl_amd", "fortran", "_build", "scarl_stat")


@pytest.mark.skipif(not os.path.exists(STAT_PROGRAM), reason="Fortran statistics program not built")
def test_fortran_stat_program_matches_python(hip, tmp_path):
    """scarl_stat (Fortran) and main_Stat.main (Python) on copies of one written sample: the
    same statistics on the sample's root and the same STAT_ file."""
    import io
    import shutil
    from scarl_amd import main_ScaRL, main_Stat, write_output as W
    print("case: Fortran statistics program vs Python, 3_TEST sample", flush=True)
    d = _copy_deck("3_TEST", tmp_path / "gen")
    main_ScaRL.main(str(d), nb_procs=2)
    src = d / "SAMPLES" / "Sample_Test_1.h5"
    for sub in ("f", "p"):
        (tmp_path / sub).mkdir()
        shutil.copyfile(src, tmp_path / sub / "Sample_Test_1.h5")
    fpath, ppath = tmp_path / "f" / "Sample_Test_1.h5", tmp_path / "p" / "Sample_Test_1.h5"
    r = subprocess.run([STAT_PROGRAM], input=f'1\n1\n0\n"{fpath}"\n"samples"\n', capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "EXIT CODE: OK" in r.stdout, r.stdout + r.stderr
    main_Stat.main(io.StringIO(f'1\n1\n0\n"{ppath}"\n"samples"\n'))
    for a, b in ((fpath, ppath), (fpath.parent / "STAT_Sample_Test_1.h5", ppath.parent / "STAT_Sample_Test_1.h5")):
        with W.H5File(str(a)) as fa, W.H5File(str(b)) as fb:
            na, nb = sorted(fa.attr_names("")), sorted(fb.attr_names(""))
            assert na == nb and "corrL_out" in na and "Sk_3" in na
            for n in na:
                kind = "i" if n in INT_ATTRS else "d"
                np.testing.assert_array_equal(fa.read_attr("", n, kind), fb.read_attr("", n, kind), err_msg=n)
