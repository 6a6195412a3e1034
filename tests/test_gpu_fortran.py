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
