"""The allocation-free REDFT01 drop-in (scarl_hip_gen_std_gauss_fft_dev_noalloc) against the
allocating one (scarl_hip_gen_std_gauss_fft_dev), through the C ABI: the same kernels, so the
outputs must be identical bit for bit, on two calls in a row (the second smaller), at the largest
accepted field (256^3), and a larger field must be refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffer from the HIP runtime the library itself uses (libamdhip64 through ctypes)."""

    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), 1) == 0

    def download(self):
        a = np.empty(self.nbytes // 8, np.float64)
        assert self.hip.hipMemcpy(C.c_void_p(a.ctypes.data), self.p, C.c_size_t(self.nbytes), 2) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _run(lib, rt, name, Np, xr, cl, seed, phi, flags):
    """One call of the named entry point into a fresh device array; returns (status, host copy)."""
    out = _Dev(rt, 8 * int(np.prod(Np)))
    np_c = np.ascontiguousarray(Np, np.int32)
    xr_c = np.ascontiguousarray(xr, np.float64)
    cl_c = np.ascontiguousarray(cl, np.float64)
    rc = getattr(lib, name)(out.p, np_c.ctypes.data, xr_c.ctypes.data, cl_c.ctypes.data, 1, seed, 0,
                            None if phi is None else phi.p, flags)
    res = None
    if rc == 0:
        assert lib.scarl_hip_batch_sync(None) == 0
        res = out.download()
    out.free()
    return rc, res


def test_noalloc_matches_dev(hip):
    from scarl_amd import _capi as A
    lib, rt = A.lib(), _hip()
    A.call("scarl_hip_set_marginal", 2, 1.0, 0.3)
    big = ((48, 40, 36), (47.0, 39.0, 35.0), (6.0, 5.0, 4.0), 11)
    small = ((20, 18, 10), (19.0, 17.0, 9.0), (3.0, 3.0, 2.0), 5)
    phi = _Dev(rt, 8 * 24 * 20 * 18)
    phi.upload(np.random.default_rng(3).random(24 * 20 * 18))
    ref = []
    for (Np, xr, cl, seed), p, flags in ((big, phi, 0), (small, None, 1)):
        rc, o = _run(lib, rt, "scarl_hip_gen_std_gauss_fft_dev", Np, xr, cl, seed, p, flags)
        A.check(rc)
        ref.append(o)
    got = []
    for (Np, xr, cl, seed), p, flags in ((big, phi, 0), (small, None, 1)):
        print(f"noalloc call Np={Np} phi={'exported' if p is not None else 'device'}", flush=True)
        rc, o = _run(lib, rt, "scarl_hip_gen_std_gauss_fft_dev_noalloc", Np, xr, cl, seed, p, flags)
        A.check(rc)
        got.append(o)
    phi.free()
    for g, r in zip(got, ref):
        assert np.all(np.isfinite(r)) and np.abs(r).max() > 0
        np.testing.assert_array_equal(g, r)


def test_noalloc_largest_field_and_refusal(hip):
    from scarl_amd import _capi as A
    lib, rt = A.lib(), _hip()
    case = ((256, 256, 256), (255.0, 255.0, 255.0), (16.0, 16.0, 8.0), 7)
    print("noalloc call Np=(256, 256, 256) device phi", flush=True)
    rc, got = _run(lib, rt, "scarl_hip_gen_std_gauss_fft_dev_noalloc", *case, None, 0)
    A.check(rc)
    rc, ref = _run(lib, rt, "scarl_hip_gen_std_gauss_fft_dev", *case, None, 0)
    A.check(rc)
    np.testing.assert_array_equal(got, ref)
    del got, ref
    # beyond the static scratch: refused on the host, nothing launched
    # (320, 256, 256): workspaces of (320*128*128 + 320*256*128) doubles = 120 MiB > 96 MiB
    rc, _ = _run(lib, rt, "scarl_hip_gen_std_gauss_fft_dev_noalloc", (320, 256, 256), (319.0, 255.0, 255.0),
                 (16.0, 16.0, 8.0), 7, None, 0)
    assert rc == A.SCARL_E_INVALID
    assert "scratch" in A.last_error()
