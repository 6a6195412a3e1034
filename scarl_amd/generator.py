"""Hermitian spectral generator (complex white noise with Hermitian symmetry -> C2R), single
fields, ensembles of realisations and blended overlapping subdomains, on the GPU.

The field is  f(x) = sum_k a(k) W(k) exp(i kappa.x)  over the full discrete spectrum, with
a(k) = sigma * sqrt(S(kappa) / sum S) (variance normalised in FP64) and W Hermitian complex
N(0,1) noise; the marginal transform of SRC_F/main_ScaRL.f90:1355-1395 follows.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _capi as A


@dataclass
class FieldSpec:
    n: Sequence[int]
    model: int = A.CM_GAUSSIAN
    corrL: Sequence[float] = (1.0, 1.0, 1.0)
    dx: Sequence[float] = (1.0, 1.0, 1.0)
    nu: float = 0.5
    sigma: float = 1.0
    margi: int = A.FOM_NONE
    avg: float = 0.0
    std_dev: float = 1.0
    seed: int = 0
    stream: int = 0
    normalize: bool = True

    def to_c(self) -> A.FieldDesc:
        d = A.FieldDesc()
        n = list(self.n) + [1] * (3 - len(self.n))
        for a in range(3):
            d.n[a] = int(n[a])
            d.dx[a] = float(self.dx[a])
            d.corrL[a] = float(self.corrL[a])
        d.model = int(self.model)
        d.nu = float(self.nu)
        d.sigma = float(self.sigma)
        d.margi = int(self.margi)
        d.avg = float(self.avg)
        d.std_dev = float(self.std_dev)
        d.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        d.stream = int(self.stream) & 0xFFFFFFFF
        d.normalize = 1 if self.normalize else 0
        return d

    @property
    def shape(self):
        n = list(self.n) + [1] * (3 - len(self.n))
        return tuple(int(v) for v in n)

    @property
    def half_shape(self):
        n = self.shape
        return (n[0] // 2 + 1, n[1], n[2])


def _dtype(precision):
    return np.float64 if precision == A.PREC_FP64 else np.float32


def _cdtype(precision):
    return np.complex128 if precision == A.PREC_FP64 else np.complex64


class Batch:
    """A set of fields generated together (one launch per pass, optionally one hipGraph)."""

    def __init__(self, handle, precision, shapes, global_shape=None):
        self._h = handle
        self.precision = precision
        self.shapes = shapes
        self.global_shape = global_shape

    @classmethod
    def ensemble(cls, specs: Sequence[FieldSpec], precision=A.PREC_FP32):
        arr = (A.FieldDesc * len(specs))(*[s.to_c() for s in specs])
        h = C.c_void_p()
        A.call("scarl_hip_batch_create", arr, len(specs), int(precision), C.byref(h))
        return cls(h, precision, [s.shape for s in specs])

    @classmethod
    def blend(cls, gdim, starts, exts, params: FieldSpec, precision=A.PREC_FP32):
        """starts/exts: three per-axis lists of subdomain starts and extents."""
        g = np.ascontiguousarray(np.asarray(gdim, np.int32))
        nb = np.ascontiguousarray(np.asarray([len(s) for s in starts], np.int32))
        st = np.ascontiguousarray(np.concatenate([np.asarray(s, np.int32) for s in starts]))
        ex = np.ascontiguousarray(np.concatenate([np.asarray(e, np.int32) for e in exts]))
        d = params.to_c()
        h = C.c_void_p()
        A.call("scarl_hip_blend_create", A.ptr(g), A.ptr(nb), A.ptr(st), A.ptr(ex), C.byref(d), int(precision),
               C.byref(h))
        shapes = [(int(a), int(b), int(c)) for c in exts[2] for b in exts[1] for a in exts[0]]
        return cls(h, precision, shapes, tuple(int(v) for v in gdim))

    def set_noise(self, i, W: Optional[np.ndarray]):
        if W is None:
            A.call("scarl_hip_batch_set_noise", self._h, int(i), None)
            return
        n = self.shapes[i]
        W = np.asfortranarray(W, dtype=_cdtype(self.precision))
        if W.shape != (n[0] // 2 + 1, n[1], n[2]):
            raise ValueError("noise must have the half-spectrum shape (n1/2+1, n2, n3)")
        A.call("scarl_hip_batch_set_noise", self._h, int(i), A.ptr(W))

    def run(self, use_graph=False):
        A.call("scarl_hip_batch_run", self._h, 1 if use_graph else 0)

    def sync(self):
        A.call("scarl_hip_batch_sync", self._h)

    def output(self, i=-1):
        shape = self.global_shape if i < 0 else self.shapes[i]
        out = np.empty(shape, _dtype(self.precision), order="F")
        A.call("scarl_hip_batch_copy_output", self._h, int(i), A.ptr(out), int(out.nbytes))
        return out

    def device_output(self, i=-1):
        p = C.c_void_p()
        dims = np.zeros(3, np.int64)
        A.call("scarl_hip_batch_output", self._h, int(i), C.byref(p), A.ptr(dims))
        return p.value, tuple(int(v) for v in dims)

    def moments(self):
        """(nslots, 5) array of {sum, sum of squares, count, mean, std} per moment slot."""
        out = np.empty((max(len(self.shapes), 1), 5))
        cnt = C.c_int32()
        A.call("scarl_hip_batch_moments", self._h, A.ptr(out), out.shape[0], C.byref(cnt))
        return out[:cnt.value]

    def lag_covariance(self, nlag):
        """Ensemble batches, after run(): (nfields, 3, nlag) per-axis covariance of every field at
        lags 0..nlag-1, centred on the field's own sample mean (FP64, compensated)."""
        out = np.empty((len(self.shapes), 3, int(nlag)))
        A.call("scarl_hip_batch_lag_cov", self._h, int(nlag), A.ptr(out))
        return out

    def field_info(self, i):
        a = C.c_double()
        s = C.c_double()
        A.call("scarl_hip_batch_field_info", self._h, int(i), C.byref(a), C.byref(s))
        return a.value, s.value

    def plan_report(self):
        """One dict per plan of the batch (pass, N, kernel, radices, T, tpp, ldp, skew, choice and,
        when the batch was autotuned, tuned_ms / default_ms / candidates, or cached)."""
        buf = C.create_string_buffer(1 << 16)
        A.call("scarl_hip_batch_plan_report", self._h, buf, len(buf))
        rows = []
        for line in buf.value.decode().splitlines():
            row = {}
            for tok in line.split():
                k, _, v = tok.partition("=")
                row[k] = v if v else True
            rows.append(row)
        return rows

    def nbytes(self):
        b = C.c_int64()
        A.call("scarl_hip_batch_bytes", self._h, C.byref(b))
        return b.value

    def close(self):
        if self._h:
            A.lib().scarl_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def gen_field_h(spec: FieldSpec, precision=A.PREC_FP64, W: Optional[np.ndarray] = None):
    """One Hermitian field (+ marginal); W: host-exported half-spectrum noise or None."""
    out = np.empty(spec.shape, _dtype(precision), order="F")
    d = spec.to_c()
    if W is not None:
        W = np.asfortranarray(W, dtype=_cdtype(precision))
        if W.shape != spec.half_shape:
            raise ValueError("noise must have the half-spectrum shape (n1/2+1, n2, n3)")
    A.call("scarl_hip_gen_field_h", A.ptr(out), C.byref(d), int(precision), A.ptr(W))
    return out


def timing(on=True):
    A.call("scarl_hip_set_timing", 1 if on else 0)


def timing_report():
    buf = C.create_string_buffer(8192)
    A.call("scarl_hip_timing_report", buf, len(buf))
    rep = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        rep[name] = (int(cnt), float(ms))
    return rep


def timing_reset():
    A.call("scarl_hip_timing_reset")


def seeded_tiling(gdim, nb, overlap, choices, seed):
    """Per-axis subdomain starts and extents that tile [0, gdim) with neighbours overlapping by
    `overlap` points: extents are drawn by seed from `choices` (sorted, evenly spaced), then fixed up
    one step at a time (a random subdomain moves one choice up or down) until the extents sum to
    gdim + (nb - 1) * overlap.  Returns (starts, exts), each a list of three lists."""
    rng = np.random.default_rng(seed)
    ch = sorted(int(c) for c in choices)
    step = ch[1] - ch[0] if len(ch) > 1 else 0
    starts, exts = [], []
    for a in range(3):
        n, g = int(nb[a]), int(gdim[a])
        if n == 1:
            starts.append([0])
            exts.append([g])
            continue
        target = g + (n - 1) * overlap
        if not (n * ch[0] <= target <= n * ch[-1]) or (step and (target - n * ch[0]) % step):
            raise ValueError(f"axis {a}: {n} extents from {ch} cannot tile {g} with overlap {overlap}")
        e = [int(v) for v in rng.choice(ch, size=n)]
        while sum(e) != target:
            i = int(rng.integers(n))
            if sum(e) > target and e[i] > ch[0]:
                e[i] -= step
            elif sum(e) < target and e[i] < ch[-1]:
                e[i] += step
        s = [0]
        for i in range(1, n):
            s.append(int(s[-1] + e[i - 1] - overlap))
        starts.append(s)
        exts.append([int(v) for v in e])
    return starts, exts


def set_tuning(key, value):
    """Plan tuning for batches created afterwards: A.TUNE_FP32_ELEMS (16 or 32 complex elements
    per thread), A.TUNE_MFMA_DFT16 (1: FP32 radix-16 first stages as split-precision bf16 MFMA
    products; 0: FP32 Stockham butterflies), A.TUNE_AUTOTUNE (1: time every plan's candidates at
    batch creation and keep the fastest) or A.TUNE_SCHEDULE (batches of several groups: 0 by
    Infinity Cache fit, 1 pipelined, 2 one group after another)."""
    A.call("scarl_hip_set_tuning", int(key), int(value))
