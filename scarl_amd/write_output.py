"""HDF5 and XDMF output of a ScaRL sample, through libscarl_h5.so.

Mirrors the reference's ``write_output`` module (SRC_F/write_output.f90) and its attribute
helpers (SRC_F/hdf5_ScaRL.f90):

* ``write_hdf5_multi_proc_3D_1ds`` (:312-451): one file with one dataset, "samples", of the
  global extents L. Every subdomain block goes into it as a hyperslab at pos_0 - 1. The dataset
  gets the attributes xMin (rank 0's corner), xMax (the last rank's far corner, written only
  when there are several ranks), pos_0 = (1,1,1) and pos_N = L. Then the .xmf is written.
* ``write_hdf5_multi_proc_3D`` (:13-147): one file with one dataset per subdomain,
  "sample_1_p<rank>". Each has xMin/xMax attributes; rank 0's has no pos_0/pos_N (as written).
* ``write_hdf5_single_proc_3D`` (:191-263): one file per rank and its .xmf.
* ``write_HDF5_attributes`` (:722-813), ``write_time_on_hdf5`` (:819-911),
  ``write_info_file`` (:269-305), ``write_constant_field`` (:149-186),
  ``write_XMF_elements`` (:455-556) and ``write_XMF_global`` (:558-656).

The reference gathers each rank's block at rank 0 over MPI. Here all blocks are already in one
process, so the functions take the gathered blocks (a list in rank order) directly.

A field is x fastest. It can be a numpy array of shape (n1, n2, n3) with unit x stride (Fortran
order, or an x-fastest view into a larger array), or a torch tensor of shape (n3, n2, n1) on any
device. A device tensor is streamed to the file in z slabs.
"""
from __future__ import annotations

import ctypes as C
import datetime as _dt
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _h5 as H

DS_NAME = "samples"
FFT = 1  # SRC_F/constants_ScaRL.f90:12 (written as the "method" attribute)
_SLAB_BYTES = 64 << 20


def _i3(v):
    a = np.ascontiguousarray(np.asarray(v, np.int32).reshape(3))
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _d3(v):
    a = np.ascontiguousarray(np.asarray(v, np.float64).reshape(3))
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _prec(dtype) -> int:
    if dtype == np.float64:
        return 64
    if dtype == np.float32:
        return 32
    raise TypeError(f"fields are float64 or float32, not {dtype}")


@dataclass
class Block:
    """One subdomain's block as the reference's rank holds it: the field (extents Np), its
    1-based position pos_0 in the global grid, and its corner coordinates coord_0 / coord_N
    (SRC_F/main_ScaRL.f90:455-460)."""
    field: object
    pos_0: Sequence[int]
    coord_0: Sequence[float]
    coord_N: Sequence[float]

    @property
    def shape(self):
        return field_shape(self.field)

    @property
    def pos_N(self):
        return [int(p) + int(n) - 1 for p, n in zip(self.pos_0, self.shape)]


def field_shape(a):
    """Fortran extents (n1, n2, n3) of an x-fastest field."""
    if _is_torch(a):
        n3, n2, n1 = a.shape
        return (int(n1), int(n2), int(n3))
    return tuple(int(v) for v in a.shape)


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


class H5File:
    """An open HDF5 file (scarl_h5 handle)."""

    def __init__(self, path, mode="r"):
        self.path = str(path)
        h = C.c_void_p()
        if mode == "w":
            H.call("scarl_h5_create", H.b(self.path), C.byref(h))
        elif mode in ("r", "r+"):
            H.call("scarl_h5_open", H.b(self.path), 1 if mode == "r+" else 0, C.byref(h))
        else:
            raise ValueError("mode is 'w', 'r' or 'r+'")
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            H.call("scarl_h5_close", self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- writing
    def create_dataset(self, name, dims, precision=64):
        _, p = _i3(dims)
        H.call("scarl_h5_create_dataset", self._h, H.b(name), p, int(precision))

    def write_block(self, name, field, offset=(0, 0, 0)):
        """Writes an x-fastest field at the 0-based Fortran offset."""
        if _is_torch(field):
            self._write_torch(name, field, offset)
            return
        a = field
        if not isinstance(a, np.ndarray) or a.ndim != 3:
            raise TypeError("field must be a 3D numpy array (n1, n2, n3) or a torch tensor (n3, n2, n1)")
        prec = _prec(a.dtype)
        it = a.itemsize
        if a.strides[0] != it or a.strides[1] % it or a.strides[2] % it:
            raise ValueError("field must be x fastest (unit stride along its first axis)")
        ld1 = a.strides[1] // it if a.shape[1] > 1 else a.shape[0]
        ld2 = a.strides[2] // it if a.shape[2] > 1 else ld1 * a.shape[1]
        _, po = _i3(offset)
        _, pc = _i3(a.shape)
        H.call("scarl_h5_write_block", self._h, H.b(name), C.c_void_p(a.ctypes.data), prec, int(ld1), int(ld2), po,
               pc)

    def _write_torch(self, name, t, offset):
        n3, n2, n1 = (int(v) for v in t.shape)
        plane = n1 * n2 * t.element_size()
        step = max(1, _SLAB_BYTES // max(plane, 1))
        for z0 in range(0, n3, step):
            z1 = min(n3, z0 + step)
            slab = t[z0:z1].detach().contiguous().cpu().numpy()
            self.write_block(name, slab.T, (offset[0], offset[1], offset[2] + z0))

    def attr_int(self, obj, name, v):
        """write_h5attr_int / write_h5attr_int_vec (SRC_F/hdf5_ScaRL.f90:31-49, 105-124)."""
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, np.int32)))
        n = 0 if np.ndim(v) == 0 else a.size
        H.call("scarl_h5_attr_i32", self._h, H.b(obj), H.b(name), a.ctypes.data_as(C.POINTER(C.c_int32)), n)

    def attr_real(self, obj, name, v):
        """write_h5attr_real / write_h5attr_real_vec (SRC_F/hdf5_ScaRL.f90:53-72, 155-172)."""
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, np.float64)))
        n = 0 if np.ndim(v) == 0 else a.size
        H.call("scarl_h5_attr_f64", self._h, H.b(obj), H.b(name), a.ctypes.data_as(C.POINTER(C.c_double)), n)

    def write_times(self, labels, times):
        """times: (ntimes, nprocs) (Fortran time_all); labels: ntimes strings."""
        t = np.asarray(times, np.float64)
        if t.ndim == 1:
            t = t[:, None]
        nt, npr = t.shape
        if len(labels) != nt:
            raise ValueError("one label per time")
        lab = b"".join(_fstring(s, H.LABEL_LEN) for s in labels)
        tf = np.asfortranarray(t)
        H.call("scarl_h5_write_times", self._h, lab, tf.ctypes.data_as(C.POINTER(C.c_double)), nt, npr)

    # ---- reading
    def dataset_info(self, name):
        dims = np.zeros(3, np.int32)
        prec = C.c_int32()
        H.call("scarl_h5_dataset_info", self._h, H.b(name), dims.ctypes.data_as(C.POINTER(C.c_int32)),
               C.byref(prec))
        return tuple(int(v) for v in dims), prec.value

    def read(self, name=DS_NAME, offset=(0, 0, 0), count=None, dtype=np.float64):
        """Reads a block (default: the whole dataset) as a Fortran-ordered (n1, n2, n3) array."""
        dims, _ = self.dataset_info(name)
        if count is None:
            count = [d - o for d, o in zip(dims, offset)]
        out = np.empty(tuple(int(c) for c in count), dtype, order="F")
        _, po = _i3(offset)
        _, pc = _i3(count)
        H.call("scarl_h5_read_block", self._h, H.b(name), C.c_void_p(out.ctypes.data), _prec(out.dtype),
               int(out.shape[0]), int(out.shape[0] * out.shape[1]), po, pc)
        return out

    def attr_names(self, obj=""):
        n = C.c_int32()
        H.call("scarl_h5_attr_count", self._h, H.b(obj), C.byref(n))
        names = []
        buf = C.create_string_buffer(256)
        for i in range(n.value):
            H.call("scarl_h5_attr_name", self._h, H.b(obj), i, buf, len(buf))
            names.append(buf.value.decode())
        return names

    def read_attr(self, obj, name, kind):
        """kind 'i' (int32) or 'd' (double); a scalar attribute comes back as a 1-element array."""
        cap = 64
        n = C.c_int32()
        if kind == "i":
            v = np.zeros(cap, np.int32)
            H.call("scarl_h5_read_attr_i32", self._h, H.b(obj), H.b(name), v.ctypes.data_as(C.POINTER(C.c_int32)),
                   cap, C.byref(n))
        else:
            v = np.zeros(cap, np.float64)
            H.call("scarl_h5_read_attr_f64", self._h, H.b(obj), H.b(name), v.ctypes.data_as(C.POINTER(C.c_double)),
                   cap, C.byref(n))
        return v[:n.value].copy()

    def read_times(self):
        nt, npr = C.c_int32(), C.c_int32()
        H.call("scarl_h5_read_times", self._h, None, 0, None, 0, C.byref(nt), C.byref(npr))
        t = np.zeros(nt.value * npr.value)
        lab = C.create_string_buffer(nt.value * H.LABEL_LEN)
        H.call("scarl_h5_read_times", self._h, t.ctypes.data_as(C.POINTER(C.c_double)), t.size, lab, nt.value,
               C.byref(nt), C.byref(npr))
        labels = [lab.raw[i * H.LABEL_LEN:(i + 1) * H.LABEL_LEN].decode() for i in range(nt.value)]
        return labels, t.reshape((nt.value, npr.value), order="F")


def _fstring(s, n) -> bytes:
    """A Fortran character(len=n) value: truncated or blank padded."""
    e = H.b(s)[:n]
    return e + b" " * (n - len(e))


def fortran_fixed(v, w=25, d=10) -> str:
    """The Fw.d edit of one value, as the .xmf records carry it."""
    buf = C.create_string_buffer(w + 1)
    H.call("scarl_fortran_fixed", float(v), int(w), int(d), buf, len(buf))
    return buf.value.decode()


def _join(folder, name):
    return f"{folder.rstrip('/') if folder else '.'}/{name}"


# ---- the reference's writers -----------------------------------------------------------------

def write_pos_on_HDF5_dataset(f: H5File, ds_name, pos_0, pos_N):
    """write_output.f90:700-716."""
    f.attr_int(ds_name, "pos_0", list(pos_0))
    f.attr_int(ds_name, "pos_N", list(pos_N))


def write_XMF_elements(HDF5_name, xMin, xMax, xNStep, XMF_name, XMF_folder, H5_TO_XMF_path, ds_name,
                       precision=64):
    """write_output.f90:455-556 (precision 32 marks an FP32 dataset; the reference writes 8)."""
    _, pmin = _d3(xMin)
    _, pmax = _d3(xMax)
    _, pn = _i3(xNStep)
    H.call("scarl_xmf_write_elements", H.b(XMF_folder), H.b(XMF_name), H.b(HDF5_name), H.b(H5_TO_XMF_path),
           H.b(ds_name), pmin, pmax, pn, int(precision))


def write_XMF_global(globXMF_path, HDF5_list, coord_0_list, xStep, xNStep):
    """write_output.f90:558-656. coord_0_list: (nb_procs, 3) origins in rank order."""
    nb = len(HDF5_list)
    c0 = np.ascontiguousarray(np.asarray(coord_0_list, np.float64).reshape(nb, 3))
    names = (C.c_char_p * nb)(*[H.b(s) for s in HDF5_list])
    _, ps = _d3(xStep)
    _, pn = _i3(xNStep)
    H.call("scarl_xmf_write_global", H.b(globXMF_path), names, c0.ctypes.data_as(C.POINTER(C.c_double)), ps, pn,
           nb)


def write_hdf5_multi_proc_3D_1ds(blocks: Sequence[Block], L, HDF5_name, XMF_name, res_folder, precision=64):
    """write_output.f90:312-451 with every rank's block gathered (rank order)."""
    if not blocks:
        raise ValueError("no blocks")
    L = [int(v) for v in L]
    path = _join(res_folder, HDF5_name)
    xMin = list(blocks[0].coord_0)
    xMax_temp = list(blocks[0].coord_N)
    with H5File(path, "w") as f:
        f.create_dataset(DS_NAME, L, precision)
        for i, blk in enumerate(blocks):
            f.write_block(DS_NAME, blk.field, [int(p) - 1 for p in blk.pos_0])
            if i == 0:
                f.attr_real(DS_NAME, "xMin", xMin)
            else:
                xMax_temp = list(blk.coord_N)
                if i == len(blocks) - 1:
                    f.attr_real(DS_NAME, "xMax", xMax_temp)
        write_pos_on_HDF5_dataset(f, DS_NAME, [1, 1, 1], L)
    write_XMF_elements(HDF5_name, xMin, xMax_temp, L, XMF_name, res_folder, ".", DS_NAME, precision)
    return path


def write_hdf5_multi_proc_3D(blocks: Sequence[Block], HDF5_name, res_folder, precision=64):
    """write_output.f90:13-147: one dataset "sample_1_p<rank>" per block in one file."""
    path = _join(res_folder, HDF5_name)
    with H5File(path, "w") as f:
        for i, blk in enumerate(blocks):
            ds = f"sample_1_p{i}"
            f.create_dataset(ds, blk.shape, precision)
            f.write_block(ds, blk.field)
            f.attr_real(ds, "xMin", list(blk.coord_0))
            f.attr_real(ds, "xMax", list(blk.coord_N))
            if i > 0:
                write_pos_on_HDF5_dataset(f, ds, blk.pos_0, blk.pos_N)
    return path


def write_hdf5_single_proc_3D(blk: Block, HDF5_name, XMF_name, res_folder, rank: Optional[int] = None,
                              precision=64):
    """write_output.f90:191-263: one block in its own file, plus its .xmf."""
    ds = DS_NAME if rank is None else f"sample_1_p{int(rank)}"
    path = _join(res_folder, HDF5_name)
    with H5File(path, "w") as f:
        f.create_dataset(ds, blk.shape, precision)
        f.write_block(ds, blk.field)
        f.attr_real(ds, "xMin", list(blk.coord_0))
        f.attr_real(ds, "xMax", list(blk.coord_N))
        write_pos_on_HDF5_dataset(f, ds, blk.pos_0, blk.pos_N)
    write_XMF_elements(HDF5_name, blk.coord_0, blk.coord_N, blk.shape, XMF_name, res_folder, ".", ds, precision)
    return path


def write_HDF5_attributes(HDF5Path, nb_procs, nDim, Nmc, method, seedStart, corrMod, margiFirst, nFields,
                          xMinGlob, xMaxGlob, xStep, corrL, overlap, L, Np, Np_ovlp):
    """write_output.f90:722-813: root attributes of an existing file."""
    with H5File(HDF5Path, "r+") as f:
        for name, v in (("nb_procs", nb_procs), ("nDim", nDim), ("Nmc", Nmc), ("method", method),
                        ("seedStart", seedStart), ("corrMod", corrMod), ("margiFirst", margiFirst)):
            f.attr_int("", name, int(v))
        for name, v in (("nFields", nFields), ("Np", Np), ("Np_ovlp", Np_ovlp), ("L", L)):
            f.attr_int("", name, list(v))
        for name, v in (("xMinGlob", xMinGlob), ("xMaxGlob", xMaxGlob), ("xStep", xStep), ("corrL", corrL),
                        ("overlap", overlap)):
            f.attr_real("", name, list(v))


def write_time_on_hdf5(time_label, time_trace, HDF5_name, res_folder):
    """write_output.f90:819-911. time_trace: (ntimes,) or (ntimes, nprocs)."""
    with H5File(_join(res_folder, HDF5_name), "r+") as f:
        f.write_times(list(time_label), time_trace)


def write_info_file(info_file_name, res_folder):
    """write_output.f90:269-305: an empty file that write_HDF5_attributes then fills."""
    with H5File(_join(res_folder, info_file_name), "w"):
        pass


def write_constant_field(xMinGlob, xMaxGlob, avg, output_name, res_folder):
    """write_output.f90:149-186: the std_dev = 0 sample, a 2x2x2 constant field."""
    field = np.full((2, 2, 2), float(avg), order="F")
    HDF5_name, XMF_name = f"{output_name}.h5", f"{output_name}.xmf"
    blk = Block(field, [1, 1, 1], list(xMinGlob), list(xMaxGlob))
    write_hdf5_single_proc_3D(blk, HDF5_name, XMF_name, res_folder)
    xStep = [b - a for a, b in zip(xMinGlob, xMaxGlob)]
    write_HDF5_attributes(_join(res_folder, HDF5_name), 1, 3, 1, -1, -1, -1, -1, [1, 1, 1], xMinGlob, xMaxGlob,
                          xStep, [-1.0, -1.0, -1.0], [0.0, 0.0, 0.0], [2, 2, 2], [2, 2, 2], [0, 0, 0])


def output_names(output_name_in, res_folder, one_file, rank=0, now: Optional[_dt.datetime] = None):
    """File names of one sample (main_ScaRL.f90:411-446): "<name>.h5" (one file) or
    "<name>_proc_<rank:05d>.h5". When that file exists, the name gets the suffix
    "-YYMMDD_hhmmss" of the current time. Returns (output_name, HDF5_name, XMF_name)."""

    def names(o):
        if one_file:
            return f"{o}.h5", f"{o}.xmf"
        return f"{o}_proc_{int(rank):05d}.h5", f"{o}_proc_{int(rank):05d}.xmf"

    output_name = output_name_in
    h5, xmf = names(output_name)
    if os.path.exists(_join(res_folder, h5)):
        t = now or _dt.datetime.now()
        output_name = f"{output_name_in}-{t.strftime('%y%m%d_%H%M%S')}"
        h5, xmf = names(output_name)
    return output_name, h5, xmf
