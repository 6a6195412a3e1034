"""``program main_ScaRL`` (SRC_F/main_ScaRL.f90): GPU versions of its hot element routines, and
the program itself (``main``), which reads the decks and writes every sample's HDF5/XDMF files.

The reference keeps the routines as internal procedures of the program; here they are callable.
Run the program with ``python -m scarl_amd.main_ScaRL <folder> --nb-procs N``.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _capi as A


def apply_UnityPartition_mtx(Np, Np_ovlp, partitionType, topo_pos, topo_shape, randField_3D):
    """apply_UnityPartition_mtx (main_ScaRL.f90:1216-1301): randField_3D *= sqrt(PoU), in place."""
    f = randField_3D
    if not (f.dtype == np.float64 and f.flags.f_contiguous):
        raise ValueError("randField_3D must be a Fortran-ordered float64 array")
    i3 = lambda v: np.ascontiguousarray(np.asarray(v, np.int32).reshape(3))
    A.call("scarl_hip_apply_unity_partition", A.ptr(f), A.ptr(i3(Np)), A.ptr(i3(Np_ovlp)), int(partitionType),
           A.ptr(i3(topo_pos)), A.ptr(i3(topo_shape)))
    return f


def normalize_field(topo_pos, topo_shape, Np, Np_ovlp, rank, comm_group, randField):
    """normalize_field (main_ScaRL.f90:903-953) on one rank, in place; returns (avg, std_dev).
    The reference reduces over every rank of comm_group: here topo_shape must be (1, 1, 1)."""
    del rank, comm_group  # one rank: the ALLREDUCE is the identity
    f = randField
    if not (f.dtype == np.float64 and f.flags.f_contiguous):
        raise ValueError("randField must be a Fortran-ordered float64 array")
    i3 = lambda v: np.ascontiguousarray(np.asarray(v, np.int32).reshape(3))
    avg, sd = C.c_double(), C.c_double()
    A.call("scarl_hip_normalize_field", A.ptr(f), A.ptr(i3(Np)), A.ptr(i3(Np_ovlp)), A.ptr(i3(topo_pos)),
           A.ptr(i3(topo_shape)), C.byref(avg), C.byref(sd))
    return avg.value, sd.value


def multiVariateTransformation(avg, std_dev, margiFirst, randField):
    """multiVariateTransformation (main_ScaRL.f90:1355-1395), in place."""
    f = randField
    if not (f.dtype == np.float64 and (f.flags.f_contiguous or f.flags.c_contiguous)):
        raise ValueError("randField must be a contiguous float64 array")
    A.call("scarl_hip_multivariate_transformation", float(avg), float(std_dev), int(margiFirst), A.ptr(f),
           int(f.size))
    return f


# ---- the program: decks in, HDF5/XDMF files out -------------------------------------------------

TIME_LABELS = ("Read_input", "Topology", "Generation", "Apply_unit", "Add_overlap", "Normalize",
               "Multivar_trans", "Write_file", "Total_Wall_time")  # main_ScaRL.f90:83-87,187-191,527-583,670-680


def _stochastic_seed():
    """seedBase < 0 (main_ScaRL.f90:499-507): a clock/entropy-based positive seed."""
    import os as _os
    s = int.from_bytes(_os.urandom(4), "little") & 0x7FFFFFFF
    return s if s != 0 else 1


def create_fields(sample, g, nb_procs, gen_meth, config, res_folder, precision=A.PREC_FP32, file_precision=64,
                  times=None, t_init=None, now=None):
    """create_fields (main_ScaRL.f90:343-739) for every rank of the sample at once.

    The subdomains the reference spreads over nb_procs ranks are generated as one batch on the
    GPU: FFT blends them (partition of unity and overlap sum fused into the last pass); FFT_MPI
    generates the one global field. The field then goes through the reference's writers, one
    block per rank. Generation, the partition of unity, the overlap sum and the marginal are one
    fused launch sequence here, so their time is reported under "Generation" and the other
    three stage times are 0.
    """
    import datetime as _dt
    import glob as _glob
    import os as _os
    import time as _time

    from . import geometry as G
    from . import write_output as W
    from .generator import Batch, FieldSpec

    times = dict(times or {})
    t_init = _time.perf_counter() if t_init is None else t_init
    tic = _time.perf_counter()
    output_name, HDF5_name, XMF_name = W.output_names(sample.output_name, res_folder, config.one_file, 0, now)
    seedBase = sample.seedBase
    seed_used = seedBase + 1 if seedBase >= 0 else _stochastic_seed()
    if gen_meth == G.FFT_MPI:
        spec = FieldSpec(n=tuple(g["L"]), model=sample.corrMod, corrL=tuple(sample.corrL), dx=tuple(g["xStep"]),
                         margi=sample.margiFirst, avg=sample.avg, std_dev=sample.std_dev, seed=seed_used)
        batch = Batch.ensemble([spec], precision)
        out_index = 0
    else:
        gdim, starts, exts = G.blend_layout(g)
        spec = FieldSpec(n=(2, 2, 2), model=sample.corrMod, corrL=tuple(sample.corrL), dx=tuple(g["xStep"]),
                         margi=sample.margiFirst, avg=sample.avg, std_dev=sample.std_dev, seed=seed_used)
        batch = Batch.blend(gdim, starts, exts, spec, precision)
        out_index = -1
    with batch:
        batch.run()
        batch.sync()
        times["Generation"] = _time.perf_counter() - tic
        for k in ("Apply_unit", "Add_overlap", "Normalize", "Multivar_trans"):
            times[k] = 0.0
        tic = _time.perf_counter()
        glob_field = batch.output(out_index)
    ranks = [G.rank_geometry(g, r) for r in range(nb_procs)]
    Np = g["Np"]
    blocks = []
    for r in ranks:
        p0 = [v - 1 for v in r["pos_0"]]
        view = glob_field[p0[0]:p0[0] + Np[0], p0[1]:p0[1] + Np[1], p0[2]:p0[2] + Np[2]]
        blocks.append(W.Block(view, r["pos_0"], r["coord_0"], r["coord_N"]))
    xMaxGlob = ranks[0]["xMaxGlob"]
    overlap = ranks[0]["overlap"]
    attrs = dict(nb_procs=nb_procs, nDim=3, Nmc=1, method=W.FFT, seedStart=seedBase, corrMod=sample.corrMod,
                 margiFirst=sample.margiFirst, nFields=g["topo_shape"], xMinGlob=g["xMinGlob"], xMaxGlob=xMaxGlob,
                 xStep=g["xStep"], corrL=sample.corrL, overlap=overlap, L=g["L"], Np=Np, Np_ovlp=g["Np_ovlp"])
    files = []
    if config.one_file:
        if config.one_hdf_ds:
            W.write_hdf5_multi_proc_3D_1ds(blocks, g["L"], HDF5_name, XMF_name, res_folder, file_precision)
        else:
            W.write_hdf5_multi_proc_3D(blocks, HDF5_name, res_folder, file_precision)
        W.write_HDF5_attributes(f"{res_folder}/{output_name}.h5", **attrs)
        files.append(HDF5_name)
    else:
        for rank, blk in enumerate(blocks):
            h5 = f"{output_name}_proc_{rank:05d}.h5"
            W.write_hdf5_single_proc_3D(blk, h5, f"{output_name}_proc_{rank:05d}.xmf", res_folder, rank,
                                        file_precision)
            W.write_HDF5_attributes(f"{res_folder}/{h5}", **attrs)
            files.append(h5)
    if not (config.one_file and config.one_hdf_ds):
        h5_list = [f"{output_name}.h5" if config.one_file else f"{output_name}_proc_{r:05d}.h5"
                   for r in range(nb_procs)]
        c0 = [[float((Np[a] - g["Np_ovlp"][a]) * r["topo_pos"][a]) * g["xStep"][a] + g["xMinGlob"][a]
               for a in range(3)] for r in ranks]
        W.write_XMF_global(f"{res_folder}/{output_name}.xmf", h5_list, c0, g["xStep"], Np)
    times["Write_file"] = _time.perf_counter() - tic
    times["Total_Wall_time"] = _time.perf_counter() - t_init
    labels = [k for k in TIME_LABELS if k in times]
    trace = np.array([times[k] for k in labels])
    # every rank's column holds the same times: all subdomains came from the same launches
    if config.one_file:
        W.write_time_on_hdf5(labels, np.repeat(trace[:, None], nb_procs, axis=1), HDF5_name, res_folder)
    else:
        for h5 in files:
            W.write_time_on_hdf5(labels, trace, h5, res_folder)
    info = None
    if config.c_info_file:
        info = f"INFO-{output_name}.h5"
        if _os.path.exists(f"{res_folder}/{info}"):
            t = now or _dt.datetime.now()
            info = f"INFO-{output_name}-{t.strftime('%y%m%d_%H%M%S')}.h5"
        W.write_info_file(info, res_folder)
        W.write_HDF5_attributes(f"{res_folder}/{info}", **attrs)
        W.write_time_on_hdf5(labels, np.repeat(trace[:, None], nb_procs, axis=1), info, res_folder)
    if config.d_sample:
        for p in _glob.glob(f"{res_folder}/{output_name}*"):
            _os.remove(p)
    return {"output_name": output_name, "files": files, "info": info, "seed": seed_used, "field": glob_field,
            "blocks": blocks, "times": dict(zip(labels, trace.tolist()))}


def main(cwd=".", nb_procs=1, precision=A.PREC_FP32, file_precision=64, now=None):
    """program main_ScaRL (main_ScaRL.f90:1-220) on one GPU: reads config_ScaRL.txt, the sample
    deck and extra_config_ScaRL.txt from cwd, and writes every sample under the output folder.
    nb_procs is the number of ranks the reference would run on; it fixes the subdomain grid."""
    import os as _os
    import time as _time

    from . import geometry as G
    from . import read_input as R
    from . import write_output as W

    t_init = _time.perf_counter()
    cwd = _os.path.abspath(cwd)
    config = R.read_config_ScaRL(f"{cwd}/config_ScaRL.txt")
    deck = R.read_input_ScaRL(config.input_file, cwd)
    extra, topo_in = R.read_extra_config_ScaRL(f"{cwd}/extra_config_ScaRL.txt")
    res_folder = _os.path.join(cwd, deck.output_folder)
    _os.makedirs(res_folder, exist_ok=True)
    results = []
    tic = t_init
    for s in deck.samples:
        times = {"Read_input": _time.perf_counter() - tic}
        tic = _time.perf_counter()
        R.verify_inputs(s, config.gen_meth)
        xStep = [c / float(p - 1) for c, p in zip(s.corrL, s.pointsPerCorrL)]
        L0 = [1 + math.ceil((b - a) / h) for a, b, h in zip(s.xMinGlob, s.xMaxGlob, xStep)]
        nov = [0, 0, 0] if config.gen_meth == G.FFT_MPI else \
            [1 + math.ceil(o * float(p - 1)) for o, p in zip(s.overlap, s.pointsPerCorrL)]
        if extra:
            if int(np.prod(topo_in)) != nb_procs:
                raise R.DeckError("ERROR! extra_config_ScaRL.txt found but procs grid dimensions do not agree "
                                  "with the total number of procs in MPI process")
            topo = list(topo_in)
        else:
            topo = G.decide_topo_shape(nb_procs, L0, nov, 1, config.gen_meth)
        used = int(np.prod(topo))
        if s.std_dev == 0.0:
            W.write_constant_field(s.xMinGlob, s.xMaxGlob, s.avg, s.output_name, res_folder)
            results.append({"output_name": s.output_name, "constant": True})
            continue
        g = G.sample_geometry(s.xMinGlob, s.xMaxGlob, s.corrL, s.overlap, s.pointsPerCorrL, topo, config.gen_meth)
        times["Topology"] = _time.perf_counter() - tic
        r = create_fields(s, g, used, config.gen_meth, config, res_folder, precision, file_precision, times, t_init,
                          now)
        r.update(topo_shape=topo, nb_procs=used, geometry=g)
        results.append(r)
        tic = _time.perf_counter()
    return results


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="Generate the samples of a ScaRL deck on one GPU")
    ap.add_argument("folder", nargs="?", default=".")
    ap.add_argument("--nb-procs", type=int, default=1, help="rank count the subdomain grid is chosen for")
    ap.add_argument("--fp64", action="store_true", help="generate in FP64 (default FP32)")
    ap.add_argument("--dataset-fp32", action="store_true", help="store FP32 datasets (default FP64, as the reference)")
    a = ap.parse_args()
    A.call("scarl_hip_init", 0)
    for res in main(a.folder, a.nb_procs, A.PREC_FP64 if a.fp64 else A.PREC_FP32, 32 if a.dataset_fp32 else 64):
        print(res.get("output_name"), res.get("files"), res.get("topo_shape"))
