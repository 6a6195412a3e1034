"""``program main_Stat`` (SRC_F/Stat_ScaRL/main_Stat.f90) for one process.

The statistics of a generated sample, re-read from its HDF5 file:
* ``read_RF_h5_File_Attributes`` (:124-170) reads the sample's root attributes;
* ``set_Local_Range`` (:282-316): with one rank the local range is the whole field;
* ``read_RF_h5_File_Table`` (:567-608) reads the dataset;
* ``calculate_average_and_stdVar_MPI`` (statistics_RF.f90:18-174), ``rebuild_Sk``
  (:180-252) and ``rebuild_corrL`` (:468-486) run on the GPU (scarl_amd.statistics);
* ``write_StatisticsOnH5`` (:200-279) adds globalAvg, globalStdDev, evntAvg, evntStdDev and,
  when corrL is estimated, corrL_out and Sk_1..Sk_nDim to a file's root;
* ``makeStatFile`` (:610-787) writes STAT_<file> next to the sample. It holds a copy of the
  sample's attributes, old_file_mb_size and the statistics.

The input comes on stdin as in the reference (:38-71): the number of files, calculateCorrL,
deleteSampleInTheEnd, then a path and a dataset name per file, each read list-directed from
its own line. With one process, estimate_nFields_2 picks one block (:322-392), so the whole
field is read, and the MPI reductions and the average over ranks are the identity.
"""
from __future__ import annotations

import glob as _glob
import os
import sys
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import statistics as S
from . import write_output as W
from .read_input import DeckError, list_read

_INT_ATTRS = ("nDim", "Nmc", "method", "corrMod", "margiFirst")
_INT_VEC = ("Np", "Np_ovlp", "L")
_REAL_VEC = ("xMinGlob", "xMaxGlob", "xStep", "corrL", "overlap")


@dataclass
class STAT:
    """The fields of type(STAT) (SRC_F/Stat_ScaRL/type_STAT.f90:7-41) that one process uses."""
    nDim: int = -1
    Nmc: int = -1
    method: int = -1
    corrMod: int = -1
    margiFirst: int = -1
    Np: List[int] = field(default_factory=list)
    Np_ovlp: List[int] = field(default_factory=list)
    L: List[int] = field(default_factory=list)
    xMinGlob: List[float] = field(default_factory=list)
    xMaxGlob: List[float] = field(default_factory=list)
    xStep: List[float] = field(default_factory=list)
    corrL: List[float] = field(default_factory=list)
    overlap: List[float] = field(default_factory=list)
    L_Loc: List[int] = field(default_factory=list)
    localRange: Optional[np.ndarray] = None
    randField: Optional[np.ndarray] = None
    evntAvg: Optional[np.ndarray] = None
    evntStdDev: Optional[np.ndarray] = None
    globalAvg: float = 0.0
    globalStdDev: float = 0.0
    SkTot_Dir: Optional[np.ndarray] = None
    nhalf: List[int] = field(default_factory=list)
    corrL_out: Optional[np.ndarray] = None


def read_RF_h5_File_Attributes(resPath) -> STAT:
    """main_Stat.f90:124-170."""
    STA = STAT()
    with W.H5File(resPath) as f:
        for n in _INT_ATTRS:
            setattr(STA, n, int(f.read_attr("", n, "i")[0]))
        for n in _INT_VEC:
            setattr(STA, n, [int(v) for v in f.read_attr("", n, "i")])
        for n in _REAL_VEC:
            setattr(STA, n, [float(v) for v in f.read_attr("", n, "d")])
    return STA


def set_Local_Range(STA: STAT, local_sample=False):
    """main_Stat.f90:282-316 for the one block of a one-process run."""
    nPoints = STA.Np if local_sample else STA.L
    STA.localRange = np.array([[1, n] for n in nPoints[:STA.nDim]], np.int64)
    STA.L_Loc = [int(b - a + 1) for a, b in STA.localRange]


def read_RF_h5_File_Table(STA: STAT, file_path, dset):
    """main_Stat.f90:567-608: the local range of the dataset as randField(:, 1)."""
    off = [int(a - 1) for a, _ in STA.localRange] + [0] * (3 - STA.nDim)
    cnt = list(STA.L_Loc) + [1] * (3 - STA.nDim)
    with W.H5File(file_path) as f:
        STA.randField = f.read(dset, off, cnt)


def calculate_statistics(STA: STAT, calculateCorrL: bool):
    """calculate_average_and_stdVar_MPI, then (calculateCorrL) rebuild_Sk and rebuild_corrL."""
    ea, es, ga, gs = S.calculate_average_and_stdVar(STA.randField.reshape(-1, 1, order="F"))
    STA.evntAvg, STA.evntStdDev, STA.globalAvg, STA.globalStdDev = ea, es, ga, gs
    if calculateCorrL:
        STA.SkTot_Dir, STA.nhalf = S.rebuild_Sk(STA.randField, STA.L_Loc, STA.nDim, ga, gs)
        STA.corrL_out = S.rebuild_corrL(STA.SkTot_Dir, STA.nhalf, STA.xStep)


def write_StatisticsOnH5(STA: STAT, resPath, calculateCorrL: bool):
    """main_Stat.f90:200-279 (an attribute that exists is replaced)."""
    with W.H5File(resPath, "r+") as f:
        f.attr_real("", "globalAvg", float(STA.globalAvg))
        f.attr_real("", "globalStdDev", float(STA.globalStdDev))
        f.attr_real("", "evntAvg", list(STA.evntAvg))
        f.attr_real("", "evntStdDev", list(STA.evntStdDev))
        if calculateCorrL:
            f.attr_real("", "corrL_out", list(STA.corrL_out))
            i0 = 0
            for d in range(STA.nDim):
                f.attr_real("", f"Sk_{d + 1}", list(STA.SkTot_Dir[i0:i0 + STA.nhalf[d]]))
                i0 += STA.nhalf[d]


def makeStatFile(resPath, STA: STAT, deleteSampleInTheEnd: int, calculateCorrL: bool):
    """main_Stat.f90:610-787: STAT_<name> in the sample's folder."""
    folder, name = os.path.split(resPath)
    newFile_Path = os.path.join(folder, "STAT_" + name)
    old_file_mb_size = float(os.path.getsize(resPath)) / 1024.0 ** 2
    with W.H5File(resPath) as old, W.H5File(newFile_Path, "w") as new:
        for n in ("nb_procs", "nDim", "Nmc", "method", "seedStart", "corrMod", "margiFirst"):
            new.attr_int("", n, int(old.read_attr("", n, "i")[0]))
        new.attr_real("", "old_file_mb_size", old_file_mb_size)
        for n in ("nFields", "Np", "Np_ovlp", "L"):
            new.attr_int("", n, [int(v) for v in old.read_attr("", n, "i")])
        for n in _REAL_VEC:
            new.attr_real("", n, [float(v) for v in old.read_attr("", n, "d")])
    write_StatisticsOnH5(STA, newFile_Path, calculateCorrL)
    if deleteSampleInTheEnd == 1:
        # "rm <path up to the file name's first dot>*" (:650-660, 765-772)
        base = os.path.join(folder, name.split(".", 1)[0]) if "." in name else resPath
        for p in _glob.glob(base + "*"):
            if p != newFile_Path:
                os.remove(p)
    return newFile_Path


def main(stream=None):
    """The program: reads its input from `stream` (default stdin); returns the STAT of every
    file in order."""
    lines = iter((stream or sys.stdin).read().splitlines())

    def nxt(kind):
        try:
            return list_read(next(lines), kind)[0]
        except StopIteration:
            raise DeckError("end of input") from None

    nFiles = nxt("i")
    calculateCorrL = nxt("i") == 1
    deleteSampleInTheEnd = nxt("i")
    out = []
    for _ in range(nFiles):
        resPath = nxt("a")
        dset_name = nxt("a")
        STA = read_RF_h5_File_Attributes(resPath)
        set_Local_Range(STA)
        read_RF_h5_File_Table(STA, resPath, dset_name)
        calculate_statistics(STA, calculateCorrL)
        if calculateCorrL:
            write_StatisticsOnH5(STA, resPath, True)
        makeStatFile(resPath, STA, deleteSampleInTheEnd, calculateCorrL)
        out.append(STA)
    return out


if __name__ == "__main__":
    from . import _capi as A
    A.call("scarl_hip_init", 0)
    for st in main():
        print(f"globalAvg = {st.globalAvg:.10g}, globalStdDev = {st.globalStdDev:.10g}, corrL_out = {st.corrL_out}")
