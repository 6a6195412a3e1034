"""Subdomain geometry of a ScaRL sample and its mapping onto one batched blend on the GPU.

Restates the geometry chain of ``program main_ScaRL`` that feeds every rank's generator call:

* ``xStep = corrL/(pointsPerCorrL - 1)``, ``L = 1 + ceiling((xMax - xMin)/xStep)``,
  ``Np_ovlp = 1 + ceiling(overlap (pointsPerCorrL - 1))`` (0 for FFT_MPI)
  (SRC_F/main_ScaRL.f90:100-111);
* ``Np = ceiling((L + (topo_shape - 1) Np_ovlp)/topo_shape)``, expanded to ``Np_ovlp`` (two
  subdomains on an axis) or ``2 Np_ovlp`` (three or more) when smaller, then
  ``L = Np topo_shape - (topo_shape - 1) Np_ovlp`` (:157-176);
* per rank: ``pos_0 = (Np - Np_ovlp) topo_pos + 1``, ``pos_N = pos_0 + Np - 1``, the corner
  coordinates, ``xRange``, the overlap length and ``xMaxGlob`` (:455-468);
* ``seedStart = seedBase + 1 + rank`` (:499-507); a negative seedBase (clock-based seed) is the
  caller's to resolve;
* ``get_topo_pos`` / ``get_topo_rank``: x-fastest rank order (:747-782);
* ``decide_topo_shape`` / ``find_factors``: the subdomain grid for nb_procs ranks when no
  extra_config_ScaRL.txt gives one (:786-978).

The reference gives each subdomain to its own MPI rank. Here all the subdomains of a sample go
into one blended batch on one GPU. Subdomain id = rank, so device noise is keyed by
``(seedBase + 1, rank)`` where the reference keys ``random_seed`` by ``seedBase + 1 + rank``.
"""
from __future__ import annotations

import math
from typing import Sequence

from . import _capi as A

FFT, FFT_MPI = 1, 2  # SRC_F/constants_ScaRL.f90:12-13


def get_topo_pos(rank: int, topo_shape: Sequence[int]):
    """SRC_F/main_ScaRL.f90:747-764."""
    s1, s2 = int(topo_shape[0]), int(topo_shape[1])
    p3 = rank // (s1 * s2)
    rest = rank - p3 * s1 * s2
    p2 = rest // s1
    return (rest - p2 * s1, p2, p3)


def get_topo_rank(topo_shape: Sequence[int], topo_pos: Sequence[int]) -> int:
    """SRC_F/main_ScaRL.f90:770-782."""
    return int(topo_shape[0]) * int(topo_shape[1]) * int(topo_pos[2]) + int(topo_shape[0]) * int(topo_pos[1]) + int(
        topo_pos[0])


def find_factors(n: int):
    """Prime factors of n in ascending order (SRC_F/main_ScaRL.f90:959-978)."""
    d, div, nxt, rest = [], 2, 3, int(n)
    while rest != 1:
        while rest % div == 0:
            d.append(div)
            rest //= div
        div, nxt = nxt, nxt + 2
    return d


def decide_topo_shape(nb_procs: int, L, nPointsOvlp, pointsPerBlockIdeal: int = 1, gen_meth=FFT):
    """SRC_F/main_ScaRL.f90:786-897.

    FFT: grow the block count on the axis with the most points per block until a block holds at
    most pointsPerBlockIdeal points (1 in the reference, :29) or there are nb_procs blocks. Then,
    for every rank count np from ceiling(0.75 nb_procs) to nb_procs, spread np's prime factors
    (largest first) over the axes with the most blocks, and keep the grid with the best
    volume/surface ratio prod(n) / (2 (n1 n2 + n2 n3 + n3 n1)). The reference compares against
    vol_surf_factor before setting it; it is read here as 0, so the first grid is always taken.
    Later grids replace it on ties.

    FFT_MPI: FFTW-MPI slabs along z, and topo_shape = (1, 1, ranks with planes).
    fftw_mpi_local_size_3d's default block is ceiling(L3 / nb_procs) planes per rank.
    """
    L = [int(v) for v in L]
    if gen_meth == FFT_MPI:
        block = -(-L[2] // int(nb_procs))
        return [1, 1, -(-L[2] // block)]
    nov = [int(v) for v in nPointsOvlp]
    nBlocks = [1, 1, 1]
    while True:
        nPoints = [L[a] + (nBlocks[a] - 1) * nov[a] for a in range(3)]
        per = [math.ceil(float(nPoints[a]) / float(nBlocks[a])) for a in range(3)]
        if per[0] * per[1] * per[2] <= pointsPerBlockIdeal:
            break
        if nBlocks[0] * nBlocks[1] * nBlocks[2] >= nb_procs:
            break
        pos = per.index(max(per))
        nBlocks[pos] += 1
    ideal = list(nBlocks)
    vol_surf = 0.0
    chosen = [1, 1, 1]
    np_start = math.ceil(0.75 * nb_procs)  # REAL(4) 0.75 is exact
    for npr in range(np_start, nb_procs + 1):
        nb = list(ideal)
        nf = [1, 1, 1]
        for fct in reversed(find_factors(npr)):
            if all(v == 1 for v in nb):
                break
            pos = nb.index(max(nb))
            nf[pos] *= fct
            nb[pos] = math.ceil(float(nb[pos]) / float(fct))
        f = float(nf[0] * nf[1] * nf[2]) / float(2 * (nf[0] * nf[1] + nf[1] * nf[2] + nf[2] * nf[0]))
        if f >= vol_surf:
            vol_surf, chosen = f, nf
    return chosen


def sample_geometry(xMinGlob, xMaxGlob, corrL, overlap, pointsPerCorrL, topo_shape, gen_meth=FFT):
    """Global geometry of one sample on a topology (every quantity per axis)."""
    g = {"xMinGlob": [float(v) for v in xMinGlob], "topo_shape": [int(v) for v in topo_shape]}
    g["xStep"] = [float(c) / float(p - 1) for c, p in zip(corrL, pointsPerCorrL)]
    L0 = [1 + math.ceil((float(b) - float(a)) / s) for a, b, s in zip(xMinGlob, xMaxGlob, g["xStep"])]
    if gen_meth == FFT_MPI:
        nov = [0, 0, 0]
    else:
        nov = [1 + math.ceil(float(o) * float(p - 1)) for o, p in zip(overlap, pointsPerCorrL)]
    Np, expanded = [], False
    for a in range(3):
        t = g["topo_shape"][a]
        n = math.ceil(float(L0[a] + (t - 1) * nov[a]) / float(t))
        if t == 2 and n < nov[a]:
            n, expanded = nov[a], True
        elif t >= 3 and n < 2 * nov[a]:
            n, expanded = 2 * nov[a], True
        Np.append(n)
    g["Np_ovlp"] = nov
    g["Np"] = Np
    g["L"] = [Np[a] * g["topo_shape"][a] - (g["topo_shape"][a] - 1) * nov[a] for a in range(3)]
    g["expanded"] = expanded
    return g


def rank_geometry(g, rank: int):
    """Per-rank subdomain of SRC_F/main_ScaRL.f90:455-468."""
    pos = get_topo_pos(rank, g["topo_shape"])
    last = get_topo_pos(g["topo_shape"][0] * g["topo_shape"][1] * g["topo_shape"][2] - 1, g["topo_shape"])
    r = {"topo_pos": pos}
    r["pos_0"] = [(g["Np"][a] - g["Np_ovlp"][a]) * pos[a] + 1 for a in range(3)]
    r["pos_N"] = [r["pos_0"][a] + g["Np"][a] - 1 for a in range(3)]
    r["coord_0"] = [float(r["pos_0"][a] - 1) * g["xStep"][a] + g["xMinGlob"][a] for a in range(3)]
    r["coord_N"] = [float(r["pos_N"][a] - 1) * g["xStep"][a] + g["xMinGlob"][a] for a in range(3)]
    r["xRange"] = [r["coord_N"][a] - r["coord_0"][a] for a in range(3)]
    r["overlap"] = [float(g["Np_ovlp"][a] - 1) * g["xStep"][a] for a in range(3)]
    r["xMaxGlob"] = [float((g["Np"][a] - g["Np_ovlp"][a]) * last[a]) * g["xStep"][a] + g["xMinGlob"][a] +
                     r["xRange"][a] for a in range(3)]
    return r


def seed_start(seedBase: int, rank: int) -> int:
    """SRC_F/main_ScaRL.f90:499-507 for a deterministic seedBase (>= 0)."""
    if seedBase < 0:
        raise ValueError("a negative seedBase asks the reference for a clock-based seed: resolve it first")
    return seedBase + 1 + rank


def blend_layout(g):
    """(gdim, starts, exts) of the sample's subdomains for ``generator.Batch.blend`` (0-based
    starts; every subdomain has the extent Np)."""
    starts = [[(g["Np"][a] - g["Np_ovlp"][a]) * p for p in range(g["topo_shape"][a])] for a in range(3)]
    exts = [[g["Np"][a]] * g["topo_shape"][a] for a in range(3)]
    return tuple(g["L"]), starts, exts


def blend_batch(g, corrL, corrMod, seedBase, margiFirst=A.FOM_NONE, avg=0.0, std_dev=1.0, nu=0.5,
                precision=A.PREC_FP32):
    """One batched blend holding every rank's subdomain of the sample (grid spacing xStep)."""
    from .generator import Batch, FieldSpec
    gdim, starts, exts = blend_layout(g)
    spec = FieldSpec(n=(2, 2, 2), model=corrMod, corrL=tuple(float(c) for c in corrL), dx=tuple(g["xStep"]), nu=nu,
                     margi=margiFirst, avg=avg, std_dev=std_dev, seed=seed_start(seedBase, 0))
    return Batch.blend(gdim, starts, exts, spec, precision)
