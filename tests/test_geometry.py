"""Sample geometry (SRC_F/main_ScaRL.f90:100-111, 157-176, 455-468, 499-507, 747-782) on the
reference's own input decks, against values worked out by hand from those lines."""
import pytest

from scarl_amd import geometry as G


def test_deck_3_test_sample_1_one_rank():
    # test/3_TEST/input_ScaRL.txt sample 1: x in [0, 10], corrL 1, overlap 1, 5 points per corrL
    g = G.sample_geometry((0, 0, 0), (10, 10, 10), (1, 1, 1), (1, 1, 1), (5, 5, 5), (1, 1, 1))
    assert g["xStep"] == [0.25] * 3              # 1 / (5 - 1)
    assert g["Np_ovlp"] == [5] * 3               # 1 + ceiling(1 * 4)
    assert g["Np"] == [41] * 3 and g["L"] == [41] * 3  # 1 + ceiling(10 / 0.25)
    r = G.rank_geometry(g, 0)
    assert r["pos_0"] == [1] * 3 and r["pos_N"] == [41] * 3
    assert r["xRange"] == [10.0] * 3 and r["xMaxGlob"] == [10.0] * 3 and r["overlap"] == [1.0] * 3


def test_deck_3_test_sample_1_eight_ranks():
    g = G.sample_geometry((0, 0, 0), (10, 10, 10), (1, 1, 1), (1, 1, 1), (5, 5, 5), (2, 2, 2))
    assert g["Np"] == [23] * 3                   # ceiling((41 + 5) / 2)
    assert g["L"] == [41] * 3                    # 23 * 2 - 5
    r = G.rank_geometry(g, 5)                    # rank 5 -> topo_pos (1, 0, 1), x fastest
    assert r["topo_pos"] == (1, 0, 1)
    assert r["pos_0"] == [19, 1, 19] and r["pos_N"] == [41, 23, 41]
    assert r["coord_0"] == [4.5, 0.0, 4.5] and r["xRange"] == [5.5] * 3
    assert r["xMaxGlob"] == [10.0] * 3
    assert G.get_topo_rank((2, 2, 2), (1, 0, 1)) == 5
    gdim, starts, exts = G.blend_layout(g)
    assert gdim == (41, 41, 41) and starts[0] == [0, 18] and exts[0] == [23, 23]


def test_deck_sample_2_three_ranks():
    # sample_2: x in [0, 20], overlap 5: L = 1 + ceiling(20 / 0.25) = 81, Np_ovlp = 1 + ceiling(5 * 4) = 21;
    # three ranks on x: ceiling((81 + 2 * 21) / 3) = 41 < 2 * 21, so the domain expands to Np = 42, L = 84
    g = G.sample_geometry((0, 0, 0), (20, 20, 20), (1, 1, 1), (5, 5, 5), (5, 5, 5), (3, 1, 1))
    assert g["Np_ovlp"] == [21] * 3
    assert g["Np"] == [42, 81, 81] and g["L"] == [84, 81, 81] and g["expanded"]
    r = G.rank_geometry(g, 2)
    assert r["pos_0"] == [43, 1, 1] and r["xMaxGlob"] == [20.75, 20.0, 20.0]


def test_domain_expansion():
    # x in [0, 2], overlap 5, 5 points per corrL: L = 9, Np_ovlp = 21
    g2 = G.sample_geometry((0, 0, 0), (2, 2, 2), (1, 1, 1), (5, 5, 5), (5, 5, 5), (2, 1, 1))
    assert g2["Np"][0] == 21 and g2["L"][0] == 21 and g2["expanded"]   # ceiling(30/2) = 15 < 21
    g3 = G.sample_geometry((0, 0, 0), (2, 2, 2), (1, 1, 1), (5, 5, 5), (5, 5, 5), (3, 1, 1))
    assert g3["Np"][0] == 42 and g3["L"][0] == 84                      # ceiling(51/3) = 17 < 42
    gm = G.sample_geometry((0, 0, 0), (2, 2, 2), (1, 1, 1), (5, 5, 5), (5, 5, 5), (2, 1, 1), gen_meth=G.FFT_MPI)
    assert gm["Np_ovlp"] == [0, 0, 0] and gm["Np"][0] == 5 and not gm["expanded"]


def test_seed_start():
    assert G.seed_start(1, 0) == 2 and G.seed_start(20, 3) == 24
    with pytest.raises(ValueError):
        G.seed_start(-1, 0)
