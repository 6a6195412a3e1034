"""CPU tests of the deck parser (scarl_amd.read_input) on the reference's own decks
(tests/golden/decks, copied from test/ and bash_ScaRL/), and of the topology choice.

Expected values are read off the decks by eye; the template decks are filled in the way the
reference's scripts fill them (sed of #SIZE# / #MAX_X, bash_ScaRL/MAX_Test/speedUp.sh:74-77,
bash_ScaRL/bash_Scarl.sh:80-82).
"""
import os
import shutil

import numpy as np

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECKS = os.path.join(ROOT, "tests", "golden", "decks")


def _deck(name, tmp_path, subst=None):
    d = tmp_path / name
    d.mkdir(parents=True)
    for f in os.listdir(os.path.join(DECKS, name)):   # plain copies: writable whatever the source's mode
        shutil.copyfile(os.path.join(DECKS, name, f), d / f)
    if subst:
        p = d / "input_ScaRL.txt"
        txt = p.read_text()
        for k, v in subst.items():
            txt = txt.replace(k, v)
        p.write_text(txt)
    return d


def test_3_test_config_and_sample(tmp_path):
    from scarl_amd import read_input as R
    d = _deck("3_TEST", tmp_path)
    c = R.read_config_ScaRL(str(d / "config_ScaRL.txt"))
    assert (c.gen_meth, c.c_info_file, c.d_sample, c.one_file, c.one_hdf_ds) == (1, True, False, True, True)
    assert c.input_file == "./input_ScaRL.txt"
    deck = R.read_input_ScaRL(c.input_file, str(d))
    # "1  #Number of Samples": the second sample block of the deck is never read
    assert deck.output_folder == "SAMPLES" and deck.nSamples == 1
    (s1,) = deck.samples
    # the commented-out xMax and corrL alternatives are skipped by getLine
    assert s1.output_name == "Sample_Test_1"
    assert s1.xMinGlob == [0.0, 0.0, 0.0] and s1.xMaxGlob == [10.0, 10.0, 10.0]
    assert s1.corrL == [1.0, 1.0, 1.0] and s1.overlap == [1.0, 1.0, 1.0] and s1.pointsPerCorrL == [5, 5, 5]
    assert (s1.corrMod, s1.margiFirst, s1.avg, s1.std_dev, s1.seedBase) == (1, 2, 1.0, 1.0, 1)
    assert R.verify_inputs(s1, c.gen_meth) == []
    # with nSamples = 2 the second block is read too
    p = d / "input_ScaRL.txt"
    p.write_text(p.read_text().replace("1               #Number of Samples", "2               #Number of Samples"))
    s2 = R.read_input_ScaRL(c.input_file, str(d)).samples[1]
    assert s2.output_name == "sample_2" and s2.xMaxGlob == [20.0, 20.0, 20.0] and s2.overlap == [5.0, 5.0, 5.0]
    assert s2.seedBase == 20


def test_fusion_config_with_inline_notes(tmp_path):
    from scarl_amd import read_input as R
    d = _deck("testFusion", tmp_path)
    c = R.read_config_ScaRL(str(d / "config_ScaRL.txt"))
    assert (c.gen_meth, c.c_info_file, c.d_sample, c.one_file, c.one_hdf_ds) == (2, True, False, True, True)
    deck = R.read_input_ScaRL(c.input_file, str(d))
    (s,) = deck.samples
    assert s.output_name == "BASIC_Sample" and s.xMaxGlob == [30.0, 30.0, 30.0] and s.margiFirst == 1


def test_points_decks(tmp_path):
    from scarl_amd import read_input as R
    for name, ppc, folder in (("5p_3TEST", [5, 5, 5], "5p_SAMPLES"), ("10p_3TEST", [10, 10, 10], "10p_SAMPLES")):
        d = _deck(name, tmp_path)
        deck = R.read_input_ScaRL("./input_ScaRL.txt", str(d))
        assert deck.output_folder == folder
        assert deck.samples[0].pointsPerCorrL == ppc


def test_template_decks_after_substitution(tmp_path):
    from scarl_amd import read_input as R
    d = _deck("GENERAL_BASH", tmp_path, {"#SIZE#": "40"})
    c = R.read_config_ScaRL(str(d / "config_ScaRL.txt"))
    assert (c.d_sample, c.one_file, c.one_hdf_ds) == (True, False, False)
    assert R.read_extra_config_ScaRL(str(d / "extra_config_ScaRL.txt")) == (True, [2, 2, 2])
    (s,) = R.read_input_ScaRL(c.input_file, str(d)).samples
    assert s.xMaxGlob == [40.0, 40.0, 40.0] and s.avg == 0.0 and s.margiFirst == 1
    d2 = _deck("MAX_Test", tmp_path, {"#SIZE#": "115"})
    (s2,) = R.read_input_ScaRL("./input_ScaRL.txt", str(d2)).samples
    assert s2.xMaxGlob == [115.0, 115.0, 115.0] and s2.output_name == "BIG_Sample"
    d3 = _deck("src_folder", tmp_path, {"#MAX_X": "60", "#MAX_Y": "50", "#MAX_Z": "40"})
    (s3,) = R.read_input_ScaRL("./input_ScaRL.txt", str(d3)).samples
    # comment-first layout with list-directed E-format numbers
    assert s3.output_name == "Sample_1" and s3.xMaxGlob == [60.0, 50.0, 40.0]
    assert s3.corrL == [1.0, 1.0, 1.0] and s3.overlap == [5.0, 5.0, 5.0] and s3.seedBase == -1


def test_unfilled_template_fails_like_the_reference_read(tmp_path):
    """Unsubstituted, the xMax line starts with '#' and is skipped, so every later read is off
    by one line; the three integers of pointsPerCorrL then meet "1  #Correlation Model"."""
    from scarl_amd import read_input as R
    d = _deck("GENERAL_BASH", tmp_path)
    with pytest.raises(R.DeckError, match="bad integer"):
        R.read_input_ScaRL("./input_ScaRL.txt", str(d))


def test_missing_files():
    from scarl_amd import read_input as R
    c = R.read_config_ScaRL("/nonexistent/config_ScaRL.txt")
    assert (c.gen_meth, c.c_info_file, c.d_sample, c.one_file, c.one_hdf_ds, c.input_file, c.found) == \
        (1, False, False, True, True, "./input_ScaRL.txt", False)
    assert R.read_extra_config_ScaRL("/nonexistent/x.txt") == (False, [-1, -1, -1])
    with pytest.raises(FileNotFoundError):
        R.read_input_ScaRL("./input_ScaRL.txt", "/nonexistent")


@pytest.mark.parametrize("buf,kinds,want", [
    ("3*1d0  # three", "ddd", [1.0, 1.0, 1.0]),
    ("1, 2 ,3", "iii", [1, 2, 3]),
    ("1.0+3 2.5D-1 .5", "ddd", [1000.0, 0.25, 0.5]),
    ("'it''s' rest", "a", ["it's"]),
    ('"a b,c/d"', "a", ["a b,c/d"]),
    ("plain/slash", "a", ["plain"]),
    ("0.000000000000000E+000  1.00000000000000", "dd", [0.0, 1.0]),
    ("-7", "i", [-7]),
])
def test_list_directed_read(buf, kinds, want):
    from scarl_amd import read_input as R
    assert R.list_read(buf, kinds) == want


@pytest.mark.parametrize("buf,kinds,msg", [
    ("5.0", "i", "bad integer"),
    ("1#note", "i", "bad integer"),
    ('"x"', "d", "character constant"),
    ("1 2", "iii", "end of record"),
    ("1 / 2 3", "iii", "end of record"),
    ("1,,3", "iii", "null value"),
    ("eof_gl", "i", "end of file"),
    ("", "i", "end of record"),
])
def test_list_directed_read_errors(buf, kinds, msg):
    from scarl_amd import read_input as R
    with pytest.raises(R.DeckError, match=msg):
        R.list_read(buf, kinds)


def test_getline_skips_left_adjusted_comments(tmp_path):
    from scarl_amd import read_input as R
    p = tmp_path / "d.txt"
    p.write_text("# a\n   # b\n  7 # c\n\n")
    src = R.LineSource(str(p))
    assert R.getLine(src) == "7 # c"
    assert R.getLine(src) == ""          # a blank line is data, not a comment
    assert R.getLine(src) == "eof_gl"


def test_verify_inputs_stops_where_the_reference_stops():
    from scarl_amd import read_input as R
    base = dict(output_name="s", xMinGlob=[0.0] * 3, xMaxGlob=[1.0] * 3, corrL=[1.0] * 3, overlap=[1.0] * 3,
                pointsPerCorrL=[5] * 3, corrMod=1, margiFirst=2, avg=1.0, std_dev=1.0, seedBase=1)
    for change, msg in ((dict(xMaxGlob=[1.0, 0.0, 1.0]), "xMinGlob >= xMaxGlob"), (dict(corrL=[1.0, -1.0, 1.0]), "corrL"),
                        (dict(corrMod=3), "Correlation model"), (dict(margiFirst=0), "marginal"),
                        (dict(avg=0.0), "lognormal"), (dict(std_dev=-1.0), "Standard deviation"),
                        (dict(pointsPerCorrL=[5, 1, 5]), "PointsPerCorrL")):
        with pytest.raises(R.DeckError, match=msg):
            R.verify_inputs(R.Sample(**{**base, **change}), 1)
    with pytest.raises(R.DeckError, match="Generation Method"):
        R.verify_inputs(R.Sample(**base), 3)
    assert R.verify_inputs(R.Sample(**{**base, "overlap": [-1.0, 1.0, 1.0]}), 1) == \
        ["WARNING: overlap < 0d0, the mesh will be discontinuous"]


def test_find_factors_and_topology_choice():
    from scarl_amd import geometry as G
    assert G.find_factors(1) == [] and G.find_factors(12) == [2, 2, 3] and G.find_factors(97) == [97]
    # L = 51^3, 5-point overlaps, 16 ranks (by hand): the block search stops at (3, 3, 2) = 18 >= 16;
    # np = 12..16 give (3,4,1) 0.316, (13,1,1) 0.241, (7,2,1) 0.304, (5,3,1) 0.326, (4,4,1) 0.333.
    assert G.decide_topo_shape(16, [51, 51, 51], [5, 5, 5]) == [4, 4, 1]
    assert G.decide_topo_shape(8, [51, 51, 51], [5, 5, 5]) == [2, 2, 2]
    assert G.decide_topo_shape(1, [51, 51, 51], [5, 5, 5]) == [1, 1, 1]
    # a long x axis takes the blocks: 4 ranks on 100 x 40 x 20
    assert G.decide_topo_shape(4, [100, 40, 20], [5, 5, 5]) == [4, 1, 1]
    # FFT_MPI: FFTW slabs of ceiling(51 / 16) = 4 planes leave 13 ranks with planes
    assert G.decide_topo_shape(16, [51, 51, 51], [0, 0, 0], gen_meth=G.FFT_MPI) == [1, 1, 13]
    assert G.decide_topo_shape(8, [51, 51, 51], [0, 0, 0], gen_meth=G.FFT_MPI) == [1, 1, 8]


def test_stat_inputs(tmp_path):
    """The Stat program's stdin decks: testFusion's holds the dataset name; the pointsTest one
    stops after the path, so the dataset read meets the end of the input, as the reference's
    READ would."""
    import io
    from scarl_amd import main_Stat, read_input as R
    lines = open(os.path.join(DECKS, "testFusion", "input_stat.txt")).read().splitlines()
    assert [R.list_read(lines[0], "i")[0], R.list_read(lines[1], "i")[0], R.list_read(lines[2], "i")[0]] == [1, 1, 0]
    assert R.list_read(lines[3], "a")[0] == "/workdir/carvalho/RF_Test/BASIC_TEST_ScaRL/SAMPLES/BASIC_Sample.h5"
    assert R.list_read(lines[4], "a")[0] == "samples"
    with pytest.raises(R.DeckError, match="end of input"):
        main_Stat.main(open(os.path.join(DECKS, "5p_3TEST", "input_stat")))
    with pytest.raises(R.DeckError, match="bad integer"):
        main_Stat.main(io.StringIO("one\n"))


def test_constant_field_sample_needs_no_gpu(tmp_path):
    """std_dev = 0 (main_ScaRL.f90:145-155): a 2x2x2 constant field through
    write_constant_field, before any generation."""
    from scarl_amd import main_ScaRL, write_output as W
    d = _deck("3_TEST", tmp_path)
    p = d / "input_ScaRL.txt"
    p.write_text(p.read_text().replace("1d0             #Standard Deviation", "0d0             #Standard Deviation"))
    (res,) = main_ScaRL.main(str(d), nb_procs=4)
    assert res == {"output_name": "Sample_Test_1", "constant": True}
    with W.H5File(str(d / "SAMPLES" / "Sample_Test_1.h5")) as f:
        assert np.array_equal(f.read(), np.full((2, 2, 2), 1.0))
        assert f.read_attr("", "xStep", "d").tolist() == [10.0, 10.0, 10.0]
