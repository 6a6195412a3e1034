"""CPU tests of the HDF5/XDMF output (libscarl_h5.so, scarl_amd.write_output).

The layout is pinned three ways:
* against the HDF5 tools (h5dump -H for extents, types and string padding; h5dump -b for the
  raw dataset bytes), which do not share code with the writer's own reader;
* against .xmf files written by hand from the reference's WRITE statements
  (write_output.f90:455-556, 558-656): tests/golden/xmf;
* against the Fortran F editing rules for the F25.10 records.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "scarl_h5.h")
GOLD = os.path.join(ROOT, "tests", "golden", "xmf")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
needs_h5dump = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^int (scarl_(?:h5|xmf|fortran)_\w+)\(", txt, flags=re.M)))


def test_header_binding_and_exports_agree():
    from scarl_amd import _h5 as H
    assert declared() == sorted(H.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", H.LIB_PATH], capture_output=True, text=True).stdout
    assert set(re.findall(r" T (scarl_\w+)", out)) == set(declared())


@pytest.mark.parametrize("v,text", [
    (0.0, "0.0000000000"),
    (0.1, "0.1000000000"),
    (-1.5, "-1.5000000000"),
    (2.0 / 3.0, "0.6666666667"),
    (-1e-12, "-0.0000000000"),
    (12345678901234.0, "12345678901234.0000000000"),   # 25 characters: fits
    (123456789012345.0, "*" * 25),                      # 26: does not fit
    (float("inf"), "Infinity"),
    (float("-inf"), "-Infinity"),
    (float("nan"), "NaN"),
])
def test_fortran_f25_10_edit(v, text):
    from scarl_amd.write_output import fortran_fixed
    s = fortran_fixed(v, 25, 10)
    assert len(s) == 25
    assert s == " " * (25 - len(text)) + text


def test_xmf_elements_matches_hand_written_golden(tmp_path):
    from scarl_amd import write_output as W
    W.write_XMF_elements("case.h5", [0.5, -1.25, 2.0], [3.5, 0.75, 5.0], [7, 5, 4], "case.xmf", str(tmp_path), ".",
                         "samples")
    got = open(tmp_path / "case.xmf", "rb").read()
    assert got == open(os.path.join(GOLD, "elements_7x5x4.xmf"), "rb").read()


def test_xmf_global_matches_hand_written_golden(tmp_path):
    from scarl_amd import write_output as W
    p = str(tmp_path / "g.xmf")
    W.write_XMF_global(p, ["g_proc_00000.h5", "g_proc_00001.h5"], [[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]],
                       [0.25, 0.5, 1.0], [7, 5, 4])
    assert open(p, "rb").read() == open(os.path.join(GOLD, "global_2procs.xmf"), "rb").read()


def _two_blocks(dtype=np.float64):
    rng = np.random.default_rng(3)
    g = np.asfortranarray(rng.standard_normal((7, 5, 4)).astype(dtype))
    # two subdomains along x with a one-point overlap, as x-fastest views of the global array
    b0 = {"field": g[0:4], "pos_0": [1, 1, 1], "coord_0": [0.0, 0.0, 0.0], "coord_N": [1.5, 2.0, 3.0]}
    b1 = {"field": g[3:7], "pos_0": [4, 1, 1], "coord_0": [1.5, 0.0, 0.0], "coord_N": [3.0, 2.0, 3.0]}
    return g, b0, b1


def _h5dump(args):
    r = subprocess.run([H5DUMP] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


@needs_h5dump
def test_1ds_layout_attributes_and_bytes(tmp_path):
    from scarl_amd import write_output as W
    g, b0, b1 = _two_blocks()
    blocks = [W.Block(**b0), W.Block(**b1)]
    W.write_hdf5_multi_proc_3D_1ds(blocks, [7, 5, 4], "s.h5", "s.xmf", str(tmp_path))
    path = str(tmp_path / "s.h5")
    hdr = _h5dump(["-H", path])
    # Fortran extents (7, 5, 4) are recorded reversed, as h5screate_simple_f records them
    assert re.search(r'DATASET "samples" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( 4, 5, 7 \)', hdr)
    for a in ("pos_0", "pos_N"):
        assert re.search(r'ATTRIBUTE "%s" \{\s*DATATYPE\s+H5T_STD_I32LE\s*DATASPACE\s+SIMPLE \{ \( 3 \)' % a, hdr)
    for a in ("xMin", "xMax"):
        assert re.search(r'ATTRIBUTE "%s" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( 3 \)' % a, hdr)
    # raw bytes in HDF5's C order: z slowest, x fastest
    raw = tmp_path / "raw.bin"
    _h5dump(["-d", "/samples", "-b", "LE", "-o", str(raw), path])
    flat = np.fromfile(raw, "<f8")
    assert np.array_equal(flat.reshape(4, 5, 7), np.ascontiguousarray(g.transpose(2, 1, 0)))
    with W.H5File(path) as f:
        assert np.array_equal(f.read(), g)
        assert f.dataset_info("samples") == ((7, 5, 4), 64)
        assert sorted(f.attr_names("samples")) == ["pos_0", "pos_N", "xMax", "xMin"]
        assert list(f.read_attr("samples", "xMin", "d")) == b0["coord_0"]
        assert list(f.read_attr("samples", "xMax", "d")) == b1["coord_N"]
        assert list(f.read_attr("samples", "pos_0", "i")) == [1, 1, 1]
        assert list(f.read_attr("samples", "pos_N", "i")) == [7, 5, 4]
        # a block read at an offset
        assert np.array_equal(f.read("samples", (2, 1, 1), (3, 2, 2)), g[2:5, 1:3, 1:3])
    # the .xmf of the 1ds file uses rank 0's corner and the last rank's far corner
    txt = open(tmp_path / "s.xmf").read()
    assert 'Dimensions="4 5 7"' in txt and "./s.h5:/samples" in txt


def test_1ds_single_rank_has_no_xmax_attribute(tmp_path):
    from scarl_amd import write_output as W
    g, b0, _ = _two_blocks()
    W.write_hdf5_multi_proc_3D_1ds([W.Block(g, [1, 1, 1], [0, 0, 0], [3, 2, 3])], [7, 5, 4], "one.h5", "one.xmf",
                                   str(tmp_path))
    with W.H5File(str(tmp_path / "one.h5")) as f:
        assert sorted(f.attr_names("samples")) == ["pos_0", "pos_N", "xMin"]


def test_multi_dataset_file_and_single_proc_files(tmp_path):
    from scarl_amd import write_output as W
    g, b0, b1 = _two_blocks()
    blocks = [W.Block(**b0), W.Block(**b1)]
    W.write_hdf5_multi_proc_3D(blocks, "m.h5", str(tmp_path))
    with W.H5File(str(tmp_path / "m.h5")) as f:
        assert np.array_equal(f.read("sample_1_p0"), g[0:4])
        assert np.array_equal(f.read("sample_1_p1"), g[3:7])
        assert sorted(f.attr_names("sample_1_p0")) == ["xMax", "xMin"]          # as written: no pos on rank 0
        assert sorted(f.attr_names("sample_1_p1")) == ["pos_0", "pos_N", "xMax", "xMin"]
        assert list(f.read_attr("sample_1_p1", "pos_N", "i")) == [7, 5, 4]
    W.write_hdf5_single_proc_3D(blocks[1], "p1.h5", "p1.xmf", str(tmp_path), rank=1)
    with W.H5File(str(tmp_path / "p1.h5")) as f:
        assert np.array_equal(f.read("sample_1_p1"), g[3:7])
        assert list(f.read_attr("sample_1_p1", "pos_0", "i")) == [4, 1, 1]
    assert "./p1.h5:/sample_1_p1\n" in open(tmp_path / "p1.xmf").read()


@needs_h5dump
def test_root_attributes_and_times(tmp_path):
    from scarl_amd import write_output as W
    path = str(tmp_path / "r.h5")
    W.H5File(path, "w").close()
    W.write_HDF5_attributes(path, 8, 3, 1, 1, 7, 1, 2, [2, 2, 2], [0.0, 0.0, 0.0], [10.0, 10.0, 10.0],
                            [0.25, 0.25, 0.25], [1.0, 1.0, 1.0], [1.0, 1.0, 1.0], [45, 45, 45], [25, 25, 25],
                            [5, 5, 5])
    labels = ["Read_input", "Topology", "Total_Wall_time"]
    trace = np.array([[0.5, 0.25], [0.125, 1.0], [2.0, 3.0]])
    W.write_time_on_hdf5(labels, trace, "r.h5", str(tmp_path))
    hdr = _h5dump(["-H", path])
    for a in ("nb_procs", "nDim", "Nmc", "method", "seedStart", "corrMod", "margiFirst"):
        assert re.search(r'ATTRIBUTE "%s" \{\s*DATATYPE\s+H5T_STD_I32LE\s*DATASPACE\s+SCALAR' % a, hdr), a
    for a in ("nFields", "Np", "Np_ovlp", "L"):
        assert re.search(r'ATTRIBUTE "%s" \{\s*DATATYPE\s+H5T_STD_I32LE\s*DATASPACE\s+SIMPLE \{ \( 3 \)' % a, hdr), a
    for a in ("xMinGlob", "xMaxGlob", "xStep", "corrL", "overlap"):
        assert re.search(r'ATTRIBUTE "%s" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( 3 \)' % a, hdr), a
    # times(ntimes, nprocs) in Fortran order is recorded as {nprocs, ntimes}
    assert re.search(r'DATASET "times" \{\s*DATATYPE\s+H5T_IEEE_F64LE\s*DATASPACE\s+SIMPLE \{ \( 2, 3 \)', hdr)
    lab = re.search(r'DATASET "time_labels" \{(.*?)DATASPACE\s+SIMPLE \{ \( 3 \)', hdr, re.S)
    assert lab and "STRSIZE 20;" in lab.group(1) and "H5T_STR_SPACEPAD" in lab.group(1)
    with W.H5File(path) as f:
        assert list(f.read_attr("", "seedStart", "i")) == [7]
        assert list(f.read_attr("", "Np_ovlp", "i")) == [5, 5, 5]
        got_l, got_t = f.read_times()
        assert got_l == [s.ljust(20) for s in labels]
        assert np.array_equal(got_t, trace)


def test_fp32_dataset_and_conversion(tmp_path):
    from scarl_amd import write_output as W
    g, b0, b1 = _two_blocks(np.float32)
    W.write_hdf5_multi_proc_3D_1ds([W.Block(**b0), W.Block(**b1)], [7, 5, 4], "f.h5", "f.xmf", str(tmp_path),
                                   precision=32)
    with W.H5File(str(tmp_path / "f.h5")) as f:
        assert f.dataset_info("samples") == ((7, 5, 4), 32)
        assert np.array_equal(f.read(dtype=np.float32), g)
    assert 'Precision="4" Dimensions="4 5 7"' in open(tmp_path / "f.xmf").read()
    # FP32 data into the reference's FP64 dataset: HDF5 converts exactly
    W.write_hdf5_multi_proc_3D_1ds([W.Block(**b0), W.Block(**b1)], [7, 5, 4], "d.h5", "d.xmf", str(tmp_path))
    with W.H5File(str(tmp_path / "d.h5")) as f:
        assert f.dataset_info("samples")[1] == 64
        assert np.array_equal(f.read(), g.astype(np.float64))


def test_invalid_blocks_are_errors(tmp_path):
    from scarl_amd import _h5 as H
    from scarl_amd import write_output as W
    with W.H5File(str(tmp_path / "e.h5"), "w") as f:
        f.create_dataset("samples", [4, 4, 4])
        with pytest.raises(H.H5Error, match="exceeds"):
            f.write_block("samples", np.zeros((3, 4, 4), order="F"), (2, 0, 0))
        with pytest.raises(H.H5Error, match="no dataset"):
            f.write_block("nope", np.zeros((2, 2, 2), order="F"))
        with pytest.raises(ValueError, match="x fastest"):
            f.write_block("samples", np.zeros((2, 2, 2), order="C"))
        with pytest.raises(H.H5Error, match="exists"):
            f.create_dataset("samples", [4, 4, 4])
    with pytest.raises(H.H5Error, match="cannot open"):
        W.H5File(str(tmp_path / "missing.h5"))


def test_constant_field_and_output_names(tmp_path):
    import datetime
    from scarl_amd import write_output as W
    W.write_constant_field([0.0, 0.0, 0.0], [4.0, 2.0, 1.0], 3.5, "c", str(tmp_path))
    with W.H5File(str(tmp_path / "c.h5")) as f:
        assert np.array_equal(f.read(), np.full((2, 2, 2), 3.5))
        assert list(f.read_attr("", "xStep", "d")) == [4.0, 2.0, 1.0]
        assert list(f.read_attr("", "method", "i")) == [-1]
        assert list(f.read_attr("", "L", "i")) == [2, 2, 2]
    when = datetime.datetime(2016, 3, 9, 14, 5, 7)
    assert W.output_names("c", str(tmp_path), True, now=when) == ("c-160309_140507", "c-160309_140507.h5",
                                                                  "c-160309_140507.xmf")
    assert W.output_names("new", str(tmp_path), True, now=when) == ("new", "new.h5", "new.xmf")
    assert W.output_names("new", str(tmp_path), False, rank=12, now=when) == ("new", "new_proc_00012.h5",
                                                                              "new_proc_00012.xmf")
