"""The Chombo HDF5 reader (mfm_chombo_*, mfm_h5_*), on the CPU.

What fixes it, independently of its code:
  - a file written by the HDF5 library itself (scipy's test file
    testhdf5_7.4_GLNX86.mat, a MATLAB 7.3 file: 512-byte user block, version 0
    superblock, symbol-table groups) reads back with the contents scipy's
    tests give it: one variable "testdouble" = pi/4 * arange(9) as a 9 x 1
    double array with the attribute MATLAB_class = "double";
  - Chombo-layout files written by tests/chombo_h5.py (levels, boxes,
    offsets, FAB values, ghosts, component names, with and without a user
    block) read back exactly;
  - damaged or foreign files give an error code, never a crash.
"""
import os

import numpy as np
import pytest

import mfmhd
from chombo_h5 import write_chombo

try:
    import scipy
    MAT = os.path.join(os.path.dirname(scipy.__file__), "io", "matlab", "tests", "data",
                       "testhdf5_7.4_GLNX86.mat")
except ImportError:
    MAT = ""


@pytest.mark.skipif(not os.path.exists(MAT), reason="scipy's HDF5 test file is absent")
def test_file_written_by_the_hdf5_library():
    f = mfmhd.H5File(MAT)
    assert f.members("/") == ["testdouble"]
    shape, cls, size = f.dataset_info("/testdouble")
    assert shape == (9, 1) and cls == mfmhd.H5File.CLASS_FLOAT and size == 8
    cls, raw = f.attribute("/testdouble", "MATLAB_class")
    assert cls == mfmhd.H5File.CLASS_STRING and raw.rstrip(b"\0") == b"double"
    a = f.read("/testdouble", np.float64)
    assert np.allclose(a[:, 0], np.pi / 4 * np.arange(9), rtol=1e-15, atol=0)
    f.close()


def _levels(rng, nc, ghost):
    levels = []
    for l, (boxes, dx) in enumerate([
            ([((0, 0, 0), (7, 7, 7)), ((8, 0, 0), (15, 7, 3))], 0.1),
            ([((4, 4, 4), (11, 9, 7)), ((12, 4, 4), (15, 7, 5)), ((0, 0, 0), (1, 1, 1))], 0.05)]):
        fabs = []
        for lo, hi in boxes:
            n = [hi[d] - lo[d] + 1 + 2 * ghost[d] for d in range(3)]
            fabs.append(rng.standard_normal((nc, n[2], n[1], n[0])))
        levels.append({"dx": dx, "dt": 0.01 / (l + 1), "ref_ratio": 2,
                       "domain": ((0, 0, 0), (15 * (l + 1) + l, 7 * (l + 1) + l, 7 * (l + 1) + l)),
                       "boxes": boxes, "fabs": fabs})
    return levels


@pytest.mark.parametrize("ghost,userblock", [((0, 0, 0), 0), ((1, 1, 1), 0), ((2, 1, 0), 512)])
def test_chombo_round_trip(tmp_path, ghost, userblock):
    rng = np.random.default_rng(5)
    nc = 23
    names = ["comp-%02d" % c for c in range(nc)]
    levels = _levels(rng, nc, ghost)
    path = str(tmp_path / "chk.3d.hdf5")
    write_chombo(path, levels, names, time=12.5, ghost=ghost, userblock=userblock)
    f = mfmhd.ChomboFile(path)
    assert (f.num_levels, f.num_components, f.space_dim, f.time) == (2, nc, 3, 12.5)
    assert f.component_names == names
    for l, lev in enumerate(levels):
        info = f.level(l)
        assert info["nboxes"] == len(lev["boxes"])
        assert (info["dx"], info["dt"], info["ref_ratio"]) == (lev["dx"], lev["dt"], 2)
        assert info["prob_domain"] == lev["domain"]
        assert info["ghost"] == ghost
        assert f.boxes(l) == lev["boxes"]
        for b, fab in enumerate(lev["fabs"]):
            a, glo = f.read_box(l, b)
            assert glo == tuple(lev["boxes"][b][0][d] - ghost[d] for d in range(3))
            assert np.array_equal(a, fab)
    f.close()


def test_generic_layer_on_a_chombo_file(tmp_path):
    rng = np.random.default_rng(1)
    levels = _levels(rng, 2, (0, 0, 0))
    path = str(tmp_path / "plt.hdf5")
    write_chombo(path, levels, ["density", "energy"])
    h = mfmhd.H5File(path)
    assert sorted(h.members("/")) == ["Chombo_global", "level_0", "level_1"]
    assert sorted(h.members("/level_1")) == ["boxes", "data:datatype=0", "data:offsets=0",
                                             "data_attributes"]
    off = h.read("/level_1/data:offsets=0", np.int64)
    sizes = [fab.size for fab in levels[1]["fabs"]]
    assert list(off) == [0] + list(np.cumsum(sizes))
    cls, raw = h.attribute("/level_0/data_attributes", "objectType")
    assert raw.rstrip(b"\0") == b"FArrayBox"
    h.close()


def test_errors_not_crashes(tmp_path):
    rng = np.random.default_rng(2)
    levels = _levels(rng, 3, (1, 1, 1))
    path = str(tmp_path / "chk.hdf5")
    write_chombo(path, levels, ["a", "b", "c"], ghost=(1, 1, 1))
    f = mfmhd.ChomboFile(path)
    with pytest.raises(mfmhd.MfmError) as e:
        f.level(5)
    assert e.value.code == mfmhd.ERR_ARG
    lib = mfmhd.load()
    buf = np.zeros(10)
    rc = lib.mfm_chombo_read_box(f.h, 0, 0, buf.ctypes.data_as(mfmhd.C.POINTER(mfmhd.C.c_double)),
                                 10)
    assert rc == mfmhd.ERR_SHAPE
    f.close()
    # not an HDF5 file
    junk = tmp_path / "junk.hdf5"
    junk.write_bytes(b"not hdf5" * 100)
    with pytest.raises(mfmhd.MfmError):
        mfmhd.ChomboFile(str(junk))
    # truncated files: every prefix either opens and then fails cleanly, or
    # is refused
    data = open(path, "rb").read()
    for cut in range(64, len(data), max(1, len(data) // 40)):
        p = tmp_path / ("cut%d.hdf5" % cut)
        p.write_bytes(data[:cut])
        try:
            g = mfmhd.ChomboFile(str(p))
        except mfmhd.MfmError:
            continue
        try:
            for l in range(g.num_levels):
                for b in range(g.level(l)["nboxes"]):
                    g.read_box(l, b)
        except mfmhd.MfmError:
            pass
        g.close()
