"""On-disk formats of the reference's programs through the C ABI (host only, no device).

  * SEPlib headers and binaries (sepRead / sepWrite, src/include/IO.hpp:14-176): round trip
    of a 3-axis model with origins, samplings and labels; the binary is xdr_float
    (big-endian) as SEPlib's srite writes by default; a header in the layout the
    reference's own Python writer produces (src/python/seppy.py:224-312: "sets next:"
    prose, several pairs per line, quoted in=) is read; a later definition of a key
    overrides an earlier one; native and 8-byte data convert; a binary of the wrong size
    is an error.
  * The plain binary descriptor (binRead / binWrite, IO.hpp:180-430): round trip, native
    byte order, '#' comments.
  * Trailing unit axes are dropped unless ndim0 asks for them (IO.hpp:121-138).
  * readCoord (src/cpp/param.cpp:4-77) on the reference's own example geometry file
    (examples/par/srcoord.txt, kept as tests/golden/srcoord.txt): 2 shots of 5 receivers.
  * The regular acquisition (param.cpp:221-244) in single precision, and the inputs of the
    reference's published FWI example (examples/elastic_fwi.ipynb cells 3-8) written as
    its notebook writes them and read back as FWI2D.x reads them.
"""
import os

import numpy as np
import pytest

from fwi2d_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fwi2d_amd", "libfwi2d_hip.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


@pytest.fixture(scope="module")
def io():
    from fwi2d_amd import lib
    return lib


def test_sep_round_trip(io, tmp_path):
    rng = np.random.default_rng(0)
    m = rng.standard_normal((3, 17, 11)).astype(np.float32)
    path = str(tmp_path / "model.H")
    io.write(path, m, o=[0.0, 0.5, 0.0], d=[0.01, 0.02, 1.0], label=["z", "x", "par"])
    a, h = io.read(path)
    assert a.shape == (3, 17, 11) and np.array_equal(a, m)
    assert h["n"] == [11, 17, 3] and h["o"] == [0.0, 0.5, 0.0] and h["d"] == [0.01, 0.02, 1.0]
    assert h["label"] == ["z", "x", "par"] and h["esize"] == 4 and h["xdr"] == 1
    assert h["in_"] == str(tmp_path / "model.H@")
    raw = np.fromfile(h["in_"], dtype=">f4")
    assert np.array_equal(raw, m.ravel())


def test_binary_round_trip(io, tmp_path):
    rng = np.random.default_rng(1)
    d = rng.standard_normal((2, 5, 40)).astype(np.float32)
    path = str(tmp_path / "data.txt")
    (tmp_path / "bin").mkdir()
    io.write(path, d, o=[0.0, 0.0, 0.0], d=[0.004, 1.0, 1.0], format=io.FORMAT_BINARY, datapath=str(tmp_path / "bin"))
    text = open(path).read()
    assert "n1=40\n" in text and "esize=4\n" in text and "in=" + str(tmp_path / "bin" / "data.txt@") in text
    with open(path, "a") as f:
        f.write("# n1=7 is a comment\n")
    a, h = io.read(path, format=io.FORMAT_BINARY)
    assert np.array_equal(a, d) and h["n"] == [40, 5, 2] and h["d"][0] == pytest.approx(0.004)
    assert np.array_equal(np.fromfile(h["in_"], dtype="<f4"), d.ravel())


def seppy_layout(path, binpath, arr_c, o, d, form="xdr"):
    """A header and binary in the layout of the reference's Python writer: the array in C
    order [nN]..[n1], written with axis 1 fastest as 4-byte floats."""
    n = arr_c.shape[::-1]
    with open(path, "w") as f:
        f.write("\nwrite_file() on host by user\n")
        f.write('\t\tsets next: in="%s"\n' % binpath)
        for k in range(len(n)):
            f.write("\t\tn%d=%d o%d=%f d%d=%.12f\n" % (k + 1, n[k], k + 1, o[k], k + 1, d[k]))
        f.write('\t\tdata_format="%s" esize=4\n' % ("xdr_float" if form == "xdr" else "native_float"))
    arr_c.astype(">f4" if form == "xdr" else "<f4").tofile(binpath)


def test_reads_the_reference_python_writer_layout(io, tmp_path):
    rng = np.random.default_rng(2)
    m = rng.standard_normal((3, 9, 7)).astype(np.float32)
    for form in ("xdr", "native"):
        path, binp = str(tmp_path / ("m_%s.H" % form)), str(tmp_path / ("m_%s.H@" % form))
        seppy_layout(path, binp, m, [0.0, 0.0, 0.0], [0.01, 0.01, 1.0], form)
        a, h = io.read(path)
        assert np.array_equal(a, m) and h["xdr"] == (form == "xdr") and h["d"][:2] == [0.01, 0.01]
    # a later definition overrides an earlier one (header history)
    with open(path, "a") as f:
        f.write("\t\tn3=1 n2=27\n")
    a, h = io.read(path)
    assert h["n"] == [7, 27] and a.shape == (27, 7) and np.array_equal(a.ravel(), m.ravel())
    # a relative in= is found next to the header
    rel_h = str(tmp_path / "rel.H")
    with open(rel_h, "w") as f:
        f.write('n1=7 n2=27 esize=4 in="m_native.H@"\n')
    a, _ = io.read(rel_h)
    assert np.array_equal(a.ravel(), m.ravel())


def test_esize8_size_check_and_axes(io, tmp_path):
    v = np.linspace(-1, 1, 24)
    binp = str(tmp_path / "v.H@")
    v.astype("<f8").tofile(binp)
    path = str(tmp_path / "v.H")
    with open(path, "w") as f:
        f.write('n1=4 n2=6 n3=1 n4=1 esize=8 in="%s"\n' % binp)
    a, h = io.read(path)
    assert a.shape == (6, 4) and np.array_equal(a.ravel(), v.astype(np.float32)) and h["esize"] == 8
    a3, h3 = io.read(path, ndim0=3)
    assert h3["n"] == [4, 6, 1] and a3.shape == (1, 6, 4)
    with open(path, "w") as f:
        f.write('n1=5 n2=6 esize=8 in="%s"\n' % binp)
    with pytest.raises(io.Fwi2dError, match="contains 192 bytes"):
        io.read(path)
    with open(path, "w") as f:
        f.write('n1=4 esize=2 in="%s"\n' % binp)
    with pytest.raises(io.Fwi2dError, match="esize"):
        io.read(path)
    with pytest.raises(io.Fwi2dError, match="could not be opened"):
        io.read(str(tmp_path / "missing.H"))


def test_read_coord_reference_example(io):
    sxz, rxz, dips = io.read_coord(os.path.join(ROOT, "tests", "golden", "srcoord.txt"))
    assert np.array_equal(sxz, [[2.5, 3.5], [1.5, 4.0]])
    x = [1.5, 2.0, 2.5, 3.0, 3.5]
    for r, dp in zip(rxz, dips):
        assert np.array_equal(r, np.stack([x, np.zeros(5)], 1)) and np.array_equal(dp, np.zeros(5))


def test_read_coord_rules(io, tmp_path):
    path = str(tmp_path / "geom.txt")
    with open(path, "w") as f:
        f.write("# header\n3 1.0 0.1 0.5 0.0 0.25\n3.7 9 9 0.6 0.0\n\n4 2.0 0.2 0.7 0.1\n3 1.0 0.1 0.8 0.0 0\n")
    sxz, rxz, dips = io.read_coord(path)
    # id 3.7 reads as 3 (same shot); a new id starts a shot even if it repeats an earlier one
    assert np.allclose(sxz, np.array([[1.0, 0.1], [2.0, 0.2], [1.0, 0.1]], np.float32))
    assert [len(r) for r in rxz] == [2, 1, 1]
    assert np.allclose(rxz[0], np.array([[0.5, 0.0], [0.6, 0.0]], np.float32)) and dips[0][0] == np.float32(0.25)
    assert dips[1][0] == 0.0


def test_regular_geometry_and_the_published_example_inputs(io, tmp_path):
    """The notebook's files, written as its Python writer does, read as FWI2D.x reads them
    (source=wavelet.H model=model0.H; shot parameters ns=10 sx0=2 sz0=0.1 sxinc=0.4 nr=401
    rx0=2 rz0=0.1 rxinc=0.01)."""
    prob, true, start, allsrc = W.notebook_elastic_fwi()
    sxz, rxz = io.regular_geometry(10, 2.0, 0.1, 0.4, 0.0, 401, 2.0, 0.1, 0.01, 0.0)
    assert np.array_equal(sxz, prob["sxz"])
    assert all(np.array_equal(rxz[s, :, :2], prob["rxz"][s]) for s in range(10))
    # single precision: x = x0 + float(r) * inc rounds differently from double arithmetic
    assert np.array_equal(rxz[0, :, 0], (np.float32(2) + np.arange(401, dtype=np.float32) * np.float32(0.01)))
    nt, dt = 1001, 0.004
    wav = allsrc[1, 0] / np.float32(np.sin(np.float32(1.5708)))
    seppy_layout(str(tmp_path / "wavelet.H"), str(tmp_path / "wavelet.H@"), wav.astype(np.float64), [0.0], [dt])
    seppy_layout(str(tmp_path / "model0.H"), str(tmp_path / "model0.H@"), start.astype(np.float64), [0.0] * 3,
                 [0.01, 0.01, 1.0])
    w, hw = io.read(str(tmp_path / "wavelet.H"))
    assert hw["n"] == [nt] and hw["d"] == [dt] and np.allclose(w, wav, rtol=0, atol=1e-6)
    m0, hm = io.read(str(tmp_path / "model0.H"))
    assert hm["n"] == [201, 801, 3] and m0.shape == (3, 801, 201) and np.array_equal(m0, start)
