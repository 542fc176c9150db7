"""ANA f0 cube I/O (snapi_ana_*) against files the reference's ana_fzwrite
wrote in write_file's layouts (tests/golden/ana_ref_3d.f0, ana_ref_4d.f0):
the values read back in the reference's index order, the files this build
writes byte for byte equal to the reference's, every data type round trip,
a big-endian file, a long header over several blocks, and the errors for a
file that is not ANA and for compressed data.  Host code only: no GPU."""
import os
import struct

import numpy as np
import pytest

import _golden
from snapi_amd import cube
from snapi_amd._native import SnapiError

REF3 = os.path.join(_golden.HERE, "ana_ref_3d.f0")
REF4 = os.path.join(_golden.HERE, "ana_ref_4d.f0")


def test_reads_reference_files():
    g = _golden.load("ana.npz")
    t3, d3, c3 = cube.ana_info(REF3)
    assert (t3, d3, c3) == (3, [3, 4, 5], False)
    a3 = cube.ana_read(REF3)
    # write_file(3-D) stores q[x3][x2][x1] = data[x1][x2][x3] as float
    assert a3.dtype == np.float32 and np.array_equal(a3, g["a3"].transpose(2, 1, 0).astype(np.float32))
    t4, d4, _ = cube.ana_info(REF4)
    assert (t4, d4) == (4, [6, 4, 3, 2])
    assert np.array_equal(cube.ana_read(REF4), g["a4"])


def test_writes_reference_bytes(tmp_path):
    g = _golden.load("ana.npz")
    p3, p4 = str(tmp_path / "w3.f0"), str(tmp_path / "w4.f0")
    cube.ana_write(p3, g["a3"].transpose(2, 1, 0).astype(np.float32))
    cube.ana_write(p4, g["a4"])
    assert open(p3, "rb").read() == open(REF3, "rb").read()
    assert open(p4, "rb").read() == open(REF4, "rb").read()


def test_round_trip_every_type_and_long_header(tmp_path):
    rng = np.random.default_rng(2)
    for dt in (np.int8, np.int16, np.int32, np.float32, np.float64, np.int64):
        a = (rng.standard_normal((2, 3, 7)) * 100).astype(dt)
        p = str(tmp_path / ("t_%s.f0" % np.dtype(dt).name))
        cube.ana_write(p, a, header="x" * 600)  # three header blocks
        assert os.path.getsize(p) == 3 * 512 + a.nbytes
        b = cube.ana_read(p)
        assert b.dtype == a.dtype and np.array_equal(a, b)


def test_big_endian_file(tmp_path):
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 1.5
    head = bytearray(512)
    struct.pack_into("<I", head, 0, 0x5555aaaa)
    head[4] = 128  # subf: big-endian data
    head[6], head[7], head[8] = 1, 4, 3  # nhb, datyp FLOAT64, ndim
    struct.pack_into("<3i", head, 192, 4, 3, 2)
    p = str(tmp_path / "be.f0")
    open(p, "wb").write(bytes(head) + a.astype(">f8").tobytes())
    assert np.array_equal(cube.ana_read(p), a)


def test_errors(tmp_path):
    p = str(tmp_path / "junk.f0")
    open(p, "wb").write(b"\0" * 600)
    with pytest.raises(SnapiError):
        cube.ana_info(p)
    head = bytearray(open(REF3, "rb").read())
    head[4] |= 1  # compressed
    q = str(tmp_path / "comp.f0")
    open(q, "wb").write(bytes(head))
    assert cube.ana_info(q)[2]
    with pytest.raises(SnapiError):
        cube.ana_read(q)
