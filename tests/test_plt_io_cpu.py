"""The tests' Tecplot reader and writer (tests/plt_io.py) agree with each other on a two-zone file, byte for byte
in the header fields the reference writes (FieldIO.h:108-218)."""
import struct

import numpy as np

import plt_io


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    names = ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"]
    zones = [rng.standard_normal((10, 3, 4, 5)), rng.standard_normal((10, 2, 2, 6))]
    path = tmp_path / "f.plt"
    plt_io.write(str(path), names, zones, time=0.5)
    f = plt_io.read(str(path))
    assert f["names"] == names and f["title"] == "Solution file"
    assert [z["dims"] for z in f["zones"]] == [(5, 4, 3), (6, 2, 2)]
    for z, ref in zip(f["zones"], zones):
        assert z["time"] == 0.5
        np.testing.assert_array_equal(z["data"], ref)
        np.testing.assert_array_equal(z["minmax"], np.stack([ref.reshape(10, -1).min(1), ref.reshape(10, -1).max(1)], 1))
    raw = path.read_bytes()
    assert raw[:16] == b"#!TDV112" + struct.pack("<ii", 1, 0)
    # title characters as int32
    assert struct.unpack_from("<14i", raw, 16) == tuple(ord(c) for c in "Solution file") + (0,)
