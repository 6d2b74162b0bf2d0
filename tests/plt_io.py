"""Tecplot binary (#!TDV112) flowfield files as the reference writes its output/flowfield.plt (FieldIO.h:89-430):
an independent reader and writer for the tests, from the published layout of that format.

Header: "#!TDV112", int32 1 (byte order), int32 0 (full file), the title and the variable names as int32 characters
ending in 0, then per zone float 299, its name, int32 parent -1, strand -2, float64 solution time, int32 colour
-1, zone type 0 (ordered), variable location 0, two face-neighbour counts 0, IMax, JMax, KMax, and 0 (no auxiliary
data); float 357 ends the header.  Data: per zone float 299, int32 2 (double) per variable, int32 0 (no passive),
0 (no sharing), -1 (no shared connectivity), a (min, max) float64 pair per variable, then each variable's
IMax * JMax * KMax values with i fastest."""
import struct

import numpy as np


def _str(t):
    return b"".join(struct.pack("<i", ord(c)) for c in t) + struct.pack("<i", 0)


def read(path):
    """Returns dict(title, names, zones=[dict(name, time, dims, minmax (nv, 2), data (nv, K, J, I))])."""
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from("<" + fmt, raw, pos)
        pos += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def text():
        out = []
        while True:
            c = take("i")
            if c == 0:
                return "".join(out)
            out.append(chr(c))

    assert raw[:8] == b"#!TDV112", raw[:8]
    pos = 8
    assert take("i") == 1 and take("i") == 0
    title = text()
    nv = take("i")
    names = [text() for _ in range(nv)]
    zones = []
    marker = take("f")
    while abs(marker - 299.0) < 1e-3:
        name = text()
        parent, strand = take("ii")
        time = take("d")
        colour, ztype, loc, raw_nb, misc = take("iiiii")
        assert (parent, strand, colour, ztype, loc, raw_nb, misc) == (-1, -2, -1, 0, 0, 0, 0)
        dims = take("iii")
        assert take("i") == 0
        zones.append(dict(name=name, time=time, dims=dims))
        marker = take("f")
    assert abs(marker - 357.0) < 1e-3
    for z in zones:
        assert abs(take("f") - 299.0) < 1e-3
        assert all(take("i") == 2 for _ in range(nv))
        assert take("iii") == (0, 0, -1)
        z["minmax"] = np.array([take("dd") for _ in range(nv)])
        n = int(np.prod(z["dims"]))
        I, J, K = z["dims"]
        z["data"] = np.frombuffer(raw, dtype="<f8", count=nv * n, offset=pos).reshape(nv, K, J, I).copy()
        pos += 8 * nv * n
    assert pos == len(raw), (pos, len(raw))
    return dict(title=title, names=names, zones=zones)


def write(path, names, zones, time=0.0):
    """zones: list of (nv, K, J, I) arrays in the order of names."""
    out = [b"#!TDV112", struct.pack("<ii", 1, 0), _str("Solution file"), struct.pack("<i", len(names))]
    out += [_str(n) for n in names]
    for q, z in enumerate(zones):
        K, J, I = z.shape[1:]
        out += [struct.pack("<f", 299.0), _str(f"zone {q}"), struct.pack("<ii", -1, -2), struct.pack("<d", time),
                struct.pack("<iiiii", -1, 0, 0, 0, 0), struct.pack("<iii", I, J, K), struct.pack("<i", 0)]
    out.append(struct.pack("<f", 357.0))
    for z in zones:
        out += [struct.pack("<f", 299.0), struct.pack("<" + "i" * len(names), *([2] * len(names))),
                struct.pack("<iii", 0, 0, -1)]
        out += [struct.pack("<dd", float(v.min()), float(v.max())) for v in z]
        out.append(np.ascontiguousarray(z, dtype="<f8").tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))
