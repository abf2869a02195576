"""A small independent reader and writer of Tecplot binary (V112) files with
ordered zones and nodal double data, from the published format description
(header: magic "#!TDV112", byte order 1, file type, title, variable names;
per zone a 299.0 marker, name, parent zone, strand id, solution time,
colour, zone type, data location, face-neighbour flags, IMax/JMax/KMax and an
auxiliary-data flag; 357.0 ends the header; data section per zone: 299.0,
one format code per variable, passive / sharing / connectivity flags,
min/max of each variable, then each variable's values).  Strings are one
32-bit integer per character and a terminating 0.  Test infrastructure."""
import struct

import numpy as np


def _str(b, o):
    s = []
    while True:
        c = struct.unpack_from("<i", b, o)[0]
        o += 4
        if c == 0:
            return "".join(s), o
        s.append(chr(c))


def read(path):
    b = open(path, "rb").read()
    assert b[:8] == b"#!TDV112", b[:8]
    o = 8
    order, ftype = struct.unpack_from("<ii", b, o)
    o += 8
    title, o = _str(b, o)
    nv = struct.unpack_from("<i", b, o)[0]
    o += 4
    names = []
    for _ in range(nv):
        n, o = _str(b, o)
        names.append(n)
    zones = []
    while True:
        marker = struct.unpack_from("<f", b, o)[0]
        o += 4
        if marker == 357.0:
            break
        assert marker == 299.0
        zname, o = _str(b, o)
        parent, strand = struct.unpack_from("<ii", b, o)
        o += 8
        time = struct.unpack_from("<d", b, o)[0]
        o += 8
        colour, ztype, loc, raw, misc = struct.unpack_from("<5i", b, o)
        o += 20
        dims = struct.unpack_from("<3i", b, o)
        o += 12
        aux = struct.unpack_from("<i", b, o)[0]
        o += 4
        zones.append(dict(name=zname, parent=parent, strand=strand, time=time, zone_type=ztype, location=loc,
                          dims=dims, aux=aux))
    for z in zones:
        n = z["dims"][0] * z["dims"][1] * z["dims"][2]
        marker = struct.unpack_from("<f", b, o)[0]
        o += 4
        assert marker == 299.0
        z["formats"] = struct.unpack_from("<%di" % nv, b, o)
        o += 4 * nv
        z["passive"], z["sharing"], z["connect"] = struct.unpack_from("<3i", b, o)
        o += 12
        mm = np.frombuffer(b, dtype="<f8", count=2 * nv, offset=o).reshape(nv, 2)
        o += 16 * nv
        z["minmax"] = mm.copy()
        data = {}
        for name in names:
            data[name] = np.frombuffer(b, dtype="<f8", count=n, offset=o).copy()
            o += 8 * n
        z["data"] = data
    assert o == len(b), (o, len(b))
    return dict(order=order, file_type=ftype, title=title, names=names, zones=zones)


def _wstr(s):
    return b"".join(struct.pack("<i", ord(c)) for c in s) + struct.pack("<i", 0)


def write(path, names, zones, time=0.0):
    """zones: list of (dims, {name: values})."""
    out = [b"#!TDV112", struct.pack("<ii", 1, 0), _wstr("Solution file"), struct.pack("<i", len(names))]
    out += [_wstr(n) for n in names]
    for q, (dims, _) in enumerate(zones):
        out += [struct.pack("<f", 299.0), _wstr("zone %d" % q), struct.pack("<ii", -1, -2), struct.pack("<d", time),
                struct.pack("<5i", -1, 0, 0, 0, 0), struct.pack("<3i", *dims), struct.pack("<i", 0)]
    out.append(struct.pack("<f", 357.0))
    for dims, data in zones:
        out += [struct.pack("<f", 299.0), struct.pack("<%di" % len(names), *([2] * len(names))),
                struct.pack("<3i", 0, 0, -1)]
        for n in names:
            v = np.asarray(data[n], dtype="<f8")
            out.append(struct.pack("<2d", v.min(), v.max()))
        for n in names:
            out.append(np.asarray(data[n], dtype="<f8").tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(out))
