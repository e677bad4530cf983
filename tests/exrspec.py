"""A reader for single-part scanline OpenEXR files without compression (the OpenEXR file layout:
magic 20000630, version 2, attributes until an empty name, the scanline offset table, then blocks
of int32 y, int32 byte count and the channels' rows in alphabetical channel order)."""
import struct

import numpy as np


def read(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630, "not an OpenEXR file"
    assert version & 0xff == 2 and not version & 0x200, "single-part scanline file expected"
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\0", name_end + 1)
        typ = data[name_end + 1:type_end].decode()
        size = struct.unpack_from("<i", data, type_end + 1)[0]
        attrs[name] = (typ, data[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    chans, cl = [], attrs["channels"][1]
    p = 0
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        ptype, _plinear, xs, ys = struct.unpack_from("<iB3xii", cl, e + 1)
        chans.append((cl[p:e].decode(), ptype, xs, ys))
        p = e + 1 + 16
    assert attrs["compression"][1] == b"\0", "uncompressed file expected"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from("<%dQ" % H, data, pos)
    assert all(c[1] == 2 and c[2] == 1 and c[3] == 1 for c in chans), "float channels, no subsampling"
    planes = {c[0]: np.empty((H, W), np.float32) for c in chans}
    for row, off in enumerate(offsets):
        y, n = struct.unpack_from("<ii", data, off)
        assert y == y0 + row and n == 4 * W * len(chans)
        block = np.frombuffer(data, np.float32, W * len(chans), off + 8)
        for k, c in enumerate(sorted(planes)):
            planes[c][row] = block[k * W:(k + 1) * W]
    return attrs, [c[0] for c in chans], planes
