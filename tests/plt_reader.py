"""Reader of Tecplot binary data files (format version 112, ordered zones,
FP32 node data), written from the published layout of the format; test
infrastructure for the output writer."""
import struct

import numpy as np


class _Cursor:
    def __init__(self, b):
        self.b, self.at = b, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.at)
        self.at += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def string(self):
        out = []
        while True:
            c = self.take("i")
            if c == 0:
                return "".join(out)
            out.append(chr(c))


def read_plt(path):
    """Returns dict(title, file_type, names, zones=[dict(name, parent, strand,
    time, dims, minmax (nvar, 2), data (nvar, points) float32)])."""
    cur = _Cursor(open(path, "rb").read())
    magic = cur.b[:8].decode()
    cur.at = 8
    assert magic == "#!TDV112", magic
    assert cur.take("i") == 1   # byte order
    file_type = cur.take("i")
    title = cur.string()
    nvar = cur.take("i")
    names = [cur.string() for _ in range(nvar)]
    zones = []
    while True:
        marker = cur.take("f")
        if marker == 357.0:
            break
        assert marker == 299.0, marker
        z = dict(name=cur.string(), parent=cur.take("i"), strand=cur.take("i"), time=cur.take("d"))
        assert cur.take("i") == -1
        assert cur.take("i") == 0   # ordered zone
        assert cur.take("i") == 0   # values at the nodes
        assert cur.take("i") == 0   # no face neighbours
        assert cur.take("i") == 0   # no miscellaneous face connections
        z["dims"] = tuple(cur.take("i") for _ in range(3))
        assert cur.take("i") == 0   # no auxiliary data
        zones.append(z)
    for z in zones:
        assert cur.take("f") == 299.0
        fmts = [cur.take("i") for _ in range(nvar)]
        assert fmts == [1] * nvar   # FP32
        assert cur.take("i") == 0   # no passive variables
        assert cur.take("i") == 0   # no variable sharing
        assert cur.take("i") == -1  # no connectivity sharing
        z["minmax"] = np.array([cur.take("dd") for _ in range(nvar)])
        npt = z["dims"][0] * z["dims"][1] * z["dims"][2]
        z["data"] = np.frombuffer(cur.b, dtype="<f4", count=nvar * npt, offset=cur.at).reshape(nvar, npt)
        cur.at += 4 * nvar * npt
    assert cur.at == len(cur.b), (cur.at, len(cur.b))
    return dict(title=title, file_type=file_type, names=names, zones=zones)
