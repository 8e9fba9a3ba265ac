"""SAC binary files (version 6 layout): 632-byte header of 70 float words,
40 integer words and 24 eight-byte strings (kevnm takes two), then npts
float32 samples, in the byte order of the machine that wrote them (the
reference's write_sac and semsks_uvw2sac write native little-endian here).
Undefined header values are -12345."""
from __future__ import annotations

import struct

import numpy as np

FLOATS = ["delta", "depmin", "depmax", "scale", "odelta", "b", "e", "o", "a", "internal1",
          "t0", "t1", "t2", "t3", "t4", "t5", "t6", "t7", "t8", "t9", "f",
          "resp0", "resp1", "resp2", "resp3", "resp4", "resp5", "resp6", "resp7", "resp8", "resp9",
          "stla", "stlo", "stel", "stdp", "evla", "evlo", "evel", "evdp", "mag",
          "user0", "user1", "user2", "user3", "user4", "user5", "user6", "user7", "user8", "user9",
          "dist", "az", "baz", "gcarc", "internal2", "internal3", "depmen", "cmpaz", "cmpinc",
          "xminimum", "xmaximum", "yminimum", "ymaximum"] + ["unused%d" % i for i in range(7)]
INTS = ["nzyear", "nzjday", "nzhour", "nzmin", "nzsec", "nzmsec", "nvhdr", "norid", "nevid", "npts",
        "internal4", "nwfid", "nxsize", "nysize", "unused8", "iftype", "idep", "iztype", "unused9", "iinst",
        "istreg", "ievreg", "ievtyp", "iqual", "isynth", "imagtyp", "imagsrc"] + \
       ["unused%d" % i for i in range(10, 18)] + ["leven", "lpspol", "lovrok", "lcalda", "unused18"]
assert len(FLOATS) == 70 and len(INTS) == 40


def read(path: str, endian: str = "<"):
    """(header dict, samples) of one SAC file."""
    raw = open(path, "rb").read()
    f = struct.unpack(endian + "70f", raw[:280])
    i = struct.unpack(endian + "40i", raw[280:440])
    h = dict(zip(FLOATS, f))
    h.update(zip(INTS, i))
    h["kstnm"] = raw[440:448].decode("ascii", "replace").strip()
    h["kevnm"] = raw[448:464].decode("ascii", "replace").strip()
    n = h["npts"]
    y = np.frombuffer(raw[632:632 + 4 * n], endian + "f4").astype(np.float32)
    return h, y
