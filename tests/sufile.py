"""Minimal numpy reader/writer for Seismic Unix trace files (tests only).

240-byte SEG-Y rev0 trace header with the SU extension fields at their byte
offsets, followed by ns native-endian float32 samples.
"""
import numpy as np

OFFSETS = {"tracl": (0, "i4"), "tracr": (4, "i4"), "fldr": (8, "i4"), "cdp": (20, "i4"), "trid": (28, "i2"),
           "offset": (36, "i4"), "sdepth": (48, "i4"), "gdel": (52, "i4"), "sdel": (56, "i4"), "scalel": (68, "i2"),
           "scalco": (70, "i2"), "sx": (72, "i4"), "gx": (80, "i4"), "ns": (114, "u2"), "dt": (116, "u2"),
           "d1": (180, "f4"), "f1": (184, "f4"), "ntr": (204, "i4")}


def read(path):
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        h = raw[pos:pos + 240]
        hdr = {k: np.frombuffer(h[o:o + np.dtype(t).itemsize], dtype=t)[0].item() for k, (o, t) in OFFSETS.items()}
        ns = hdr["ns"]
        data = np.frombuffer(raw[pos + 240:pos + 240 + 4 * ns], dtype="f4").copy()
        out.append((hdr, data))
        pos += 240 + 4 * ns
    return out


def write(path, traces):
    """traces: list of (header dict, float32 data)."""
    with open(path, "wb") as f:
        for hdr, data in traces:
            h = bytearray(240)
            for k, v in hdr.items():
                o, t = OFFSETS[k]
                b = np.array([v], dtype=t).tobytes()
                h[o:o + len(b)] = b
            f.write(bytes(h))
            f.write(np.ascontiguousarray(data, dtype="f4").tobytes())
