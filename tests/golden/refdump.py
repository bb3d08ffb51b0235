"""Reader for the tagged binary dumps written by oracle/ref_hooks.cpp."""
import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            tag = f.read(16)
            if len(tag) < 16:
                break
            kind = f.read(1).decode()
            elem, n = np.frombuffer(f.read(16), dtype=np.uint64)
            raw = f.read(int(elem) * int(n))
            dt = np.float64 if kind == "d" else np.int32
            out[tag.rstrip(b"\0").decode()] = np.frombuffer(raw, dtype=dt).copy()
    return out
