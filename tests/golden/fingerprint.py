"""Content fingerprint of a CSR, stored in every golden fixture so a changed
generator is caught before any count is compared."""
import numpy as np


def fingerprint(offsets, cols):
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    col = np.ascontiguousarray(cols, dtype=np.uint32)
    w = (np.arange(len(col), dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    with np.errstate(over="ignore"):
        s1 = int(np.sum(col.astype(np.uint64) * w, dtype=np.uint64))
        s2 = int(np.sum(off * ((np.arange(len(off), dtype=np.uint64) % np.uint64(999983)) + np.uint64(7)),
                        dtype=np.uint64))
    return {"n": int(len(off) - 1), "m": int(off[-1]) if len(off) else 0, "cols_sum": s1, "off_sum": s2}
