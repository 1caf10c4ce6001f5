"""Position-sensitive 64-bit digests of arrays, for fixtures too large to
commit whole (full-size AMG hierarchies): two wrapping weighted sums of the
raw 64-bit words (values by their bit pattern, so equal digests mean equal
bits up to a vanishing collision chance).  Shared by the fixture generator and
the GPU tests that compare against it."""
import numpy as np

_M = np.uint64(0x9E3779B97F4A7C15)


def digest(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        u = a.view(np.uint64)
    else:
        u = a.astype(np.int64).view(np.uint64)
    k = np.arange(len(u), dtype=np.uint64)
    with np.errstate(over="ignore"):
        h1 = int(np.sum(u * (k * np.uint64(2) + np.uint64(1)), dtype=np.uint64))
        h2 = int(np.sum(u * ((k + np.uint64(1)) * _M | np.uint64(1)), dtype=np.uint64))
    return f"{h1:016x}{h2:016x}"


def csr_digest(rp, ci, v):
    return {"rp": digest(rp), "ci": digest(ci), "v": digest(v), "v_sum": float(np.sum(v)),
            "v_abs": float(np.sum(np.abs(v)))}
