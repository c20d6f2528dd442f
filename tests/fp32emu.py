"""Float32 emulation helpers of the tests: correctly rounded float32 FMA
in numpy, for reproducing the kernels' FP32 arithmetic on the host."""
import numpy as np

F32 = np.float32


def fma32(a, b, c):
    """Correctly rounded float32 a*b + c: the float64 product is exact,
    the float64 sum is made round-to-odd (exact residual by TwoSum), and
    the cast to float32 then rounds once correctly."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    c = np.broadcast_to(np.asarray(c, dtype=np.float64), a.shape)
    p = a * b
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)
    bits = s.view(np.int64).copy()
    fix = (e != 0) & ((bits & 1) == 0)
    # move to the neighbour on the residual's side (away from or towards 0)
    step = np.where((e > 0) == (s > 0), 1, -1)
    bits[fix] += step[fix]
    return bits.view(np.float64).astype(F32)
