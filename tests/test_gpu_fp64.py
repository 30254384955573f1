"""Float32 rates against an FP64 recomputation of the same table: the same
build in float64 (its layers are bit-exact to the reference's float64 run),
compared state by state over every layer.  The stated bound is 1e-6
absolute; the measured maximum is printed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def max_abs_error(pkg, pattern, target):
    r32 = pkg.build_in_memory(pattern, target, success_rate_dtype="float32")
    r64 = pkg.build_in_memory(pattern, target, success_rate_dtype="float64")
    worst = 0.0
    try:
        for i in range(r32.steps):
            k32, v32 = r32.book(i)
            k64, v64 = r64.book(i)
            keys = np.union1d(k32, k64)
            a = np.zeros(keys.size)
            b = np.zeros(keys.size)
            a[np.searchsorted(keys, k32)] = v32.astype(np.float64)
            b[np.searchsorted(keys, k64)] = v64
            if keys.size:
                worst = max(worst, float(np.abs(a - b).max()))
    finally:
        r32.close()
        r64.close()
    return worst


@pytest.mark.parametrize("pattern,target", [("442", 9), ("442", 8), ("L3", 7)])
def test_float32_within_1e6_of_fp64(pkg, pattern, target):
    err = max_abs_error(pkg, pattern, target)
    print(f"case fp64 {pattern}@{2**target}: max |float32 - float64| = {err:.3e}", flush=True)
    assert err <= 1e-6
