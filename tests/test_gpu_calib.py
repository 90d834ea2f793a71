"""The calibration kernels that the probe phases are measured against
(kg_calib_*): they run, return a positive rate, order the way the memory
hierarchy does (a table that fits the caches reads faster than one that does
not; atomics on one word are slower than spread ones), and refuse arguments
outside their contract."""
import pytest

from gpu_util import ctx, say
import kgpu

pytestmark = pytest.mark.gpu


def test_random_read_rates():
    say("calib: random reads")
    c = ctx()
    small = kgpu.calib_random_read(c, 1 << 20, 32, 1 << 24, reps=2)
    big = kgpu.calib_random_read(c, 1 << 30, 32, 1 << 24, reps=2)
    assert small > 0 and big > 0
    assert small > big, (small, big)
    with pytest.raises(kgpu.KgError, match="read_bytes"):
        kgpu.calib_random_read(c, 1 << 20, 12, 1 << 10)
    with pytest.raises(kgpu.KgError, match="power of two"):
        kgpu.calib_random_read(c, 3 << 20, 8, 1 << 10)


def test_atomic_rates():
    say("calib: atomics")
    c = ctx()
    one = kgpu.calib_atomic_add(c, 1, 1 << 18, reps=2)
    spread = kgpu.calib_atomic_add(c, 1 << 20, 1 << 24, reps=2)
    assert one > 0 and spread > one, (one, spread)
    with pytest.raises(kgpu.KgError, match="power of two"):
        kgpu.calib_atomic_add(c, 3, 1 << 10)
