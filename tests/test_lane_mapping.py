"""The pair kernels split a cell's nI sites over lanes as lane = phase * nI +
site, with phase = lane / nI computed as (lane + 0.5) * rcp(nI) in FP32 and
truncated. Exhaustive check over lane < 64, nI <= 64 that this equals the
integer quotient, also when the hardware reciprocal is off by up to 2 ulp."""
import numpy as np


def test_float_reciprocal_quotient_is_exact():
    up, down = np.float32(np.inf), np.float32(-np.inf)
    for n in range(1, 65):
        r0 = np.float32(1.0) / np.float32(n)
        rs = [r0, np.nextafter(r0, up), np.nextafter(np.nextafter(r0, up), up),
              np.nextafter(r0, down), np.nextafter(np.nextafter(r0, down), down)]
        for r in rs:
            lanes = np.arange(64, dtype=np.float32)
            q = ((lanes + np.float32(0.5)) * np.float32(r)).astype(np.int32)
            np.testing.assert_array_equal(q, np.arange(64) // n)
