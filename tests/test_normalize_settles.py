"""Repeated quaternion normalisation (Quaternion::normalize, q scaled by
1/sqrt(|q|^2), Quaternion.h:40-42) settles: ls1 normalises every molecule at
each container update, and for one-centre systems the device applies those
k normalisations at once when q is read, stopping at the fixed point or
two-cycle the sequence reaches. This checks on the CPU, in the same FP64
operation order, that the sequence settles within a few applications for
random quaternions of any scale, far below the device's cap of 64."""
import numpy as np


def normalise(q):
    qs = 1.0 / np.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2]
                       + q[:, 3] * q[:, 3])
    return q * qs[:, None]


def test_repeated_normalisation_settles():
    rng = np.random.default_rng(3)
    q = rng.normal(size=(200_000, 4)) * 10.0 ** rng.uniform(-3, 3, size=(200_000, 1))
    settled_at = np.full(len(q), -1)
    prev = None
    cur = q.copy()
    for t in range(16):
        nxt = normalise(cur)
        fixed = np.all(nxt == cur, axis=1)
        cycle = np.zeros(len(q), bool) if prev is None else np.all(nxt == prev, axis=1)
        newly = (fixed | cycle) & (settled_at < 0)
        settled_at[newly] = t
        prev, cur = cur, nxt
    assert (settled_at >= 0).all()
    assert settled_at.max() <= 6
