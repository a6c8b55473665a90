"""Pins of the oracle's threshold-tie classifier (tie_bits, the per-candidate
FP32-vs-FP64 check behind the per-result tie counts).

The reference's narrow phase fires when the current distance is below the band
with a strict compare (src/metal/bruteforce.metal:149-151).  A candidate is a
decision tie when FP32 and FP64 evaluations of the same FP32 inputs disagree,
and a band tie when its FP64 distance lies within 4 FP32 ulps of the band.
Both are checked against the textbook point-to-plane distance in numpy float64
on points just above the interior of a tilted triangle, where the decision
reduces to distance < band.
"""
import numpy as np

from oracle import oracle as O

F = np.float32


def plane_distance(q, t0, t1, t2):
    q, t0, t1, t2 = (np.asarray(a, np.float64) for a in (q, t0, t1, t2))
    n = np.cross(t1 - t0, t2 - t0)
    return float(np.dot(n, q - t0) / np.linalg.norm(n))


def test_tie_bits_against_plane_distance():
    t0 = np.array([0.1, 0.2, 0.1], F)
    t1 = np.array([0.9, 0.23, 0.12], F)
    t2 = np.array([0.15, 0.27, 0.85], F)
    n = np.cross(t1.astype(np.float64) - t0, t2.astype(np.float64) - t0)
    n /= np.linalg.norm(n)
    band = F(0.0123)
    ulp = float(np.spacing(band))
    rng = np.random.default_rng(5)
    counts = np.zeros(4, int)
    for delta in rng.uniform(-40, 40, 3000) * ulp:
        b = rng.uniform(0.2, 0.4, 2)  # an interior point, new rounding each time
        c = t0.astype(np.float64) + b[0] * (t1.astype(np.float64) - t0) + b[1] * (t2.astype(np.float64) - t0)
        q = (c + n * (float(band) + delta)).astype(F)
        qp = (c + n * 0.01).astype(F)  # the previous position further out: not crossed
        fired32, _ = O.narrow(np.array([[0, 0]], np.uint32), q.reshape(1, 3), qp.reshape(1, 3),
                              np.stack([t0, t1, t2]), np.array([[0, 1, 2]], np.uint32), False, 0.0, band)
        f32 = bool(fired32[0])
        d64 = plane_distance(q, t0, t1, t2)
        assert d64 > 0
        bits = O.tie_bits(q, qp, t0, t1, t2, band, f32)
        assert bool(bits & 1) == (f32 != (d64 < float(band))), (delta, d64, f32)
        assert bool(bits & 2) == (abs(d64 - float(band)) <= 4 * ulp), (delta, d64)
        counts[bits] += 1
    # the sample holds clear cases, band ties and FP32/FP64 decision ties
    assert counts[0] > 0 and (counts[2] + counts[3]) > 0 and (counts[1] + counts[3]) > 0


def test_far_candidates_are_no_ties():
    t0, t1, t2 = (np.array(a, F) for a in ([0, 0, 0], [1, 0, 0], [0, 0, 1]))
    qp = np.array([0.2, 0.5, 0.2], F)
    for y, fired in ((0.3, False), (0.001, True)):
        q = np.array([0.2, y, 0.2], F)
        assert O.tie_bits(q, qp, t0, t1, t2, 0.0123, fired) == 0
    # a wrong FP32 decision passed in is reported as a decision tie
    q = np.array([0.2, 0.3, 0.2], F)
    assert O.tie_bits(q, qp, t0, t1, t2, 0.0123, True) & 1
