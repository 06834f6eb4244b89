"""Seeded synthetic workloads: deterministic, with the stated zero fractions."""
import numpy as np

from fzgpu import datagen as D


def test_generators_are_deterministic_and_shaped():
    a = D.waves((200, 100, 1), 5)
    b = D.waves((200, 100, 1), 5)
    assert a.shape == (1, 100, 200) and a.dtype == np.float32
    assert np.array_equal(a, b)
    assert not np.array_equal(a, D.waves((200, 100, 1), 6))


def test_zero_fractions():
    w = D.waves((400, 300, 1), 1)
    assert 0.15 <= np.mean(w == 0) < 0.2 and w.min() == 0 and w.max() <= 1
    p = D.peaks((4000, 60, 1), 2)
    assert abs(np.mean(p == 0) - 0.97) < 0.002
    b = D.blobs((60, 50, 40), 3, ncentres=20)
    assert abs(np.mean(b == 0) - 0.70) < 0.01 and b.max() <= 1e-3 + 1e-9
