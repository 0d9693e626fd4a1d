"""Pins oracle/philox.py: the Random123 known-answer vectors of
Philox4x32-10 and the distribution of the normal and uniform draws."""
import numpy as np

import philox


def test_known_answer_vectors():
    # Random123 kat_vectors, philox4x32 with 10 rounds: (counter, key) -> output
    kats = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kats:
        got = philox.philox4x32_10(*ctr, *key)
        assert tuple(int(x) for x in got) == want


def test_vectorised_equals_scalar():
    steps = np.arange(5, dtype=np.uint64)[:, None]
    ids = np.arange(7, dtype=np.uint64)[None, :]
    z = philox.normal(123456789, steps, ids)
    for s in range(5):
        for i in range(7):
            assert z[s, i] == philox.normal(123456789, s, i)
    # seeds above 2^32 use both key words
    assert philox.normal(1 << 40, 3, 4) != philox.normal(0, 3, 4)


def test_normal_and_uniform_statistics():
    z = philox.ring_noise(3001, 0, 200, 1000).ravel()
    assert abs(z.mean()) < 5 / np.sqrt(z.size)
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0 / z.size)
    assert abs(np.mean(z ** 4) - 3.0) < 5 * np.sqrt(96.0 / z.size)  # var(z^4) = 105 - 9
    u, w = philox.uniform2(7, np.arange(100000), 11)
    for a in (u, w):
        assert a.min() >= 0.0 and a.max() < 1.0
        assert abs(a.mean() - 0.5) < 5 * np.sqrt(1 / 12 / a.size)
    assert abs(np.corrcoef(u, w)[0, 1]) < 0.02
