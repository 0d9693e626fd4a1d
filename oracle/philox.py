"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC'11) and the draws the device noise streams build on
it, restated in numpy so that the oracle can be fed the exact noise a device
run used.  Test infrastructure: only the tests import it.

Streams (counter words c0..c3, key k0..k1):
  normal(seed, step, idx):   c = (step lo, step hi, idx, 0x6d617669), k = seed;
      u1 = (((c0 << 21) ^ (c1 >> 11)) + 1) 2^-53 in (0, 1],
      u2 = (((c2 << 21) ^ (c3 >> 11)) & (2^53 - 1)) 2^-53 in [0, 1),
      z = sqrt(-2 log u1) cos(2 pi u2)                   (Box-Muller)
  uniform2(seed, step, idx): c = (step lo, step hi, idx, 0x72746d62), k = seed;
      u = (((c0 << 21) ^ (c1 >> 11)) & (2^53 - 1)) 2^-53, w likewise from c2, c3

The rounds are pinned by the Random123 known-answer vectors
(tests/test_oracle_philox.py)."""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)
TAG_NORMAL = 0x6D617669
TAG_UNIFORM = 0x72746D62


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Ten rounds over broadcastable uint32 counter words and key words;
    returns the four output words as uint64 arrays holding 32-bit values."""
    c = [np.asarray(x, dtype=np.uint64) & MASK32 for x in (c0, c1, c2, c3)]
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


def _words(seed, step, idx, tag):
    step = np.asarray(step, dtype=np.uint64)
    idx = np.asarray(idx, dtype=np.uint64)
    step, idx = np.broadcast_arrays(step, idx)
    seed = int(seed)
    return philox4x32_10(step & MASK32, step >> np.uint64(32), idx, np.full(idx.shape, tag, dtype=np.uint64),
                         seed & 0xFFFFFFFF, seed >> 32)


def _u53(hi, lo):
    return ((hi << np.uint64(21)) ^ (lo >> np.uint64(11))) & np.uint64((1 << 53) - 1)


def normal(seed, step, idx):
    """Standard normal draw of (seed, step, idx); step and idx broadcast."""
    c0, c1, c2, c3 = _words(seed, step, idx, TAG_NORMAL)
    u1 = (_u53(c0, c1).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = _u53(c2, c3).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(np.pi * (2.0 * u2))


def uniform2(seed, step, idx):
    c0, c1, c2, c3 = _words(seed, step, idx, TAG_UNIFORM)
    return _u53(c0, c1).astype(np.float64) * 2.0 ** -53, _u53(c2, c3).astype(np.float64) * 2.0 ** -53


def ring_noise(seed, first_step, nsteps, nrings):
    """noise[s, r] = normal(seed, first_step + s, r): the rotational noise a
    device rings run draws for ring slot r at step counter first_step + s.
    A device step advances its counter before the ring update draws, so a run
    that starts at counter c draws its steps at c + 1, c + 2, ..."""
    steps = np.arange(first_step, first_step + nsteps, dtype=np.uint64)[:, None]
    rings = np.arange(nrings, dtype=np.uint64)[None, :]
    return normal(seed, steps, rings)
