"""The throughput mode's random numbers (device Philox4x32-10; the oracle's
numpy restatement, oracle/oracle.py, is the checker the GPU test compares
the device stream with bit for bit):

* known answers: the published Random123 known-answer vectors of
  philox4x32-10 (zero counter and key, all-ones, and the pi-digit case);
* uniformity: chi-square of 2^21 uniforms over 256 bins and of the pairs over
  a 32 x 32 grid, and a Kolmogorov-Smirnov test;
* independence: correlations between the two uniforms of a draw, between
  neighbouring counter indices, between the draw streams (shift, axes,
  reactions) and between collision steps, all within 5 standard errors of 0.

The inputs are fixed, so each statistic is a fixed number; the bounds leave
it a wide margin."""
from __future__ import annotations

import numpy as np
from scipy import stats

from oracle import oracle as O

KAT = [  # (counter, key, output), Random123 kat_vectors for philox4x32 with 10 rounds
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]
N = 1 << 20
SEED = 4357787


def test_philox_known_answers():
    for ctr, key, out in KAT:
        got = O.philox4x32_10([ctr], key)[0]
        assert [int(v) for v in got] == list(out)


def test_philox_uniform2_packs_53_bits_of_two_words():
    """The uniform is (w0 << 21 ^ w1 >> 11) / 2^53 of the block's words, so
    it lies in [0, 1) on the 2^-53 grid; a changed index changes the block."""
    u = O.philox_uniform2(SEED, 1, 3, np.arange(4096))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(u * 2.0 ** 53 == np.floor(u * 2.0 ** 53))
    ctr = np.array([[5, 0, 3, 1 << 16]], np.uint32)
    w = O.philox4x32_10(ctr, (SEED & 0xFFFFFFFF, SEED >> 32)).astype(np.uint64)[0]
    a = float((int(w[0]) << 21) ^ (int(w[1]) >> 11)) * 2.0 ** -53
    assert u[5, 0] == a


def test_philox_uniformity():
    u = O.philox_uniform2(SEED, 1, 11, np.arange(N))
    flat = u.reshape(-1)
    h = np.bincount((flat * 256).astype(np.int64), minlength=256)
    assert stats.chisquare(h).pvalue > 1e-3
    h2 = np.bincount((u[:, 0] * 32).astype(np.int64) * 32 + (u[:, 1] * 32).astype(np.int64), minlength=1024)
    assert stats.chisquare(h2).pvalue > 1e-3
    assert stats.kstest(flat, "uniform").pvalue > 1e-3
    assert abs(flat.mean() - 0.5) < 5 * np.sqrt(1.0 / 12.0 / flat.size)


def test_philox_independence():
    idx = np.arange(N)
    a = O.philox_uniform2(SEED, 1, 11, idx)
    bound = 5.0 / np.sqrt(N)
    corr = lambda x, y: abs(np.corrcoef(x, y)[0, 1])
    assert corr(a[:, 0], a[:, 1]) < bound                   # the two uniforms of one draw
    assert corr(a[:-1, 0], a[1:, 0]) < bound                # neighbouring counter indices (cells)
    for stream in (0, 2):                                   # the other draw streams
        assert corr(a[:, 0], O.philox_uniform2(SEED, stream, 11, idx)[:, 0]) < bound
    assert corr(a[:, 0], O.philox_uniform2(SEED, 1, 12, idx)[:, 0]) < bound   # the next collision
    assert corr(a[:, 0], O.philox_uniform2(SEED + 1, 1, 11, idx)[:, 0]) < bound   # another replica's key
