"""Seeded synthetic inputs of the benchmark workloads (host side, numpy).

15-puzzle disk segment (SlidingTilePuzzleOptimized<4,4>, segment = 2^32
consecutive packed indexes, Puzzles/SlidingTilePuzzleOptimized.cpp:12-26,142-159):
  * segment: the first valid id among splitmix64(seed + k) mod 3805, k = 0, 1, ...
    (valid: t12 <= 12, t13 <= 13, t14 <= 14 for seg = t12 | t13 << 4 | t14 << 8)
  * blank cell b = CROSS_BLANKS[splitmix64(seed ^ BLANK_SALT) mod 6], one of the
    six cells whose vertical moves change the segment (down from 9..11, up
    from 13..15; Puzzles/SlidingTilePuzzleOptimized.cpp:247-285), so the step
    both marks in-segment children and buckets cross-segment ones
  * in-segment ids i = b + 16 m, m < 12!/2 (every such id is a valid state);
    id i is in the frontier iff splitmix64(seed ^ i) < floor(ppm * 2^64 / 10^6)
  * its op bits: splitmix64(seed ^ i ^ OPS_SALT) mod 15 (a frontier entry always
    has a move left, so never 15)
"""
import numpy as np

MASK64 = (1 << 64) - 1
FIFTEEN_SEGMENTS = 3805           # 0xEDC + 1
FIFTEEN_M = 239500800             # 12!/2 packed tile codes per blank cell
BLANK_SALT = 0x626C616E6B
CROSS_BLANKS = (9, 10, 11, 13, 14, 15)
OPS_SALT = 0x6F70626974730000


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def splitmix64_np(x):
    x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def fifteen_segment(seed):
    k = 0
    while True:
        seg = splitmix64((seed + k) & MASK64) % FIFTEEN_SEGMENTS
        if (seg & 15) <= 12 and ((seg >> 4) & 15) <= 13 and (seg >> 8) <= 14:
            return seg, CROSS_BLANKS[splitmix64(seed ^ BLANK_SALT) % len(CROSS_BLANKS)]
        k += 1


def threshold(ppm):
    return (int(ppm) << 64) // 1000000


def bernoulli_frontier(seed, ppm, blank, m_stride=1, m_limit=FIFTEEN_M, chunk=1 << 24):
    """Frontier (ids ascending, op bits) over m = 0, m_stride, 2 m_stride, ... < m_limit.
    A stride above 1 keeps test-sized frontiers spread over every tile digit."""
    thr = np.uint64(threshold(ppm))
    s = np.uint64(seed & MASK64)
    ids, ops = [], []
    for j0 in range(0, (m_limit + m_stride - 1) // m_stride, chunk):
        m = np.arange(j0, min((m_limit + m_stride - 1) // m_stride, j0 + chunk), dtype=np.uint64) * np.uint64(m_stride)
        m = m[m < np.uint64(m_limit)]
        i = np.uint64(blank) + np.uint64(16) * m
        keep = splitmix64_np(s ^ i) < thr
        i = i[keep]
        ids.append(i.astype(np.uint32))
        ops.append((splitmix64_np(s ^ i ^ np.uint64(OPS_SALT)) % np.uint64(15)).astype(np.uint8))
    return np.concatenate(ids), np.concatenate(ops)
