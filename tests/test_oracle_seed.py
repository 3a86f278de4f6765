"""Pin the oracle's SeedGenerator keys (random/seed.hpp:54-341) against the
reference's own seed KAT files example/random/random_seed_{K}x{bits}.data:
column 1 is SeedGenerator<result_type, ID, false> (plain), column 2 the
randomized generator, for get() = 1, 2, ... .  The key types of the three
engines fall in these classes: Philox4x32's u32[2] as 2x32 and 1x64
(Threefry2x32), ARS's u32[4] as 4x32 and 2x64 (Threefry2x64), Threefry4x64's
u64[4] as 4x64 and 8x32 (Threefry4x64).  CPU only."""
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [
    # file, engine kind, words per get(), word bits
    ("2x32", orc.PHILOX4X32, 2, 32),
    ("1x64", orc.PHILOX4X32, 1, 64),
    ("4x32", orc.ARS, 4, 32),
    ("2x64", orc.ARS, 2, 64),
    ("4x64", orc.THREEFRY4X64, 4, 64),
    ("8x32", orc.THREEFRY4X64, 8, 32),
]


def load_kat(name):
    plain, rand = [], []
    with open(os.path.join(GOLDEN, "random_seed_%s.data" % name)) as f:
        for line in f:
            a, b = line.split()
            plain.append(int(a))
            rand.append(int(b))
    return plain, rand


def key_words(keys, kind, k, bits):
    """Engine-layout keys (n, 4) as the K words of the KAT result type."""
    out = []
    for key in keys:
        key = [int(v) for v in key]
        if kind == orc.THREEFRY4X64:
            w64 = key
        else:  # u32 words in key[0..]
            nw = 2 if kind == orc.PHILOX4X32 else 4
            w64 = [key[2 * i] | (key[2 * i + 1] << 32) for i in range(nw // 2)]
        if bits == 64:
            out.extend(w64[:k])
        else:
            for w in w64:
                out.extend([w & 0xFFFFFFFF, w >> 32])
    return out


@pytest.mark.parametrize("name,kind,k,bits", CASES)
def test_seed_kat(name, kind, k, bits):
    plain, rand = load_kat(name)
    n = len(plain) // k
    for randomize, want in ((False, plain), (True, rand)):
        keys = orc.seed_keys(kind, 1, n, randomize=randomize)
        got = key_words(keys, kind, k, bits)
        assert got == want, (name, randomize)


def test_seed_partition_closed_form():
    """Partitioned generators (random/seed.hpp:171-199): np ranks of a
    one-word seed interleave, s -> (s % maxs) * np + rank."""
    np_, rank = 4, 3
    keys = orc.seed_keys(orc.PHILOX4X32, 5, 10, np_=np_, rank=rank,
                         randomize=False)
    s = np.arange(5, 15, dtype=np.uint64) * np_ + rank
    assert np.array_equal(keys[:, 0], s & 0xFFFFFFFF)
    assert np.array_equal(keys[:, 1], s >> np.uint64(32))
    # two-word seeds carry the rank in the last word
    keys = orc.seed_keys(orc.ARS, 5, 3, np_=np_, rank=rank, randomize=False)
    assert list(keys[0]) == [5, 0, 3, 0]
    keys = orc.seed_keys(orc.THREEFRY4X64, 9, 2, np_=np_, rank=rank,
                         randomize=False)
    assert list(keys[1]) == [10, 0, 0, 3]
