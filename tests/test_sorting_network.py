"""The 16-key compare-exchange network of the learned upsampler's kNN
selection (kOddEvenMerge16 in granular_gpu_amd/csrc/mlp.hip) sorts every
input: by the 0-1 principle it suffices that it sorts all 2^16 zero-one
vectors.  CPU only; reads the table from the kernel source."""
import re
from pathlib import Path

import numpy as np

SRC = Path(__file__).resolve().parents[1] / "granular_gpu_amd" / "csrc" / "mlp.hip"


def network():
    m = re.search(r"kOddEvenMerge16\[126\]\s*=\s*\{([^}]*)\}", SRC.read_text())
    vals = [int(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == 126
    return list(zip(vals[0::2], vals[1::2]))


def test_network_shape():
    net = network()
    assert len(net) == 63
    assert all(0 <= a < b < 16 for a, b in net)


def test_sorts_every_zero_one_input():
    net = network()
    v = ((np.arange(1 << 16)[:, None] >> np.arange(16)[None, :]) & 1).astype(np.int8)
    for a, b in net:
        lo = np.minimum(v[:, a], v[:, b])
        hi = np.maximum(v[:, a], v[:, b])
        v[:, a], v[:, b] = lo, hi
    assert (np.diff(v, axis=1) >= 0).all()


def test_sorts_keys_with_ties():
    # equal keys (the "no candidate" key repeats) and random 64-bit keys
    rng = np.random.default_rng(1)
    for _ in range(200):
        keys = rng.integers(0, 1 << 62, 16, dtype=np.uint64)
        keys[rng.random(16) < 0.3] = np.uint64(2 ** 64 - 1)
        v = list(keys)
        for a, b in network():
            if v[b] < v[a]:
                v[a], v[b] = v[b], v[a]
        assert v == sorted(keys)


def network8():
    m = re.search(r"kOddEvenMerge8\[38\]\s*=\s*\{([^}]*)\}", SRC.read_text())
    vals = [int(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == 38
    return list(zip(vals[0::2], vals[1::2]))


def test_eight_key_network_sorts_every_zero_one_input():
    net = network8()
    assert len(net) == 19 and all(0 <= a < b < 8 for a, b in net)
    v = ((np.arange(1 << 8)[:, None] >> np.arange(8)[None, :]) & 1).astype(np.int8)
    for a, b in net:
        lo = np.minimum(v[:, a], v[:, b])
        hi = np.maximum(v[:, a], v[:, b])
        v[:, a], v[:, b] = lo, hi
    assert (np.diff(v, axis=1) >= 0).all()


def test_merge_of_eight_keeps_the_sixteen_smallest():
    # the merge step of the 8-key selection: minima against the reversed
    # padded chunk, then the bitonic merger, on random keys with ties
    rng = np.random.default_rng(2)
    none = np.uint64(2 ** 64 - 1)
    for _ in range(300):
        top = np.sort(rng.integers(0, 1 << 20, 16, dtype=np.uint64))
        top[rng.random(16) < 0.2] = none
        top = np.sort(top)
        c = np.sort(rng.integers(0, 1 << 20, 8, dtype=np.uint64))
        t = list(top)
        for i in range(8):
            t[8 + i] = min(t[8 + i], c[7 - i])
        j = 8
        while j > 0:
            for i in range(16):
                l = i ^ j
                if l > i and t[l] < t[i]:
                    t[i], t[l] = t[l], t[i]
            j >>= 1
        assert t == sorted(np.concatenate([top, c]))[:16]
