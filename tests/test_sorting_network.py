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
