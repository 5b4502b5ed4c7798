"""Host-side logic of the library that needs no GPU, against the oracle."""
import numpy as np

import gxy
import oracle_py as O


def test_best_chain_matches_oracle():
    """gxy_best_chain (O(n log n)) equals the oracle's restatement of the
    reference's O(n^2) scan, ties included, from 1 to 1500 HSPs."""
    rng = np.random.default_rng(7)
    for trial in range(120):
        n = int(rng.choice([1, 2, 5, 30, 63, 64, 65, 200, 1500]))
        span = int(rng.choice([50, 500, 100000]))
        hs = [(int(rng.integers(0, span)), int(rng.integers(0, span)), int(rng.integers(1, 60)),
               int(rng.choice([100, 200, 300]) if trial % 3 == 0 else rng.integers(1, 5000)))
              for _ in range(n)]
        assert gxy.best_chain(hs) == O.best_chain(hs), (trial, n)


def test_best_chain_zero_length_and_nonpositive():
    rng = np.random.default_rng(8)
    for trial in range(60):
        n = int(rng.integers(1, 120))
        hs = [(int(rng.integers(0, 80)), int(rng.integers(0, 80)), int(rng.integers(0, 5)),
               int(rng.integers(-50, 200))) for _ in range(n)]
        assert gxy.best_chain(hs) == O.best_chain(hs), trial
