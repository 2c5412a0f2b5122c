"""Host-side functions of the library that need no GPU."""
import numpy as np


def test_cubic_matches_oracle(oracle):
    import nekbone_amd as nb
    # cubic, src/driver.f:196-247 (the oracle's restatement is pinned in
    # test_oracle_math); every element count up to 5000
    for n in range(1, 5001):
        assert nb.cubic(n) == oracle.cubic(n), n


def test_cubic_factorizes():
    import nekbone_amd as nb
    for n in (1, 7, 64, 97, 360, 4096, 8191, 8192):
        mx, my, mz = nb.cubic(n)
        assert mx * my * mz == n and mx >= my >= mz >= 1
