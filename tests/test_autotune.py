"""Host logic of the chunk-length autotuner (no GPU): the fill length is the smallest
chunk length whose chunk count fits the given rounds of resident wavefronts."""
import numpy as np


def test_fill_chunk_len_is_the_smallest_fitting_length():
    from bayesnewton_amd.autotune import fill_chunk_len
    rng = np.random.default_rng(5)
    n = np.floor(np.exp(rng.uniform(np.log(256), np.log(16384), 500))).astype(np.int64)
    for n_cu, rounds in [(8, 1), (8, 2), (32, 1), (3, 4)]:
        cap = n_cu * 4 * 2 * rounds * 64
        c = fill_chunk_len(n, n_cu, rounds=rounds)
        brute = next(k for k in range(16, 4097) if np.sum(-(-n // k)) <= cap)
        assert c == brute, (n_cu, rounds, c, brute)


def test_fill_chunk_len_bounds():
    from bayesnewton_amd.autotune import fill_chunk_len
    assert fill_chunk_len(np.array([5, 100, 3000]), 256) == 16          # everything fits at the floor
    assert fill_chunk_len(np.full(10_000, 1 << 20), 1) is None           # nothing in [16, 4096] fits
