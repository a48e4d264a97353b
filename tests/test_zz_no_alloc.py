"""The allocation-free sampler (kq_sample_probabilities_noalloc) against the allocating one
(kq_sample_probabilities) and the oracle: seeded indices bit-identical, on consecutive calls that share
the library's fixed workspace (the second with fewer probabilities and shots), and its size limits."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from kettle_amd import _lib
from oracle import oracle as orc

MAX_PROBS = 1 << 20
MAX_SHOTS = 1 << 20


def _probs(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.random(n) ** 3
    p[rng.integers(0, n, size=max(1, n // 7))] = 0.0   # zero-probability entries inside the CDF
    return p / p.sum()


def _sample(fn_name: str, p: np.ndarray, shots: int, seed: int) -> np.ndarray:
    out = np.zeros(shots, dtype=np.uint64)
    fn = getattr(_lib.load(), fn_name)
    _lib.check(fn(_lib.ptr(p), p.size, shots, 1, seed, _lib.ptr(out) if shots else None))
    return out


@pytest.mark.gpu
def test_noalloc_matches_allocating_sampler_on_consecutive_calls():
    cases = [(MAX_PROBS, 200_000, 11), (1 << 12, 3000, 12)]   # the second call is smaller
    for n, shots, seed in cases:
        p = _probs(n, seed)
        got = _sample("kq_sample_probabilities_noalloc", p, shots, seed)
        ref = _sample("kq_sample_probabilities", p, shots, seed)
        assert np.array_equal(got, ref), (n, shots)
        assert np.array_equal(got, orc.sample_probabilities(p, shots, seed)), (n, shots)
        assert np.all(p[got.astype(np.int64)] > 0.0)


@pytest.mark.gpu
def test_noalloc_limits_and_zero_shots():
    p = _probs(64, 3)
    assert _sample("kq_sample_probabilities_noalloc", p, 0, 1).size == 0
    lib = _lib.load()
    out = np.zeros(4, dtype=np.uint64)
    big = np.full(MAX_PROBS + 1, 1.0 / (MAX_PROBS + 1))
    st = lib.kq_sample_probabilities_noalloc(_lib.ptr(big), big.size, 4, 1, 0, _lib.ptr(out))
    assert st == _lib.KQ_ERR_INVALID and b"at most" in lib.kq_last_error()
    st = lib.kq_sample_probabilities_noalloc(_lib.ptr(p), p.size, MAX_SHOTS + 1, 1, 0, _lib.ptr(out))
    assert st == _lib.KQ_ERR_INVALID
    # the workspace still serves a valid call after the rejected ones
    got = _sample("kq_sample_probabilities_noalloc", p, 1000, 5)
    assert np.array_equal(got, _sample("kq_sample_probabilities", p, 1000, 5))
