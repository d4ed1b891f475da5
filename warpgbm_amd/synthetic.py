"""Seeded synthetic data of the benchmark workloads (host NumPy arrays).

Shared by ``bench.py`` and the GPU tests so both use the same inputs:

* ``regression``: X ~ N(0, 1) float32 [N, F]; y = tanh(X[:, :20] w) + N(0, 0.1);
* ``era_workload``: 500k x 400 pre-binned int8 codes in [0, 5), 200
  contiguous eras of lognormal(7.8, 0.5) sizes, a quintile-quantised target
  from a linear signal on 30 features whose sign flips on 10 of them in a
  random 40 % of the eras;
* ``multiclass_workload``: 200k x 100 float32 ~N(0, 1), 5 classes = argmax of
  5 random projections of 30 features plus Gumbel noise.
"""
from __future__ import annotations

import numpy as np


def regression(n, f, seed=0, informative=20):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f), dtype=np.float32)
    w = rng.standard_normal(informative)
    y = (np.tanh(X[:, :informative] @ w) + rng.normal(0, 0.1, n)).astype(np.float32)
    return X, y


def era_workload(seed=0, n=500_000, f=400, eras=200):
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 5, (n, f)).astype(np.int8)
    s = rng.lognormal(7.8, 0.5, eras)
    s = np.floor(s / s.sum() * n).astype(int)
    s[np.argmax(s)] += n - s.sum()
    era = np.repeat(np.arange(eras), s).astype(np.int32)
    c = rng.standard_normal(30)
    flip = np.ones((eras, 30))
    flip[np.ix_(rng.choice(eras, int(0.4 * eras), replace=False), np.arange(10))] = -1
    sig = ((X[:, :30].astype(np.float32) - 2.0) * (c * flip[era])).sum(1)
    y = (np.searchsorted(np.quantile(sig, [0.2, 0.4, 0.6, 0.8]), sig) / 4.0).astype(np.float32)
    return X, y, era


def multiclass_workload(seed=0, n=200_000, f=100, k=5):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f)).astype(np.float32)
    lab = np.argmax(X[:, :30] @ rng.standard_normal((30, k)) + rng.gumbel(size=(n, k)), axis=1)
    return X, lab
