"""Pins of the oracle's threshold-sparse demag (ApplyThresholdDense /
FindThresholdFraction, DemagAuxFunctions.f90:20-122,353-428;
ConvertDenseToSparse_s, util_micromag.f90:81-138):

* the fraction threshold leaves a count of positive entries <= t within
  +-1% of the requested fraction (the reference's stopping rule), found
  by brute-force counting here;
* fraction >= 1 drops everything (zero demag field);
* threshold 0 keeps the dense result; a threshold above max|K| gives a zero
  demag field;
* the thresholded demag field equals the FP64 product over the entries with
  |K| > t computed with numpy from the dense tensor.
"""

import numpy as np
import pytest

RES, L = [8, 4, 2], [40e-9, 20e-9, 6e-9]


def _m(n, seed=0):
    rng = np.random.default_rng(seed)
    m = rng.normal(size=(3, n))
    return (m / np.linalg.norm(m, axis=0)).reshape(-1)


@pytest.mark.parametrize("frac", [0.3, 0.9, 0.995])
def test_fraction_threshold_brute_force(oracle, frac):
    K = oracle.tensor_uniform(RES, L).astype(np.float32)
    t = oracle.threshold_fraction(K, frac)
    pos = K[K > 0]
    cnt = np.count_nonzero(pos <= np.float32(t))
    assert (frac - 0.01) * pos.size <= cnt <= (frac + 0.01) * pos.size
    assert pos.min() <= t <= pos.max()
    assert oracle.threshold_fraction(K, 1.0) == np.finfo(np.float32).max


def test_threshold_field_equals_masked_product(oracle):
    n = int(np.prod(RES))
    K = oracle.tensor_uniform(RES, L).astype(np.float32)
    m = _m(n)
    ms = 8e5
    dense = oracle.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=4.42e3)
    _, Hd = dense.dmdt(0.0, m)
    same = oracle.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=4.42e3, demag_approx=2, dem_thres=0.0)
    _, Hs = same.dmdt(0.0, m)
    assert np.array_equal(Hs[2], Hd[2])
    big = oracle.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=4.42e3, demag_approx=2, dem_thres=1.0)
    _, Hz = big.dmdt(0.0, m)
    assert np.all(Hz[2] == 0.0)
    thr = np.float32(np.quantile(np.abs(K), 0.8))
    q = oracle.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=4.42e3, demag_approx=2, dem_thres=float(thr))
    _, Ht = q.dmdt(0.0, m)
    Km = np.where(np.abs(K) > thr, K, 0).astype(np.float64)
    m32 = m.astype(np.float32).astype(np.float64).reshape(3, n)
    comp = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    ref = np.zeros((3, n))
    for a in range(3):
        for b in range(3):
            ref[a] -= Km[comp[a][b]] @ m32[b]
    ref *= np.float32(ms)
    assert np.abs(Ht[2] - ref.reshape(-1)).max() <= 1e-6 * np.abs(ref).max()
    assert np.abs(Ht[2] - Hd[2]).max() > 1e-4 * np.abs(Hd[2]).max()  # the threshold changed the field
    f = oracle.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=4.42e3, demag_approx=4, dem_thres=1.0)
    _, Hf = f.dmdt(0.0, m)
    assert np.all(Hf[2] == 0.0)
