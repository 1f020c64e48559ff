"""Host logic of the randomized SVD's rank certificate (CPU only).

The certificate receives the captured eigenvalues lam of a projection of the
theta matrix and the uncaptured weight tail.  Every spectrum it must cover
satisfies sigma_i^2 >= lam_i (interlacing) and sum_i (sigma_i^2 - lam_i) +
sum_{i>=l} sigma_i^2 = tail.  These tests build such spectra explicitly,
including the adversarial ones that put all missing weight on one index, and
check the certified ranks against the exact FP64 rule and against the float32
rule of the reference (the oracle's choose_rank_from_sigma)."""
import numpy as np
import pytest
import torch

from atlas_q.rsvd import MAX_SPREAD, RULE_BAND, certify_rank
from oracle import ref_mps as R


def exact_rank(sig2, eps, cap):
    P = np.cumsum(sig2)
    hit = np.nonzero(P >= (1 - eps * eps) * P[-1])[0]
    return min(int(hit[0]) + 1, cap)


def f32_rank(sig2, eps, cap):
    S = torch.from_numpy(np.sqrt(sig2)).to(torch.float32)
    return R.choose_rank_from_sigma(S, eps, cap)[0]


def spectra(rng, n=400, l=150):
    """Decaying spectrum and captured eigenvalues with missing weight placed
    uniformly, at the top, at the crossing region, or beyond l only."""
    tau = rng.uniform(4.0, 30.0)
    sig2 = np.exp(-np.arange(n) / tau) * rng.uniform(0.5, 1.5, n)
    sig2 = np.sort(sig2)[::-1]
    sig2 /= sig2.sum()
    beyond = sig2[l:].sum()
    extra = 10.0 ** rng.uniform(-10, -4)
    mode = rng.integers(4)
    d = np.zeros(l)
    if mode == 1:
        d[0] = extra
    elif mode == 2:
        j = int(rng.integers(l))
        d[j] = min(extra, sig2[j])
    elif mode == 3:
        w = rng.random(l)
        d = np.minimum(extra * w / w.sum(), sig2[:l])
    lam = sig2[:l] - d
    tail = beyond + d.sum()
    return sig2, np.sort(lam)[::-1], tail


@pytest.mark.parametrize("eps", [3e-2, 1e-2, 3e-3, 1e-3])
def test_certified_ranks_contain_exact_and_float32_rule(eps):
    rng = np.random.default_rng(int(-np.log10(eps) * 100))
    certified = 0
    for _ in range(300):
        sig2, lam, tail = spectra(rng)
        for cap in (16, 64, 1024):
            r = certify_rank(lam, tail, 1.0, eps, cap)
            if r is None:
                continue
            certified += 1
            lo, hi = r
            assert hi - lo <= MAX_SPREAD
            assert lo <= exact_rank(sig2, eps, cap) <= hi
            assert lo <= f32_rank(sig2, eps, cap) <= hi
    assert certified > 300          # the certificate is not vacuous


def test_adversarial_tail_on_crossing_index_is_covered():
    # all uncaptured weight sits on the element where the threshold is crossed
    sig2 = np.exp(-np.arange(200) / 8.0)
    sig2 /= sig2.sum()
    eps = 1e-2
    k = exact_rank(sig2, eps, 1000)
    for shift in (0.0, 1e-9, 1e-8, 5e-8):
        lam = sig2[:120].copy()
        lam[k - 1] -= shift
        tail = sig2[120:].sum() + shift
        r = certify_rank(lam, tail, 1.0, eps, 1000)
        if r is not None:
            assert r[0] <= k <= r[1]


def test_cap_below_every_crossing_certifies_cap():
    sig2 = np.full(300, 1.0 / 300)
    lam = sig2[:100]
    assert certify_rank(lam, sig2[100:].sum(), 1.0, 1e-3, 64) == (64, 64)
    # the crossing may lie beyond the captured width: no certificate
    assert certify_rank(lam, sig2[100:].sum(), 1.0, 1e-3, 512) is None


def test_threshold_inside_band_yields_close_ranks_or_none():
    sig2 = np.array([0.5, 0.3, 0.2 - 1e-6, 1e-6 - 2e-8, 2e-8])
    eps = 1e-3                            # thresh = 1 - 1e-6: crossing at index 2
    r = certify_rank(sig2[:4], sig2[4:].sum(), 1.0, eps, 10)
    assert r is not None and r[0] <= 3 <= r[1]
    # a tail larger than the gap to the threshold leaves the crossing open
    assert certify_rank(sig2[:2], sig2[2:].sum(), 1.0, eps, 10) is None


def test_band_is_one_float32_ulp_of_the_total():
    assert RULE_BAND == np.spacing(np.float32(1.0)).item()
