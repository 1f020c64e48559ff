"""Host logic of the randomized SVD's rank certificate (CPU only).

The certificate receives the captured eigenvalues lam of a projection of the
theta matrix and the uncaptured weight tail.  Every spectrum it must cover
satisfies sigma_i^2 >= lam_i (interlacing) and sum_i (sigma_i^2 - lam_i) +
sum_{i>=l} sigma_i^2 = tail.  These tests pin the float32 emulation of the
rank rule to the oracle's rule, then build admissible spectra explicitly --
including adversarial ones that put all missing weight on one index -- and
check the certified rank against the oracle's float32 rule on the true
spectrum."""
import numpy as np
import pytest
import torch

from atlas_q.rsvd import certify_rank, rule_rank_f32
from oracle import ref_mps as R


def oracle_rank(sig2, eps, cap=10 ** 9):
    S = torch.from_numpy(np.sqrt(np.asarray(sig2, np.float64))).to(torch.float32)
    return R.choose_rank_from_sigma(S, eps, cap)[0]


def decaying(rng, n):
    tau = rng.uniform(3.0, 40.0)
    sig2 = np.exp(-np.arange(n) / tau) * rng.uniform(0.5, 1.5, n) + 1e-9 * rng.random(n)
    sig2 = np.sort(sig2)[::-1]
    return sig2 / sig2.sum()


@pytest.mark.parametrize("eps", [1e-1, 1e-2, 1e-3, 3e-4, 1e-4, 1e-6, 1e-8])
def test_float32_rule_emulation_matches_oracle(eps):
    rng = np.random.default_rng(int(-np.log10(eps) * 10))
    for trial in range(200):
        n = int(rng.integers(1, 300))
        kind = trial % 4
        if kind == 0:
            sig2 = decaying(rng, n)
        elif kind == 1:
            sig2 = np.sort(rng.random(n))[::-1]                      # flat-ish
        elif kind == 2:
            sig2 = np.repeat(np.sort(rng.random(max(1, n // 5)))[::-1], 5)  # ties
        else:
            sig2 = np.concatenate([decaying(rng, n), np.zeros(int(rng.integers(0, 20)))])
        S = np.sqrt(sig2 * rng.uniform(0.1, 10.0))
        assert rule_rank_f32(S, eps) == oracle_rank(S ** 2, eps), (trial, eps)


def admissible(rng, sig2, l, mode):
    """Captured eigenvalues and tail for a true spectrum sig2; the missed weight
    inside the captured range is spread, concentrated on the top value, or
    concentrated on a random index."""
    d = np.zeros(l)
    extra = 10.0 ** rng.uniform(-14, -7)
    if mode == 1:
        d[0] = min(extra, sig2[0])
    elif mode == 2:
        j = int(rng.integers(l))
        d[j] = min(extra, sig2[j])
    elif mode == 3:
        w = rng.random(l)
        d = np.minimum(extra * w / w.sum(), sig2[:l])
    lam = np.sort(sig2[:l] - d)[::-1]
    return lam, sig2[l:].sum() + d.sum()


@pytest.mark.parametrize("eps", [1e-2, 1e-3, 1e-6, 1e-8])
def test_certified_rank_is_the_float32_rule_on_the_true_spectrum(eps):
    rng = np.random.default_rng(7 + int(-np.log10(eps)))
    certified = agree = 0
    for trial in range(300):
        sig2 = decaying(rng, 500)
        lam, tail = admissible(rng, sig2, 350, trial % 4)
        for cap in (16, 1024):
            k = certify_rank(lam, tail, eps, cap)
            if k is None:
                continue
            certified += 1
            want = min(oracle_rank(sig2, eps), cap)
            assert abs(k - want) <= 1, (trial, cap, k, want)
            agree += k == want
    assert certified >= 300                 # the certificate is not vacuous
    assert agree >= 0.99 * certified, (agree, certified)


def test_both_extremes_are_needed():
    """All missed weight on the first value moves the crossing earlier; a
    certificate that looked only at the spectrum with the tail last would
    accept a rank the true spectrum does not give."""
    sig2 = np.exp(-np.arange(60) / 4.0)
    sig2 /= sig2.sum()
    eps = 1e-2
    k_true = oracle_rank(sig2, eps)
    P = np.cumsum(sig2)
    gap = P[k_true - 1] - (1 - eps * eps)          # margin of the true crossing
    lam = sig2[:40].copy()
    lam[0] -= 2 * gap + 1e-12                       # missed weight larger than the margin
    tail = sig2[40:].sum() + 2 * gap + 1e-12
    lam = np.sort(lam)[::-1]
    k_last = rule_rank_f32(np.sqrt(np.concatenate([lam, [tail]])), eps)
    assert k_last != k_true                          # the one-sided view is wrong here
    assert certify_rank(lam, tail, eps, 1000) is None


def test_cap_below_every_crossing_certifies_cap():
    sig2 = np.full(300, 1.0 / 300)
    lam = sig2[:100]
    assert certify_rank(lam, sig2[100:].sum(), 1e-3, 64) == 64
    # the crossing may lie beyond the captured width: no certificate
    assert certify_rank(lam, sig2[100:].sum(), 1e-3, 512) is None


def test_exact_capture_reproduces_the_rule():
    rng = np.random.default_rng(11)
    for eps in (1e-2, 1e-4, 1e-8):
        sig2 = decaying(rng, 200)
        assert certify_rank(sig2, 0.0, eps, 10 ** 6) == rule_rank_f32(np.sqrt(sig2), eps)
