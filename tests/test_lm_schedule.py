"""The LM control of lm_mils (milosUtils.c:466-635) pinned by what the
mathematics fixes, on the oracle (CPU).

With only S0 and S1 free the spectra are exactly linear in the parameters,
so the LM loop is damped linear least squares whose every trip has a closed
form (tests/lm_linear.py).  The oracle's inversion must follow that
trajectory trip by trip:
- accept path: the model after each of 1..6 trips (factor 5 on accept,
  :585) and the exit on the toplim test (:579);
- reject path: a fixed point of check() (S0 clamped at its 1e-4 bound while
  the data want it negative) makes every trial equal the current model, so
  every trip is rejected and the damping grows x10 per trip (:612) until it
  passes 1e7 (:617): 8 trips from float(0.1), 9 from 0.1;
- logclambda: the factor log10f(chi2)/log10f(chi2_0) (:621) on both updates,
  down to the negative damping that ends the loop once chi2 drops below 1
  (lambda < 1e-25, :617).
A copy of the closed form with the factors 5 and 10 swapped must fail,
which shows that the comparison can tell them apart.
"""
import numpy as np
import pytest

import oracle as O
import lm_linear as LL
from fixtures import FE6173, TRIPLET

LAM = O.grid(FE6173, -290, 20, 290)
PSF = O.gauss_psf(60.0, LAM)
# a weak, narrow line: S0 and S1 strongly correlated (slow LM convergence)
BASE = np.array([6., 50., 0.3, 0.03, 0.1, 40., 70., 0.3, 0.65, 0, 1.])
FIX_S = [0] * 7 + [1, 1] + [0, 0]          # S0, S1 free
FIX_S0 = [0] * 7 + [1, 0] + [0, 0]         # S0 free
TRUTH = (0.28, 0.69)
GUESS_S = (0.35, 0.5)
F01 = float(np.float32(0.1))               # InitialDiagonalElement read through a float


def setup(fix, noise=1e-6, **kw):
    return O.make_setup(FE6173, TRIPLET, LAM, psf=PSF, fix=fix, noise=noise, **kw)


def basis():
    s = setup(FIX_S)
    return LL.linear_basis(lambda m: O.synth(s, m), BASE)


def observed(truth, seed, sigma=1e-3):
    u, v = basis()
    rng = np.random.default_rng(seed)
    return (truth[0] * u + truth[1] * v + rng.normal(0, sigma, u.shape)).astype(np.float32)


def model_of(s0, s1):
    m = BASE.copy()
    m[7], m[8] = s0, s1
    return m


def test_spectra_linear_in_s0_s1():
    """The premise: u is the unit I vector and S(S0, S1) = S0 u + S1 v to
    rounding, with the PSF (continuum trick) and without."""
    for psf in (PSF, None):
        s = O.make_setup(FE6173, TRIPLET, LAM, psf=psf, fix=FIX_S)
        u, v = LL.linear_basis(lambda m: O.synth(s, m), BASE)
        for s0, s1 in ((0.7, 1.3), (0.05, 0.2), (1.9, 0.01)):
            S = O.synth(s, model_of(s0, s1))
            assert np.abs(S - (s0 * u + s1 * v)).max() < 1e-14
        S_a = O.synth(s, model_of(0.3, 0.65))
        S_b = O.synth(s, model_of(0.1, 0.65))
        assert np.abs((S_a - S_b) / 0.2 - u).max() < 1e-14


def _closed(obs, fix_cols, m0, noise, miter, **kw):
    u, v = basis()
    J = np.stack([u.ravel(), v.ravel()], 1)[:, fix_cols]
    return LL.trajectory(J, obs.astype(np.float64).ravel(), np.asarray(m0), noise, kw.pop("ilambda", F01),
                         miter, nfree=4 * len(LAM) - 9, **kw)


# noise seeds of the accept-path case (any seed gives the same pattern)
ACCEPT_SEEDS = [20, 23]
# an exit threshold between the 5th and 6th trips' chi2 changes: with the
# reference's 1e-12 the loop ends once chi2 stalls at the rounding level,
# a trip that depends on the rounding of the synthesis, not on the schedule
TOPLIM_PIN = 5e-5


def accept_case(seed):
    """The observations and the closed-form run of the accept-path case;
    asserts the margins that make its trip counts independent of rounding."""
    obs = observed(TRUTH, seed)
    ref = _closed(obs, [0, 1], GUESS_S, 1e-6, 50)
    trips = ref["trips"]
    # trips 1..6 are accepted with chi2 changes far above rounding
    assert ref["iters"] >= 7
    assert all(t["accepted"] and t["rel"] > 1e-9 for t in trips[:6])
    # the 6th trip's change is 20x below TOPLIM_PIN (in the reference's
    # units, |dchi2| * 100 / chi2), the 5th's 20x above
    assert trips[5]["rel"] * 100 < TOPLIM_PIN / 20 and trips[4]["rel"] * 100 > TOPLIM_PIN * 20
    return obs, ref


@pytest.mark.parametrize("seed", ACCEPT_SEEDS)
def test_accept_path_trip_by_trip(seed):
    obs, ref = accept_case(seed)
    trips = ref["trips"]
    for k in range(1, 7):
        s = setup(FIX_S, miter=k)
        r = O.lm(s, obs, model_of(*GUESS_S))
        np.testing.assert_allclose(r["model"][7:9], trips[k - 1]["model"], rtol=1e-12, atol=0)
        assert r["iters"] == k and r["accepted"] == k
        np.testing.assert_allclose(r["chi2"], np.float32(trips[k - 1]["chi2"]), rtol=1e-6)
        assert np.array_equal(r["model"][:7], BASE[:7])
    # the toplim exit: the whole run stops after the 6th trip
    pin = _closed(obs, [0, 1], GUESS_S, 1e-6, 50, toplim=TOPLIM_PIN)
    assert pin["iters"] == 6 and pin["accepted"] == 6
    r = O.lm(setup(FIX_S, toplim=TOPLIM_PIN), obs, model_of(*GUESS_S))
    assert r["iters"] == 6 and r["accepted"] == 6
    np.testing.assert_allclose(r["model"][7:9], pin["model"], rtol=1e-12)
    # with the reference's toplim the run goes on until chi2 stalls; the
    # model moves by less than 1e-8 after the 6th trip
    r = O.lm(setup(FIX_S), obs, model_of(*GUESS_S))
    assert 7 <= r["iters"] <= 50 and r["accepted"] >= 6
    np.testing.assert_allclose(r["model"][7:9], ref["model"], rtol=1e-8)


def test_swapped_factors_are_told_apart():
    """A schedule with the accept and reject factors swapped gives another
    trajectory by the second trip: the comparison above would fail it."""
    obs = observed(TRUTH, ACCEPT_SEEDS[0])
    good = _closed(obs, [0, 1], GUESS_S, 1e-6, 3)
    bad = _closed(obs, [0, 1], GUESS_S, 1e-6, 3, better=10.0, worst=5.0)
    assert np.abs(good["trips"][1]["model"] - bad["trips"][1]["model"]).max() > 1e-3
    r = O.lm(setup(FIX_S, miter=2), obs, model_of(*GUESS_S))
    assert np.abs(r["model"][7:9] - bad["trips"][1]["model"]).max() > 1e-3


@pytest.mark.parametrize("ilambda,trips", [(F01, 8), (0.1, 9)])
def test_reject_path_damping_grows_to_the_bound(ilambda, trips):
    """The data want S0 = -0.3; the guess sits on check()'s bound S0 = 1e-4,
    so every trial is clamped back onto the guess: same spectra, same chi2,
    rejected.  lambda = ilambda 10^k exceeds 1e7 after 8 trips from the
    float-read 0.1 (0.100000001...), 9 from 0.1 itself."""
    obs = observed((-0.3, 0.65), 3)
    g = model_of(1e-4, 0.65)
    # closed form without check(): the first step takes S0 below the bound
    # (and every later step from the bound too: the data want S0 = -0.3),
    # so with the clamp every trial is the guess itself
    u, v = basis()
    free = _closed(obs - 0.65 * v.astype(np.float32), [0], [1e-4], 1e-6, 1, ilambda=ilambda,
                   toplim=0.0)
    assert free["trips"][0]["model"][0] < -0.1
    lam = ilambda
    for k in range(trips):
        lam = lam * 10.0
    assert lam > 1e7 and lam / 10.0 <= 1e7
    s = setup(FIX_S0, ilambda=ilambda, toplim=0.0)
    r = O.lm(s, obs, g)
    assert r["iters"] == trips and r["accepted"] == 0
    assert np.array_equal(r["model"], g)
    chi0 = np.sum((obs.astype(np.float64) - O.synth(s, g)) ** 2) / 1e-6 / (4 * len(LAM) - 9)
    np.testing.assert_allclose(r["chi2"], np.float32(chi0), rtol=1e-6)
    # the same with the reference's toplim: an unchanged chi2 ends the loop
    # on the first trip; with miter below the bound, miter trips
    assert O.lm(setup(FIX_S0, ilambda=ilambda), obs, g)["iters"] == 1
    assert O.lm(setup(FIX_S0, ilambda=ilambda, toplim=0.0, miter=5), obs, g)["iters"] == 5


@pytest.mark.parametrize("seed", [5, 7])
def test_logclambda_factor_trip_by_trip(seed):
    """logclambda = 1: the damping is divided by 5 f and multiplied by 10 f
    with f = log10f(chi2)/log10f(chi2_0) of the previous trip; once chi2
    drops below 1 (chi2_0 > 1), f < 0, the damping turns negative and the
    loop ends on lambda < 1e-25.  log10f differs between libraries by an
    ulp or two, hence the tolerance."""
    obs = observed(TRUTH, seed)
    ref = _closed(obs, [0, 1], GUESS_S, 1e-6, 50, logclambda=1)
    plain = _closed(obs, [0, 1], GUESS_S, 1e-6, 50)
    trips = ref["trips"]
    # 5 accepted trips, each changing chi2 far above rounding; the 4th
    # brings chi2 below 1, the 5th's update makes the damping negative
    assert len(trips) == 5 and all(t["accepted"] and t["rel"] > 1e-4 for t in trips)
    assert trips[-1]["damping"] < 0 and ref["chi2_0"] > 1
    assert trips[-2]["chi2"] < 1 and all(t["chi2"] > 1 for t in trips[:-2])
    # the factor changes the trajectory from the third trip on
    assert np.abs(trips[2]["model"] - plain["trips"][2]["model"]).max() > 1e-4
    for k in range(1, len(trips) + 1):
        r = O.lm(setup(FIX_S, miter=k, logclambda=1), obs, model_of(*GUESS_S))
        np.testing.assert_allclose(r["model"][7:9], trips[k - 1]["model"], rtol=1e-8)
        assert r["iters"] == k
    r = O.lm(setup(FIX_S, logclambda=1), obs, model_of(*GUESS_S))
    assert r["iters"] == ref["iters"] and r["accepted"] == ref["accepted"]
