"""The LM control of the HIP kernels (pm_invert, through the C ABI) pinned
by the closed-form damped least-squares trajectories of tests/lm_linear.py:
with only S0 and S1 free the spectra are linear in the parameters, so every
trip of lm_mils (milosUtils.c:466-635) has a closed form.  The same cases
as tests/test_lm_schedule.py runs on the oracle:
- accept path, model after each of 1..6 trips (factor 5, :585) and the exit
  on the toplim test (:579);
- reject path, damping x10 per trip (:612) up to the 1e7 bound (:617);
- logclambda factor (:621), down to the negative damping that ends the loop.
FP64 to 1e-10 relative; the mixed form (FP32 spectra) to 1e-4 on the
trips whose chi2 change exceeds its FP32 rounding, exactly on the reject
path.  Pixels start from the guess
(no classical estimates, the .per path of milos.c:741-775)."""
import sys

import numpy as np
import pytest

import pmilos
import lm_linear as LL
import test_lm_schedule as S
from fixtures import FE6173, TRIPLET

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ctx_for(fix, guess, precision=pmilos.PM_FP64, **kw):
    c = pmilos.Context(FE6173, TRIPLET, S.LAM, psf=S.PSF, fix=fix, guess=guess, classical=False,
                       precision=precision, **kw)
    if precision == pmilos.PM_MIXED:
        c.set_mixed_toplim(0.0)
    return c


# (form, relative tolerance on the model, trips compared).  The mixed form's
# chi2, summed from FP32 spectra, carries a relative noise of ~5e-6 here:
# its 6th trip (a 2e-9 change) and its exit are decided by that noise
FORMS = [(pmilos.PM_FP64, 1e-10, 6), (pmilos.PM_MIXED, 1e-4, 5)]


@pytest.mark.parametrize("precision,rtol,ntrips", FORMS)
def test_accept_path_trip_by_trip_gpu(precision, rtol, ntrips):
    cases = [S.accept_case(seed) for seed in S.ACCEPT_SEEDS]
    cube = np.stack([obs for obs, _ in cases])
    guess = S.model_of(*S.GUESS_S)
    for k in range(1, ntrips + 1):
        say("case lm accept precision=%d miter=%d" % (precision, k))
        r = ctx_for(S.FIX_S, guess, precision, miter=k).invert(cube)
        for i, (obs, ref) in enumerate(cases):
            t = ref["trips"][k - 1]
            np.testing.assert_allclose(r["model"][i, 7:9], t["model"], rtol=rtol, atol=0)
            assert r["iters"][i] == k and r["accepted"][i] == k
            np.testing.assert_allclose(r["chi2"][i], t["chi2"], rtol=max(rtol, 1e-6))
            assert np.array_equal(r["model"][i, :7], S.BASE[:7])
    if precision != pmilos.PM_FP64:
        return
    say("case lm accept precision=%d toplim exit" % precision)
    r = ctx_for(S.FIX_S, guess, precision, toplim=S.TOPLIM_PIN).invert(cube)
    for i, (obs, ref) in enumerate(cases):
        pin = S._closed(obs, [0, 1], S.GUESS_S, 1e-6, 50, toplim=S.TOPLIM_PIN)
        assert r["iters"][i] == pin["iters"] == 6 and r["accepted"][i] == 6
        np.testing.assert_allclose(r["model"][i, 7:9], pin["model"], rtol=rtol)


@pytest.mark.parametrize("precision", [pmilos.PM_FP64, pmilos.PM_MIXED])
@pytest.mark.parametrize("ilambda,trips", [(S.F01, 8), (0.1, 9)])
def test_reject_path_damping_grows_to_the_bound_gpu(precision, ilambda, trips):
    say("case lm reject precision=%d ilambda=%r" % (precision, ilambda))
    obs = S.observed((-0.3, 0.65), 3)
    g = S.model_of(1e-4, 0.65)
    cube = np.stack([obs, obs])
    r = ctx_for(S.FIX_S0, g, precision, ilambda=ilambda, toplim=0.0).invert(cube, want_fit=True)
    assert np.all(r["iters"] == trips) and np.all(r["accepted"] == 0)
    assert np.array_equal(r["model"], np.stack([g, g]))
    # chi2 = the guess's, unchanged; the fit = the last trial = the guess's
    # spectra
    s = S.setup(S.FIX_S0)
    syn = S.O.synth(s, g)
    chi0 = np.sum((obs.astype(np.float64) - syn) ** 2) / 1e-6 / (4 * len(S.LAM) - 9)
    np.testing.assert_allclose(r["chi2"], np.float32(chi0), rtol=1e-5)
    np.testing.assert_allclose(r["fit"][0], syn, atol=1e-6)
    r = ctx_for(S.FIX_S0, g, precision, ilambda=ilambda).invert(cube)
    assert np.all(r["iters"] == 1)
    r = ctx_for(S.FIX_S0, g, precision, ilambda=ilambda, toplim=0.0, miter=5).invert(cube)
    assert np.all(r["iters"] == 5)


@pytest.mark.parametrize("precision,rtol", [(pmilos.PM_FP64, 1e-8), (pmilos.PM_MIXED, 1e-4)])
def test_logclambda_factor_trip_by_trip_gpu(precision, rtol):
    seeds = [5, 7]
    cube = np.stack([S.observed(S.TRUTH, sd) for sd in seeds])
    refs = [S._closed(o, [0, 1], S.GUESS_S, 1e-6, 50, logclambda=1) for o in cube]
    guess = S.model_of(*S.GUESS_S)
    for k in range(1, 6):
        say("case lm logclambda precision=%d miter=%d" % (precision, k))
        r = ctx_for(S.FIX_S, guess, precision, miter=k, logclambda=1).invert(cube)
        for i, ref in enumerate(refs):
            np.testing.assert_allclose(r["model"][i, 7:9], ref["trips"][k - 1]["model"], rtol=rtol)
            assert r["iters"][i] == k
    r = ctx_for(S.FIX_S, guess, precision, logclambda=1).invert(cube)
    for i, ref in enumerate(refs):
        assert r["iters"][i] == ref["iters"] == 5 and r["accepted"][i] == 5
