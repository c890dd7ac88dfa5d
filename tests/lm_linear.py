"""Closed-form Levenberg-Marquardt trajectories for a model whose spectra are
exactly linear in its free parameters.

With only S0 and S1 free (every absorption parameter fixed) the emergent
Stokes vector of a Milne-Eddington atmosphere is S = S0 u + S1 v with
u = (1 on I, 0 on Q, U, V) and v the (S1-scaled) line profile: the source
function enters linearly (mil_sinrf.c:268-278) and the PSF keeps it linear,
since the continuum trick convolves I - Ic and the kernel sums to 1.  The LM
loop is then damped linear least squares and every trip has a closed form:

    A = J^T W J,  beta(m) = J^T W (J m - obs)
    delta = (A + lambda diag(A))^-1 beta(m),  trial = m - delta

with the reference's control rules (lm_mils, milosUtils.c:466-635):
- damping starts at ilambda; a trial is accepted iff its chi2 is below the
  current one; accept -> lambda / (5 f), reject -> lambda * 10 f
  (PARBETA_better / PARBETA_worst, :475-477, :585, :612);
- the loop ends after the trip in which |ochi2 - chi2| * 100 / chi2 <
  toplim or chi2 < 1e-4 (:579-580), or lambda > 1e7 or lambda < 1e-25
  (:617-618), or after miter trips (:623), counting rejected trips;
- f = 1, or with logclambda f = log10f(chi2) / log10f(chi2_0) of the trip's
  trial chi2 (float), taking effect from the next trip (:621);
- chi2 = sum w_s (S - obs)^2 / sigma^2 / nfree (fchisqr, lib.c:126-152);
  the reported chi2 is the last accepted one (:625).

This module is test infrastructure: it only checks the oracle and the HIP
library, and is written from the rules above, not from either of them.
"""
import numpy as np


def linear_basis(synth, base, s1=0.65):
    """(u, v) of S = S0 u + S1 v from one synthesis: u is the unit I
    vector, v = (S - S0 u) / S1 at the model `base` with S1 = s1.
    synth(model) -> [4][L]."""
    m = np.array(base, dtype=np.float64)
    m[8] = s1
    S = synth(m)
    L = S.shape[1]
    u = np.zeros((4, L))
    u[0] = 1.0
    v = (S - m[7] * u) / m[8]
    return u, v


def log10f(x):
    return np.log10(np.float32(x), dtype=np.float32)


def trajectory(J, obs, m0, noise, ilambda, miter, toplim=1e-12, logclambda=0, nfree=None,
               weight=(1.0, 1.0, 1.0, 1.0), better=5.0, worst=10.0):
    """The LM run on spectra J @ m (J: [4L][k], obs: [4L] float64 of the
    float32 observations).  Returns a dict: model, chi2 (last accepted),
    iters, accepted, and per trip (accepted?, relative chi2 change
    (ochi2 - chi2) / ochi2, damping after the trip, model after the trip)."""
    J = np.asarray(J, dtype=np.float64)
    obs = np.asarray(obs, dtype=np.float64)
    n4 = J.shape[0]
    L = n4 // 4
    wv = np.repeat(np.asarray(weight, dtype=np.float64), L) / noise
    if nfree is None:
        nfree = n4 - J.shape[1]

    def chi2(m):
        r = J @ m - obs
        return float(np.sum(r * r * wv) / nfree)

    A = J.T @ (J * wv[:, None])
    m = np.array(m0, dtype=np.float64)
    och = chi2(m)
    chi0 = och
    lam = float(ilambda)
    f = 1.0
    it = nacc = 0
    trips = []
    while True:
        beta = J.T @ ((J @ m - obs) * wv)
        cov = A.copy()
        cov[np.diag_indices_from(cov)] *= 1.0 + lam
        delta = np.linalg.solve(cov, beta)
        trial = m - delta
        ch = chi2(trial)
        rel = (och - ch) / och
        stop = abs((och - ch) * 100 / ch) < toplim or ch < 1e-4
        acc = ch - och < 0.0
        if acc:
            lam = lam / (better * f)
            m = trial
            och = ch
            nacc += 1
        else:
            lam = lam * worst * f
        if lam > 1e7 or lam < 1e-25:
            stop = True
        it += 1
        if logclambda:
            f = float(np.float32(log10f(ch) / log10f(chi0)))
        trips.append(dict(accepted=acc, rel=rel, chi2=ch, damping=lam, model=m.copy()))
        if not (it < miter and not stop):
            break
    return dict(model=m, chi2=och, chi2_0=chi0, iters=it, accepted=nacc, trips=trips)
