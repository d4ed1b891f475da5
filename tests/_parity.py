"""Per-element error bounds against the FP64 oracle, and a session record of
the largest error each check saw (printed at the end of the test session).

A histogram bin is a sum, so its error is judged against the sum of the
magnitudes of its terms (its condition scale): a device bin G must satisfy
|G - G64| <= tol * sum_{r in bin} |g_r|, and a bin with no nonzero term must
be exactly zero.  Gains and directions are judged per threshold against the
error that such per-bin histogram errors can propagate into them.
"""
from __future__ import annotations

import numpy as np

HIST_TOL = 1e-5  # per-bin, relative to the bin's sum of |terms|

RECORD: dict[str, tuple[float, float]] = {}


def record(name: str, value: float, bound: float) -> None:
    """Keep the largest observed error of a named check (and its bound)."""
    old = RECORD.get(name)
    if old is None or value > old[0]:
        RECORD[name] = (float(value), float(bound))


def per_bin_error(got, ref, abs_sum):
    """Largest |got - ref| / abs_sum over bins with abs_sum > 0, and whether
    every bin with abs_sum == 0 is exactly zero on the device."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    abs_sum = np.asarray(abs_sum, np.float64)
    nz = abs_sum > 0
    err = float(np.max(np.abs(got[nz] - ref[nz]) / abs_sum[nz])) if nz.any() else 0.0
    zeros_exact = bool(np.all(got[~nz] == 0.0))
    return err, zeros_exact


def check_hist(name, got, ref, abs_sum, tol=HIST_TOL):
    err, zeros_exact = per_bin_error(got, ref, abs_sum)
    record(name, err, tol)
    assert zeros_exact, f"{name}: a bin with no rows is not exactly zero"
    assert err <= tol, f"{name}: per-bin relative error {err:.3e} > {tol:.1e}"
    return err


def gain_error_bound(G, H, AG, AH, lam, eps=HIST_TOL):
    """Per-threshold bound on |gain_device - gain_fp64| and on the error of
    the direction test GL/(HL+lam) - GR/(HR+lam), for device histograms whose
    bins are within eps * AG (gradients) and eps * AH (hessians) of the FP64
    bins G, H ([..., B] arrays).  First-order propagation through
    x^2/(y+lam) (|d| <= 2|x|dx/(y+lam) + x^2 dy/(y+lam)^2) for the left, right
    and parent terms, plus float32 rounding of the stored gain (2^-23 of the
    terms' magnitudes) and a factor 2 of slack for second-order terms."""
    G = np.asarray(G, np.float64)
    H = np.asarray(H, np.float64)
    AG = np.asarray(AG, np.float64)
    AH = np.asarray(AH, np.float64)
    GL = np.cumsum(G, -1)[..., :-1]
    HL = np.cumsum(H, -1)[..., :-1]
    Gt, Ht = G.sum(-1, keepdims=True), H.sum(-1, keepdims=True)
    GR, HR = Gt - GL, Ht - HL
    aGL = np.cumsum(AG, -1)[..., :-1]
    aHL = np.cumsum(AH, -1)[..., :-1]
    aGt, aHt = AG.sum(-1, keepdims=True), AH.sum(-1, keepdims=True)
    aGR, aHR = aGt - aGL, aHt - aHL

    def term(x, y, ax, ay):
        d = y + lam
        return (2 * np.abs(x) * eps * ax / d + x * x * eps * ay / (d * d)), np.abs(x * x / d)

    with np.errstate(divide="ignore", invalid="ignore"):
        eL, mL = term(GL, HL, aGL, aHL)
        eR, mR = term(GR, HR, aGR, aHR)
        eP, mP = term(Gt, Ht, aGt, aHt)
        gain_bound = 2.0 * (eL + eR + eP) + 2.0 ** -23 * (mL + mR + mP)
        dL = HL + lam
        dR = HR + lam
        dir_bound = 2.0 * (eps * aGL / dL + np.abs(GL) * eps * aHL / (dL * dL) + eps * aGR / dR +
                           np.abs(GR) * eps * aHR / (dR * dR))
        margin = np.abs(GL / dL - GR / dR)
    return gain_bound, dir_bound, margin
