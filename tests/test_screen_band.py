"""The list build's FP32 screen (neighbor.hip, `screen_band`): in the
expanded form v = |X_j|^2 - 2 X_i . X_j (three fma, X the FP32 image
coordinates relative to the brick centre), compared with
(rc+skin)^2 -/+ band - |X_i|^2, must never decide a pair the exact FP64
predicate would decide the other way. This emulates the screen's FP32
arithmetic (fma as one rounding of the exact product-sum) on pairs near the
cutoff, for halo extents E from 4 to 20 Å, and checks that the error in r^2
stays far inside the band half-width 8.6e-6 E^2 + 1.2e-7 rc^2 + 1e-6."""
import numpy as np
import pytest

f32 = np.float32


def fma32(a, b, c):
    # float32 products are exact in float64; the sum is rounded once to float32
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)


def band(E, rcs):
    return 8.6e-6 * E * E + 1.2e-7 * rcs * rcs + 1e-6


@pytest.mark.parametrize("E,rcs", [(4.0, 2.8), (6.4, 2.8), (7.4, 3.3), (10.0, 4.4), (20.0, 8.8)])
def test_screen_error_far_inside_band(E, rcs):
    rng = np.random.default_rng(int(E * 10))
    n = 200_000
    xi = rng.uniform(-E, E, (n, 3))
    d = rng.normal(size=(n, 3))
    d *= (rcs * (1.0 + rng.uniform(-1e-3, 1e-3, (n, 1)))) / np.linalg.norm(d, axis=1)[:, None]
    xj = xi + d
    keep = np.all(np.abs(xj) <= E, axis=1)
    xi, xj = xi[keep], xj[keep]
    exact = np.sum((xi - xj) ** 2, axis=1)  # FP64, the reference's arithmetic up to its own rounding
    Xi, Xj = xi.astype(f32), xj.astype(f32)

    def norm2(X):
        return fma32(X[:, 2], X[:, 2], fma32(X[:, 1], X[:, 1], (X[:, 0] * X[:, 0]).astype(f32)))

    Si, Sj = norm2(Xi), norm2(Xj)
    A = f32(-2.0) * Xi
    v = fma32(A[:, 0], Xj[:, 0], fma32(A[:, 1], Xj[:, 1], fma32(A[:, 2], Xj[:, 2], Sj)))
    lo = (f32(rcs * rcs) - f32(band(E, rcs))).astype(f32)
    tlo = (lo - Si).astype(f32)  # the per-atom threshold as the kernel forms it
    # the screen's r^2 estimate and its offset from the exact value
    err = (v.astype(np.float64) + Si.astype(np.float64)) - exact
    assert np.abs(err).max() < 0.25 * band(E, rcs)
    # and no pair beyond the cutoff by more than the band is accepted
    accepted = v < tlo
    assert not np.any(accepted & (exact >= rcs * rcs))
