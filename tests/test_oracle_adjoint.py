"""Pins the oracle's reverse-mode derivative (estimate_backward) by central finite
differences of its FP64 forward (estimate_fp64), for every stencil, wrap rule, weights and
both bin kinds.  A dropped term (e.g. the mean-density dependence of the reference
stencil, whose corner weights do not sum to one), a wrong sign or a transposed axis fails
here at the 1e-6 level."""
import numpy as np
import pytest

from oracle import ref_fp64 as R


@pytest.mark.parametrize("stencil,mode", [("CIC_REF", R.WRAP_VEC_TRUE), ("CIC_REF", R.WRAP_SCAN_FALSE),
                                          ("CIC", 0), ("TSC", 0)])
@pytest.mark.parametrize("kind", ["kf", "fundamental"])
def test_backward_matches_central_differences(stencil, mode, kind):
    rng = np.random.default_rng(11 + mode + len(stencil) + len(kind))
    n, box, N = 12, 90.0, 250
    p = rng.random((N, 3)) * box
    p[:4, 0] = box - 0.2 * box / n          # last cell: exercises the no-wrap rules
    w = 0.5 + rng.random(N)
    kF = 2 * np.pi / box
    edges = None if kind == "fundamental" else kF * (0.5 + np.arange(n // 2 + 1))
    P, cnt, _ = R.estimate_fp64(*p.T, w, box, n, edges, stencil, 2, mode)
    c = rng.standard_normal(P.shape)
    c[cnt == 0] = 0.0
    gp, gw = R.estimate_backward(*p.T, w, box, n, edges, c, stencil, 2, mode)
    loss = lambda q, ww: np.nansum(c * R.estimate_fp64(*q.T, ww, box, n, edges, stencil, 2, mode)[0])
    eps = 1e-6 * box / n
    for i in list(rng.choice(N, 5, replace=False)) + [0]:
        for ax in range(3):
            q = p.copy()
            q[i, ax] += eps
            lp = loss(q, w)
            q[i, ax] -= 2 * eps
            lm = loss(q, w)
            assert abs((lp - lm) / (2 * eps) - gp[i, ax]) <= 1e-6 * np.abs(gp).max()
        ww = w.copy()
        ww[i] += 1e-6
        lp = loss(p, ww)
        ww[i] -= 2e-6
        lm = loss(p, ww)
        assert abs((lp - lm) / 2e-6 - gw[i]) <= 1e-6 * np.abs(gw).max()


def test_reference_stencil_mean_density_term_matters():
    # without the gbar term the reference stencil's gradient is wrong (its weights do not sum to 1)
    rng = np.random.default_rng(5)
    n, box, N = 12, 90.0, 200
    p = rng.random((N, 3)) * box
    kF = 2 * np.pi / box
    edges = kF * (0.5 + np.arange(n // 2 + 1))
    P, cnt, mesh = R.estimate_fp64(*p.T, None, box, n, edges, "CIC_REF")
    c = rng.standard_normal(P.shape)
    gp, _ = R.estimate_backward(*p.T, None, box, n, edges, c, "CIC_REF")
    pos = [p[:, a] * n / box for a in range(3)]
    b = np.zeros((N, 3))
    for flat, valid, W, dW in R.corner_stencil("CIC_REF", pos, n):
        for ax in range(3):
            b[:, ax] += dW[ax] * valid
    assert np.abs(b).max() > 0.05      # sum_c dW_c / dpos does not vanish for the :149 stencil
