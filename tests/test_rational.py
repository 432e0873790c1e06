"""The best rational approximation of x^gamma behind the fractional SPDE
precision (statX/inferQ/PrecisionMatrixRational.ipp:300-344 and the baryrat
interface it calls). The best approximation is unique and characterised by
equioscillation (Chebyshev's alternation theorem): the error reaches its
maximum modulus with alternating signs at 2m + 2 points."""
import numpy as np
import pytest

from statfem_amd import rational as R


@pytest.mark.parametrize("gamma,m", [(0.5, 4), (0.25, 4), (0.75, 4), (-0.25, 4), (-0.5, 4), (0.02, 4), (0.98, 2),
                                     (-0.95, 6), (-0.05, 6)])
def test_brasil_equioscillates(gamma, m):
    top, bot, info = R.brasil_factors(gamma, m)
    delta = R._delta(gamma, m)
    X = info["extrema"]
    assert len(X) == 2 * m + 2 and np.all(np.diff(X) > 0) and X[0] >= delta and X[-1] <= 1.0
    # the barycentric approximation: equal local maxima (to 1e-7)
    assert info["spread"] < 1e-7
    # the factor form (roots from the barycentric pencil) at the extrema:
    # alternating signs, equal moduli to baryrat's own tolerance (roots near 0
    # carry a little less relative accuracy than the barycentric form)
    e = X ** gamma - R.factor_form(top, bot, X)
    assert np.all(np.sign(e[1:]) == -np.sign(e[:-1]))
    assert (np.abs(e).max() - np.abs(e).min()) / np.abs(e).max() < 1e-4
    # and that modulus is the maximum over the whole interval
    x = np.r_[np.geomspace(delta, 1.0, 20001), np.linspace(delta, 1.0, 20001)]
    assert np.abs(x ** gamma - R.factor_form(top, bot, x)).max() <= np.abs(e).max() * (1 + 1e-4)
    # zeros and poles are real, negative and interlace
    c, d = top[1:], bot[1:]
    assert np.all(c < 0) and np.all(d < 0)
    allr = np.sort(np.r_[c, d])
    kinds = np.isin(allr, c)
    assert np.all(kinds[1:] != kinds[:-1])


def test_brasil_error_decreases_with_degree():
    errs = []
    for m in (2, 3, 4, 5):
        top, bot, info = R.brasil_factors(0.5, m)
        errs.append(info["errors"].max())
    assert all(b < 0.5 * a for a, b in zip(errs, errs[1:]))


def test_no_convergence_is_an_error():
    with pytest.raises(RuntimeError):
        R.brasil_factors(0.5, 4, maxiter=3)


def test_exponent_outside_range_is_refused():
    for g in (0.0, 1.0, -1.0, 1.5):
        with pytest.raises(ValueError):
            R.brasil_factors(g, 4)
