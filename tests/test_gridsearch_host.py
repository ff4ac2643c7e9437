"""Host-only parts of the grid-search autofocus driver (no GPU needed):

* sarf_fit_linear / sarf_fit_quadratic -- the reference's bestFit / quadFit
  (src/autofocus/autofocus_gpu.cu:315-416) -- pinned by exact lines and
  parabolas (closed form) and by numpy.polyfit on noisy tracks;
* sarf_grid_covariance -- the covariance estimate of the multi-resolution
  search -- pinned by exactly quadratic objectives, whose inverse Hessian is
  known in closed form, on boxes with unequal spacings and cross terms so a
  transposed or mis-scaled term fails;
* reference_grid_box -- the first-level box of the reference's driver.
"""
import numpy as np
import pytest

from uncc_sar_focusing_amd import autofocus as af
from uncc_sar_focusing_amd._lib import SarfError


def test_fit_linear_exact_line_and_polyfit():
    i = np.arange(150, dtype=np.float64)
    v = (-812.25 + 3.5 * i).astype(np.float32)
    c = af.fit_linear(v)
    assert c.dtype == np.float32
    np.testing.assert_allclose(c, [-812.25, 3.5], rtol=1e-6)
    rng = np.random.default_rng(3)
    v = (12.0 - 0.37 * i + rng.normal(0, 0.5, i.size)).astype(np.float32)
    ref = np.polyfit(i, v.astype(np.float64), 1)  # [slope, intercept]
    np.testing.assert_allclose(af.fit_linear(v), [ref[1], ref[0]], rtol=1e-5, atol=1e-5)
    # skip: every skip-th sample only
    sk = 3
    ref = np.polyfit(i[::sk], v[::sk].astype(np.float64), 1)
    np.testing.assert_allclose(af.fit_linear(v, sk), [ref[1], ref[0]], rtol=1e-5, atol=1e-5)


def test_fit_quadratic_exact_parabola_and_polyfit():
    i = np.arange(180, dtype=np.float64)
    v = (7.0 - 0.25 * i + 0.0625 * i * i).astype(np.float32)
    np.testing.assert_allclose(af.fit_quadratic(v), [7.0, -0.25, 0.0625], rtol=2e-4, atol=2e-4)
    rng = np.random.default_rng(4)
    v = (-3.0 + 0.8 * i - 0.004 * i * i + rng.normal(0, 0.2, i.size)).astype(np.float32)
    ref = np.polyfit(i, v.astype(np.float64), 2)  # [quad, lin, const]
    np.testing.assert_allclose(af.fit_quadratic(v), ref[::-1], rtol=1e-3, atol=1e-4)
    ref = np.polyfit(i[::2], v[::2].astype(np.float64), 2)
    np.testing.assert_allclose(af.fit_quadratic(v, 2), ref[::-1], rtol=1e-3, atol=1e-4)


def test_fits_reject_too_few_samples():
    with pytest.raises(SarfError):
        af.fit_linear(np.zeros(1, np.float32))
    with pytest.raises(SarfError):
        af.fit_quadratic(np.zeros(2, np.float32))
    with pytest.raises(SarfError):
        af.fit_linear(np.zeros(10, np.float32), 0)


def _box_values(start, end, num, f):
    axes = [np.linspace(s, e, n) if n > 1 else np.array([s]) for s, e, n in zip(start, end, num)]
    P = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, len(num))
    return P, np.array([f(p) for p in P])


@pytest.mark.parametrize("centre", ["interior", "edge"])
def test_grid_covariance_is_inverse_hessian_of_quadratic(centre):
    # SPD Hessian with cross terms, unequal axis spacings
    H = np.array([[40.0, 6.0, -3.0], [6.0, 9.0, 1.5], [-3.0, 1.5, 25.0]])
    start, end, num = [-0.3, -1.0, 0.2], [0.3, 2.0, 0.6], [7, 5, 9]
    P, _ = _box_values(start, end, num, lambda p: 0.0)
    m = P[len(P) // 2 + 3] if centre == "interior" else P[0]
    g = np.array([0.01, -0.02, 0.005])
    f = lambda p: 3.0 + g @ (p - m) + 0.5 * (p - m) @ H @ (p - m)  # noqa: E731
    P, v = _box_values(start, end, num, f)
    best = int(np.argmin([np.linalg.norm(p - m) for p in P]))
    cov = af.grid_covariance(start, end, num, v, best)
    np.testing.assert_allclose(cov, np.linalg.inv(H), rtol=1e-8, atol=1e-10)


def test_grid_covariance_fixed_two_sample_and_saddle_axes():
    H = np.diag([4.0, 10.0])
    start, end = [0.0, -1.0, 5.0, 0.0], [1.0, 1.0, 5.0, 1.0]
    num = [5, 7, 1, 2]
    f = lambda p: 0.5 * np.array([p[0] - 0.5, p[1]]) @ H @ np.array([p[0] - 0.5, p[1]])  # noqa: E731
    P, v = _box_values(start, end, num, f)
    best = int(np.argmin(v))
    cov = af.grid_covariance(start, end, num, v, best)
    # free axes 0, 1: inverse Hessian; fixed axis 2: zeros; two-sample axis 3: NaN
    np.testing.assert_allclose(cov[:2, :2], np.linalg.inv(H), rtol=1e-9)
    assert np.all(cov[2, :3] == 0) and np.all(cov[:3, 2] == 0)
    assert np.all(np.isnan(cov[3, :])) and np.all(np.isnan(cov[:, 3]))
    # a saddle is not positive definite: NaN on the free axes
    f2 = lambda p: (p[0] - 0.5) ** 2 - p[1] ** 2  # noqa: E731
    P, v = _box_values(start[:2], end[:2], num[:2], f2)
    cov = af.grid_covariance(start[:2], end[:2], num[:2], v, int(np.argmin(np.abs(v))))
    assert np.all(np.isnan(cov))


def test_reference_grid_box_centres_on_the_fitted_track():
    Np = 120
    i = np.arange(Np)
    ant = np.stack([-350.0 + 5.9 * i, np.full(Np, 10000.0), 7500.0 - 0.01 * i], axis=1)
    start, end, num = af.reference_grid_box(ant, 7)
    cx, cy, cz = (af.fit_linear(ant[:, a].astype(np.float32)) for a in range(3))
    np.testing.assert_allclose(start[[0, 2, 4]], [cx[0], cy[0], cz[0]])
    np.testing.assert_allclose(end[[0, 2, 4]], [cx[0], cy[0], cz[0]])
    np.testing.assert_allclose((start[[1, 3, 5]] + end[[1, 3, 5]]) / 2, [cx[1], cy[1], cz[1]],
                               atol=1e-5)
    np.testing.assert_allclose(end[[1, 3, 5]] - start[[1, 3, 5]], 5.2, atol=1e-5)
    assert list(num) == [1, 23, 1, 23, 1, 23, 5]
    assert (start[6], end[6]) == (0.9, 1.1)
    start, end, num = af.reference_grid_box(ant, 10)
    assert list(num) == [1, 5, 5, 1, 5, 5, 1, 5, 5, 5]
    np.testing.assert_allclose(end[[2, 5, 8]] - start[[2, 5, 8]], 1.3, atol=1e-6)


def test_detrended_basis_has_no_constant_or_linear_part():
    B = af.poly_basis(257)
    Bd = af.detrended_basis(B)
    t = np.linspace(-1, 1, 257)
    assert np.abs(Bd @ np.ones_like(t)).max() < 1e-9
    assert np.abs(Bd @ t).max() < 1e-9
    # t^2 and t^4 lose their mean, t^3 its linear part (3/5 t for the continuum)
    np.testing.assert_allclose(B - Bd, (B - Bd)[:, [0]] + np.outer((B - Bd)[:, -1] - (B - Bd)[:, 0], (t + 1) / 2),
                               atol=1e-9)
