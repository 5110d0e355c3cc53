"""Pinning of the CPU oracle (oracle/oracle.c) - CPU only.

The reference holds no golden vectors for these paths (its tests recompute
the numpy oracle on random inputs every run), so every oracle function is
pinned here by what the mathematics fixes: closed forms on polynomial
fields, hand-computed known answers (the hdiff delta spike), and an
independent textbook computation (a dense linear solve of the tridiagonal
systems whose entries are derived by hand for structured inputs).  Each
check is chosen so that a dropped term, a wrong sign, a wrong neighbour or
a transposed axis fails it.
"""

import numpy as np
import pytest

from conftest import inner
from oracle import oracle
from stencil_benchmarks_amd.tools import array

LAYOUTS = [(2, 1, 0), (0, 1, 2)]


def _field(domain, halo, func, layout=(2, 1, 0), dtype=np.float64):
    """Halo-inclusive field with value func(i, j, k) at interior index (i, j, k)."""
    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    a = array.alloc_array(shape, dtype, layout, 0, halo)
    i, j, k = np.meshgrid(*[np.arange(n) - h for n, h in zip(shape, halo)], indexing="ij")
    a[...] = func(i.astype(np.float64), j.astype(np.float64), k.astype(np.float64))
    return a


def _grid(domain):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in domain], indexing="ij")


# ---- copy / averages / Laplacian -------------------------------------------------

@pytest.mark.parametrize("layout", LAYOUTS)
def test_copy_is_identity(layout):
    domain, halo = (7, 5, 4), (1, 2, 1)
    rng = np.random.default_rng(0)
    a = _field(domain, halo, lambda i, j, k: rng.random(i.shape), layout)
    assert np.array_equal(oracle.copy(a, domain, halo), a[inner(domain, halo)])


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_onesided_average_of_affine_field(axis, layout):
    # (f(x + e_a) + f(x)) / 2 = f(x) + slope_a / 2 for f = 2i + 4j + 8k + 1
    domain, halo = (6, 5, 4), (1, 1, 1)
    a = _field(domain, halo, lambda i, j, k: 2 * i + 4 * j + 8 * k + 1, layout)
    i, j, k = _grid(domain)
    expected = 2 * i + 4 * j + 8 * k + 1 + [1.0, 2.0, 4.0][axis]
    assert np.array_equal(oracle.onesided_average(a, axis, domain, halo), expected)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_symmetric_average_of_quadratic_field(axis):
    # ((x+1)^2 + (x-1)^2) / 2 = x^2 + 1 along the axis, others unchanged
    domain, halo = (6, 5, 4), (1, 1, 1)
    sq = [lambda i, j, k: i * i, lambda i, j, k: j * j, lambda i, j, k: k * k][axis]
    a = _field(domain, halo, lambda i, j, k: sq(i, j, k) + 3 * (i + j + k))
    i, j, k = _grid(domain)
    expected = sq(i, j, k) + 1 + 3 * (i + j + k)
    assert np.array_equal(oracle.symmetric_average(a, axis, domain, halo), expected)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_laplacian_known_answers(layout):
    domain, halo = (6, 5, 4), (1, 1, 1)
    zero = np.zeros(domain)
    affine = _field(domain, halo, lambda i, j, k: 3 * i - 5 * j + 7 * k + 2, layout)
    for along in [(1, 1, 0), (1, 0, 0), (0, 0, 1), (1, 1, 1)]:
        assert np.array_equal(oracle.laplacian(affine, along, domain, halo), zero), along
    # sum over enabled axes of (2 f - f(x-e) - f(x+e)): -2 per enabled quadratic axis
    quad = _field(domain, halo, lambda i, j, k: i * i + 2 * j * j + 3 * k * k, layout)
    for along, value in [((1, 0, 0), -2), ((0, 1, 0), -4), ((0, 0, 1), -6), ((1, 1, 0), -6),
                         ((1, 1, 1), -12)]:
        got = oracle.laplacian(quad, along, domain, halo)
        assert np.array_equal(got, np.full(domain, float(value))), along


# ---- horizontal diffusion --------------------------------------------------------

HALO2 = (2, 2, 0)


def test_hdiff_constant_field_and_zero_coeff_are_identity():
    domain = (9, 8, 3)
    rng = np.random.default_rng(1)
    const = _field(domain, HALO2, lambda i, j, k: np.full(i.shape, 0.375))
    coeff = _field(domain, HALO2, lambda i, j, k: rng.random(i.shape))
    assert np.array_equal(oracle.hdiff(const, coeff, domain, HALO2), np.full(domain, 0.375))
    inp = _field(domain, HALO2, lambda i, j, k: rng.random(i.shape))
    zero = _field(domain, HALO2, lambda i, j, k: np.zeros(i.shape))
    assert np.array_equal(oracle.hdiff(inp, zero, domain, HALO2), inp[inner(domain, HALO2)])


@pytest.mark.parametrize("axis", [0, 1])
def test_hdiff_quartic_closed_form(axis):
    # f = (x + 10)^4 along one horizontal axis: lap = -(12 (x+10)^2 + 2), the
    # flux lap(x+1) - lap(x) = -12 (2 (x+10) + 1) has the opposite sign of the
    # gradient (kept by the limiter) and its divergence is -24, so
    # out = in + 24 coeff exactly.
    domain = (7, 6, 2)
    rng = np.random.default_rng(2)
    quartic = [lambda i, j, k: (i + 10) ** 4, lambda i, j, k: (j + 10) ** 4][axis]
    inp = _field(domain, HALO2, quartic)
    coeff = _field(domain, HALO2, lambda i, j, k: np.round(rng.random(i.shape) * 64) / 64)
    expected = inp[inner(domain, HALO2)] + 24 * coeff[inner(domain, HALO2)]
    assert np.array_equal(oracle.hdiff(inp, coeff, domain, HALO2), expected)


def test_hdiff_delta_spike():
    # unit spike at the centre, uniform coeff c: the limiter zeroes the two
    # fluxes next to the spike along each axis (flux and gradient share their
    # sign there), so the centre keeps 1; the four neighbours get -3c, the
    # points two away along an axis +c, the diagonal neighbours +2c.
    domain, c = (9, 9, 1), 0.125
    ci = cj = 4
    inp = _field(domain, HALO2, lambda i, j, k: ((i == ci) & (j == cj)).astype(float))
    coeff = _field(domain, HALO2, lambda i, j, k: np.full(i.shape, c))
    got = oracle.hdiff(inp, coeff, domain, HALO2)[:, :, 0]
    expected = np.zeros(domain[:2])
    expected[ci, cj] = 1.0
    for di, dj in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
        expected[ci + di, cj + dj] = -3 * c
    for di, dj in [(2, 0), (-2, 0), (0, 2), (0, -2)]:
        expected[ci + di, cj + dj] = c
    for di, dj in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
        expected[ci + di, cj + dj] = 2 * c
    assert np.array_equal(got, expected)


# ---- vertical advection ----------------------------------------------------------

DTR = 3.0 / 20.0


def _vadv_fields(domain, halo, stage, pos, tens, ts, wcon):
    return [_field(domain, halo, f) for f in (stage, pos, tens, ts, wcon)]


def test_vadv_without_wind_is_explicit():
    # wcon = 0: a = c = 0, b = dtr, corr = 0, so x = d / dtr and
    # tensstage = dtr (x - pos) = tens + tensstage (to rounding)
    domain, halo = (5, 4, 9), (1, 1, 0)
    rng = np.random.default_rng(3)
    fields = _vadv_fields(domain, halo, *[lambda i, j, k: rng.random(i.shape)] * 4,
                          lambda i, j, k: np.zeros(i.shape))
    got = oracle.vadv(*fields, domain, halo)
    expected = fields[2][inner(domain, halo)] + fields[3][inner(domain, halo)]
    np.testing.assert_allclose(got, expected, rtol=1e-14, atol=1e-15)


def _dense_solution(w_ip, w_i, w_ip_next, w_i_next, stage, pos, tens, ts):
    """Solve one column from hand-derived entries: gav(k) = -(w(k)+w'(k))/4,
    gcv(k) = (w(k+1)+w'(k+1))/4, a = gav/2, c = gcv/2, b = dtr - a - c
    (dtr - c at k = 0, dtr - a at the top), d = dtr pos + tens + ts + corr,
    corr = -a (st(k-1) - st(k)) - c (st(k+1) - st(k)); returns dtr (x - pos)."""
    nz = len(pos)
    A = np.zeros((nz, nz))
    d = np.zeros(nz)
    for k in range(nz):
        gav = -0.25 * (w_ip[k] + w_i[k])
        gcv = 0.25 * (w_ip_next[k] + w_i_next[k]) if k < nz - 1 else 0.0
        a, c = gav / 2, gcv / 2
        corr = 0.0
        if k > 0:
            A[k, k - 1] = a
            corr -= a * (stage[k - 1] - stage[k])
        if k < nz - 1:
            A[k, k + 1] = c
            corr -= c * (stage[k + 1] - stage[k])
        A[k, k] = DTR - (a if k > 0 else 0.0) - (c if k < nz - 1 else 0.0)
        d[k] = DTR * pos[k] + tens[k] + ts[k] + corr
    return DTR * (np.linalg.solve(A, d) - pos)


@pytest.mark.parametrize("shift", [(1, 0), (0, 1), (0, 0)])
def test_vadv_matches_dense_solve(shift):
    # structured, diagonally dominant columns: wcon varies along i, j and k;
    # stage = k^2 makes the correction term non-trivial
    domain, halo = (4, 3, 12), (1, 1, 0)
    rng = np.random.default_rng(4)
    fields = _vadv_fields(
        domain, halo,
        lambda i, j, k: k * k / 64.0,
        lambda i, j, k: rng.random(i.shape),
        lambda i, j, k: rng.random(i.shape),
        lambda i, j, k: rng.random(i.shape),
        lambda i, j, k: 0.05 * np.sin(i + 2 * j + 0.3 * k) + 0.06,
    )
    got = oracle.vadv(*fields, domain, halo, shift)
    stage, pos, tens, ts, wcon = fields
    hi, hj, _ = halo
    for i in range(domain[0]):
        for j in range(domain[1]):
            col = lambda f, di=0, dj=0: f[hi + i + di, hj + j + dj, :]
            w = col(wcon)
            wn = col(wcon, *shift)
            expected = _dense_solution(wn, w, np.roll(wn, -1), np.roll(w, -1), col(stage),
                                       col(pos), col(tens), col(ts))
            np.testing.assert_allclose(got[i, j], expected, rtol=1e-12, atol=1e-13)


def test_vadv_long_double_agrees_on_random_columns():
    # FP64 against x87 extended precision on the benchmark's U[0,1) inputs:
    # the columns are not diagonally dominant, yet the two stay close on the
    # bulk of the points (error bar of the FP64 oracle itself)
    domain, halo = (16, 8, 40), (1, 1, 0)
    rng = np.random.default_rng(5)
    fields = _vadv_fields(domain, halo, *[lambda i, j, k: rng.random(i.shape)] * 5)
    f64 = oracle.vadv(*fields, domain, halo)
    ld = oracle.vadv(*fields, domain, halo, arithmetic="longdouble")
    rel = np.abs(f64 - ld) / np.maximum(np.abs(ld), 1e-3)
    assert np.median(rel) < 1e-13
    assert np.quantile(rel, 0.99) < 1e-9


# ---- STREAM ------------------------------------------------------------------------

@pytest.mark.parametrize("ntimes", [1, 3])
def test_stream_closed_form(ntimes):
    # c = a; b = s c; c = a + b; a = b + s c from (1, 2, 0), s = 3:
    # a = 15^n, b = 3 15^(n-1), c = 4 15^(n-1) (the reference's check)
    n = 1000
    a, b, c = oracle.stream(np.ones(n), np.full(n, 2.0), np.zeros(n), ntimes)
    assert np.all(a == 15.0 ** ntimes)
    assert np.all(b == 3 * 15.0 ** (ntimes - 1))
    assert np.all(c == 4 * 15.0 ** (ntimes - 1))


def test_float32_oracle_follows_float32_rounding():
    # the float32 hdiff restatement rounds every operation to float32: on the
    # quartic field its result differs from the float64 one by float32
    # rounding only
    domain = (7, 6, 2)
    inp = _field(domain, HALO2, lambda i, j, k: (i + 10) ** 4 / 4096.0, dtype=np.float32)
    coeff = _field(domain, HALO2, lambda i, j, k: np.full(i.shape, 0.25), dtype=np.float32)
    got = oracle.hdiff(inp, coeff, domain, HALO2)
    assert got.dtype == np.float32
    expected = inp[inner(domain, HALO2)].astype(np.float64) + 24 * 0.25 / 4096.0
    np.testing.assert_allclose(got, expected, rtol=2e-7)
