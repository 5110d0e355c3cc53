"""Writes the known-answer fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_fixtures.py``).

The reference holds no golden vectors for its stencils (its tests recompute
the numpy oracle on random inputs), so these fixtures come from what the
mathematics fixes, computed here without the oracle:

* ``hdiff_quartic.npz``: f = (x + 10)^4 along i (and, second case, along j)
  with coeff on a 1/64 grid; lap = -(12 (x+10)^2 + 2), the flux has the
  opposite sign of the gradient (the limiter keeps it) and its divergence is
  -24, so out = in + 24 coeff exactly (stencils/base.py:276-311);
* ``hdiff_spike.npz``: a unit spike with uniform coeff c: the limiter zeroes
  the fluxes next to the spike, so the centre keeps 1, its four neighbours
  get -3c, the points two away +c, the diagonal neighbours +2c;
* ``laplacian_quadratic.npz``: f = i^2 + 3 j^2 + 5 k^2 + 7 ijk: each axis
  term 2 f - f(-e) - f(+e) is -2, -6, -10 (the mixed term cancels), so the
  x/y Laplacian is -8 and the x/y/z one -18 (stencils/base.py:225-254);
* ``vadv_dense.npz``: structured columns (wcon varying along i, j and k,
  stage = k^2/64) and the tridiagonal system of stencils/base.py:415-473
  written out entry by entry, solved with numpy.linalg.solve (a textbook
  dense solve, not the Thomas recurrence): expected utensstage to ~1e-13.

Inputs are halo-inclusive C-ordered float64 arrays with index (i, j, k);
the interior starts at the halo width.
"""

from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DTR = 3.0 / 20.0


def _grid(domain, halo):
    shape = tuple(d + 2 * h for d, h in zip(domain, halo))
    return np.meshgrid(*[np.arange(n, dtype=np.float64) - h for n, h in zip(shape, halo)],
                       indexing="ij")


def _interior(a, domain, halo):
    return a[tuple(slice(h, h + d) for d, h in zip(domain, halo))]


def hdiff_quartic():
    domain, halo = (12, 10, 3), (2, 2, 0)
    rng = np.random.default_rng(11)
    i, j, _ = _grid(domain, halo)
    coeff = np.round(rng.random(i.shape) * 64) / 64
    out = {"domain": np.array(domain), "halo": np.array(halo), "coeff": coeff}
    for name, x in (("x", i), ("y", j)):
        inp = (x + 10) ** 4
        out[f"inp_{name}"] = inp
        out[f"out_{name}"] = _interior(inp, domain, halo) + 24 * _interior(coeff, domain, halo)
    return out


def hdiff_spike():
    domain, halo, c = (9, 9, 2), (2, 2, 0), 0.125
    ci = cj = 4
    i, j, _ = _grid(domain, halo)
    inp = ((i == ci) & (j == cj)).astype(np.float64)
    coeff = np.full(i.shape, c)
    expected = np.zeros(domain)
    expected[ci, cj, :] = 1.0
    for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        expected[ci + di, cj + dj, :] = -3 * c
    for di, dj in ((2, 0), (-2, 0), (0, 2), (0, -2)):
        expected[ci + di, cj + dj, :] = c
    for di, dj in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
        expected[ci + di, cj + dj, :] = 2 * c
    return {"domain": np.array(domain), "halo": np.array(halo), "inp": inp, "coeff": coeff,
            "out": expected}


def laplacian_quadratic():
    domain, halo = (11, 7, 5), (1, 1, 1)
    i, j, k = _grid(domain, halo)
    inp = i ** 2 + 3 * j ** 2 + 5 * k ** 2 + 7 * i * j * k
    return {"domain": np.array(domain), "halo": np.array(halo), "inp": inp,
            "out_xy": np.full(domain, -8.0), "out_xyz": np.full(domain, -18.0)}


def _dense_column(w_n, w, stage, pos, tens, ts):
    """utensstage of one column from the system written out entry by entry:
    gav(k) = -(w_n(k) + w(k))/4, gcv(k) = (w_n(k+1) + w(k+1))/4, a = gav/2,
    c = gcv/2, b = dtr - a - c (dtr - c at the bottom, dtr - a at the top),
    d = dtr pos + tens + ts - a (st(k-1) - st(k)) - c (st(k+1) - st(k))."""
    nz = len(pos)
    A = np.zeros((nz, nz))
    d = np.zeros(nz)
    for k in range(nz):
        a = -0.25 * (w_n[k] + w[k]) / 2 if k > 0 else 0.0
        c = 0.25 * (w_n[k + 1] + w[k + 1]) / 2 if k < nz - 1 else 0.0
        corr = 0.0
        if k > 0:
            A[k, k - 1] = a
            corr -= a * (stage[k - 1] - stage[k])
        if k < nz - 1:
            A[k, k + 1] = c
            corr -= c * (stage[k + 1] - stage[k])
        A[k, k] = DTR - a - c
        d[k] = DTR * pos[k] + tens[k] + ts[k] + corr
    return DTR * (np.linalg.solve(A, d) - pos)


def vadv_dense():
    domain, halo = (6, 4, 14), (1, 1, 0)
    rng = np.random.default_rng(12)
    i, j, k = _grid(domain, halo)
    stage = k * k / 64.0
    pos, tens, ts = (rng.random(i.shape) for _ in range(3))
    wcon = 0.05 * np.sin(i + 2 * j + 0.3 * k) + 0.06
    expected = np.zeros(domain)
    hi, hj, _ = halo
    for ii in range(domain[0]):
        for jj in range(domain[1]):
            col = lambda f, di=0: f[hi + ii + di, hj + jj, :]
            expected[ii, jj] = _dense_column(col(wcon, 1), col(wcon), col(stage), col(pos),
                                             col(tens), col(ts))
    return {"domain": np.array(domain), "halo": np.array(halo), "stage": stage, "pos": pos,
            "tens": tens, "tensstage": ts, "wcon": wcon, "out": expected}


FIXTURES = {"hdiff_quartic": hdiff_quartic, "hdiff_spike": hdiff_spike,
            "laplacian_quadratic": laplacian_quadratic, "vadv_dense": vadv_dense}

if __name__ == "__main__":
    for name, make in FIXTURES.items():
        np.savez_compressed(HERE / f"{name}.npz", **make())
        print("wrote", HERE / f"{name}.npz")
