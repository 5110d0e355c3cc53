"""numpy hdiff computed on the fly from shifted windows.

Counterpart of the reference's numpy ``OnTheFly`` benchmark
(stencils/numpy/horizontal_diffusion.py:39-56): it writes every point at
least two points away from the array border in i and j (all k), and uses
the oracle's operation order, so FP64 results equal the oracle's bit for
bit.  Here each Laplacian and flux needed by a point is evaluated from
windows of the input shifted relative to the output region, instead of
materialising grown temporaries.
"""

import numpy as np

from ....tools import timing
from ..base import HorizontalDiffusionStencil


class OnTheFly(HorizontalDiffusionStencil):
    @timing.return_time
    def run_stencil(self, data):
        inp, coeff, out = data
        ni, nj = inp.shape[0], inp.shape[1]

        def win(a, di=0, dj=0):
            return a[2 + di : ni - 2 + di, 2 + dj : nj - 2 + dj, :]

        def lap(di, dj):
            return 4 * win(inp, di, dj) - (
                win(inp, di + 1, dj) + win(inp, di - 1, dj) + win(inp, di, dj + 1) + win(inp, di, dj - 1)
            )

        def limited(flux, gradient):
            return np.where(flux * gradient > 0, 0, flux)

        centre = lap(0, 0)
        flx = limited(lap(1, 0) - centre, win(inp, 1, 0) - win(inp))
        flx_m = limited(centre - lap(-1, 0), win(inp) - win(inp, -1, 0))
        fly = limited(lap(0, 1) - centre, win(inp, 0, 1) - win(inp))
        fly_m = limited(centre - lap(0, -1), win(inp) - win(inp, 0, -1))
        out[2 : ni - 2, 2 : nj - 2, :] = win(inp) - win(coeff) * (flx - flx_m + fly - fly_m)
