"""Shared test helpers: turn synthetic inputs into the oracle's Problem."""
import numpy as np

import ocean_oracle as orc


def to_problem(inp, e0=True, do_mult=True, long_range=True):
    m = inp.mult
    M = orc.Multiplet(m.lmin, m.lmax, list(m.nproj), m.lvl, m.lvh, m.c, m.mpm, list(m.X),
                      m.somelr, m.someli, m.have_spin_orbit)
    return orc.Problem(inp.nb, inp.kmesh, inp.xmesh, inp.lc, inp.nspn, inp.celvol,
                       inp.inter_scale, inp.u, inp.W, inp.energy, inp.allow, M,
                       e0=e0, do_mult=do_mult, long_range=long_range)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
