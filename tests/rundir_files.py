"""Synthetic run directories for the reader tests (the writer lives in the package),
and the tables ocean.x would form from them, by the oracle's restatement of
nbsemkcmel."""
import numpy as np

import ocean_oracle as orc
from ocean_amd.rundir import EV2HARTREE as EV2H
from ocean_amd.rundir import write_synthetic_rundir as write_rundir


def _printed(x, sig):
    """x as read back from a Fortran E-format field of `sig` significant digits."""
    return np.array([float("%.*E" % (sig - 1, v)) for v in np.ravel(x)]).reshape(np.shape(x))


def expected_tables(gen):
    """(rtab, ptab, mpm in Hartree) that ocean.x reads back after nbsemkcmel wrote
    rpottrim (E24.16: 16 digits) and cmel<l> (e15.8: 8 digits)."""
    v, rv = gen["rpot"]
    lmin, lmax, nproj = gen["lmin"], gen["lmax"], gen["nproj"]
    mpm = np.zeros((max(nproj), max(nproj), lmax - lmin + 1), order="F")
    vtrim = None
    for li, l in enumerate(range(lmin, lmax + 1)):
        vtrim, cmel, _ = orc.nbsemkcmel(v, rv, gen["rmax"], gen["rad"], gen["ae"][l])
        npl = nproj[li]
        mpm[:npl, :npl, li] = _printed(cmel, 8)
    return _printed(rv, 16), _printed(vtrim, 16), mpm * EV2H


__all__ = ["EV2H", "write_rundir", "expected_tables"]
