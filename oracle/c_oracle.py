"""ctypes wrapper of the C/OpenMP restatement (oracle/ocean_oracle.c).

TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline leg).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libocean_oracle.so")
_L = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        _L = C.CDLL(LIB)
        common = [_IP, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _IP, _DP, _DP, _DP, _DP, _DP, _DP, _DP]
        _L.oracle_xact.argtypes = common + [_DP, _DP, _DP, _DP, C.c_int, C.c_int, C.c_int]
        _L.oracle_xact.restype = C.c_int
        _L.oracle_haydock.argtypes = common + [_DP, _DP, C.c_int, _DP, _DP, _DP, C.c_int]
        _L.oracle_haydock.restype = C.c_int
        _L.oracle_max_threads.restype = C.c_int
        _L.oracle_fft3.argtypes = [_DP, _DP, _IP, C.c_int]
    return _L


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


class CSystem:
    """Holds F-ordered planes of one synthetic problem for the C restatement."""

    def __init__(self, inp, e0=True, mult=True, lr=True):
        m = inp.mult
        self.keep = []
        self.N = inp.nb * inp.nk * inp.nalpha
        self.shape = (inp.nb, inp.nk, inp.nalpha)
        mhr, mhi, _, _ = m.packed()
        arrs = dict(ur=_f(inp.u.real), ui=_f(inp.u.imag), W=_f(inp.W), er=_f(inp.energy.real),
                    ei=_f(inp.energy.imag), allow=_f(inp.allow), mpcr=_f(m.c.real), mpci=_f(m.c.imag),
                    mpm=_f(m.mpm), mhr=_f(mhr), mhi=_f(mhi), sr=_f(m.somelr), si=_f(m.someli))
        self.a = arrs
        self.dims = np.array([inp.nb, *inp.kmesh, inp.nx, inp.lc, inp.nspn, m.lmin, m.lmax,
                              max(m.nproj), m.lvl, m.lvh, int(m.have_spin_orbit), int(e0),
                              int(mult), int(lr)], dtype=np.int32)
        self.scal = np.array([inp.celvol, inp.inter_scale])
        self.nproj = np.array(m.nproj, dtype=np.int32)

    def _args(self):
        a = self.a
        P = lambda x: x.ctypes.data_as(_DP)  # noqa: E731
        return [self.dims.ctypes.data_as(_IP), P(self.scal), P(a["ur"]), P(a["ui"]), P(a["W"]),
                P(a["er"]), P(a["ei"]), P(a["allow"]), self.nproj.ctypes.data_as(_IP), P(a["mpcr"]),
                P(a["mpci"]), P(a["mpm"]), P(a["mhr"]), P(a["mhi"]), P(a["sr"]), P(a["si"])]

    def xact(self, psi, backwards=False, nthreads=None, nrep=1):
        nthreads = nthreads or lib().oracle_max_threads()
        pr, pi = _f(psi.real), _f(psi.imag)
        hr = np.zeros(self.shape, order="F")
        hi = np.zeros(self.shape, order="F")
        P = lambda x: x.ctypes.data_as(_DP)  # noqa: E731
        rc = lib().oracle_xact(*self._args(), P(pr), P(pi), P(hr), P(hi), int(backwards), nthreads, nrep)
        assert rc == 0
        return hr + 1j * hi

    def haydock(self, psi0, niter, nthreads=None):
        nthreads = nthreads or lib().oracle_max_threads()
        pr, pi = _f(psi0.real), _f(psi0.imag)
        ra, rb, ia = np.zeros(niter + 1), np.zeros(niter + 1), np.zeros(niter + 1)
        P = lambda x: x.ctypes.data_as(_DP)  # noqa: E731
        rc = lib().oracle_haydock(*self._args(), P(pr), P(pi), niter, P(ra), P(rb), P(ia), nthreads)
        return ra, rb, ia, rc


def max_threads():
    return lib().oracle_max_threads()
