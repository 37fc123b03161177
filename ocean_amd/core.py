"""ctypes binding of libocean_gpu.so and a Python mirror of OCEAN2's call surface.

    OCEAN_xact(sys, inter_scale, psi, new_psi, ierr, backwards, nhflag)
        OCEAN2/OCEAN_action.f90:286            -> OceanGPU.xact
    OCEAN_haydock_do(sys, hay_vec, restartBSE, newEps, ierr)
        OCEAN2/OCEAN_haydock.f90:167           -> OceanGPU.haydock
    OCEAN_gmres_do_recycle(sys, hay_vec, ierr)
        OCEAN2/OCEAN_gmres.f90:508             -> OceanGPU.gmres
    write_core(fh, iter, kpref, ...)
        OCEAN2/OCEAN_haydock.f90:1801          -> spectrum

Errors raise ``OceanGPUError`` carrying the library's integer code (the
reference's ierr).  There is no CPU fallback: without the built library every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libocean_gpu.so")
EV2HARTREE = 0.036749322175666
HARTREE2EV = 27.211386245980

FP64 = 0
MIXED = 1


class OceanGPUError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = _lib().ocean_gpu_strerror(code).decode() if _LIB is not None else str(code)
        super().__init__(f"{where}: ierr={code} ({msg})")


class Desc(C.Structure):
    _fields_ = [
        ("num_bands", C.c_int32), ("kmesh", C.c_int32 * 3), ("xmesh", C.c_int32 * 3),
        ("nspn", C.c_int32), ("lc", C.c_int32), ("lmin", C.c_int32), ("lmax", C.c_int32),
        ("nproj", C.c_int32 * 8), ("lvl", C.c_int32), ("lvh", C.c_int32),
        ("celvol", C.c_double), ("inter_scale", C.c_double), ("e0", C.c_int32),
        ("mult", C.c_int32), ("long_range", C.c_int32), ("precision", C.c_int32),
    ]


_LIB = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_H = C.c_void_p

_SIGS = {
    "ocean_gpu_create": (C.c_int, [C.POINTER(Desc), C.POINTER(_H)]),
    "ocean_gpu_create_multi": (C.c_int, [C.POINTER(Desc), C.c_int32, C.POINTER(_H)]),
    "ocean_gpu_nrhs": (C.c_int, [_H, _IP]),
    "ocean_gpu_haydock_multi": (C.c_int, [_H, _DP, _DP, C.c_int32, _DP, _DP, _DP, _IP]),
    "ocean_gpu_set_tiles": (C.c_int, [_H, _IP]),
    "ocean_gpu_destroy": (C.c_int, [_H]),
    "ocean_gpu_strerror": (C.c_char_p, [C.c_int]),
    "ocean_gpu_set_precision": (C.c_int, [_H, C.c_int32]),
    "ocean_gpu_load_bloch": (C.c_int, [_H, _DP, _DP]),
    "ocean_gpu_load_bloch_raw": (C.c_int, [_H, _DP, _DP, _IP]),
    "ocean_gpu_load_obf": (C.c_int, [_H, C.c_int32, _DP, _DP, _DP, _DP, _IP]),
    "ocean_gpu_load_w": (C.c_int, [_H, _DP]),
    "ocean_gpu_build_w": (C.c_int, [_H, _DP, _DP, _DP, C.c_int32, C.c_double, C.c_double]),
    "ocean_gpu_load_mult": (C.c_int, [_H, _DP, _DP, _DP, _DP, _DP, _DP, _DP, C.c_int32]),
    "ocean_gpu_load_energies": (C.c_int, [_H, _DP, _DP, _DP]),
    "ocean_gpu_mult_setup": (C.c_int, [_H, _DP, C.c_double, C.c_int32, C.c_double, _DP, _DP, _DP, _DP, _DP,
                                       _IP]),
    "ocean_gpu_xact": (C.c_int, [_H, _DP, _DP, _DP, _DP, C.c_int32, _IP]),
    "ocean_gpu_vec_new": (C.c_int, [_H, _IP]),
    "ocean_gpu_vec_free": (C.c_int, [_H, C.c_int32]),
    "ocean_gpu_vec_put": (C.c_int, [_H, C.c_int32, _DP, _DP]),
    "ocean_gpu_vec_get": (C.c_int, [_H, C.c_int32, _DP, _DP]),
    "ocean_gpu_vec_dot": (C.c_int, [_H, C.c_int32, C.c_int32, _DP, _DP]),
    "ocean_gpu_vec_nrm": (C.c_int, [_H, C.c_int32, _DP]),
    "ocean_gpu_vec_axpy": (C.c_int, [_H, C.c_double, C.c_double, C.c_int32, C.c_int32]),
    "ocean_gpu_vec_scal": (C.c_int, [_H, C.c_double, C.c_int32]),
    "ocean_gpu_vec_copy": (C.c_int, [_H, C.c_int32, C.c_int32]),
    "ocean_gpu_vec_axmy": (C.c_int, [_H, C.c_double, C.c_double, C.c_int32, C.c_int32]),
    "ocean_gpu_vec_axmz": (C.c_int, [_H, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32]),
    "ocean_gpu_vec_divide": (C.c_int, [_H, C.c_int32, C.c_double, C.c_double, C.c_int32]),
    "ocean_gpu_xact_vec": (C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32, _IP]),
    "ocean_gpu_xact_vec_repeat": (C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32]),
    "ocean_gpu_profile_xact": (C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32]),
    "ocean_gpu_haydock": (C.c_int, [_H, _DP, _DP, C.c_int32, _DP, _DP, _DP, _IP]),
    "ocean_gpu_haydock_early": (C.c_int, [_H, _DP, _DP, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_int32, _DP, _DP,
                                          _DP, _IP]),
    "ocean_gpu_gmres": (C.c_int, [_H, _DP, _DP, C.c_int32, _DP, C.c_double, C.c_double,
                                  C.c_double, C.c_int32, C.c_double, _DP, _DP, _DP]),
    "ocean_gpu_spectrum": (C.c_int, [_DP, _DP, _DP, C.c_int32, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int32, C.c_int32, _DP]),
    "ocean_gpu_haydock_abc": (C.c_int, [_H, _DP, _DP, C.c_int32, _DP, _DP, _DP, _DP, _DP, _DP, _IP]),
    "ocean_gpu_spectrum_abc": (C.c_int, [_DP, _DP, _DP, _DP, _DP, _DP, C.c_int32, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int32, C.c_int32, _DP]),
    "ocean_gpu_spectrum_noalloc": (C.c_int, [_DP, _DP, _DP, C.c_int32, C.c_double, C.c_double,
                                             C.c_double, C.c_double, C.c_int32, C.c_int32, _DP]),
    "ocean_gpu_write_core": (C.c_int, [_DP, _DP, _DP, C.c_int32, C.c_double, C.c_void_p, _DP, _IP]),
    "ocean_gpu_tridiag_eigs": (C.c_int, [_DP, _DP, C.c_int32, _DP]),
    "ocean_gpu_ospect": (C.c_int, [_DP, _DP, C.c_int32, C.c_double, C.c_void_p, _DP, _IP, _IP]),
    "ocean_gpu_last_timing": (C.c_int, [_H, C.c_char_p, _DP]),
    "ocean_gpu_autotune": (C.c_int, [_H, C.c_int32]),
    "ocean_gpu_get_tiles": (C.c_int, [_H, _IP]),
    "ocean_gpu_load_bloch_dev": (C.c_int, [_H, C.c_void_p, C.c_void_p]),
    "ocean_gpu_spectral_grad": (C.c_int, [_H, _DP, _DP, C.c_int32, _DP, C.c_double, C.c_double,
                                          C.c_double, C.c_int32, C.c_double, _DP, _DP, _DP, _DP,
                                          _DP, _DP, _DP]),
    "ocean_gpu_batch_create": (C.c_int, [C.POINTER(_H), C.c_int32, C.POINTER(_H)]),
    "ocean_gpu_batch_destroy": (C.c_int, [_H]),
    "ocean_gpu_batch_ngroups": (C.c_int, [_H, _IP]),
    "ocean_gpu_batch_group_info": (C.c_int, [_H, C.c_int32, _IP, _DP]),
    "ocean_gpu_batch_autotune": (C.c_int, [_H, C.c_int32]),
    "ocean_gpu_batch_xact": (C.c_int, [_H, _IP, _IP, C.c_int32]),
    "ocean_gpu_batch_last_ms": (C.c_int, [_H, _DP]),
    "ocean_gpu_batch_haydock": (C.c_int, [_H, C.POINTER(_DP), C.POINTER(_DP), C.c_int32,
                                          C.POINTER(_DP), C.POINTER(_DP), C.POINTER(_DP), _IP]),
    "ocean_gpu_batch_gmres": (C.c_int, [_H, C.POINTER(_DP), C.POINTER(_DP), _IP, C.POINTER(_DP), C.c_double,
                                        C.c_double, C.c_double, C.c_int32, _DP, C.POINTER(_DP), _IP]),
}


def _lib():
    """Load libocean_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m ocean_amd.build` "
                               "(the GPU path has no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def exported_symbols():
    return list(_SIGS)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), "need F-ordered float64"
    return a.ctypes.data_as(_DP)


def _planes(z, shape=None, name="array"):
    z = np.asarray(z)
    if shape is not None:
        _need(z, shape, name)
    return (np.asfortranarray(z.real, dtype=np.float64), np.asfortranarray(z.imag, dtype=np.float64))


def _need(a, shape, name):
    """The library reads and writes exactly prod(shape) elements through raw pointers:
    refuse anything else before the call."""
    if tuple(np.shape(a)) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(np.shape(a))}, expected {tuple(shape)}")


def _chk(rc, where):
    if rc != 0:
        raise OceanGPUError(rc, where)


def spectrum(real_a, real_b, niter, kpref, gam, el, eh, ne, imag_a=None, literal_cmplx=True):
    """write_core continued fraction on the GPU; returns (ne+1, 5) columns."""
    lib = _lib()
    ra = np.ascontiguousarray(real_a, np.float64)
    rb = np.ascontiguousarray(real_b, np.float64)
    ia = None if imag_a is None else np.ascontiguousarray(imag_a, np.float64)
    out = np.zeros((ne + 1, 5))
    _chk(lib.ocean_gpu_spectrum(_p(ra), _p(rb), _p(ia), niter, kpref, gam, el, eh, ne,
                                1 if literal_cmplx else 0, out.ctypes.data_as(_DP)), "spectrum")
    return out


def spectrum_abc(real_a, imag_a, real_b, imag_b, real_c, imag_c, niter, kpref, gam, el, eh, ne,
                 literal_cmplx=True):
    """write_core's fraction with complex b and c (the bi-Lanczos coefficients of
    OceanGPU.haydock_abc) on the GPU; returns (ne+1, 5) columns."""
    lib = _lib()
    arrs = [np.ascontiguousarray(v, np.float64) for v in (real_a, imag_a, real_b, imag_b, real_c, imag_c)]
    for v in arrs:
        if v.shape[0] < niter + 1:
            raise ValueError("coefficient arrays need niter + 1 entries")
    out = np.zeros((ne + 1, 5))
    _chk(lib.ocean_gpu_spectrum_abc(*[_p(v) for v in arrs], niter, kpref, gam, el, eh, ne,
                                    1 if literal_cmplx else 0, out.ctypes.data_as(_DP)), "spectrum_abc")
    return out


class SpectDesc(C.Structure):
    """ocean_spect_desc of include/ocean_gpu.h."""
    _fields_ = [("gam0", C.c_double), ("el", C.c_double), ("eh", C.c_double), ("ne", C.c_int32),
                ("ebase_ev", C.c_double), ("gam_gauss", C.c_double), ("old_broaden", C.c_int32),
                ("nval", C.c_double), ("eps", C.c_double), ("ucvol", C.c_double),
                ("literal_cmplx", C.c_int32)]


def write_core(real_a, real_b, niter, kpref, gam0, el, eh, ne, ebase_eV=0.0, old_broaden=False,
               nval=0.0, eps=1.0, ucvol=1.0, gam_gauss=0.0, imag_a=None, literal_cmplx=True):
    """write_core (nine columns, or the five of its Voigt form) on the GPU."""
    lib = _lib()
    ra = np.ascontiguousarray(real_a, np.float64)
    rb = np.ascontiguousarray(real_b, np.float64)
    ia = None if imag_a is None else np.ascontiguousarray(imag_a, np.float64)
    sd = SpectDesc(gam0, el, eh, ne, ebase_eV, gam_gauss, 1 if old_broaden else 0, nval, eps, ucvol,
                   1 if literal_cmplx else 0)
    out = np.zeros(9 * (ne + 1))
    ncols = C.c_int32(0)
    _chk(lib.ocean_gpu_write_core(_p(ra), _p(rb), _p(ia), niter, kpref, C.byref(sd),
                                  out.ctypes.data_as(_DP), C.byref(ncols)), "write_core")
    nc = ncols.value
    return out[:nc * (ne + 1)].reshape(ne + 1, nc)


class OspectDesc(C.Structure):
    """ocean_ospect_desc of include/ocean_gpu.h."""
    _fields_ = [("ne", C.c_int32), ("el_ev", C.c_double), ("eh_ev", C.c_double), ("gam0_ev", C.c_double),
                ("ebase_ev", C.c_double), ("eps0", C.c_double), ("nval", C.c_double), ("ucvol", C.c_double),
                ("is_val", C.c_int32)]


def ospect(real_a, real_b, it, kpref, ne, el_eV, eh_eV, gam0_eV, ebase_eV=0.0, eps0=1.0, nval=0.0,
           ucvol=1.0, is_val=False):
    """o_spect.x's spectrum for one Lanczos run on the GPU (rows as o_spect writes them)."""
    lib = _lib()
    ra = np.ascontiguousarray(real_a, np.float64)
    rb = np.ascontiguousarray(real_b, np.float64)
    assert ra.size >= it + 1 and rb.size >= it + 1
    od = OspectDesc(ne, el_eV, eh_eV, gam0_eV, ebase_eV, eps0, nval, ucvol, 1 if is_val else 0)
    out = np.zeros(8 * (ne + 1))
    nr, nc = C.c_int32(0), C.c_int32(0)
    _chk(lib.ocean_gpu_ospect(_p(ra), _p(rb), it, kpref, C.byref(od), out.ctypes.data_as(_DP), C.byref(nr),
                              C.byref(nc)), "ospect")
    return out[:nr.value * nc.value].reshape(nr.value, nc.value)


def tridiag_eigs(real_a, real_b, n):
    """Ascending eigenvalues of the (n+1)-square Lanczos matrix (a(0:n), b(1:n))."""
    lib = _lib()
    ra = np.ascontiguousarray(real_a, np.float64)
    rb = np.ascontiguousarray(real_b, np.float64)
    assert ra.size >= n + 1 and rb.size >= n + 1
    w = np.zeros(n + 1)
    _chk(lib.ocean_gpu_tridiag_eigs(_p(ra), _p(rb), n, w.ctypes.data_as(_DP)), "tridiag_eigs")
    return w


class OceanGPU:
    """One device-resident core-level BSE problem (the handle of the C-ABI)."""

    def __init__(self, nb, kmesh, xmesh, lc, nspn, celvol, inter_scale=1.0, lmin=0, lmax=3,
                 nproj=(1, 1, 1, 1), lvl=None, lvh=None, e0=True, mult=True, long_range=True,
                 precision=FP64, nrhs=1):
        """nrhs > 1: a handle whose vectors stack nrhs right-hand sides sharing this
        Hamiltonian (ocean_gpu_create_multi); energies and multiplet inputs are one
        right-hand side's (shape0), vectors are (nb, nk, nrhs * nalpha)."""
        lib = _lib()
        d = Desc()
        d.num_bands = nb
        d.kmesh[:] = list(kmesh)
        d.xmesh[:] = list(xmesh)
        d.nspn = nspn
        d.lc = lc
        d.lmin, d.lmax = lmin, lmax
        npj = list(nproj) + [0] * (8 - len(nproj))
        d.nproj[:] = npj[:8]
        d.lvl = lmin if lvl is None else lvl
        d.lvh = lmax if lvh is None else lvh
        d.celvol = celvol
        d.inter_scale = inter_scale
        d.e0, d.mult, d.long_range = int(e0), int(mult), int(long_range)
        d.precision = precision
        self.desc = d
        self.nb, self.nk, self.nx = nb, int(np.prod(kmesh)), int(np.prod(xmesh))
        self.na0 = 4 * (2 * lc + 1)
        self.nrhs = int(nrhs)
        self.na = self.nrhs * self.na0
        self.nspn = nspn
        self.shape = (nb, self.nk, self.na)
        self.shape0 = (nb, self.nk, self.na0)
        h = _H()
        if self.nrhs == 1:
            _chk(lib.ocean_gpu_create(C.byref(d), C.byref(h)), "ocean_gpu_create")
        else:
            _chk(lib.ocean_gpu_create_multi(C.byref(d), self.nrhs, C.byref(h)), "ocean_gpu_create_multi")
        self.h = h
        self.kpref = None

    @classmethod
    def from_inputs(cls, inp, precision=FP64, e0=True, mult=True, long_range=True, nrhs=1):
        """Create a handle and load every piece of state from synthetic inputs."""
        m = inp.mult
        g = cls(inp.nb, inp.kmesh, inp.xmesh, inp.lc, inp.nspn, inp.celvol, inp.inter_scale,
                m.lmin, m.lmax, tuple(m.nproj), m.lvl, m.lvh, e0, mult, long_range, precision, nrhs)
        g.load_inputs(inp)
        return g

    def load_inputs(self, inp):
        if inp.u is not None:  # else the caller loads u itself (load_bloch_dev)
            self.load_bloch(inp.u)
        self.load_w(inp.W)
        self.load_mult(inp.mult)
        self.load_energies(inp.energy, inp.allow)
        self.kpref = inp.kpref

    def close(self):
        if getattr(self, "h", None):
            _lib().ocean_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- loads (module state) ----
    def set_precision(self, precision):
        _chk(_lib().ocean_gpu_set_precision(self.h, precision), "set_precision")

    def load_bloch(self, u):
        re, im = _planes(u, (self.nb, self.nk, self.nx, self.nspn), "u")
        _chk(_lib().ocean_gpu_load_bloch(self.h, _p(re), _p(im)), "load_bloch")

    def load_bloch_planes(self, re, im):
        """u as two F-ordered float64 planes (nb, nk, nx, nspn)."""
        for a, nm in ((re, "u_re"), (im, "u_im")):
            _need(a, (self.nb, self.nk, self.nx, self.nspn), nm)
        _chk(_lib().ocean_gpu_load_bloch(self.h, _p(re), _p(im)), "load_bloch")

    def load_bloch_dev(self, re_ptr, im_ptr):
        """u from two device planes (integer device addresses, same layout as load_bloch)."""
        _chk(_lib().ocean_gpu_load_bloch_dev(self.h, C.c_void_p(re_ptr), C.c_void_p(im_ptr)),
             "load_bloch_dev")

    def load_bloch_raw(self, u_raw, xoff):
        re, im = _planes(u_raw, (self.nb, self.nk, self.nx, self.nspn), "u_raw")
        xo = np.ascontiguousarray(xoff, np.int32)
        _need(xo, (self.nx, 3), "xoff")
        _chk(_lib().ocean_gpu_load_bloch_raw(self.h, _p(re), _p(im), xo.ctypes.data_as(_IP)),
             "load_bloch_raw")

    def load_obf(self, obf, obf2u, xoff):
        """The OBF form (ocean_gpu_load_obf): obf[o, x], obf2u[n, o, k] complex, xoff
        [x, 3] from setup_host.obf_x_offsets.  The long-range term then runs
        lr_act_obf2 (OCEAN2/long_range.f90:262-455) until load_bloch* is called."""
        obf = np.asarray(obf, np.complex128)
        no = obf.shape[0]
        fre = np.asfortranarray(obf.real)
        fim = np.asfortranarray(obf.imag)
        c = np.asarray(obf2u, np.complex128)
        if c.shape != (self.nb, no, self.nk) or obf.shape != (no, self.nx):
            raise ValueError("obf must be (num_obf, nx) and obf2u (nb, num_obf, nk)")
        cre = np.asfortranarray(c.real)
        cim = np.asfortranarray(c.imag)
        xo = np.ascontiguousarray(xoff, np.int32)
        if xo.shape != (self.nx, 3):
            raise ValueError("xoff must be (nx, 3)")
        _chk(_lib().ocean_gpu_load_obf(self.h, no, _p(fre), _p(fim), _p(cre), _p(cim),
                                       xo.ctypes.data_as(_IP)), "load_obf")

    def load_w(self, W):
        _need(W, (self.nk, self.nx), "W")
        Wf = np.asfortranarray(W, np.float64)
        _chk(_lib().ocean_gpu_load_w(self.h, _p(Wf)), "load_w")

    def build_w(self, rdist, rtab, ptab, epsilon0, scale=1.0):
        _need(rdist, (self.nk, self.nx), "rdist")
        rd = np.asfortranarray(rdist, np.float64)
        rt = np.ascontiguousarray(rtab, np.float64)
        pt = np.ascontiguousarray(ptab, np.float64)
        if rt.ndim != 1 or pt.shape != rt.shape:
            raise ValueError("rtab and ptab must be 1-D of the same length")
        self.nptab = len(rt)
        _chk(_lib().ocean_gpu_build_w(self.h, _p(rd), _p(rt), _p(pt), len(rt), epsilon0, scale),
             "build_w")

    def mult_setup(self, slater, scfac, xi_eV, life_time_eV=None, tdlda=False):
        """Interaction blocks on the device (ocean_gpu_mult_setup).  slater: per valence
        l = lvl..lvh a pair (fk, gk) of dicts k -> (nproj, nproj) matrices already
        multiplied by their scale factors; life_time_eV the two values of
        core_broaden.ipt or None; tdlda: fk holds the k = 0 kk matrix instead of the
        F^k.  Returns (X list of (mham, mham) complex blocks in Hartree, somelr,
        someli, have_so)."""
        d = self.desc
        lc = d.lc
        flat, sizes = [], []
        for i, l in enumerate(range(d.lvl, d.lvh + 1)):
            npl = d.nproj[l - d.lmin] if d.lmin <= l <= d.lmax else 0
            fk, gk = slater[i]
            if npl:
                kf = [0] if tdlda else list(range(2, min(2 * lc, 2 * l) + 1, 2))
                ks = [("f", k) for k in kf] + \
                     [("g", k) for k in range(abs(lc - l), lc + l + 1, 2)]
                for which, k in ks:
                    m = (fk if which == "f" else gk)[k]
                    flat.append(np.asfortranarray(m, np.float64).ravel(order="F"))
            sizes.append(4 * (2 * lc + 1) * (2 * l + 1) * npl)
        sl = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(1))
        ntot = max(sum(m * m for m in sizes), 1)
        mhr, mhi = np.zeros(ntot), np.zeros(ntot)
        na = self.na0
        sr, si = np.zeros((na, na), order="F"), np.zeros((na, na), order="F")
        hso = C.c_int32()
        lt = None if life_time_eV is None else np.ascontiguousarray(life_time_eV, np.float64).ravel()
        if lt is not None and lt.size < 2:
            raise ValueError("life_time_eV needs the two values of core_broaden.ipt")
        lt = None if lt is None else np.ascontiguousarray(lt[:2])
        _chk(_lib().ocean_gpu_mult_setup(self.h, _p(sl), scfac, int(bool(tdlda)), xi_eV,
                                         None if lt is None else _p(lt), _p(mhr),
                                         _p(mhi), _p(sr), _p(si), C.byref(hso)), "mult_setup")
        X, j = [], 0
        for m in sizes:
            X.append(np.asfortranarray((mhr[j:j + m * m] + 1j * mhi[j:j + m * m]).reshape(m, m, order="F")))
            j += m * m
        return X, sr, si, bool(hso.value)

    def load_mult(self, m):
        npmax = max(m.nproj)
        nl = m.lmax - m.lmin + 1
        c = m.c
        _need(c, (self.nb, self.nk, npmax, 2 * m.lmax + 1, nl, self.nspn), "projector array c")
        _need(m.mpm, (npmax, npmax, nl), "mpm")
        _need(m.somelr, (self.na0, self.na0), "somelr")
        _need(m.someli, (self.na0, self.na0), "someli")
        d = self.desc
        if (m.lmin, m.lmax, m.lvl, m.lvh) != (d.lmin, d.lmax, d.lvl, d.lvh) or \
                list(m.nproj) != list(d.nproj[:nl]):
            raise ValueError("multiplet l ranges / projector counts differ from the handle's")
        for l, Xl in zip(range(m.lvl, m.lvh + 1), m.X):
            npl = m.nproj[l - m.lmin] if m.lmin <= l <= m.lmax else 0
            mham = 4 * (2 * d.lc + 1) * (2 * l + 1) * npl
            _need(Xl, (mham, mham), f"X[l={l}]")
        cr, ci = _planes(c)
        mpm = np.asfortranarray(m.mpm, np.float64)
        mhr, mhi, _, _ = m.packed()
        sr = np.asfortranarray(m.somelr, np.float64)
        si = np.asfortranarray(m.someli, np.float64)
        _chk(_lib().ocean_gpu_load_mult(self.h, _p(cr), _p(ci), _p(mpm), _p(mhr), _p(mhi), _p(sr),
                                        _p(si), int(bool(m.have_spin_orbit))), "load_mult")

    def load_energies(self, energy, allow):
        """One right-hand side's energies and allow mask (shape0), shared by all."""
        er, ei = _planes(energy, self.shape0, "energy")
        _need(allow, self.shape0, "allow")
        al = np.asfortranarray(allow, np.float64)
        _chk(_lib().ocean_gpu_load_energies(self.h, _p(er), _p(ei), _p(al)), "load_energies")

    # ---- the action ----
    def xact(self, psi, backwards=False, hflag=None):
        pr, pi = _planes(psi, self.shape, "psi")
        hr = np.zeros(self.shape, order="F")
        hi = np.zeros(self.shape, order="F")
        hf = None if hflag is None else np.ascontiguousarray(hflag, np.int32)
        if hf is not None:
            _need(hf, (6,), "hflag")
        _chk(_lib().ocean_gpu_xact(self.h, _p(pr), _p(pi), _p(hr), _p(hi), int(bool(backwards)),
                                   None if hf is None else hf.ctypes.data_as(_IP)), "xact")
        return hr + 1j * hi

    # ---- vectors ----
    def vec_new(self, z=None):
        i = C.c_int32()
        _chk(_lib().ocean_gpu_vec_new(self.h, C.byref(i)), "vec_new")
        if z is not None:
            self.vec_put(i.value, z)
        return i.value

    def vec_put(self, vid, z):
        re, im = _planes(z, self.shape, "vector")
        _chk(_lib().ocean_gpu_vec_put(self.h, vid, _p(re), _p(im)), "vec_put")

    def vec_get(self, vid):
        re = np.zeros(self.shape, order="F")
        im = np.zeros(self.shape, order="F")
        _chk(_lib().ocean_gpu_vec_get(self.h, vid, _p(re), _p(im)), "vec_get")
        return re + 1j * im

    def vec_dot(self, p, q):
        r, i = C.c_double(), C.c_double()
        _chk(_lib().ocean_gpu_vec_dot(self.h, p, q, C.byref(r), C.byref(i)), "vec_dot")
        return r.value, i.value

    def vec_nrm(self, p):
        v = C.c_double()
        _chk(_lib().ocean_gpu_vec_nrm(self.h, p, C.byref(v)), "vec_nrm")
        return v.value

    def vec_axpy(self, a, x, y):
        a = complex(a)
        _chk(_lib().ocean_gpu_vec_axpy(self.h, a.real, a.imag, x, y), "vec_axpy")

    def vec_scal(self, a, x):
        _chk(_lib().ocean_gpu_vec_scal(self.h, float(a), x), "vec_scal")

    def vec_copy(self, dst, src):
        _chk(_lib().ocean_gpu_vec_copy(self.h, dst, src), "vec_copy")

    def vec_axmy(self, a, x, y):
        """y = a x - y (OCEAN_psi_axmy)."""
        _chk(_lib().ocean_gpu_vec_axmy(self.h, a.real, a.imag, x, y), "vec_axmy")

    def vec_axmz(self, a, x, y, z):
        """y = a x - z (OCEAN_psi_axmz)."""
        _chk(_lib().ocean_gpu_vec_axmz(self.h, a.real, a.imag, x, y, z), "vec_axmz")

    def vec_divide(self, x, rval=None, ival=None):
        """x = x / (rval + i ival) with the reference's optional-argument branches."""
        which = (1 if rval is not None else 0) + (2 if ival is not None else 0)
        _chk(_lib().ocean_gpu_vec_divide(self.h, x, rval or 0.0, ival or 0.0, which), "vec_divide")

    def xact_vec(self, vin, vout, backwards=False, hflag=None):
        hf = None if hflag is None else np.ascontiguousarray(hflag, np.int32)
        if hf is not None:
            _need(hf, (6,), "hflag")
        _chk(_lib().ocean_gpu_xact_vec(self.h, vin, vout, int(bool(backwards)),
                                       None if hf is None else hf.ctypes.data_as(_IP)), "xact_vec")

    def xact_repeat(self, vin, vout, nrep):
        """nrep back-to-back actions; returns the event-timed ms per action."""
        _chk(_lib().ocean_gpu_xact_vec_repeat(self.h, vin, vout, nrep), "xact_vec_repeat")
        return self.timing("total")

    def autotune(self, reps=5):
        """Tile shapes of this problem's action picked by the autotuner; returns them."""
        _chk(_lib().ocean_gpu_autotune(self.h, reps), "autotune")
        return self.tiles()

    def set_tiles(self, proj, adj, G=0, adj_mf=-1):
        t = np.array([proj, adj, G, adj_mf], np.int32)
        _chk(_lib().ocean_gpu_set_tiles(self.h, t.ctypes.data_as(_IP)), "set_tiles")

    def tiles(self):
        t = np.zeros(4, np.int32)
        _chk(_lib().ocean_gpu_get_tiles(self.h, t.ctypes.data_as(_IP)), "get_tiles")
        return tuple(int(v) for v in t)

    def timing(self, name):
        v = C.c_double()
        _chk(_lib().ocean_gpu_last_timing(self.h, name.encode(), C.byref(v)), "last_timing")
        return v.value

    def profile_xact(self, vin, vout, nrep):
        """Per-kernel-family average durations (ms) with events between the families."""
        _chk(_lib().ocean_gpu_profile_xact(self.h, vin, vout, nrep), "profile_xact")
        out = {}
        for k in ("proj", "fft", "adj", "mult"):
            v = C.c_double()
            _chk(_lib().ocean_gpu_last_timing(self.h, k.encode(), C.byref(v)), "last_timing")
            out[k] = v.value
        return out

    # ---- solvers ----
    def haydock(self, psi0, niter, return_status=False):
        """Returns (real_a, real_b, imag_a) with niter+1 entries each (index 0..niter).
        With return_status the library's code is not raised: (a, b, ai, niter_done, rc)
        (rc = -1 and niter_done = the failing iteration on a NaN, as the reference's
        OCEAN_hay_ab_twoterm, OCEAN2/OCEAN_haydock.f90:912-920)."""
        pr, pi = _planes(psi0, self.shape, "psi0")
        a = np.zeros(niter + 1)
        b = np.zeros(niter + 1)
        ai = np.zeros(niter + 1)
        done = C.c_int32()
        rc = _lib().ocean_gpu_haydock(self.h, _p(pr), _p(pi), niter, _p(a), _p(b), _p(ai),
                                      C.byref(done))
        if return_status:
            return a, b, ai, done.value, rc
        _chk(rc, "haydock")
        return a, b, ai

    def haydock_abc(self, psi0, niter, return_status=False):
        """Non-Hermitian (bi-Lanczos) recursion (ocean_gpu_haydock_abc).  Returns
        (real_a, imag_a, real_b, imag_b, real_c, imag_c), niter+1 entries each; with
        return_status also (niter_done, rc) instead of raising."""
        pr, pi = _planes(psi0, self.shape, "psi0")
        outs = [np.zeros(niter + 1) for _ in range(6)]
        done = C.c_int32()
        rc = _lib().ocean_gpu_haydock_abc(self.h, _p(pr), _p(pi), niter, *[_p(v) for v in outs],
                                          C.byref(done))
        if return_status:
            return (*outs, done.value, rc)
        _chk(rc, "haydock_abc")
        return tuple(outs)

    def haydock_multi(self, psi0s, niter, return_status=False):
        """The recursions of all nrhs right-hand sides at once (ocean_gpu_haydock_multi).
        psi0s: (nrhs, nb, nk, nalpha) starting vectors, or the stacked (nb, nk, nrhs *
        nalpha) vector.  Returns (a, b, ai), each (nrhs, niter + 1); with return_status
        (a, b, ai, niter_done[nrhs], rc)."""
        z = np.asarray(psi0s)
        if z.shape == (self.nrhs,) + self.shape0:
            z = np.concatenate(list(z), axis=2)
        pr, pi = _planes(z, self.shape, "psi0s")
        R, n1 = self.nrhs, niter + 1
        a, b, ai = np.zeros(R * n1), np.zeros(R * n1), np.zeros(R * n1)
        done = np.zeros(R, np.int32)
        rc = _lib().ocean_gpu_haydock_multi(self.h, _p(pr), _p(pi), niter, _p(a), _p(b), _p(ai),
                                            done.ctypes.data_as(_IP))
        a, b, ai = a.reshape(R, n1), b.reshape(R, n1), ai.reshape(R, n1)
        if return_status:
            return a, b, ai, done, rc
        _chk(rc, "haydock_multi")
        return a, b, ai

    def spectral_grad(self, hay_vec, energies_eV, gres_eV, gprc_eV, ffff, depth, target, kpref=None,
                      ptab=True, scale=True, gamma=True, scissor=True):
        """Misfit L = sum_E (sigma(E) - target)^2 of the GMRES spectrum and its gradient
        (ocean_gpu_spectral_grad): dict(loss, sigma, g_ptab, g_scale, g_gamma_eV, g_scissor_eV)."""
        kpref = self.kpref if kpref is None else kpref
        pr, pi = _planes(hay_vec, self.shape, "hay_vec")
        E = np.ascontiguousarray(energies_eV, np.float64)
        tg = np.ascontiguousarray(target, np.float64)
        if E.ndim != 1 or tg.shape != E.shape:
            raise ValueError("energies and target must be 1-D of the same length")
        sig = np.zeros(len(E))
        loss = C.c_double()
        gp = np.zeros(getattr(self, "nptab", 0)) if ptab else None
        gs, gg, gd = C.c_double(), C.c_double(), C.c_double()
        rc = _lib().ocean_gpu_spectral_grad(
            self.h, _p(pr), _p(pi), len(E), _p(E), gres_eV, gprc_eV, ffff, depth, kpref, _p(tg),
            C.byref(loss), _p(sig), _p(gp), C.byref(gs) if scale else None,
            C.byref(gg) if gamma else None, C.byref(gd) if scissor else None)
        _chk(rc, "spectral_grad")
        return dict(loss=loss.value, sigma=sig, g_ptab=gp, g_scale=gs.value if scale else None,
                    g_gamma_eV=gg.value if gamma else None, g_scissor_eV=gd.value if scissor else None)

    def haydock_early(self, psi0, niter, spacing, threshold, gam, el, eh, ne, kpref=None):
        """Haydock with the reference's early exit; returns (a, b, ai, niter_done)."""
        kpref = self.kpref if kpref is None else kpref
        pr, pi = _planes(psi0, self.shape, "psi0")
        a, b, ai = np.zeros(niter + 1), np.zeros(niter + 1), np.zeros(niter + 1)
        done = C.c_int32()
        _chk(_lib().ocean_gpu_haydock_early(self.h, _p(pr), _p(pi), niter, spacing, threshold, kpref,
                                            gam, el, eh, ne, _p(a), _p(b), _p(ai), C.byref(done)),
             "haydock_early")
        return a, b, ai, done.value

    def gmres(self, hay_vec, energies_eV, gres_eV, gprc_eV, ffff, depth, kpref=None,
              return_x=False, return_status=False):
        """Rows (E_eV, 1 - Re<b|x> kpref, -Im<b|x> kpref, rel_err, iters, global_iters).
        With return_status the library's code is returned beside the rows instead of
        raised."""
        kpref = self.kpref if kpref is None else kpref
        pr, pi = _planes(hay_vec, self.shape, "hay_vec")
        E = np.ascontiguousarray(energies_eV, np.float64)
        if E.ndim != 1:
            raise ValueError("energies must be 1-D")
        out = np.zeros((len(E), 6))
        xr = xi = None
        if return_x:
            xr = np.zeros(self.shape + (len(E),), order="F")
            xi = np.zeros(self.shape + (len(E),), order="F")
        rc = _lib().ocean_gpu_gmres(self.h, _p(pr), _p(pi), len(E), _p(E), gres_eV, gprc_eV, ffff,
                                    depth, kpref, out.ctypes.data_as(_DP), _p(xr), _p(xi))
        if return_status:
            return (out, xr + 1j * xi, rc) if return_x else (out, rc)
        _chk(rc, "gmres")
        if return_x:
            return out, xr + 1j * xi
        return out


class OceanBatch:
    """Many OceanGPU problems whose actions and Haydock recursions run together
    (one launch per kernel stage per group; see ocean_gpu_batch_* in the header).
    The problems must stay open while the batch is used."""

    def __init__(self, problems):
        self.problems = list(problems)
        n = len(self.problems)
        arr = (_H * n)(*[p.h for p in self.problems])
        self.b = _H()
        _chk(_lib().ocean_gpu_batch_create(arr, n, C.byref(self.b)), "batch_create")

    def close(self):
        if self.b:
            _lib().ocean_gpu_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def groups(self):
        """[(problems, precision, alpha template, proj tile, adj tile, G, adjoint form, tune_ms[3]), ...]"""
        ng = C.c_int32()
        _chk(_lib().ocean_gpu_batch_ngroups(self.b, C.byref(ng)), "batch_ngroups")
        out = []
        for g in range(ng.value):
            info = np.zeros(7, np.int32)
            ms = np.zeros(3)
            _chk(_lib().ocean_gpu_batch_group_info(self.b, g, info.ctypes.data_as(_IP), _p(ms)),
                 "batch_group_info")
            out.append(tuple(int(v) for v in info) + (ms.tolist(),))
        return out

    def autotune(self, reps=5):
        _chk(_lib().ocean_gpu_batch_autotune(self.b, reps), "batch_autotune")
        return self.groups()

    def xact(self, in_ids, out_ids, nrep=1):
        """nrep batched actions; returns ms per batched action (all problems)."""
        ii = np.ascontiguousarray(in_ids, np.int32)
        oo = np.ascontiguousarray(out_ids, np.int32)
        _need(ii, (len(self.problems),), "in_ids")
        _need(oo, (len(self.problems),), "out_ids")
        _chk(_lib().ocean_gpu_batch_xact(self.b, ii.ctypes.data_as(_IP), oo.ctypes.data_as(_IP), nrep),
             "batch_xact")
        ms = C.c_double()
        _chk(_lib().ocean_gpu_batch_last_ms(self.b, C.byref(ms)), "batch_last_ms")
        return ms.value

    def haydock(self, psi0s, niter):
        """Per problem (real_a, real_b, imag_a); also returns the recursion's ms."""
        n = len(self.problems)
        if len(psi0s) != n:
            raise ValueError("one starting vector per problem")
        planes = [_planes(p, g.shape, "psi0") for p, g in zip(psi0s, self.problems)]
        pr = (_DP * n)(*[_p(x[0]) for x in planes])
        pi = (_DP * n)(*[_p(x[1]) for x in planes])
        A = [np.zeros(niter + 1) for _ in range(n)]
        Bv = [np.zeros(niter + 1) for _ in range(n)]
        AI = [np.zeros(niter + 1) for _ in range(n)]
        pa = (_DP * n)(*[_p(x) for x in A])
        pb = (_DP * n)(*[_p(x) for x in Bv])
        pai = (_DP * n)(*[_p(x) for x in AI])
        done = np.zeros(n, np.int32)
        rc = _lib().ocean_gpu_batch_haydock(self.b, pr, pi, niter, pa, pb, pai, done.ctypes.data_as(_IP))
        _chk(rc, "batch_haydock")
        ms = C.c_double()
        _chk(_lib().ocean_gpu_batch_last_ms(self.b, C.byref(ms)), "batch_last_ms")
        return [(A[q], Bv[q], AI[q]) for q in range(n)], ms.value

    def gmres(self, hay_vecs, energies_eV, gres_eV, gprc_eV, ffff, depth, kprefs=None, return_status=False):
        """GMRES of every problem at once (ocean_gpu_batch_gmres).  energies_eV: one
        list per problem.  Returns (rows per problem, wall ms) -- and the per-problem
        codes with return_status (no exception then)."""
        n = len(self.problems)
        if len(hay_vecs) != n or len(energies_eV) != n:
            raise ValueError("one starting vector and one energy list per problem")
        planes = [_planes(v, g.shape, "hay_vec") for v, g in zip(hay_vecs, self.problems)]
        Es = [np.ascontiguousarray(e, np.float64) for e in energies_eV]
        kp = np.ascontiguousarray([g.kpref if kprefs is None else kprefs[q] for q, g in enumerate(self.problems)],
                                  np.float64)
        outs = [np.zeros((len(e), 6)) for e in Es]
        nE = np.ascontiguousarray([len(e) for e in Es], np.int32)
        pr = (_DP * n)(*[_p(x[0]) for x in planes])
        pi = (_DP * n)(*[_p(x[1]) for x in planes])
        pe = (_DP * n)(*[e.ctypes.data_as(_DP) for e in Es])
        po = (_DP * n)(*[o.ctypes.data_as(_DP) for o in outs])
        rcs = np.zeros(n, np.int32)
        rc = _lib().ocean_gpu_batch_gmres(self.b, pr, pi, nE.ctypes.data_as(_IP), pe, gres_eV, gprc_eV, ffff,
                                          depth, kp.ctypes.data_as(_DP), po, rcs.ctypes.data_as(_IP))
        ms = C.c_double()
        _chk(_lib().ocean_gpu_batch_last_ms(self.b, C.byref(ms)), "batch_last_ms")
        if return_status:
            return outs, ms.value, rcs
        _chk(rc, "batch_gmres")
        return outs, ms.value

