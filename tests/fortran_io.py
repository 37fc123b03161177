"""Writer/reader for the Fortran host driver's stream files (tests only)."""
import numpy as np


def write_driver_input(path, inp, niter, ne, gam, el, eh, energies=(), gres=0.5, gprc=13.605,
                       ffff=1e-10, depth=20, precision=0):
    m = inp.mult
    nproj = list(m.nproj) + [0] * (8 - len(m.nproj))
    hdr = [inp.nb, *inp.kmesh, *inp.xmesh, inp.nspn, inp.lc, m.lmin, m.lmax, m.lvl, m.lvh, *nproj,
           int(m.have_spin_orbit), niter, ne, len(energies), depth, precision]
    mhr, mhi, _, _ = m.packed()
    f64 = lambda a: np.asarray(a, dtype="<f8").ravel(order="F")  # noqa: E731
    with open(path, "wb") as f:
        f.write(np.asarray(hdr, dtype="<i4").tobytes())
        f.write(np.asarray([inp.celvol, inp.inter_scale, inp.kpref, gam, el, eh, gres, gprc, ffff],
                           dtype="<f8").tobytes())
        for a in (inp.u.real, inp.u.imag, inp.W, inp.energy.real, inp.energy.imag, inp.allow,
                  m.c.real, m.c.imag, m.mpm, mhr, mhi, m.somelr, m.someli, inp.psi0.real,
                  inp.psi0.imag):
            f.write(f64(a).tobytes())
        if len(energies):
            f.write(f64(energies).tobytes())


def read_driver_output(path, shape, niter, nen):
    raw = open(path, "rb").read()
    ierr, done = np.frombuffer(raw[:8], dtype="<i4")
    d = np.frombuffer(raw[8:], dtype="<f8")
    N = int(np.prod(shape))
    hr, hi = d[:N], d[N:2 * N]
    a = d[2 * N:2 * N + niter + 1]
    b = d[2 * N + niter + 1:2 * N + 2 * (niter + 1)]
    rows = d[2 * N + 2 * (niter + 1):].reshape((nen, 6)) if nen else None
    h = (hr + 1j * hi).reshape(shape, order="F")
    return int(ierr), int(done), h, a, b, rows
