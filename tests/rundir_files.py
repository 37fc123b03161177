"""Writes a synthetic OCEAN2 core-level run directory in the file formats the
reference reads (test infrastructure): returns the arrays the files were made from.

Formats: OCEAN2/OCEAN_system.f90:166-231, :310-335, :644-726 (meshes, nspin, ZNL,
nbuse, brange, mode, cks.normal, runlist, xyz.wyck); bse.in lines
(OCEAN_multiplet.f90:298-303, OCEAN_haydock.f90:2109-2116); con.u2.dat
(OCEAN_bloch.f90:998-1115); wvfcninfo (OCEAN_energies.f90:649-687); rpottrim
(nbsemkcmel.f90:134-141 writes it, long_range.f90:1688-1696 reads it); prjfile,
cmel, fk/gk (OCEAN_multiplet.f90:309-316, :486-504, :2064-2114); parcksc
(OCEAN_multiplet.f90:426-451); mels (OCEAN_psi.f90:5964-5975).
"""
import os

import numpy as np

EV2H = 0.036749322175666


def _w(path, text):
    with open(path, "w") as f:
        f.write(text)


def _fortran_sequential(path, arrays):
    with open(path, "wb") as f:
        for a in arrays:
            b = np.ascontiguousarray(a).tobytes() if a.ndim <= 1 else np.asfortranarray(a).tobytes(order="F")
            n = np.int32(len(b)).tobytes()
            f.write(n + b + n)


def write_rundir(d, nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, nspn=1, extra_bands=2, seed=0,
                 xi_eV=0.0, lmin=0, lmax=2, nproj=(2, 1, 1), inter_scale=1.0, eps=6.0,
                 efermi_eV=4.0, eshift_eV=0.7, tau=(0.13, 0.27, 0.41), niter=30, a_bohr=5.5):
    rng = np.random.default_rng(seed)
    os.makedirs(d, exist_ok=True)
    Z, n, el, indx, photon = 26, 1 if lc == 0 else 2, "Fe", 1, 1
    nk, nx = int(np.prod(kmesh)), int(np.prod(xmesh))
    bif = nb + extra_bands
    assert bif <= nx
    _w(os.path.join(d, "xmesh.ipt"), "%d %d %d\n" % xmesh)
    _w(os.path.join(d, "kmesh.ipt"), "%d %d %d\n" % kmesh)
    _w(os.path.join(d, "nspin"), "%d\n" % nspn)
    _w(os.path.join(d, "ZNL"), "%d %d %d\n" % (Z, n, lc))
    _w(os.path.join(d, "nbuse.ipt"), "%d\n" % nb)
    _w(os.path.join(d, "brange.ipt"), "1 4\n5 %d\n" % (4 + bif))
    _w(os.path.join(d, "qinunitsofbvectors.ipt"), "0.0 0.0 0.0\n")
    avec = np.array([[a_bohr, 0.1, 0.0], [0.0, a_bohr * 1.05, 0.2], [0.1, 0.0, a_bohr * 0.95]]).T
    _w(os.path.join(d, "avecsinbohr.ipt"), "".join("%.17g %.17g %.17g\n" % tuple(avec[:, i]) for i in range(3)))
    _w(os.path.join(d, "xyz.wyck"), "2\nO 0.5 0.5 0.5\n%s %.17g %.17g %.17g\n" % ((el,) + tuple(tau)))
    _w(os.path.join(d, "runlist"), "1\n%d %d %d %s %s %d %d XAS\n" % (Z, n, lc, el, "1s" if lc == 0 else "2p",
                                                                       indx, photon))
    lvl, lvh = lmin, lmax
    _w(os.path.join(d, "bse.in"), "%d %d %d\n%.17g\nhay\n%d 2000 -5.0 40.0 0.5 0.0\n" % (lc, lvl, lvh, xi_eV, niter))
    _w(os.path.join(d, "epsilon"), "%.17g\n" % eps)
    _w(os.path.join(d, "mode"), "%.17g %d\n" % (inter_scale, niter))
    _w(os.path.join(d, "cks.normal"), ".true.\n")
    _w(os.path.join(d, "metal"), ".false.\n")
    _w(os.path.join(d, "efermiinrydberg.ipt"), "%.17g\n" % (2.0 * efermi_eV * EV2H))
    _w(os.path.join(d, "eshift.ipt"), "%.17g\n" % eshift_eV)
    # Bloch functions: orthonormal over x per (k, spin), all bands in the file
    u = np.zeros((nspn, nk, bif, nx), np.complex128)
    for s in range(nspn):
        for k in range(nk):
            q, _ = np.linalg.qr(rng.standard_normal((nx, bif)) + 1j * rng.standard_normal((nx, bif)))
            u[s, k] = q.T
    u.tofile(os.path.join(d, "con.u2.dat"))
    # band energies (Hartree), sorted per k
    E = np.sort(rng.uniform(0.0, 30.0, size=(bif, nk, nspn)), axis=0) * EV2H
    hdr = np.array([bif, nk] + ([nspn] if nspn == 2 else []), np.int32)
    _fortran_sequential(os.path.join(d, "wvfcninfo"), [hdr, np.asfortranarray(E)])
    # screened potential table (ptab, rtab rows)
    rtab = 12.0 * np.arange(1, 201) / 200.0
    ptab = np.exp(-rtab / 3.0) / (4.0 * np.maximum(rtab, 0.2))
    add14 = "%2s%04d_n%02dl%02d" % (el, indx, n, lc)
    _w(os.path.join(d, "rpottrim.z" + add14),
       "#%8d\n" % rtab.size + "".join("%24.16E %24.16E\n" % (p, r) for p, r in zip(ptab, rtab)))
    # projectors
    nl = lmax - lmin + 1
    assert len(nproj) == nl
    _w(os.path.join(d, "prjfilez%03d" % Z), "%d %d\n%s\n" % (lmin, lmax, " ".join(str(v) for v in nproj)))
    cmel = {}
    for li, l in enumerate(range(lmin, lmax + 1)):
        m = rng.uniform(-1.0, 1.0, (nproj[li], nproj[li]))
        m = 0.5 * (m + m.T)
        cmel[l] = m
        _w(os.path.join(d, "cmel%d" % l), " ".join("%.17g" % v for v in m.ravel(order="F")) + "\n")
    add10 = "z%03dn%02dl%02d" % (Z, n, lc)
    slater = {}
    for li, lv in enumerate(range(lvl, lvh + 1)):
        npl = nproj[lv - lmin]
        for pre, ks in (("fk", range(2, min(2 * lc, 2 * lv) + 1, 2)), ("gk", range(abs(lc - lv), lc + lv + 1, 2))):
            for k in ks:
                m = rng.uniform(0.5, 2.0, (npl, npl))
                m = 0.5 * (m + m.T)
                sc = 0.8
                slater[(pre, lv, k)] = (m, sc)
                _w(os.path.join(d, "%2s%1d%1d%1d%s" % (pre, lc, lv, k, add10)),
                   " ".join("%.17g" % v for v in m.ravel(order="F")) + " %.17g\n" % sc)
    nptot = sum((2 * l + 1) * nproj[i] for i, l in enumerate(range(lmin, lmax + 1)))
    vol = abs(np.linalg.det(avec))
    cks = (rng.standard_normal((nptot, bif * nk, nspn)) + 1j * rng.standard_normal((nptot, bif * nk, nspn))) \
        * np.sqrt(vol / nb)
    with open(os.path.join(d, "parcksc.%2s%04d" % (el, indx)), "wb") as f:
        f.write(np.array([nptot, bif * nk, nspn], np.int32).tobytes())
        f.write(np.asfortranarray(cks).tobytes(order="F"))
    mels = rng.standard_normal((nptot, 2 * lc + 1)) + 1j * rng.standard_normal((nptot, 2 * lc + 1))
    _w(os.path.join(d, "mels.z%03dn%02dl%02dp%02d" % (Z, n, lc, photon)),
       "".join("%.17g %.17g\n" % (mels[p, m].real, mels[p, m].imag) for m in range(2 * lc + 1) for p in range(nptot)))
    return dict(u=u, E=E, rtab=rtab, ptab=ptab, cmel=cmel, slater=slater, cks=cks, mels=mels, avec=avec,
                tau=np.array(tau), nb=nb, kmesh=kmesh, xmesh=xmesh, lc=lc, nspn=nspn, bif=bif, lmin=lmin,
                lmax=lmax, nproj=list(nproj), xi_eV=xi_eV, inter_scale=inter_scale, eps=eps,
                efermi=efermi_eV * EV2H, eshift=eshift_eV * EV2H, nptot=nptot, celvol=vol, niter=niter)
