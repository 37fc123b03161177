"""Reader of an OCEAN2 run directory for a core-level (XAS) run.

Reads the files OCEAN2 reads when it sets up the core-level BSE for the first core
run of ``runlist`` and turns them into the arrays the C-ABI loaders take; the two
setup steps with a device form run on the GPU: the Bloch (or OBF) phase and W through
``ocean_gpu_load_bloch_raw`` / ``ocean_gpu_load_obf`` and ``ocean_gpu_build_w``, the
multiplet interaction blocks and the spin-orbit matrix through ``ocean_gpu_mult_setup``.  File formats and the
derived quantities follow the reference routines cited at each reader (paths
relative to the OCEAN 3.2.1 tree).

``write_synthetic_rundir`` writes a seeded synthetic run directory in the same
formats (for this reader's tests, and as input for OCEAN2's own CPU build).

Not read (a run that needs them raises ``NotImplementedError``): XES / valence /
RIXS runs, GW corrections, doping and ``findNewFermi.ipt``, occupation types other
than the default, ``isolated.inp``, ``pbc.inp``, ``epsilon3D``, ``file_brange.ipt``
and the legacy (sequential) cks files.
"""
from __future__ import annotations

import dataclasses
import os
from dataclasses import dataclass, field

import numpy as np

from . import setup_host as sh
from .synth import MultipletInputs

EV2HARTREE = sh.EV2HARTREE


def _p(d, name):
    return os.path.join(d, name)


def _tokens(path):
    with open(path) as f:
        return f.read().split()


def _fbool(tok):
    """Fortran list-directed logical: .true./.false./T/F."""
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


def _ffloat(tok):
    return float(tok.lower().replace("d", "e"))


def _refuse(d, names):
    for n in names:
        if os.path.exists(_p(d, n)):
            raise NotImplementedError(f"run directory uses {n}, which this reader does not support")


def fortran_records(path):
    """Records of a Fortran unformatted sequential file (4-byte length markers)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data, np.int32, 1, pos)[0])
        rec = data[pos + 4:pos + 4 + n]
        if int(np.frombuffer(data, np.int32, 1, pos + 4 + n)[0]) != n:
            raise ValueError(f"{path}: broken record marker at byte {pos}")
        out.append(rec)
        pos += 8 + n
    return out


@dataclass
class CoreRun:
    """One line of ``runlist`` (OCEAN2/OCEAN_system.f90:655-705)."""
    Z: int
    n: int
    l: int
    elname: str
    corelevel: str
    indx: int
    photon: int
    calc_type: str


@dataclass
class RunSystem:
    """o_system fields of a core run (OCEAN2/OCEAN_system.f90:160-471)."""
    xmesh: tuple
    kmesh: tuple
    nspn: int
    ZNL: tuple
    nb: int
    brange: tuple
    avec: np.ndarray          # columns are the lattice vectors (bohr)
    celvol: float
    run: CoreRun
    tau: np.ndarray           # site, reduced coordinates (xyz.wyck)
    inter_scale: float
    epsilon0: float
    mult: bool
    lvl: int
    lvh: int
    xi_eV: float
    solver: dict = field(default_factory=dict)
    rpa: bool = False             # rpa file: no long-range term, no direct multiplet term
    tdlda: bool = False           # tdlda file: F^0 (kk files) in place of the F^k
    xshift_override: tuple = (1, 1, 1)

    @property
    def nk(self):
        return int(np.prod(self.kmesh))

    @property
    def nx(self):
        return int(np.prod(self.xmesh))

    @property
    def nalpha(self):
        return 4 * (2 * self.ZNL[2] + 1)

    @property
    def bands_in_file(self):
        return self.brange[3] - self.brange[2] + 1


def read_system(d):
    """Meshes, spins, ZNL, bands, lattice, first core run, site, mode, epsilon,
    bse.in (OCEAN2/OCEAN_system.f90:166-231, :310-335, :644-726;
    OCEAN_multiplet.f90:298-303; OCEAN_haydock.f90:2109-2116)."""
    _refuse(d, ["force_val_ham_spin.ipt", "file_brange.ipt", "isolated.inp", "pbc.inp", "epsilon3D"])
    xmesh = tuple(int(t) for t in _tokens(_p(d, "xmesh.ipt"))[:3])
    kmesh = tuple(int(t) for t in _tokens(_p(d, "kmesh.ipt"))[:3])
    nspn = int(_tokens(_p(d, "nspin"))[0])
    ZNL = tuple(int(t) for t in _tokens(_p(d, "ZNL"))[:3])
    nb = int(_tokens(_p(d, "nbuse.ipt"))[0])
    brange = tuple(int(t) for t in _tokens(_p(d, "brange.ipt"))[:4])
    av = np.array([_ffloat(t) for t in _tokens(_p(d, "avecsinbohr.ipt"))[:9]])
    avec = np.asfortranarray(av.reshape(3, 3, order="F"))
    celvol = abs(float(np.linalg.det(avec)))
    if not _fbool(_tokens(_p(d, "cks.normal"))[0]):
        raise NotImplementedError("cks.normal .false. (XES) runs are not supported")
    inter_scale = _ffloat(_tokens(_p(d, "mode"))[0])
    epsilon0 = _ffloat(_tokens(_p(d, "epsilon"))[0])
    mult = True
    if os.path.exists(_p(d, "mult.ipt")):
        mult = _fbool(_tokens(_p(d, "mult.ipt"))[0])
    # runlist: the first core run
    with open(_p(d, "runlist")) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    nruns = int(lines[0].split()[0])
    run = None
    for ln in lines[1:1 + nruns]:
        t = ln.split()
        r = CoreRun(int(t[0]), int(t[1]), int(t[2]), t[3].strip("'\""), t[4].strip("'\""), int(t[5]),
                    int(t[6]), t[7].strip("'\"").upper())
        if r.calc_type in ("XAS",):
            run = r
            break
        raise NotImplementedError(f"run type {r.calc_type} is not supported")
    if run is None:
        raise ValueError("runlist holds no core run")
    if tuple(ZNL) != (run.Z, run.n, run.l):
        raise ValueError("ZNL and the runlist entry disagree")
    # site: the indx-th atom of that element in xyz.wyck
    with open(_p(d, "xyz.wyck")) as f:
        wl = [ln for ln in f.read().splitlines() if ln.strip()]
    ntot = int(wl[0].split()[0])
    seen = 0
    tau = None
    for ln in wl[1:1 + ntot]:
        t = ln.split()
        if t[0].strip("'\"") == run.elname:
            seen += 1
            if seen == run.indx:
                tau = np.array([_ffloat(v) for v in t[1:4]])
                break
    if tau is None:
        raise ValueError(f"site {run.elname} {run.indx} not in xyz.wyck")
    # bse.in: lc lvl lvh / xi (eV) / hay / niter ne el eh gam ebase (eV)
    with open(_p(d, "bse.in")) as f:
        bl = [ln for ln in f.read().splitlines() if ln.strip()]
    t = bl[0].split()
    lc, lvl, lvh = int(t[0]), int(t[1]), int(t[2])
    if lc != run.l:
        raise ValueError("bse.in core l disagrees with ZNL")
    xi = _ffloat(bl[1].split()[0])
    solver = {"kind": bl[2].split()[0].strip("'\"")}
    if solver["kind"] == "hay" and len(bl) > 3:
        v = bl[3].split()
        solver.update(niter=int(v[0]), ne=int(v[1]), el=_ffloat(v[2]), eh=_ffloat(v[3]),
                      gam=_ffloat(v[4]), ebase=_ffloat(v[5]) if len(v) > 5 else 0.0)
    S = RunSystem(xmesh, kmesh, nspn, ZNL, nb, brange, avec, celvol, run, tau, inter_scale,
                  epsilon0, mult, lvl, lvh, xi, solver)
    # rpa: long_range off (OCEAN2/OCEAN_system.f90:338-341) and no cmel (OCEAN_multiplet.f90:488-489)
    S.rpa = os.path.exists(_p(d, "rpa"))
    # tdlda: a logical in the file (OCEAN2/OCEAN_multiplet.f90:2065-2070)
    # (an empty file ends that read with an end-of-file error)
    S.tdlda = False
    if os.path.exists(_p(d, "tdlda")):
        tt = _tokens(_p(d, "tdlda"))
        if not tt:
            raise ValueError(f"{_p(d, 'tdlda')}: empty, a Fortran logical is expected")
        S.tdlda = _fbool(tt[0])
    # xshift_override.ipt: three 0/1 factors on the x shift, else all 1 (OCEAN2/OCEAN_bloch.f90:279-292)
    if os.path.exists(_p(d, "xshift_override.ipt")):
        xo = tuple(int(t) for t in _tokens(_p(d, "xshift_override.ipt"))[:3])
        S.xshift_override = xo if (max(xo) <= 1 and min(xo) >= 0) else (1, 1, 1)
    return S


def read_u2(d, S: RunSystem):
    """con.u2.dat: stream complex128, [spin][k][band 1..bandsInFile][x], x fastest;
    the first nb bands of each k are kept (OCEAN2/OCEAN_bloch.f90:998-1115)."""
    nb, nk, nx, ns, bif = S.nb, S.nk, S.nx, S.nspn, S.bands_in_file
    if bif < nb:
        raise ValueError("brange holds fewer conduction bands than nbuse.ipt")
    raw = np.fromfile(_p(d, "con.u2.dat"), dtype=np.complex128)
    if raw.size != ns * nk * bif * nx:
        raise ValueError(f"con.u2.dat holds {raw.size} values, expected {ns * nk * bif * nx}")
    a = raw.reshape(ns, nk, bif, nx)[:, :, :nb, :]
    return np.asfortranarray(a.transpose(2, 1, 3, 0))  # (nb, nk, nx, nspn)


def read_bloch_control(d):
    """bloch_control: use_obf, use_fake_obf (OCEAN2/long_range.f90:128-146); without
    the file both are false and the Bloch functions come from con.u2.dat."""
    if not os.path.exists(_p(d, "bloch_control")):
        return False, False
    t = _tokens(_p(d, "bloch_control"))
    return _fbool(t[0]), _fbool(t[1])


def read_obf(d, S: RunSystem):
    """The OBF form of the Bloch functions, as OCEAN_obf_init and OCEAN_obf_load read
    it (OCEAN2/OCEAN_obf.f90:424-508, :510-732):

    obf_control     num_obf, val_num_bands, con_num_bands (:465-470)
    u1.dat          Fortran sequential, one record of complex(xmesh(3), xmesh(2),
                    xmesh(1)) per basis function (:551-572), taken in the x order
                    with ix fastest
    con_eigvecs.dat stream complex128: the coefficients of k are num_obf * nb values
                    at element offset (k-1) num_obf con_num_bands + num_obf ivh2, as
                    (num_obf, nb) (:704-715); ivh2 = ivh of brange.ipt, or for a metal
                    the second entry of k's ibeg.h line minus one (:650-699)

    Returns (obf [num_obf, nx], obf2u [nb, num_obf, nk])."""
    if S.nspn != 1:
        raise NotImplementedError("the OBF form handles nspn = 1 only (OCEAN_obf.f90:669-673)")
    t = _tokens(_p(d, "obf_control"))
    no, width = int(t[0]), int(t[2])
    X1, X2, X3 = S.xmesh
    recs = fortran_records(_p(d, "u1.dat"))
    if len(recs) < no:
        raise ValueError("u1.dat holds fewer basis functions than obf_control")
    obf = np.zeros((no, S.nx), np.complex128)
    for o in range(no):
        a = np.frombuffer(recs[o], np.complex128, count=S.nx)
        obf[o] = a.reshape(X1, X2, X3).transpose(2, 1, 0).ravel()
    ivl, ivh, icl, ich = (int(v) for v in _tokens(_p(d, "brange.ipt"))[:4])
    if S.nb > 1 + (ich - icl):
        raise ValueError("loadux ... nbd mismatch -- cf brange.ipt")
    metal = _fbool(_tokens(_p(d, "metal"))[0])
    if metal:
        rows = [ln.split() for ln in open(_p(d, "ibeg.h")).read().splitlines() if ln.strip()]
        ivh2 = [int(r[1]) - 1 for r in rows[:S.nk]]
    else:
        ivh2 = [ivh] * S.nk
    raw = np.fromfile(_p(d, "con_eigvecs.dat"), dtype=np.complex128)
    obf2u = np.zeros((S.nb, no, S.nk), np.complex128, order="F")
    for k in range(S.nk):
        off = k * no * width + no * ivh2[k]
        if off + no * S.nb > raw.size:
            raise ValueError("con_eigvecs.dat is truncated")
        obf2u[:, :, k] = raw[off:off + no * S.nb].reshape(S.nb, no)
    return obf, obf2u


def read_energies(d, S: RunSystem):
    """wvfcninfo (Hartree, first nb bands), the Fermi level, the core-level shift and
    the allow mask, spread over alpha by valence spin (OCEAN2/OCEAN_energies.f90:
    303-368 load, :371-487 allow (default occupation), :629-700 read_energies,
    :777-825 read_coreLevelShift, :829-887 stubby, :1340-1430 Fermi level)."""
    _refuse(d, ["doping", "findNewFermi.ipt", "occupation.ipt", "GW", "gw_control"])
    recs = fortran_records(_p(d, "wvfcninfo"))
    hdr = np.frombuffer(recs[0], np.int32)
    nbd, nq = int(hdr[0]), int(hdr[1])
    nspn = int(hdr[2]) if S.nspn == 2 else 1
    if nbd < S.nb or nq != S.nk or nspn != S.nspn:
        raise ValueError("wvfcninfo does not match the run's bands, k points or spins")
    E = np.frombuffer(recs[1], np.float64).reshape(nbd, nq, nspn, order="F")[:S.nb]
    # Fermi level: efermiinrydberg.ipt (Rydberg) / 2
    efermi = _ffloat(_tokens(_p(d, "efermiinrydberg.ipt"))[0]) / 2.0
    # core-level shift: cls.z<El><indx>_n<nn>l<ll> (eV), else eshift.ipt unless noshift_lumo
    r = S.run
    cls = _p(d, "cls.z%2s%04d_n%02dl%02d" % (r.elname, r.indx, S.ZNL[1], S.ZNL[2]))
    if os.path.exists(cls):
        shift = _ffloat(_tokens(cls)[0]) * EV2HARTREE
    else:
        noshift = os.path.exists(_p(d, "noshift_lumo")) and _fbool(_tokens(_p(d, "noshift_lumo"))[0])
        shift = 0.0 if noshift else _ffloat(_tokens(_p(d, "eshift.ipt"))[0]) * EV2HARTREE
    # core_allow_fixed with the default occupation: XAS allows E >= E_F, else 0
    allow_b = np.where(E >= efermi, 1.0, 0.0)
    na = S.nalpha
    vs = np.zeros(na, np.int64) if S.nspn == 1 else np.arange(na) % 2
    energy = np.zeros((S.nb, S.nk, na), np.complex128, order="F")
    allow = np.zeros((S.nb, S.nk, na), np.float64, order="F")
    for a in range(na):
        energy[:, :, a] = E[:, :, vs[a]] + shift
        allow[:, :, a] = allow_b[:, :, vs[a]]
    return energy, allow, efermi, shift


def read_rpottrim(d, S: RunSystem):
    """rpottrim.z<El><indx>_n<nn>l<ll>: a header (word, count), then (ptab, rtab) rows
    (OCEAN2/long_range.f90:1688-1696)."""
    r = S.run
    name = "rpottrim.z%2s%04d_n%02dl%02d" % (r.elname, r.indx, S.ZNL[1], S.ZNL[2])
    with open(_p(d, name)) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    n = int(lines[0].split()[1])
    vals = np.array([[_ffloat(t) for t in ln.split()[:2]] for ln in lines[1:1 + n]])
    return np.ascontiguousarray(vals[:, 1]), np.ascontiguousarray(vals[:, 0])  # rtab, ptab


def read_prjfile(d, S: RunSystem):
    """prjfilez<Z>: lmin lmax, then nproj(lmin:lmax) (OCEAN2/OCEAN_multiplet.f90:309-316)."""
    t = _tokens(_p(d, "prjfilez%03d" % S.ZNL[0]))
    lmin, lmax = int(t[0]), int(t[1])
    nproj = [int(v) for v in t[2:2 + lmax - lmin + 1]]
    if S.lvl < lmin or S.lvh > lmax:
        raise ValueError("bad lvl lvh")
    return lmin, lmax, nproj


def read_parcksc(d, S: RunSystem, nptot):
    """parcksc.<El><indx>: stream int32 (nptot, ntot, nspn), then complex128
    (nptot, ntot, nspn) (OCEAN2/OCEAN_multiplet.f90:426-451, OCEAN_psi.f90:5920-5937)."""
    r = S.run
    path = _p(d, "parcksc.%2s%04d" % (r.elname, r.indx))
    hdr = np.fromfile(path, dtype=np.int32, count=3)
    if int(hdr[0]) != nptot or int(hdr[2]) != S.nspn:
        raise ValueError("parcksc header does not match the projectors or spins")
    ntot = int(hdr[1])
    nbd = ntot // S.nk
    if nbd * S.nk != ntot or nbd < S.nb:
        raise ValueError("Badly formed cks")
    cks = np.fromfile(path, dtype=np.complex128, offset=12)
    if cks.size != nptot * ntot * S.nspn:
        raise ValueError("parcksc is truncated")
    return np.asfortranarray(cks.reshape(nptot, ntot, S.nspn, order="F")), nbd


def projector_array(cks, S: RunSystem, lmin, lmax, nproj, bands_per_k):
    """mpcr/mpci(nb, nk, nu, m, l, s) from the cks projector order (l, m, nu)
    (OCEAN2/OCEAN_multiplet.f90:451-462); bands past nb are skipped."""
    npmax = max(nproj)
    nl = lmax - lmin + 1
    c = np.zeros((S.nb, S.nk, npmax, 2 * lmax + 1, nl, S.nspn), np.complex128, order="F")
    byk = cks.reshape(cks.shape[0], bands_per_k, S.nk, S.nspn, order="F")[:, :S.nb]
    ip = 0
    for li, l in enumerate(range(lmin, lmax + 1)):
        for m in range(-l, l + 1):
            for nu in range(nproj[li]):
                c[:, :, nu, m + lmax, li, :] = byk[ip]
                ip += 1
    return c


def _fortran_e(x, w, dgt):
    """x in the Fortran edit descriptor Ew.d (0.ddd..E+ee, right-justified)."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * dgt + "E+00"
    else:
        m, e = ("%.*E" % (dgt - 1, abs(x))).split("E")
        digits = m.replace(".", "")
        ex = int(e) + 1
        es = ("E%+03d" % ex) if abs(ex) < 100 else ("%+04d" % ex)
        body = ("-" if x < 0 else "") + "0." + digits + es
    return body.rjust(w)


def projector_files_present(d, S: RunSystem):
    """The inputs nbsemkcmel reads (prjfile, radfile, ae<l>, rpot) are in the directory."""
    r = S.run
    add04 = "z%03d" % S.ZNL[0]
    add14 = "%2s%04d_n%02dl%02d" % (r.elname, r.indx, S.ZNL[1], S.ZNL[2])
    return all(os.path.exists(_p(d, n)) for n in ("prjfile" + add04, "radfile" + add04, "rpot.z" + add14))


def make_cmel(d, S: RunSystem):
    """What ocean.x does at set-up through nbsemkcmel (OCEAN2/nbsemkcmel.f90:8-209,
    called by OCEAN_create_central, OCEAN_multiplet.f90:376): read prjfile<z>,
    radfile<z> (rmax, nrmax, nr), rpot.z<run> ((V, r) rows) and ae<l><z> (r,
    phi_1..phi_nproj rows), split the potential at rmax, integrate the projector
    matrices, and write rpottrim.z<run> ('#', count; E24.16 pairs), rpotdiff.z<run>
    (2f10.5) and cmel<l> (cmel then nmel, 4(1x,e15.8)) in the reference's formats, so
    the readers below see the values ocean.x reads back (eight significant digits
    in cmel)."""
    r = S.run
    add04 = "z%03d" % S.ZNL[0]
    add14 = "%2s%04d_n%02dl%02d" % (r.elname, r.indx, S.ZNL[1], S.ZNL[2])
    t = _tokens(_p(d, "prjfile" + add04))
    lmin, lmax = int(t[0]), int(t[1])
    nnu = [int(v) for v in t[2:2 + lmax - lmin + 1]]
    t = _tokens(_p(d, "radfile" + add04))
    rmax, nr = _ffloat(t[0]), int(t[2])
    if rmax > 9.9:
        raise ValueError("radfile: bad rmax (nbsemkcmel.f90:43 stops above 9.9)")
    rows = []
    with open(_p(d, "rpot.z" + add14)) as f:
        for ln in f:
            p = ln.split()
            if len(p) < 2:
                break
            try:
                rows.append((_ffloat(p[0]), _ffloat(p[1])))
            except ValueError:
                break  # the reference's read loop ends at the first unreadable record
    v = np.array([a for a, _ in rows])
    rv = np.array([b for _, b in rows])
    vtrim, vdiff = sh.trim_core_hole_potential(v, rv, rmax)
    _w(_p(d, "rpottrim.z" + add14), "#%8d\n" % v.size +
       "".join("%s %s\n" % (_fortran_e(a, 24, 16), _fortran_e(b, 24, 16)) for a, b in zip(vtrim, rv)))
    _w(_p(d, "rpotdiff.z" + add14), "".join("%10.5f%10.5f\n" % (a, b) for a, b in zip(vdiff, rv)))
    for li, l in enumerate(range(lmin, lmax + 1)):
        vals = np.array([_ffloat(x) for x in _tokens(_p(d, "ae%d%s" % (l, add04)))[:nr * (1 + nnu[li])]])
        tab = vals.reshape(nr, 1 + nnu[li])
        cmel, nmel = sh.projector_cmel(tab[:, 0], tab[:, 1:], rv, vdiff, rmax)
        out = []
        for m in (cmel, nmel):
            flat = m.ravel(order="F")
            for i0 in range(0, flat.size, 4):
                out.append("".join(" " + _fortran_e(x, 15, 8) for x in flat[i0:i0 + 4]) + "\n")
        _w(_p(d, "cmel%d" % l), "".join(out))


def read_cmel(d, lmin, lmax, nproj):
    """cmel<l>: nproj(l)^2 values (eV) in column order -> mpm (Hartree)
    (OCEAN2/OCEAN_multiplet.f90:486-504)."""
    npmax = max(nproj)
    mpm = np.zeros((npmax, npmax, lmax - lmin + 1), order="F")
    for li, l in enumerate(range(lmin, lmax + 1)):
        npl = nproj[li]
        v = np.array([_ffloat(t) for t in _tokens(_p(d, "cmel%d" % l))[:npl * npl]])
        mpm[:npl, :npl, li] = v.reshape(npl, npl, order="F")
    return mpm * EV2HARTREE


def read_slater(d, S: RunSystem, lv, npl):
    """fk<lc><lv><k>z<Z>n<nn>l<ll> / gk...: np x np Slater integrals and their scale,
    kept pre-multiplied; scfac (default 1) (OCEAN2/OCEAN_multiplet.f90:2064-2114)."""
    lc = S.ZNL[2]
    add10 = "z%03dn%02dl%02d" % S.ZNL
    scfac = 1.0
    if os.path.exists(_p(d, "scfac")):
        scfac = _ffloat(_tokens(_p(d, "scfac"))[0])

    def one(prefix, k):
        t = _tokens(_p(d, "%2s%1d%1d%1d%s" % (prefix, lc, lv, k, add10)))
        v = np.array([_ffloat(x) for x in t[:npl * npl + 1]])
        return v[:npl * npl].reshape(npl, npl, order="F") * v[npl * npl]

    if S.tdlda:
        fk = {0: one("kk", 0)}  # OCEAN2/OCEAN_multiplet.f90:2082-2094
    else:
        fk = {k: one("fk", k) for k in range(2, min(2 * lc, 2 * lv) + 1, 2)}
    gk = {k: one("gk", k) for k in range(abs(lc - lv), lc + lv + 1, 2)}
    return fk, gk, scfac


def read_mels(d, S: RunSystem, nptot):
    """mels.z<Z>n<nn>l<ll>p<photon>: (re, im) per (m_c, p), m_c outer
    (OCEAN2/OCEAN_psi.f90:5964-5975)."""
    r = S.run
    t = _tokens(_p(d, "mels.z%03dn%02dl%02dp%02d" % (S.ZNL[0], S.ZNL[1], S.ZNL[2], r.photon)))
    lc = S.ZNL[2]
    v = np.array([_ffloat(x) for x in t[:2 * nptot * (2 * lc + 1)]]).reshape(2 * lc + 1, nptot, 2)
    return np.asfortranarray((v[:, :, 0] + 1j * v[:, :, 1]).T)  # (nptot, 2lc+1)


def starting_vector(mels, cks, S: RunSystem, bands_per_k):
    """psi0 from the dipole elements and projector overlaps, then normalised with
    kpref = 4 pi |psi|^2 / (N_k Omega^2) (OCEAN2/OCEAN_psi.f90:5887-6031, :5589-5638)."""
    lc, nb, nk, na = S.ZNL[2], S.nb, S.nk, S.nalpha
    psi = np.zeros((nb, nk, na), np.complex128, order="F")
    ia = 0
    for icms in (-1, 1):
        for icml in range(-lc, lc + 1):
            for ivms in (-1, 1):
                if icms == ivms:
                    sp = 0 if S.nspn == 1 else (0 if ivms == -1 else 1)
                    byk = cks[:, :, sp].reshape(cks.shape[0], bands_per_k, nk, order="F")[:, :nb]
                    me = mels[:, icml + lc]
                    # rr - ii + i(-ri - ir) = conj(me) . conj(pc) summed over projectors
                    psi[:, :, ia] = np.einsum("p,pnk->nk", np.conj(me), np.conj(byk))
                ia += 1
    return sh.normalise_psi0(psi, nk, S.celvol)


@dataclass
class RunInputs:
    """Everything a core run's handle is loaded with, read from the directory."""
    system: RunSystem
    u_raw: np.ndarray
    xoff: np.ndarray          # int32 (3, nx) wrapped x offsets for the device phase
    tau_c: np.ndarray
    xshift: np.ndarray
    rdist: np.ndarray         # r(k, x) for the device W
    rtab: np.ndarray
    ptab: np.ndarray
    energy: np.ndarray
    allow: np.ndarray
    mult: MultipletInputs
    psi0: np.ndarray
    kpref: float
    efermi: float
    shift: float
    obf_mode: str = "bloch"   # "bloch" (con.u2.dat), "obf" (lr_act_obf2) or "fake" (use_fake_obf)
    obf: np.ndarray | None = None
    obf2u: np.ndarray | None = None
    slater: list | None = None  # per valence l: (fk, gk) dicts k -> scaled Slater matrix
    scfac: float = 1.0
    life_time: np.ndarray | None = None  # core_broaden.ipt (eV)


def read_rundir(d):
    """Read a core XAS run directory into the loader arrays."""
    S = read_system(d)
    use_obf, use_fake = read_bloch_control(d)
    obf = obf2u = None
    mode = "bloch"
    if use_obf:
        obf, obf2u = read_obf(d, S)
        mode = "fake" if use_fake else "obf"
    if mode == "obf":
        # lr_act_obf2 on the OBF arrays; OCEAN_obf_make_phase centres the site
        u_raw = None
        xoff, tau_c, xshift = sh.obf_x_offsets(S.xmesh, S.kmesh, S.tau)
    else:
        # con.u2.dat, or the Bloch functions OCEAN_obf_lrLOAD builds from the OBFs;
        # both are centred and phased as OCEAN_bloch_lrLOAD does
        u_raw = read_u2(d, S) if mode == "bloch" else sh.obf_fake_bloch(obf, obf2u)
        # OCEAN_obf_lrLOAD centres without the override (OCEAN2/OCEAN_obf.f90:165-186)
        xo = S.xshift_override if mode == "bloch" else (1, 1, 1)
        xoff, tau_c, xshift = sh.x_offsets(S.xmesh, S.kmesh, S.tau, xo)
    rdist, _ = sh.w_geometry(S.avec, S.kmesh, S.xmesh, tau_c, xshift)
    if projector_files_present(d, S):
        make_cmel(d, S)  # ocean.x writes rpottrim and cmel<l> at set-up
    rtab, ptab = read_rpottrim(d, S)
    energy, allow, efermi, shift = read_energies(d, S)
    lmin, lmax, nproj = read_prjfile(d, S)
    nptot = sum((2 * l + 1) * nproj[i] for i, l in enumerate(range(lmin, lmax + 1)))
    cks, bpk = read_parcksc(d, S, nptot)
    c = projector_array(cks, S, lmin, lmax, nproj, bpk)
    mpm = read_cmel(d, lmin, lmax, nproj) if not S.rpa else \
        np.zeros((max(nproj), max(nproj), lmax - lmin + 1), order="F")
    X = []
    slater = []
    scfac = 1.0
    for lv in range(S.lvl, S.lvh + 1):
        npl = nproj[lv - lmin]
        fk, gk, scfac = read_slater(d, S, lv, npl)
        slater.append((fk, gk))
        X.append(np.asfortranarray(sh.exchange_matrix(S.ZNL[2], lv, npl, fk, gk, scfac, S.tdlda) * EV2HARTREE))
    # core_broaden.ipt: spin-orbit-dependent core-hole lifetimes (eV), OCEAN_multiplet.f90:613-627
    life_time = None
    if os.path.exists(_p(d, "core_broaden.ipt")):
        life_time = np.array([_ffloat(t) for t in _tokens(_p(d, "core_broaden.ipt"))[:2]])
    sr, si, have_so = sh.spin_orbit(S.ZNL[2], S.xi_eV, life_time)
    mult = MultipletInputs(lmin, lmax, nproj, S.lvl, S.lvh, c, mpm, X, sr, si, have_so)
    mels = read_mels(d, S, nptot)
    psi0, kpref = starting_vector(mels, cks, S, bpk)
    return RunInputs(S, u_raw, xoff, tau_c, xshift, rdist, rtab, ptab, energy, allow, mult, psi0,
                     kpref, efermi, shift, mode, obf, obf2u, slater, scfac, life_time)


def open_rundir(d, precision=None):
    """A loaded handle for the run directory: u phased and W built on the device.
    Returns (OceanGPU, RunInputs)."""
    from . import core
    R = read_rundir(d)
    S = R.system
    m = R.mult
    prec = core.FP64 if precision is None else precision
    long_range = S.inter_scale >= 0.000001 and not S.rpa  # inter_min, OCEAN_system.f90:157; rpa :338-341
    g = core.OceanGPU(S.nb, S.kmesh, S.xmesh, S.ZNL[2], S.nspn, S.celvol, S.inter_scale, m.lmin, m.lmax,
                      tuple(m.nproj), m.lvl, m.lvh, True, S.mult, long_range, prec)
    if R.slater is not None:
        # the interaction blocks and the spin-orbit matrix formed on the device
        X, sr, si, have_so = g.mult_setup(R.slater, R.scfac, S.xi_eV, R.life_time, S.tdlda)
        m = dataclasses.replace(m, X=X, somelr=sr, someli=si, have_spin_orbit=have_so)
    if R.obf_mode == "obf":
        g.load_obf(R.obf, R.obf2u, R.xoff)
    else:
        g.load_bloch_raw(R.u_raw, R.xoff)
    g.build_w(R.rdist, R.rtab, R.ptab, S.epsilon0, S.inter_scale)
    g.load_mult(m)
    g.load_energies(R.energy, R.allow)
    g.kpref = R.kpref
    return g, R


def hartree_window(S: RunSystem):
    """The Haydock settings of bse.in line 4 (eV) converted as OCEAN_haydock.f90:
    2109-2116 does: (niter, ne, el, eh, gam, ebase) with energies in Hartree."""
    s = S.solver
    return (s["niter"], s["ne"], s["el"] * EV2HARTREE, s["eh"] * EV2HARTREE, s["gam"] * EV2HARTREE,
            s["ebase"])


# ---------------------------------------------------------------------------
# Writer: synthetic run directories (for this reader and for OCEAN2's own CPU build)
# ---------------------------------------------------------------------------
def _w(path, text):
    with open(path, "w") as f:
        f.write(text)


def _fortran_sequential(path, arrays):
    with open(path, "wb") as f:
        for a in arrays:
            b = np.ascontiguousarray(a).tobytes() if a.ndim <= 1 else np.asfortranarray(a).tobytes(order="F")
            n = np.int32(len(b)).tobytes()
            f.write(n + b + n)




def write_synthetic_rundir(d, nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, nspn=1, extra_bands=2, seed=0,
                 xi_eV=0.0, lmin=0, lmax=2, nproj=(2, 1, 1), inter_scale=1.0, eps=6.0,
                 efermi_eV=4.0, eshift_eV=0.7, tau=(0.13, 0.27, 0.41), niter=30, a_bohr=5.5,
                 obf=None, life_time_eV=None, rpa=False, tdlda=False, xshift_override=None):
    """Write a synthetic core XAS run directory in the file formats OCEAN2 reads
    (seeded: Bloch functions orthonormal over x per k, sorted band energies, the
    screened potential exp(-r/3)/(4 max(r, 0.2)), projector radial functions,
    random projector overlaps, dipole elements and Slater integrals).  Returns the
    arrays the files hold.  Like a CNBSE directory before ocean.x has run, it holds
    no rpottrim or cmel: the reader forms them as ocean.x does (make_cmel).

    Formats: OCEAN2/OCEAN_system.f90:166-231, :310-335, :644-726 (meshes, nspin, ZNL,
    nbuse, brange, mode, cks.normal, runlist, xyz.wyck); bse.in
    (OCEAN_multiplet.f90:298-303, OCEAN_haydock.f90:2109-2116); con.u2.dat
    (OCEAN_bloch.f90:998-1115); wvfcninfo (OCEAN_energies.f90:649-687); prjfile,
    radfile, ae<l>, rpot (nbsemkcmel.f90:31-87, :145-150); fk/gk
    (OCEAN_multiplet.f90:309-316, :2064-2114); parcksc
    (OCEAN_multiplet.f90:426-451); mels (OCEAN_psi.f90:5964-5975).

    obf = dict(num=..., fake=False) writes the OBF form instead of con.u2.dat:
    bloch_control, obf_control, u1.dat and con_eigvecs.dat (OCEAN2/long_range.f90:
    128-146, OCEAN_obf.f90:465-470, :551-572, :704-715), with orthogonal basis
    functions of norm N_k^-1/2 and unit-norm coefficient vectors per (n, k).
    """
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
    _w(os.path.join(d, "efermiinrydberg.ipt"), "%.17g\n" % (2.0 * efermi_eV * EV2HARTREE))
    _w(os.path.join(d, "eshift.ipt"), "%.17g\n" % eshift_eV)
    if life_time_eV is not None:
        _w(os.path.join(d, "core_broaden.ipt"), "%.17g %.17g\n" % tuple(life_time_eV))
    if rpa:
        _w(os.path.join(d, "rpa"), "\n")
    if tdlda:
        _w(os.path.join(d, "tdlda"), ".true.\n")
    if xshift_override is not None:
        _w(os.path.join(d, "xshift_override.ipt"), "%d %d %d\n" % tuple(xshift_override))
    # Bloch functions: orthonormal over x per (k, spin), all bands in the file
    u = np.zeros((nspn, nk, bif, nx), np.complex128)
    for s in range(nspn):
        for k in range(nk):
            q, _ = np.linalg.qr(rng.standard_normal((nx, bif)) + 1j * rng.standard_normal((nx, bif)))
            u[s, k] = q.T
    obf_arrays = None
    if obf is None:
        u.tofile(os.path.join(d, "con.u2.dat"))
    else:
        no, ivh = int(obf["num"]), 4
        width = ivh + bif
        _w(os.path.join(d, "bloch_control"), ".true.\n%s\n" % (".true." if obf.get("fake") else ".false."))
        _w(os.path.join(d, "obf_control"), "%d\n%d\n%d\n" % (no, ivh, width))
        q, _ = np.linalg.qr(rng.standard_normal((nx, no)) + 1j * rng.standard_normal((nx, no)))
        f = q.T / np.sqrt(nk)                                           # [o, x], ix fastest
        # record o: complex(X3, X2, X1), element (iz, iy, ix)
        recs = [f[o].reshape(xmesh[2], xmesh[1], xmesh[0]) for o in range(no)]
        _fortran_sequential(os.path.join(d, "u1.dat"), recs)
        coef = rng.standard_normal((nk, width, no)) + 1j * rng.standard_normal((nk, width, no))
        coef /= np.sqrt(np.sum(np.abs(coef) ** 2, axis=2, keepdims=True))
        coef.tofile(os.path.join(d, "con_eigvecs.dat"))                 # per k: (no, width)
        obf_arrays = dict(obf=f, obf2u=np.asfortranarray(coef[:, ivh:ivh + nb, :].transpose(1, 2, 0)))
    # band energies (Hartree), sorted per k
    E = np.sort(rng.uniform(0.0, 30.0, size=(bif, nk, nspn)), axis=0) * EV2HARTREE
    hdr = np.array([bif, nk] + ([nspn] if nspn == 2 else []), np.int32)
    _fortran_sequential(os.path.join(d, "wvfcninfo"), [hdr, np.asfortranarray(E)])
    # the inputs of nbsemkcmel, which ocean.x runs at set-up: the screened
    # potential as (V, r) rows from r = 0 to 12 bohr, the projector radius and
    # logarithmic radial grid (rmax on the grid, 60 Bode panels inside it), the
    # projector radial functions per l
    add14 = "%2s%04d_n%02dl%02d" % (el, indx, n, lc)
    rv = 12.0 * np.arange(0, 201) / 200.0
    v = np.exp(-rv / 3.0) / (4.0 * np.maximum(rv, 0.2))
    _w(os.path.join(d, "rpot.z" + add14), "".join("%.17g %.17g\n" % (a, b) for a, b in zip(v, rv)))
    rmax, dl, nr = 2.0, 0.025, 321
    rad = rmax * np.exp(dl * (np.arange(nr) - 240))
    _w(os.path.join(d, "radfilez%03d" % Z), "%.17g %d %d\n" % (rmax, nr, nr))
    nl = lmax - lmin + 1
    assert len(nproj) == nl
    _w(os.path.join(d, "prjfilez%03d" % Z), "%d %d\n%s\n" % (lmin, lmax, " ".join(str(v) for v in nproj)))
    ae = {}
    for li, l in enumerate(range(lmin, lmax + 1)):
        ph = np.stack([rad ** l * np.exp(-0.5 * (rad / (0.8 * (1.0 + 0.3 * nu))) ** 2) for nu in range(nproj[li])],
                      axis=1)
        ph = ph / np.sqrt(np.trapezoid(ph * ph * (rad ** 2)[:, None], rad, axis=0))
        ph = ph * rng.uniform(0.8, 1.2, nproj[li])
        ae[l] = ph
        _w(os.path.join(d, "ae%dz%03d" % (l, Z)),
           "".join(" ".join("%.17g" % t for t in (rad[i],) + tuple(ph[i])) + "\n" for i in range(nr)))
    add10 = "z%03dn%02dl%02d" % (Z, n, lc)
    slater = {}
    for li, lv in enumerate(range(lvl, lvh + 1)):
        npl = nproj[lv - lmin]
        fpre, fks = ("kk", [0]) if tdlda else ("fk", range(2, min(2 * lc, 2 * lv) + 1, 2))
        for pre, ks in ((fpre, fks), ("gk", range(abs(lc - lv), lc + lv + 1, 2))):
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
    return dict(u=u, E=E, rpot=(v, rv), rmax=rmax, rad=rad, ae=ae, slater=slater, cks=cks, mels=mels, avec=avec,
                tau=np.array(tau), nb=nb, kmesh=kmesh, xmesh=xmesh, lc=lc, nspn=nspn, bif=bif, lmin=lmin,
                lmax=lmax, nproj=list(nproj), xi_eV=xi_eV, inter_scale=inter_scale, eps=eps,
                efermi=efermi_eV * EV2HARTREE, eshift=eshift_eV * EV2HARTREE, nptot=nptot, celvol=vol, niter=niter,
                obf=obf_arrays, life_time=None if life_time_eV is None else np.array(life_time_eV, float),
                rpa=rpa, tdlda=tdlda, xshift_override=xshift_override)


__all__ = ["read_system", "read_u2", "read_bloch_control", "read_obf", "read_energies", "read_rpottrim", "read_prjfile", "read_parcksc",
           "projector_array", "read_cmel", "read_slater", "read_mels", "starting_vector", "read_rundir",
           "open_rundir", "fortran_records", "RunInputs", "RunSystem", "CoreRun", "hartree_window",
           "write_synthetic_rundir"]
