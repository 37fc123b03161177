"""A core run read from its run directory, phased and W-built on the device, against
the oracle's action and recursion on the same files."""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr
from rundir_files import EV2H, expected_tables, write_rundir

pytestmark = pytest.mark.gpu

SHAPES = {
    "k_edge": dict(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=1, nproj=(2, 1, 1)),
    "l23_spin": dict(nb=6, kmesh=(2, 1, 3), xmesh=(3, 2, 3), lc=1, nspn=2, seed=2, xi_eV=3.0, lmin=1, lmax=2,
                     nproj=(1, 2)),
    # core_broaden.ipt: spin-orbit-dependent lifetimes make S complex (non-Hermitian H)
    "l23_broaden": dict(nb=6, kmesh=(2, 2, 1), xmesh=(3, 3, 2), lc=1, seed=4, xi_eV=4.0, lmin=1, lmax=2,
                        nproj=(1, 1), life_time_eV=(0.3, 0.55)),
    # tdlda (F^0 from the kk files) with a partial x-shift override
    "l23_tdlda": dict(nb=6, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=1, seed=5, xi_eV=2.0, lmin=0, lmax=2,
                      nproj=(1, 1, 1), tdlda=True, xshift_override=(0, 1, 1)),
    # rpa: no long-range term and no direct multiplet term
    "k_rpa": dict(nb=6, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=6, nproj=(1, 1, 1), rpa=True),
}


def oracle_problem(gen, R):
    S = R.system
    xo = gen.get("xshift_override") or (1, 1, 1)
    u, tau_c, xshift = orc.bloch_phase(R.u_raw, S.xmesh, S.kmesh, S.tau, xo)
    # the W table and the direct projector matrix from the oracle's nbsemkcmel on the
    # directory's rpot / radfile / ae files (the reader forms them itself)
    rtab, ptab, mpm = expected_tables(gen)
    if gen.get("rpa"):
        mpm = np.zeros_like(mpm)
    W = orc.populate_W(S.avec, S.kmesh, S.xmesh, tau_c, xshift, rtab, ptab, S.epsilon0, S.inter_scale)
    energy, allow = orc.energies_core_xas(gen["E"][:gen["nb"]], gen["efermi"], gen["eshift"], gen["lc"],
                                          gen["nspn"])
    m = R.mult
    X = []
    for lv in range(m.lvl, m.lvh + 1):
        npl = m.nproj[lv - m.lmin]
        fpre, fks = ("kk", [0]) if gen.get("tdlda") else ("fk", range(2, min(2 * gen["lc"], 2 * lv) + 1, 2))
        fk = {k: gen["slater"][(fpre, lv, k)][0] for k in fks}
        sf = {k: gen["slater"][(fpre, lv, k)][1] for k in fk}
        gk = {k: gen["slater"][("gk", lv, k)][0] for k in range(abs(gen["lc"] - lv), gen["lc"] + lv + 1, 2)}
        sg = {k: gen["slater"][("gk", lv, k)][1] for k in gk}
        X.append(orc.nbsemhsetup2(gen["lc"], lv, npl, fk, sf, gk, sg, tdlda=bool(gen.get("tdlda"))) * EV2H)
    sr, si, have_so = orc.soprep(gen["lc"], gen["xi_eV"], gen.get("life_time"))
    M = orc.Multiplet(m.lmin, m.lmax, list(m.nproj), m.lvl, m.lvh, m.c, mpm, X, sr, si, have_so)
    P = orc.Problem(S.nb, S.kmesh, S.xmesh, gen["lc"], S.nspn, S.celvol, S.inter_scale, u, W, energy, allow, M,
                    long_range=not gen.get("rpa"))
    psi = orc.psi_dotter(gen["mels"], gen["cks"], gen["lc"], gen["nb"], S.nk, gen["nspn"],
                         bands_in_file=gen["bif"])
    psi0, kpref = orc.core_pnorm(psi, S.nk, S.celvol)
    return P, psi0, kpref


@pytest.mark.parametrize("name", list(SHAPES))
def test_rundir_action_and_recursion(tmp_path, name):
    from ocean_amd import rundir
    d = str(tmp_path / name)
    gen = write_rundir(d, **SHAPES[name])
    print("run directory case", name, flush=True)
    g, R = rundir.open_rundir(d)
    P, psi0, kpref = oracle_problem(gen, R)
    assert abs(g.kpref - kpref) <= 1e-14 * kpref
    for back in (False, True):
        assert relerr(g.xact(R.psi0, backwards=back), orc.xact(P, psi0, backwards=back)) < 1e-12, back
    niter = 20
    a, b, _ = g.haydock(R.psi0, niter)
    ra, rb, _, ierr = orc.haydock(P, psi0, niter)
    assert ierr == 0
    assert np.max(np.abs(a[:niter] - ra[:niter]) / np.abs(ra[:niter])) < 1e-10
    assert np.max(np.abs(b[1:niter + 1] - rb[1:niter + 1]) / np.abs(rb[1:niter + 1])) < 1e-10
    g.close()


@pytest.mark.parametrize("fake", [False, True])
def test_rundir_obf_forms(tmp_path, fake):
    """A run directory in the OBF form: lr_act_obf2 (use_obf) or the Bloch functions
    OCEAN_obf_lrLOAD builds (use_fake_obf), against the oracle."""
    from ocean_amd import rundir
    d = str(tmp_path / ("fake" if fake else "obf"))
    shape = dict(nb=5, kmesh=(3, 2, 2), xmesh=(3, 4, 2), lc=0, seed=3, nproj=(2, 1, 1), tau=(0.9, 0.27, 0.41))
    gen = write_rundir(d, obf=dict(num=7, fake=fake), **shape)
    print("OBF run directory, fake", fake, flush=True)
    g, R = rundir.open_rundir(d)
    S = R.system
    f, c = gen["obf"]["obf"], gen["obf"]["obf2u"]
    if fake:
        R2 = rundir.RunInputs(**{**R.__dict__, "u_raw": orc.obf_lrload_bloch(f, c)[..., None]})
        P, psi0, _ = oracle_problem(gen, R2)
        ref = [orc.xact(P, psi0, backwards=b) for b in (False, True)]
    else:
        phs, _, tau_c, xshift = orc.obf_phase(S.xmesh, S.kmesh, S.tau)
        R2 = rundir.RunInputs(**{**R.__dict__, "u_raw": np.zeros((S.nb, S.nk, S.nx, 1), complex)})
        P, psi0, _ = oracle_problem(gen, R2)
        P.W = orc.populate_W(S.avec, S.kmesh, S.xmesh, tau_c, xshift, R.rtab, R.ptab, S.epsilon0,
                             S.inter_scale)
        ref = [orc.xact_obf(P, psi0, f, c, phs, backwards=b) for b in (False, True)]
    for back, r in zip((False, True), ref):
        assert relerr(g.xact(R.psi0, backwards=back), r) < 1e-12, back
    g.close()
