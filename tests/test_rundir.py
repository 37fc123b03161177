"""The run-directory reader (ocean_amd.rundir) against the files a synthetic OCEAN2
core run directory holds and the oracle's restatements of the setup steps."""
import math
import os

import numpy as np
import pytest

import ocean_oracle as orc
from ocean_amd import rundir
from ocean_amd import setup_host as sh
from rundir_files import EV2H, expected_tables, write_rundir

SHAPES = {
    "k_edge": dict(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=1, nproj=(2, 1, 1)),
    "l23_spin": dict(nb=6, kmesh=(2, 1, 3), xmesh=(3, 2, 3), lc=1, nspn=2, seed=2, xi_eV=3.0, lmin=1, lmax=2,
                     nproj=(1, 2)),
}


@pytest.fixture(scope="module", params=list(SHAPES))
def case(request, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("run_" + request.param))
    gen = write_rundir(d, **SHAPES[request.param])
    return d, gen, rundir.read_rundir(d)


def test_system_and_bloch_files(case):
    d, gen, R = case
    S = R.system
    assert S.xmesh == tuple(gen["xmesh"]) and S.kmesh == tuple(gen["kmesh"]) and S.nb == gen["nb"]
    assert S.bands_in_file == gen["bif"] and S.nspn == gen["nspn"]
    assert np.array_equal(S.avec, gen["avec"]) and np.array_equal(S.tau, gen["tau"])
    assert S.run.calc_type == "XAS" and S.solver["niter"] == gen["niter"]
    # con.u2.dat: [spin][k][band][x] -> (nb, nk, nx, nspn), extra bands skipped
    ref = gen["u"][:, :, :gen["nb"], :].transpose(2, 1, 3, 0)
    assert np.array_equal(R.u_raw, ref)


def test_energies_fermi_and_shift(case):
    d, gen, R = case
    E = gen["E"][:gen["nb"]]
    en, al = orc.energies_core_xas(E, gen["efermi"], gen["eshift"], gen["lc"], gen["nspn"])
    assert np.array_equal(R.energy, en) and np.array_equal(R.allow, al)
    assert R.efermi == gen["efermi"] and abs(R.shift - gen["eshift"]) < 1e-15
    # the mask's defining inequality holds on every entry, and both values occur
    assert set(np.unique(R.allow)) == {0.0, 1.0}
    assert np.all((R.energy.real - R.shift >= R.efermi) == (R.allow == 1.0))


def test_starting_vector(case):
    d, gen, R = case
    S = R.system
    psi = orc.psi_dotter(gen["mels"], gen["cks"], gen["lc"], gen["nb"], S.nk, gen["nspn"],
                         bands_in_file=gen["bif"])
    psi0, kpref = orc.core_pnorm(psi, S.nk, S.celvol)
    assert np.abs(R.psi0 - psi0).max() < 1e-14
    assert abs(R.kpref - kpref) <= 1e-14 * kpref


def test_multiplet_files(case):
    d, gen, R = case
    m = R.mult
    lmin, lmax, nproj = gen["lmin"], gen["lmax"], gen["nproj"]
    # mpm from the cmel files the reader formed (nbsemkcmel), as ocean.x reads them back
    _, _, mpm = expected_tables(gen)
    assert np.abs(m.mpm - mpm).max() <= 1e-15 * np.abs(mpm).max()
    # projector order (l, m, nu) of the cks rows, bands of each k past nb skipped
    S = R.system
    ip = 0
    for li, l in enumerate(range(lmin, lmax + 1)):
        for mm in range(-l, l + 1):
            for nu in range(nproj[li]):
                for k in range(S.nk):
                    rows = gen["cks"][ip, k * gen["bif"]:k * gen["bif"] + gen["nb"], :]
                    assert np.array_equal(m.c[:, k, nu, mm + lmax, li, :], rows)
                ip += 1
    # exchange blocks against the oracle's nbsemhsetup2
    for i, lv in enumerate(range(m.lvl, m.lvh + 1)):
        npl = nproj[lv - lmin]
        fk = {k: gen["slater"][("fk", lv, k)][0] for k in range(2, min(2 * gen["lc"], 2 * lv) + 1, 2)}
        sf = {k: gen["slater"][("fk", lv, k)][1] for k in fk}
        gk = {k: gen["slater"][("gk", lv, k)][0] for k in range(abs(gen["lc"] - lv), gen["lc"] + lv + 1, 2)}
        sg = {k: gen["slater"][("gk", lv, k)][1] for k in gk}
        ref = orc.nbsemhsetup2(gen["lc"], lv, npl, fk, sf, gk, sg) * EV2H
        assert np.abs(m.X[i] - ref).max() <= 1e-12 * max(1e-30, np.abs(ref).max())
    sr, si, _ = orc.soprep(gen["lc"], gen["xi_eV"])
    assert np.abs(m.somelr - sr).max() < 1e-15 and np.abs(m.someli - si).max() < 1e-15


def test_w_table_and_geometry(case):
    d, gen, R = case
    S = R.system
    rtab, ptab, _ = expected_tables(gen)
    assert np.array_equal(R.rtab, rtab) and np.array_equal(R.ptab, ptab)
    _, tau_c, xshift = orc.bloch_phase(np.zeros((1, S.nk, S.nx, 1), complex, order="F"), S.xmesh, S.kmesh,
                                       S.tau)
    Wref = orc.populate_W(S.avec, S.kmesh, S.xmesh, tau_c, xshift, R.rtab, R.ptab, S.epsilon0,
                          S.inter_scale)
    W = sh.build_w(S.avec, S.kmesh, S.xmesh, R.tau_c, R.xshift, R.rtab, R.ptab, S.epsilon0, S.inter_scale)
    assert np.abs(W - Wref).max() <= 1e-14 * np.abs(Wref).max()


def test_unsupported_inputs_are_refused(tmp_path):
    d = str(tmp_path / "run")
    write_rundir(d, seed=3)
    open(os.path.join(d, "isolated.inp"), "w").write("\n")
    with pytest.raises(NotImplementedError):
        rundir.read_rundir(d)
    os.remove(os.path.join(d, "isolated.inp"))
    open(os.path.join(d, "doping"), "w").write("0.1\n")
    with pytest.raises(NotImplementedError):
        rundir.read_rundir(d)


OBF_SHAPE = dict(nb=5, kmesh=(3, 2, 2), xmesh=(3, 4, 2), lc=0, seed=3, nproj=(2, 1, 1), tau=(0.9, 0.27, 0.41))


def test_obf_files(tmp_path):
    """bloch_control/obf_control/u1.dat/con_eigvecs.dat read back into the arrays the
    writer drew; make_phase's centring (odd kmesh(1): xshift(1) from xmesh(2))."""
    d = str(tmp_path / "obf")
    gen = write_rundir(d, obf=dict(num=7), **OBF_SHAPE)
    assert not os.path.exists(os.path.join(d, "con.u2.dat"))
    R = rundir.read_rundir(d)
    assert R.obf_mode == "obf" and R.u_raw is None
    assert np.array_equal(R.obf, gen["obf"]["obf"])
    assert np.array_equal(R.obf2u, gen["obf"]["obf2u"])
    S = R.system
    _, xoff, tau_c, xshift = orc.obf_phase(S.xmesh, S.kmesh, S.tau)
    assert np.array_equal(R.xoff, xoff) and np.array_equal(R.xshift, xshift)
    assert np.allclose(R.tau_c, tau_c, rtol=0, atol=1e-15)
    assert xshift[0] == math.floor(4 * (0.9 - 0.5))  # xmesh(2) = 4, not xmesh(1) = 3
    r, _ = sh.w_geometry(S.avec, S.kmesh, S.xmesh, tau_c, xshift)
    assert np.array_equal(R.rdist, r)


def test_fake_obf_bloch_functions(tmp_path):
    """use_fake_obf: the Bloch functions OCEAN_obf_lrLOAD builds, against the
    loop-literal restatement, then the Bloch form's centring."""
    d = str(tmp_path / "fake")
    gen = write_rundir(d, obf=dict(num=6, fake=True), **OBF_SHAPE)
    R = rundir.read_rundir(d)
    assert R.obf_mode == "fake"
    ref = orc.obf_lrload_bloch(gen["obf"]["obf"], gen["obf"]["obf2u"])
    assert np.abs(R.u_raw[..., 0] - ref).max() < 1e-14 * np.abs(ref).max()
    S = R.system
    _, _, xs = orc.bloch_phase(np.zeros((1, S.nk, S.nx, 1), complex), S.xmesh, S.kmesh, S.tau)
    assert np.array_equal(R.xshift, xs)


def test_obf_lrload_closed_forms():
    """One basis function: plain normalisation.  Two functions at one x point: the
    norm accumulates |c1 f1|^2 + |c1 f1 + c2 f2|^2 (the partial sum after each)."""
    rng = np.random.default_rng(4)
    f = rng.standard_normal((1, 5)) + 1j * rng.standard_normal((1, 5))
    c = rng.standard_normal((3, 1, 2)) + 1j * rng.standard_normal((3, 1, 2))
    u = orc.obf_lrload_bloch(f, c)
    ref = c[:, 0, :, None] * f[0][None, None, :]
    ref = ref / np.linalg.norm(ref, axis=2, keepdims=True)
    assert np.abs(u - ref).max() < 1e-14
    f2 = np.array([[1.0 + 2.0j], [0.5 - 1.0j]])
    c2 = np.array([[[2.0 - 1.0j], [1.0 + 1.0j]]])  # (nb=1, no=2, nk=1)
    s1 = (2.0 - 1.0j) * (1.0 + 2.0j)
    s2 = s1 + (1.0 + 1.0j) * (0.5 - 1.0j)
    want = s2 / math.sqrt(abs(s1) ** 2 + abs(s2) ** 2)
    assert abs(orc.obf_lrload_bloch(f2, c2)[0, 0, 0] - want) < 1e-15
    assert np.abs(sh.obf_fake_bloch(f2, c2)[0, 0, 0, 0] - want) < 1e-15


def test_obf_needs_one_spin(tmp_path):
    d = str(tmp_path / "obf2")
    write_rundir(d, obf=dict(num=4), **dict(SHAPES["l23_spin"]))
    with pytest.raises(NotImplementedError):
        rundir.read_rundir(d)


def test_core_broaden_lifetimes(tmp_path):
    """core_broaden.ipt: the spin-orbit matrix gains the lifetime term, equal to the
    oracle's soprep; the j = lc +- 1/2 levels get -life_time(1) and -life_time(2)."""
    d = str(tmp_path / "broaden")
    gen = write_rundir(d, life_time_eV=(0.3, 0.55), **SHAPES["l23_spin"])
    R = rundir.read_rundir(d)
    assert np.array_equal(R.life_time, gen["life_time"])
    sr, si, have = orc.soprep(gen["lc"], gen["xi_eV"], gen["life_time"])
    assert have and R.mult.have_spin_orbit
    assert np.abs(R.mult.somelr - sr).max() == 0.0 and np.abs(R.mult.someli - si).max() == 0.0


def test_rpa_tdlda_and_xshift_override(tmp_path):
    """rpa: no cmel (mpm zero) and no long-range term; tdlda: the blocks from the kk
    matrices (F^0) against the oracle's quadrature nbsemhsetup2; xshift_override.ipt
    factors on the x shift (an invalid value falls back to 1 1 1)."""
    d = str(tmp_path / "opts")
    gen = write_rundir(d, lc=1, xi_eV=2.0, lmin=0, lmax=2, nproj=(1, 1, 1), seed=5, tdlda=True, rpa=True,
                       xshift_override=(0, 1, 1), tau=(0.6, 0.27, 0.41))
    R = rundir.read_rundir(d)
    S = R.system
    assert S.rpa and S.tdlda and S.xshift_override == (0, 1, 1)
    assert np.abs(R.mult.mpm).max() == 0.0
    xo, tc, xs = sh.x_offsets(S.xmesh, S.kmesh, S.tau, (0, 1, 1))
    assert np.array_equal(R.xoff, xo) and np.array_equal(R.xshift, xs) and xs[0] == 0
    for i, lv in enumerate(range(S.lvl, S.lvh + 1)):
        fk = {0: gen["slater"][("kk", lv, 0)][0]}
        sf = {0: gen["slater"][("kk", lv, 0)][1]}
        ks = range(abs(1 - lv), 1 + lv + 1, 2)
        gk = {k: gen["slater"][("gk", lv, k)][0] for k in ks}
        sg = {k: gen["slater"][("gk", lv, k)][1] for k in ks}
        ref = orc.nbsemhsetup2(1, lv, 1, fk, sf, gk, sg, tdlda=True) * EV2H
        assert np.abs(R.mult.X[i] - ref).max() <= 1e-14 * np.abs(ref).max()
    open(os.path.join(d, "xshift_override.ipt"), "w").write("2 1 1\n")
    assert rundir.read_system(d).xshift_override == (1, 1, 1)
    # tdlda holding .false. is the plain F^k form; an empty file is an input error
    open(os.path.join(d, "tdlda"), "w").write(".false.\n")
    assert not rundir.read_system(d).tdlda
    open(os.path.join(d, "tdlda"), "w").write("")
    with pytest.raises(ValueError, match="tdlda"):
        rundir.read_system(d)


def test_nbsemkcmel_outputs_are_formed_like_ocean_x(tmp_path):
    """A fresh directory holds only nbsemkcmel's inputs; reading it writes rpottrim,
    rpotdiff and cmel<l> in the reference's formats (OCEAN2/nbsemkcmel.f90:134-202),
    whose values the reader then takes, and a second read reproduces them."""
    import re
    d = str(tmp_path / "fresh")
    gen = write_rundir(d, seed=4, nproj=(2, 1, 1))
    names = os.listdir(d)
    assert not any(n.startswith(("cmel", "rpottrim", "rpotdiff")) for n in names)
    R1 = rundir.read_rundir(d)
    add14 = "Fe0001_n01l00"
    lines = open(os.path.join(d, "rpottrim.z" + add14)).read().splitlines()
    assert re.fullmatch(r"#\s*\d+", lines[0]) and int(lines[0][1:]) == gen["rpot"][0].size
    # (E24.16, 1X, E24.16): right-justified 0.d..dE+ee fields
    for ln in lines[1:]:
        assert len(ln) == 49 and ln[24] == " "
        for f in (ln[:24], ln[25:]):
            assert re.fullmatch(r" *-?0\.\d{16}E[+-]\d\d", f), f
    cm = open(os.path.join(d, "cmel0")).read().splitlines()
    # 2 x 2 cmel then 2 x 2 nmel, four (1X, E15.8) fields per line
    assert len(cm) == 2
    for ln in cm:
        assert len(ln) == 64
        for i in range(4):
            f = ln[16 * i:16 * i + 16]
            assert f[0] == " " and re.fullmatch(r" *-?0\.\d{8}E[+-]\d\d", f[1:]), f
    assert os.path.exists(os.path.join(d, "rpotdiff.z" + add14))
    R2 = rundir.read_rundir(d)
    assert np.array_equal(R1.mult.mpm, R2.mult.mpm) and np.array_equal(R1.ptab, R2.ptab)
    # the flat part of the table inside the projector radius and the tail outside it
    v, rv = gen["rpot"]
    inside = rv < gen["rmax"]
    assert np.all(R1.ptab[inside] == R1.ptab[inside][0])
    assert np.allclose(R1.ptab[~inside], v[~inside], rtol=1e-15, atol=0)


def test_projector_cmel_bode_rule_closed_form():
    """The product's Bode-rule quadrature on the logarithmic grid against a closed
    form: phi = r exp(-r^2/2) (one projector), vdiff = 1 - r/rmax inside rmax."""
    rmax, dl = 2.0, 0.01
    rad = rmax * np.exp(dl * (np.arange(1201) - 1000))
    phi = (rad * np.exp(-0.5 * rad ** 2))[:, None]
    rv = np.linspace(0.0, 10.0, 20001)
    vdiff = np.where(rv < rmax, 1.0 - rv / rmax, 0.0)
    cmel, nmel = sh.projector_cmel(rad, phi, rv, vdiff, rmax)
    from scipy.integrate import quad
    ref_n = quad(lambda r: r ** 4 * np.exp(-r * r), 0.0, rmax)[0]
    ref_c = quad(lambda r: r ** 4 * np.exp(-r * r) * (1.0 - r / rmax), 0.0, rmax)[0]
    assert abs(nmel[0, 0] - ref_n) < 1e-7 * ref_n
    assert abs(cmel[0, 0] / sh.HARTREE2EV - ref_c) < 1e-6 * ref_c
