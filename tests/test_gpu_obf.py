"""GPU OBF long-range form (ocean_gpu_load_obf) vs the oracle's lr_act_obf2 path.

Same bars as the Bloch form: H.psi within 1e-12 relative, Haydock a/b within 1e-10
relative over the first 50 steps.
"""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import relerr, to_problem

pytestmark = pytest.mark.gpu

from ocean_amd import synth  # noqa: E402

CASES = {
    "k_obf": dict(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=3, no=11),
    # odd kmesh(1) (make_phase's xmesh(2) shift), ragged band and basis counts
    "k_obf_odd": dict(nb=13, kmesh=(3, 4, 2), xmesh=(4, 3, 5), lc=0, seed=4, no=37),
    # L2,3 with spin-orbit; more than 64 basis functions (two o tiles)
    "l23_obf": dict(nb=16, kmesh=(4, 4, 4), xmesh=(4, 4, 5), lc=1, seed=5, xi_eV=6.0, no=70),
    "k_obf_6cube": dict(nb=32, kmesh=(6, 6, 6), xmesh=(6, 6, 6), lc=0, seed=6, no=100),
}


def _setup(name):
    c = dict(CASES[name])
    no = c.pop("no")
    inp = synth.make_inputs(**c)
    ob = synth.make_obf(inp, no, seed=c["seed"] + 100)
    tau0 = inp.tau + inp.xshift / np.array(inp.xmesh, float)
    phs, xoff, _, _ = orc.obf_phase(inp.xmesh, inp.kmesh, tau0)
    assert np.array_equal(ob.xoff, xoff)
    P = to_problem(inp)
    P.W = ob.W
    from ocean_amd import core
    g = core.OceanGPU.from_inputs(inp)
    g.load_w(ob.W)
    g.load_obf(ob.obf, ob.obf2u, ob.xoff)
    return inp, ob, phs, P, g


@pytest.mark.parametrize("name", list(CASES))
def test_xact_obf_matches_oracle(name):
    print("case", name, flush=True)
    inp, ob, phs, P, g = _setup(name)
    rng = np.random.default_rng(21)
    psi = np.asfortranarray(rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape))
    for back in (False, True):
        ref = orc.xact_obf(P, psi, ob.obf, ob.obf2u, phs, backwards=back)
        got = g.xact(psi, backwards=back)
        assert relerr(got, ref) < 1e-12, (name, back, relerr(got, ref))
    for hf in ((1, 0, 0, 0, 0, 0), (0, 0, 0, 0, 1, 0), (0, 0, 0, 0, 0, 1)):
        ref = orc.xact_obf(P, psi, ob.obf, ob.obf2u, phs, hflag=hf)
        got = g.xact(psi, hflag=hf)
        assert relerr(got, ref) < 1e-12, (name, hf, relerr(got, ref))
    g.close()


def test_haydock_obf_matches_oracle():
    inp, ob, phs, P, g = _setup("l23_obf")
    niter = 60
    act = lambda v: orc.xact_obf(P, v, ob.obf, ob.obf2u, phs)  # noqa: E731
    a_ref, b_ref, _, ierr = orc.haydock(P, inp.psi0, niter, act=act)
    a, b, _ = g.haydock(inp.psi0, niter)
    assert ierr == 0
    ra = np.abs(a[:50] - a_ref[:50]) / np.abs(a_ref[:50])
    rb = np.abs(b[1:51] - b_ref[1:51]) / np.abs(b_ref[1:51])
    print("max rel a", ra.max(), "b", rb.max())
    assert ra.max() < 1e-10 and rb.max() < 1e-10
    g.close()


def test_obf_drop_in_shape_matches_oracle():
    """The K-edge drop-in shapes (nb 64, k-mesh 8^3, x-mesh 8^3) with 128 basis functions."""
    inp = synth.make_inputs(64, (8, 8, 8), (8, 8, 8), seed=0)
    ob = synth.make_obf(inp, 128, seed=1)
    tau0 = inp.tau + inp.xshift / np.array(inp.xmesh, float)
    phs, _, _, _ = orc.obf_phase(inp.xmesh, inp.kmesh, tau0)
    P = to_problem(inp)
    P.W = ob.W
    from ocean_amd import core
    g = core.OceanGPU.from_inputs(inp)
    g.load_w(ob.W)
    g.load_obf(ob.obf, ob.obf2u, ob.xoff)
    ref = orc.xact_obf(P, inp.psi0, ob.obf, ob.obf2u, phs)
    got = g.xact(inp.psi0)
    assert relerr(got, ref) < 1e-12, relerr(got, ref)
    g.close()


def test_obf_form_switches_and_guards():
    from ocean_amd import core
    inp, ob, phs, P, g = _setup("k_obf")
    with pytest.raises(core.OceanGPUError):
        g.set_precision(core.MIXED)
    # loading Bloch functions returns the handle to the Bloch form
    g.load_bloch(inp.u)
    g.load_w(inp.W)
    P0 = to_problem(inp)
    psi = inp.psi0
    assert relerr(g.xact(psi), orc.xact(P0, psi)) < 1e-12
    # and back again
    g.load_w(ob.W)
    g.load_obf(ob.obf, ob.obf2u, ob.xoff)
    assert relerr(g.xact(psi), orc.xact_obf(P, psi, ob.obf, ob.obf2u, phs)) < 1e-12
    g.close()
