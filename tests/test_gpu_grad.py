"""Differentiation form: gradient of a GMRES spectral misfit through the C-ABI.

Checked against (1) the oracle's adjoint gradient with dense solves (itself pinned
by central differences in test_oracle.py) and (2) central finite differences
(h = 1e-6) of the GPU loss, for the W radial table, the screening scale, the
broadening and the scissor shift.  The GPU call sits behind a
torch.autograd.Function (CPU tensors: torch never touches the GPU here).
"""
import numpy as np
import pytest

import ocean_oracle as orc
from grad_common import oracle_problem, target_params

pytestmark = pytest.mark.gpu

from ocean_amd import core, setup_host as sh, synth  # noqa: E402

E = np.linspace(-2.0, 32.0, 6)
GPRC, FFFF, DEPTH = 13.605, 1e-26, 40


@pytest.fixture(scope="module")
def setup():
    inp = synth.make_inputs(nb=4, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=31, inter_scale=0.9)
    tp, ts, tg, td = target_params(inp)
    target = orc.spectral_sigma(oracle_problem(inp, tp, ts, td), inp.psi0, E, tg, inp.kpref)
    rdist = sh.w_geometry(inp.avec, inp.kmesh, inp.xmesh, inp.tau, inp.xshift)[0]
    return inp, target, rdist


def gpu_at(inp, rdist, ptab, s, scissor_eV, precision=core.FP64):
    """A handle whose state is the problem at theta (W from the device builder)."""
    m = inp.mult
    g = core.OceanGPU(inp.nb, inp.kmesh, inp.xmesh, inp.lc, inp.nspn, inp.celvol, s, m.lmin, m.lmax,
                      tuple(m.nproj), m.lvl, m.lvh, precision=precision)
    g.load_bloch(inp.u)
    g.build_w(rdist, inp.rtab, ptab, inp.epsilon0, s)
    g.load_mult(m)
    g.load_energies(np.asfortranarray(inp.energy + scissor_eV * orc.EV2HARTREE), inp.allow)
    g.kpref = inp.kpref
    return g


def run(inp, rdist, target, ptab, s, gam, d, precision=core.FP64, ffff=FFFF):
    g = gpu_at(inp, rdist, ptab, s, d, precision)
    out = g.spectral_grad(inp.psi0, E, gam, GPRC, ffff, DEPTH, target)
    g.close()
    return out


def test_gradient_matches_oracle_and_finite_differences(setup):
    inp, target, rdist = setup
    p0, s0 = inp.ptab.copy(), inp.inter_scale
    G = run(inp, rdist, target, p0, s0, 0.5, 0.0)
    ww = orc.populate_W_weights(inp.avec, inp.kmesh, inp.xmesh, inp.tau, inp.xshift, inp.rtab, s0)
    R = orc.spectral_grad(oracle_problem(inp, p0, s0, 0.0), inp.psi0, E, 0.5, inp.kpref, target,
                          wweights=ww, nptab=len(p0))
    print("loss", G["loss"], R["loss"], flush=True)
    assert abs(G["loss"] - R["loss"]) <= 1e-9 * R["loss"]
    for k in ("g_gamma_eV", "g_scissor_eV", "g_scale"):
        print(k, G[k], R[k], flush=True)
        assert abs(G[k] - R[k]) <= 1e-7 * abs(R[k]), k
    gp, rp = G["g_ptab"], R["g_ptab"]
    assert np.max(np.abs(gp - rp)) <= 1e-7 * np.max(np.abs(rp))

    h = 1e-6

    def L(p=p0, s=s0, gam=0.5, d=0.0):
        return run(inp, rdist, target, p, s, gam, d)["loss"]

    fd = {"g_gamma_eV": (L(gam=0.5 + h) - L(gam=0.5 - h)) / (2 * h),
          "g_scissor_eV": (L(d=h) - L(d=-h)) / (2 * h),
          "g_scale": (L(s=s0 + h) - L(s=s0 - h)) / (2 * h)}
    for k, v in fd.items():
        print("fd", k, G[k], v, flush=True)
        assert abs(G[k] - v) <= 1e-4 * abs(v), k
    for i in np.argsort(-np.abs(gp))[:3]:
        dp = np.zeros_like(p0)
        dp[i] = h
        v = (L(p=p0 + dp) - L(p=p0 - dp)) / (2 * h)
        print("fd ptab", i, gp[i], v, flush=True)
        assert abs(gp[i] - v) <= 1e-4 * abs(v)


def test_mixed_gradient_close_to_fp64(setup):
    inp, target, rdist = setup
    G64 = run(inp, rdist, target, inp.ptab, inp.inter_scale, 0.5, 0.0)
    # an FP32 action bounds the reachable residual: |g|^2 < 1e-12 (|g| ~ 1e-6)
    G32 = run(inp, rdist, target, inp.ptab, inp.inter_scale, 0.5, 0.0, precision=core.MIXED, ffff=1e-12)
    for k in ("g_gamma_eV", "g_scissor_eV", "g_scale"):
        print(k, G64[k], G32[k], flush=True)
        assert abs(G32[k] - G64[k]) <= 1e-3 * abs(G64[k])
    assert np.max(np.abs(G32["g_ptab"] - G64["g_ptab"])) <= 1e-3 * np.max(np.abs(G64["g_ptab"]))


def test_autograd_function(setup):
    torch = pytest.importorskip("torch")
    inp, target, rdist = setup

    class SpectralMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ptab, scale, gamma, scissor):
            out = run(inp, rdist, target, ptab.detach().numpy().astype(np.float64), float(scale),
                      float(gamma), float(scissor))
            ctx.save_for_backward(torch.from_numpy(out["g_ptab"].copy()),
                                  torch.tensor(out["g_scale"], dtype=torch.float64),
                                  torch.tensor(out["g_gamma_eV"], dtype=torch.float64),
                                  torch.tensor(out["g_scissor_eV"], dtype=torch.float64))
            return torch.tensor(out["loss"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            gp, gs, gg, gd = ctx.saved_tensors
            return gout * gp, gout * gs, gout * gg, gout * gd

    ptab = torch.tensor(inp.ptab, dtype=torch.float64, requires_grad=True)
    scale = torch.tensor(inp.inter_scale, dtype=torch.float64, requires_grad=True)
    gamma = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    scissor = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    loss = SpectralMisfit.apply(ptab, scale, gamma, scissor)
    (2.0 * loss).backward()
    ref = run(inp, rdist, target, inp.ptab, inp.inter_scale, 0.5, 0.0)
    assert np.allclose(ptab.grad.numpy(), 2.0 * ref["g_ptab"], rtol=0, atol=0)
    assert float(gamma.grad) == 2.0 * ref["g_gamma_eV"]
    # one descent step along the gradient lowers the misfit
    with torch.no_grad():
        step = 1e-3 / max(float(gamma.grad.abs()), 1e-30)
        l2 = SpectralMisfit.apply(ptab, scale, gamma - step * gamma.grad, scissor)
    assert l2.item() < loss.item()
