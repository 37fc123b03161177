"""The convention pins of test_oracle_pins.py through the device path: the device
phase (ocean_gpu_load_bloch_raw) and W builder (ocean_gpu_build_w) must give the
plane-wave Fourier sums of W, and the device action must select the valence spin of
each alpha as the alpha order says."""
import numpy as np
import pytest

import ocean_oracle as orc
from test_oracle_pins import alpha_spins

pytestmark = pytest.mark.gpu

from ocean_amd import core, setup_host as sh, synth  # noqa: E402


@pytest.mark.parametrize("lc", [0, 1])
@pytest.mark.parametrize("precision", [core.FP64, core.MIXED])
def test_device_valence_spin_of_alpha(lc, precision):
    inp = synth.make_inputs(nb=4, kmesh=(2, 2, 2), xmesh=(2, 3, 2), lc=lc, nspn=2, seed=12, xi_eV=1.0)
    inp.u[..., 1] = 0.0
    rng = np.random.default_rng(4)
    psi = np.asfortranarray(rng.standard_normal(inp.energy.shape) + 1j * rng.standard_normal(inp.energy.shape))
    with core.OceanGPU.from_inputs(inp, precision=precision, e0=False) as g:
        h = g.xact(psi, hflag=(1, 0, 0, 0, 0, 0))
    for a, (_, _, sv) in enumerate(alpha_spins(lc)):
        if sv == 1:
            assert not h[:, :, a].any(), a
        else:
            assert np.abs(h[:, :, a]).max() > 1e-6 * np.abs(h).max(), a


@pytest.mark.parametrize("kmesh,xmesh,tau", [((2, 2, 2), (4, 3, 3), (0.13, 0.27, 0.41)),
                                             ((3, 2, 1), (3, 4, 2), (0.52, 0.08, 0.77)),
                                             ((4, 1, 3), (2, 3, 5), (0.9, 0.45, 0.05))])
def test_device_plane_waves_give_the_fourier_sum_of_w(kmesh, xmesh, tau):
    """One constant band, phased on the device from the raw functions, W built on the
    device: <k|H_lr|k'> = -(N_k N_x)^-1 sum_r exp(-2 pi i (k - k').r) W(r - tau)."""
    avec = np.array([[6.0, 0.3, 0.0], [0.0, 5.5, 0.2], [0.1, 0.0, 6.3]])
    nk, nx = int(np.prod(kmesh)), int(np.prod(xmesh))
    rtab, ptab = synth.w_table()
    xoff, tau_c, xshift = sh.x_offsets(xmesh, kmesh, np.asarray(tau))
    rdist = sh.w_geometry(avec, kmesh, xmesh, tau_c, xshift)[0]
    g = core.OceanGPU(1, kmesh, xmesh, 0, 1, float(np.linalg.det(avec)), 1.0, e0=False, mult=False)
    g.load_bloch_raw(np.full((1, nk, nx, 1), 1.0 / np.sqrt(nx), complex, order="F"), xoff)
    g.build_w(rdist, rtab, ptab, 4.0, 1.0)
    g.load_energies(np.zeros(g.shape, complex, order="F"), np.ones(g.shape, order="F"))
    H = np.zeros((nk, nk), complex)
    for j in range(nk):
        e = np.zeros(g.shape, complex, order="F")
        e[0, j, 0] = 1.0
        H[:, j] = g.xact(e, hflag=(1, 0, 0, 0, 0, 0))[0, :, 0]
    g.close()
    # the closed form over the grid points at which lr_populate_W2 evaluates W
    W = orc.populate_W(avec, kmesh, xmesh, tau_c, xshift, rtab, ptab, 4.0)
    _, _, cart = orc._w_geometry(avec, kmesh, xmesh, tau_c, xshift)
    pos = np.einsum("ij,kxj->kxi", np.linalg.inv(avec), cart) + np.asarray(tau)[None, None, :]
    kv = np.array([(a / kmesh[0], b / kmesh[1], c / kmesh[2]) for a in range(kmesh[0])
                   for b in range(kmesh[1]) for c in range(kmesh[2])])
    Hc = np.zeros((nk, nk), complex)
    for k in range(nk):
        for kp in range(nk):
            Hc[k, kp] = -np.sum(np.exp(-2j * np.pi * np.einsum("i,kxi->kx", kv[k] - kv[kp], pos)) * W) / (nk * nx)
    assert np.abs(H - Hc).max() < 1e-13 * np.abs(Hc).max()
