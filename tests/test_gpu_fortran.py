"""The Fortran host (iso_c_binding module + driver) through the C-ABI on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import ocean_oracle as orc
from fortran_io import read_driver_output, write_driver_input
from helpers import relerr, to_problem
from ocean_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ocean_amd", "fortran", "ocean_gpu_driver.x")


def test_fortran_host_xact_haydock_gmres(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    inp = synth.make_inputs(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=1, seed=12, xi_eV=5.0)
    P = to_problem(inp)
    E2H = orc.EV2HARTREE
    niter, ne = 40, 300
    gam, el, eh = 0.5 * E2H, -5 * E2H, 40 * E2H
    energies = np.array([0.0, 8.0, 16.0])
    fin = tmp_path / "bse_gpu.in"
    write_driver_input(fin, inp, niter, ne, gam, el, eh, energies, ffff=1e-12, depth=30)
    print("fortran driver case: nb=8 k=2^3 x=3^3 lc=1", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ierr, done, h, a, b, rows = read_driver_output(tmp_path / "out.bin", (inp.nb, inp.nk, inp.nalpha),
                                                   niter, len(energies))
    assert ierr == 0 and done == niter
    assert relerr(h, orc.xact(P, inp.psi0)) < 1e-12
    ra, rb, ia, _ = orc.haydock(P, inp.psi0, niter)
    assert np.max(np.abs(a[:niter] - ra[:niter]) / np.abs(ra[:niter])) < 1e-10
    sp = orc.write_core_full(ra, rb, niter, inp.kpref, gam, el, eh, ne, ucvol=inp.celvol)
    txt = np.loadtxt(tmp_path / "out.absspct", comments="#")
    assert txt.shape == (ne + 1, 9)
    for c in range(1, 9):  # 8 printed digits
        assert np.abs(txt[:, c] - sp[:, c]).max() / np.abs(sp[:, c]).max() < 1e-7, c
    lanc = open(tmp_path / "out.abslanc").read().split("\n")
    assert int(lanc[0].split()[0]) == niter
    eig = np.array([float(l.split()[2]) for l in lanc[niter + 2:] if l.strip()])
    assert eig.shape == (niter + 1,)
    assert np.abs(eig - orc.redtrid_eigs(ra, rb, niter)).max() < 2e-10  # f20.10
    ref_rows = orc.gmres_do_recycle(P, inp.psi0, energies, 0.5, 13.605, 1e-12, 30, inp.kpref)
    assert np.abs(rows[:, 2] - ref_rows[:, 2]).max() / np.abs(ref_rows[:, 2]).max() < 1e-6


def test_fortran_early_exit(tmp_path):
    """OCEAN_haydock_do_gpu with the early-exit arguments stops where the reference's
    check_convergence would, and the files describe the iterations actually run."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (run __graft_entry__.build())")
    inp = synth.make_inputs(nb=8, kmesh=(2, 2, 2), xmesh=(3, 3, 3), lc=0, seed=13)
    P = to_problem(inp)
    E2H = orc.EV2HARTREE
    niter, ne = 60, 200
    gam, el, eh = 0.5 * E2H, -5 * E2H, 40 * E2H
    fin = tmp_path / "bse_gpu.in"
    write_driver_input(fin, inp, niter, ne, gam, el, eh, np.array([]), ffff=1e-12, depth=30)
    ra, rb, ia, _ = orc.haydock(P, inp.psi0, niter)
    spacing, thr, nconv = 5, 1e-2, 100
    stop = orc.haydock_early_exit(ra, rb, niter, spacing, thr, inp.kpref, gam, el, eh, nconv, ia)
    assert spacing < stop < niter  # the case exercises an actual early stop
    print("fortran early-exit case: stop expected at", stop, flush=True)
    r = subprocess.run([DRIVER, str(fin), str(tmp_path / "out"), str(spacing), repr(thr), str(nconv)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ierr, done, h, a, b, rows = read_driver_output(tmp_path / "out.bin", (inp.nb, inp.nk, inp.nalpha),
                                                   niter, 0)
    assert ierr == 0 and done == stop
    assert np.max(np.abs(a[:done] - ra[:done]) / np.abs(ra[:done])) < 1e-10
    lanc = open(tmp_path / "out.abslanc").read().split("\n")
    assert int(lanc[0].split()[0]) == done
