"""Pins of the oracle's non-Hermitian (bi-Lanczos) recursion and of its complex
continued fraction (OCEAN2/OCEAN_haydock.f90:53-160, :1250-1433, :1836-1880).

The recursion is checked against what its definition fixes, through the dense
operator rather than the vector code:
  - on a Hermitian problem whose backwards action equals the forward one it reduces to
    the Hermitian recursion (a real, b = c = the Hermitian b);
  - its vectors satisfy the three-term relations of both sides with the dense H and the
    dense backwards operator, which is H^dagger for a complex-symmetric spin-orbit
    matrix and complex band energies;
  - its first coefficients equal closed forms in the moments m_k = <v|A^k|v>:
      a(0) = m1,  c(1) = sqrt(m2 - m1^2) (principal root),  b(1) = conj(c(1)),
      a(1) = (m3 - 2 m1 m2 + m1^3) / |m2 - m1^2|
    (the last because the reference scales the back vector by r/c = c, so that
    <w2|v2> = r/|r|, not 1).
The fraction with complex b and c is checked against the (0,0) element of the
resolvent of the tridiagonal matrix it encodes, and reduces exactly to the Hermitian
write_core when c = b is real.
"""
import numpy as np
import pytest

import ocean_oracle as orc
from helpers import to_problem

from ocean_amd import setup_host as sh, synth  # noqa: E402


def _problem(seed, symmetric_complex, complex_energy=False, sparse_allow=False):
    inp = synth.make_inputs(nb=3, kmesh=(1, 1, 2), xmesh=(2, 2, 2), lc=0, seed=seed)
    rng = np.random.default_rng(seed + 50)
    na = inp.energy.shape[2]
    if symmetric_complex:
        s = rng.standard_normal((na, na)) + 1j * rng.standard_normal((na, na))
        s = 0.05 * (s + s.T)
    else:
        s = rng.standard_normal((na, na))
        s = 0.05 * (s + s.T) + 0j
    inp.mult.somelr = np.asfortranarray(s.real)
    inp.mult.someli = np.asfortranarray(s.imag)
    inp.mult.have_spin_orbit = True
    if complex_energy:
        inp.energy = np.asfortranarray(inp.energy + 1j * 0.02 * rng.uniform(size=inp.energy.shape))
    if sparse_allow:
        inp.allow = np.asfortranarray((rng.uniform(size=inp.allow.shape) > 0.25).astype(np.float64))
    z = rng.standard_normal(inp.energy.shape) + 1j * rng.standard_normal(inp.energy.shape)
    psi0 = sh.normalise_psi0(np.asfortranarray(z), inp.nk, inp.celvol)[0]
    return inp, to_problem(inp), psi0


def test_abc_reduces_to_hermitian_recursion():
    _, P, psi0 = _problem(5, symmetric_complex=False)
    H = orc.dense_H(P)
    assert np.abs(H - H.conj().T).max() < 1e-15
    n = 10
    ra, ia, rb, ib, rc, ic, ierr = orc.haydock_abc(P, psi0, n)
    ha, hb, _, herr = orc.haydock(P, psi0, n)
    assert ierr == 0 and herr == 0
    assert np.abs(ra[:n] - ha[:n]).max() < 1e-12
    assert np.abs(ia[:n]).max() < 1e-12
    assert np.abs(rb[1:] - hb[1:]).max() < 1e-12
    assert np.abs(rc[1:] - hb[1:]).max() < 1e-12
    assert np.abs(ib).max() < 1e-12 and np.abs(ic).max() < 1e-12
    assert rb[0] == rc[0] == ib[0] == ic[0] == 0.0


def _tridiag(ra, ia, rb, ib, rc, ic, n):
    """(n+1) x n matrices of the forward and back three-term recurrences."""
    a = ra + 1j * ia
    b = rb + 1j * ib
    c = rc + 1j * ic
    Tf = np.zeros((n + 1, n), complex)
    Tb = np.zeros((n + 1, n), complex)
    for j in range(n):
        Tf[j, j] = a[j]
        Tb[j, j] = np.conj(a[j])
        if j > 0:
            Tf[j - 1, j] = b[j]
            Tb[j - 1, j] = np.conj(c[j])
        Tf[j + 1, j] = c[j + 1]
        Tb[j + 1, j] = np.conj(b[j + 1])
    return Tf, Tb


@pytest.mark.parametrize("complex_energy,sparse_allow", [(False, False), (True, True)])
def test_abc_three_term_relations_with_dense_operators(complex_energy, sparse_allow):
    _, P, psi0 = _problem(6, symmetric_complex=True, complex_energy=complex_energy, sparse_allow=sparse_allow)
    H = orc.dense_H(P)
    Hb = orc.dense_H(P, backwards=True)
    assert np.abs(H - H.conj().T).max() > 1e-3  # not Hermitian
    assert np.abs(Hb - H.conj().T).max() < 1e-15  # the backwards action is H^dagger
    M = np.diag(P.allow.ravel(order="F"))
    A, Ab = M @ H @ M, M @ Hb @ M
    n = 8
    ra, ia, rb, ib, rc, ic, ierr, V, W = orc.haydock_abc(P, psi0, n, vectors=True)
    assert ierr == 0
    Vm = np.stack([v.ravel(order="F") for v in V], axis=1)
    Wm = np.stack([w.ravel(order="F") for w in W], axis=1)
    Tf, Tb = _tridiag(ra, ia, rb, ib, rc, ic, n)
    scale = np.abs(A).max()
    assert np.abs(A @ Vm[:, :n] - Vm @ Tf).max() < 1e-12 * scale
    assert np.abs(Ab @ Wm[:, :n] - Wm @ Tb).max() < 1e-12 * scale
    # a sign error in any coefficient breaks the relations
    Tbad = Tf.copy()
    Tbad[0, 1] = np.conj(Tbad[0, 1])
    assert np.abs(A @ Vm[:, :n] - Vm @ Tbad).max() > 1e-6


def test_abc_first_coefficients_from_moments():
    _, P, psi0 = _problem(7, symmetric_complex=True, complex_energy=True)
    H = orc.dense_H(P)
    M = np.diag(P.allow.ravel(order="F"))
    A = M @ H @ M
    v = M @ psi0.ravel(order="F")
    assert abs(np.vdot(v, v) - 1.0) < 1e-14
    m1, m2, m3 = (np.vdot(v, np.linalg.matrix_power(A, k) @ v) for k in (1, 2, 3))
    ra, ia, rb, ib, rc, ic, ierr = orc.haydock_abc(P, psi0, 3)
    assert ierr == 0
    r1 = m2 - m1 ** 2
    assert abs(r1.imag) > 1e-6 * abs(r1)  # a genuinely complex <w|v>
    assert abs(complex(ra[0], ia[0]) - m1) < 1e-13
    c1 = np.sqrt(r1)
    assert abs(complex(rc[1], ic[1]) - c1) < 1e-13
    assert abs(complex(rb[1], ib[1]) - np.conj(c1)) < 1e-13
    a1 = (m3 - 2 * m1 * m2 + m1 ** 3) / abs(r1)
    assert abs(complex(ra[1], ia[1]) - a1) < 1e-12 * abs(a1)


def test_write_core_abc_is_write_core_for_hermitian_coefficients():
    rng = np.random.default_rng(3)
    n = 30
    a = rng.uniform(0.0, 1.0, n + 1)
    b = rng.uniform(0.1, 0.4, n + 1)
    ia = rng.uniform(0.0, 0.01, n + 1)
    z = np.zeros(n + 1)
    args = (n, 1.3, 0.02, -0.5, 1.5, 300)
    for lit in (True, False):
        ref = orc.write_core(a, b, *args, imag_a=ia, literal_cmplx=lit)
        got = orc.write_core_abc(a, ia, b, z, b, z, *args, literal_cmplx=lit)
        assert np.array_equal(ref, got)


def test_write_core_abc_is_the_tridiagonal_resolvent():
    """With b(n) c(n) = 0 the terminator is exact and 1/delta(e) is the (0,0) element of
    (z - T)^-1 for z = e + i gam and the tridiagonal T with diagonal a - i ia and
    off-diagonal products b(j) c(j): Spect = kpref |Im G00|, SPKK = kpref Re G00."""
    rng = np.random.default_rng(4)
    n = 12
    ra, ia = rng.uniform(0, 1, n + 1), rng.uniform(-0.05, 0.05, n + 1)
    rb, ib, rc, ic = (rng.uniform(-0.3, 0.3, n + 1) for _ in range(4))
    rb[n] = ib[n] = rc[n] = ic[n] = 0.0
    kpref, gam, el, eh, ne = 2.0, 0.03, -0.5, 1.5, 200
    got = orc.write_core_abc(ra, ia, rb, ib, rc, ic, n, kpref, gam, el, eh, ne, literal_cmplx=False,
                             hartree2ev=1.0)
    T = np.diag(ra[:n] - 1j * ia[:n]).astype(complex)
    for j in range(1, n):
        T[j - 1, j] = complex(rb[j], ib[j])
        T[j, j - 1] = complex(rc[j], ic[j])
    e = got[:, 0]
    G = np.array([np.linalg.inv((ee + 1j * gam) * np.eye(n) - T)[0, 0] for ee in e])
    assert np.abs(got[:, 1] - kpref * np.abs(G.imag)).max() < 1e-12 * np.abs(G).max() * kpref
    assert np.abs(got[:, 3] - kpref * G.real).max() < 1e-12 * np.abs(G).max() * kpref
    lit = orc.write_core_abc(ra, ia, rb, ib, rc, ic, n, kpref, gam, el, eh, ne, literal_cmplx=True,
                             hartree2ev=1.0)
    assert np.abs(lit[:, 1] - got[:, 1]).max() < 1e-5 * np.abs(got[:, 1]).max()
