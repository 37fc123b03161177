"""Parameterised spectral-misfit problems for the differentiation form.

theta = (ptab[200], s, gamma_eV, scissor_eV): the W radial table, the screening
scale (W and the multiplet prefactor), the broadening and the scissor shift.
"""
import copy

import numpy as np

import ocean_oracle as orc
from helpers import to_problem

EV2H = orc.EV2HARTREE


def oracle_problem(inp, ptab, s, scissor_eV):
    """The oracle's Problem at theta (W rebuilt by the oracle's own populate_W)."""
    P = to_problem(inp)
    P = copy.copy(P)
    P.W = orc.populate_W(inp.avec, inp.kmesh, inp.xmesh, inp.tau, inp.xshift, inp.rtab, ptab,
                         inp.epsilon0, s)
    P.inter_scale = s
    P.energy = np.asfortranarray(inp.energy + scissor_eV * EV2H)
    return P


def target_params(inp):
    """Perturbed parameters of the target spectrum: W x 1.1 plus a Gaussian bump,
    gamma 0.7 eV, scissor +0.3 eV."""
    bump = 0.05 * np.exp(-((inp.rtab - 2.0) / 0.5) ** 2)
    return inp.ptab * 1.1 + bump, inp.inter_scale, 0.7, 0.3
