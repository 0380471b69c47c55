"""The oracle's forward-mode parameter derivatives (dual numbers through the LJ and SW
IFC3 loops: fco_lj_ifc3_dparam, fco_sw_ifc3_dparam) pinned on the CPU.

Pins that do not restate the derivative: exact linearity where the tensor is linear in
a parameter (LJ in epsilon; SW in A and lambda, which scale the pair and three-body parts,
and in B, which enters the pair part linearly with q = 0), and Richardson-extrapolated
central differences of the primal oracle for the nonlinear ones (LJ sigma; SW gamma and
sigma).  A dropped term, a wrong sign or a wrong power of a parameter fails these.
"""
from dataclasses import replace

import numpy as np
import pytest

from fc3 import systems as S
from oracle import fc3_oracle as O

LJ = S.ARGON_LJ
SW = S.StillingerWeberSilicon()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _richardson(f, x, h):
    d1 = (f(x + h) - f(x - h)) / (2 * h)
    d2 = (f(x + h / 2) - f(x - h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("tol", [0.0, 1e-12])
def test_lj_epsilon_is_linear(tol):
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=3)
    calc = S.AnalyticalCalculator(tol, LJ.r_cut)
    d = O.lj_third_order_dparam(sysd, LJ, calc, "epsilon")
    assert _rel(d, O.lj_third_order(sysd, LJ, calc) / LJ.epsilon) <= 1e-13


def test_lj_sigma_against_richardson():
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=4)
    calc = S.AnalyticalCalculator(0.0, LJ.r_cut)
    d = O.lj_third_order_dparam(sysd, LJ, calc, "sigma")
    f = lambda s: O.lj_third_order(sysd, S.LJ(s, LJ.epsilon, LJ.r_cut), calc)
    fd = _richardson(f, LJ.sigma, 1e-3 * LJ.sigma)
    r = _rel(d, fd)
    print(f"LJ d/dsigma vs Richardson: {r:.2e}")
    assert r <= 1e-8


def test_lj_mass_weighting_of_derivative():
    sysd = S.fcc(S.ARGON_A, (2, 2, 2), S.ARGON_MASS, 0.05, seed=5)
    calc = S.AnalyticalCalculator(1e-12, LJ.r_cut)
    m = np.random.default_rng(6).uniform(20.0, 40.0, sysd.n_atoms)
    d = O.lj_third_order_dparam(sysd, LJ, calc, "epsilon", masses=m)
    psi = O.mass_weight_third_order(O.lj_third_order(sysd, LJ, calc), m)
    assert _rel(d, psi / LJ.epsilon) <= 1e-13


def _sw(**kw):
    return S.StillingerWeberSilicon(params=replace(SW.params, **kw))


@pytest.fixture(scope="module")
def sw_case():
    sysd = S.diamond(S.SI_A, (2, 2, 2), S.SI_MASS, 0.03, seed=7)
    calc = S.AutoDiffCalculator(0.0, 3.7)    # below a sigma, so sigma can move both ways
    return sysd, calc


def test_sw_linear_parameters(sw_case):
    sysd, calc = sw_case
    psi = O.sw_third_order(sysd, SW, calc)
    pair = O.sw_third_order(sysd, _sw(lam=0.0), calc)          # lambda = 0: pair part only
    nob = O.sw_third_order(sysd, _sw(B=0.0), calc)
    p = SW.params
    assert _rel(O.sw_third_order_dparam(sysd, SW, calc, "A"), pair / p.A) <= 1e-12
    assert _rel(O.sw_third_order_dparam(sysd, SW, calc, "lambda"), (psi - pair) / p.lam) <= 1e-10
    assert _rel(O.sw_third_order_dparam(sysd, SW, calc, "B"), (psi - nob) / p.B) <= 1e-10


@pytest.mark.parametrize("name,field,step", [("gamma", "gamma", 1e-3), ("sigma", "sigma", 1e-4)])
def test_sw_nonlinear_parameters_against_richardson(sw_case, name, field, step):
    sysd, calc = sw_case
    d = O.sw_third_order_dparam(sysd, SW, calc, name)
    x = getattr(SW.params, field)
    f = lambda v: O.sw_third_order(sysd, _sw(**{field: v}), calc)
    fd = _richardson(f, x, step * x)
    r = _rel(d, fd)
    print(f"SW d/d{name} vs Richardson: {r:.2e}")
    assert r <= 1e-7
