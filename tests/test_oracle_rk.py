"""The oracle's Runge-Kutta tables and stage updates, pinned the way
FElib/test/common/timeint_rk/test_timeint_rk.f90 pins the reference's:
convergence order on the harmonic oscillator (explicit schemes, :31-34, :125-166)
and on the damped oscillator split into an implicit rotation and an explicit
damping (IMEX schemes, :37-38, :171-272), at the reference's thresholds; plus
the Butcher coefficients' order conditions in closed form.
"""
import math

import numpy as np
import pytest

import fedg_oracle as O

END, DT1, DT2 = 2.0, 0.02, 0.0025
SAVE1, SAVE2 = 1, 8


def run_ex(name, dt, save):
    tab = O.rk_tab(name, dt)
    omg = 2 * math.pi
    u, v = np.array([1.0]), np.array([0.0])
    buf = {k: np.zeros(1) for k in ("u0", "ut", "v0", "vt")}
    te_u, te_v = np.zeros(tab.tend_buf_size), np.zeros(tab.tend_buf_size)
    err = []
    for n in range(1, int(round(END / dt)) + 1):
        for s in range(1, tab.nstage + 1):
            ind = tab.tend_buf_indmap[s - 1]
            te_u[ind] = omg * v[0]
            te_v[ind] = -omg * u[0]
            O.rk_advance(tab, s, u, buf["u0"], buf["ut"], te_u, bs=1)
            O.rk_advance(tab, s, v, buf["v0"], buf["vt"], te_v, bs=1)
        if n % save == 0:
            err.append(u[0] - math.cos(omg * n * dt))
    return np.array(err)


def run_imex(name, dt, save):
    tab = O.rk_tab(name, dt)
    omg, r = 2 * math.pi, 0.1
    omgg = math.sqrt(omg ** 2 - r ** 2 / 4)
    u, v = np.array([1.0]), np.array([0.0])
    buf = {k: np.zeros(1) for k in ("u0", "ut", "v0", "vt")}
    nb = tab.tend_buf_size
    te_u, te_v, ti_u, ti_v = (np.zeros(nb) for _ in range(4))
    err = []
    for n in range(1, int(round(END / dt)) + 1):
        for s in range(1, tab.nstage + 1):
            ind = tab.tend_buf_indmap[s - 1]
            fac = tab.a_im[(s - 1) * 4 + (s - 1)] * dt
            coef = fac * omg
            if abs(fac) > 0:
                ui = (u[0] + coef * v[0]) / (1 + coef ** 2)
                vi = (v[0] - coef * u[0]) / (1 + coef ** 2)
                ti_u[ind] = (ui - u[0]) / fac
                ti_v[ind] = (vi - v[0]) / fac
            else:
                ti_u[ind] = omg * v[0]
                ti_v[ind] = -omg * u[0]
            O.rk_store_impl(tab, s, u, buf["u0"], buf["ut"], ti_u, bs=1)
            O.rk_store_impl(tab, s, v, buf["v0"], buf["vt"], ti_v, bs=1)
            te_u[ind] = -r * u[0]
            te_v[ind] = 0.0
            O.rk_advance(tab, s, u, buf["u0"], buf["ut"], te_u, ti_u, bs=1)
            O.rk_advance(tab, s, v, buf["v0"], buf["vt"], te_v, ti_v, bs=1)
        if n % save == 0:
            t = n * dt
            ans = (math.cos(omgg * t) - r / (2 * omgg) * math.sin(omgg * t)) * math.exp(-0.5 * r * t)
            err.append(u[0] - ans)
    return np.array(err)


@pytest.mark.parametrize("name,order", [("ERK_1s1o", 0.98), ("ERK_4s4o", 3.98)])
def test_explicit_order(name, order):
    e1, e2 = run_ex(name, DT1, SAVE1), run_ex(name, DT2, SAVE2)
    assert len(e1) == len(e2) == 100
    rate = math.log(np.linalg.norm(e1) / np.linalg.norm(e2)) / math.log(DT1 / DT2)
    assert rate >= order, rate


@pytest.mark.parametrize("name,order", [("IMEX_ARK232", 1.98), ("IMEX_ARK324", 2.98)])
def test_imex_order(name, order):
    e1, e2 = run_imex(name, DT1, SAVE1), run_imex(name, DT2, SAVE2)
    rate = math.log(np.linalg.norm(e1) / np.linalg.norm(e2)) / math.log(DT1 / DT2)
    assert rate >= order, rate


@pytest.mark.parametrize("name,order", [("ERK_4s4o", 4), ("IMEX_ARK232", 2), ("IMEX_ARK324", 3)])
def test_butcher_order_conditions(name, order):
    t = O.rk_tab(name, 1.0)
    s = t.nstage
    for part in ("ex", "im") if t.imex else ("ex",):
        A = np.array(getattr(t, "a_" + part)).reshape(4, 4)[:s, :s]
        b = np.array(getattr(t, "b_" + part))[:s]
        c = A.sum(axis=1)
        assert abs(b.sum() - 1) <= 1e-15
        assert abs(b @ c - 0.5) <= 1e-14
        if order >= 3:
            assert abs(b @ c ** 2 - 1 / 3) <= 1e-14
            assert abs(b @ A @ c - 1 / 6) <= 1e-14
        if order >= 4:
            assert abs(b @ c ** 3 - 0.25) <= 1e-14
            assert abs(b @ (c * (A @ c)) - 1 / 8) <= 1e-14
            assert abs(b @ A @ c ** 2 - 1 / 12) <= 1e-14
            assert abs(b @ A @ A @ c - 1 / 24) <= 1e-14
    if t.imex:
        # both tableaux share the abscissae and the weights (ARK coupling)
        Ae = np.array(t.a_ex).reshape(4, 4)[:s, :s]
        Ai = np.array(t.a_im).reshape(4, 4)[:s, :s]
        assert np.max(np.abs(Ae.sum(axis=1) - Ai.sum(axis=1))) <= 1e-15
        assert np.max(np.abs(np.array(t.b_ex)[:s] - np.array(t.b_im)[:s])) <= 1e-15
