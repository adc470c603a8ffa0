"""Pins the oracle's TableKernel (test infrastructure) against closed forms.

Checks, following src/Kernel/tests/testTableKernel.py:30-140 where it still
matches the kernel code:
  * table W, dW, d2W vs the analytic B-spline (rel tol 1e-3, 1e-2, 1e-2)
  * the parabolic fit reproduces every tabulated point exactly at x = 0 and
    x = 1 and is the unique quadratic through three points (independent fit)
  * nperh increasing, Wsum non-decreasing (TableKernel.cc:537-546 ENSUREs)
  * Wsum against an independent lattice sum of |dW/deta| with the nu-th root
    (TableKernel.cc:23-93) to 1e-10, and the inverse lookup to 1e-10.
The reference test sums W (not |dW|) and takes no root; that no longer matches
TableKernel::setNperhValues, so the lattice-sum check here follows the code.
"""
import math

import numpy as np
import pytest


def fuzzy(a, b, tol):
    return abs(a - b) / max(1.0, abs(a) + abs(b)) < tol


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_table_vs_analytic(oracle, dim):
    T = oracle.Table(dim, 1000)
    n = 1000
    deta = 2.0 / (n - 1)
    for i in range(n):
        eta = i * deta
        assert fuzzy(T.W(eta), oracle.bspline(dim, 0, eta), 1e-3)
        assert fuzzy(T.gradW(eta), oracle.bspline(dim, 1, eta), 1e-2)
        assert fuzzy(T.grad2W(eta), oracle.bspline(dim, 2, eta), 1e-2)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_bspline_closed_form(oracle, dim):
    norm = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * math.pi), 3: 1.0 / math.pi}[dim]
    for eta in np.linspace(0, 2.2, 57):
        if eta < 1:
            w = norm * (1 - 1.5 * eta ** 2 + 0.75 * eta ** 3)
            g = -norm * (3 - 2.25 * eta) * eta
        elif eta < 2:
            w = norm * 0.25 * (2 - eta) ** 3
            g = -norm * 0.75 * (2 - eta) ** 2
        else:
            w = g = 0.0
        assert oracle.bspline(dim, 0, eta) == pytest.approx(w, rel=1e-14, abs=1e-300)
        assert oracle.bspline(dim, 1, eta) == pytest.approx(g, rel=1e-14, abs=1e-300)
    # volume normalisation: integral of W over space is 1
    r = np.linspace(0, 2, 200001)
    w = np.array([oracle.bspline(dim, 0, x) for x in r[::100]])
    rr = r[::100]
    shell = {1: 2.0 * np.ones_like(rr), 2: 2 * math.pi * rr, 3: 4 * math.pi * rr ** 2}[dim]
    assert np.trapezoid(w * shell, rr) == pytest.approx(1.0, rel=1e-5)


@pytest.mark.parametrize("dim", [1, 3])
def test_parabolic_fit(oracle, dim):
    T = oracle.Table(dim, 1000)
    a = T.arrays()
    n = 1000
    step = 2.0 / (n - 1)
    t = np.array([oracle.bspline(dim, 0, i * step) for i in range(n)])
    g = np.array([oracle.bspline(dim, 1, i * step) for i in range(n)])
    for i1 in range(1, n - 1):
        t0, t1, t2 = t[i1 - 1], t[i1], t[i1 + 1]
        # independent fit: solve the 3x3 system numerically
        A = np.array([[0, 0, 1], [1, 1, 1], [4, 2, 1]], dtype=float)
        c, b, a0 = np.linalg.solve(A, [t0, t1, t2])
        assert a["Ak"][i1] == t0
        assert a["Ak"][i1] + a["Bk"][i1] + a["Ck"][i1] == pytest.approx(t1, abs=1e-15)
        assert a["Bk"][i1] == pytest.approx(b, abs=1e-13)
        assert a["Ck"][i1] == pytest.approx(c, abs=1e-13)
        assert a["Ag"][i1] == g[i1 - 1]
    # lookups hit the table points exactly at the left node
    for i in (0, 1, 17, 500, 996):
        assert T.W(i * step) == t[i] or fuzzy(T.W(i * step), t[i], 1e-15)


def lattice_grad_sum(T, dim, nperh):
    deta = 1.0 / nperh
    nx = int(2.0 * nperh) + 1
    e = np.arange(nx) * deta
    mult = np.where(np.arange(nx) > 0, 2.0, 1.0)
    if dim == 1:
        return sum(2.0 * abs(T.gradW(x)) for x in e[1:] if x < 2.0)
    if dim == 2:
        s = 0.0
        for iy in range(nx):
            for ix in range(nx):
                r = math.hypot(e[ix], e[iy])
                if r == 0 or e[ix] >= 2.0 or e[iy] >= 2.0:
                    continue
                s += abs(T.gradW(r)) * mult[ix] * mult[iy]
        return math.sqrt(s)
    s = 0.0
    for iz in range(nx):
        for iy in range(nx):
            for ix in range(nx):
                if e[ix] >= 2.0 or e[iy] >= 2.0 or e[iz] >= 2.0:
                    continue
                r = math.sqrt(e[ix] ** 2 + e[iy] ** 2 + e[iz] ** 2)
                if r == 0:
                    continue
                s += abs(T.gradW(r)) * mult[ix] * mult[iy] * mult[iz]
    return s ** (1.0 / 3.0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_wsum_tables(oracle, dim):
    T = oracle.Table(dim, 1000)
    a = T.arrays()
    nperh, wsum = a["nperh"], a["wsum"]
    assert np.all(np.diff(nperh) > 0)
    assert np.all(np.diff(wsum) >= 0)
    assert nperh[0] == 0.25 and nperh[-1] == pytest.approx(10.0, rel=1e-15)
    idx = [0, 1, 60, 61, 100, 250, 400, 700, 999] if dim == 3 else list(range(0, 1000, 37)) + [999]
    for i in idx:
        ref = lattice_grad_sum(T, dim, nperh[i])
        assert fuzzy(wsum[i], ref, 1e-10), (i, nperh[i], wsum[i], ref)
        if wsum[i] > 0 and (i == 0 or wsum[i] > wsum[i - 1]):
            assert fuzzy(T.nperh_from_wsum(wsum[i]), nperh[i], 1e-10)
