"""Exact solution of the 1-D Riemann problem for an ideal gas (test
infrastructure).

The method is Toro's (E. F. Toro, "Riemann Solvers and Numerical Methods for
Fluid Dynamics", 3rd ed., 2009, chapter 4): the star pressure p* solves
f_L(p) + f_R(p) + (u_R - u_L) = 0, where f_K is the shock (Rankine-Hugoniot)
branch for p > p_K and the rarefaction (isentropic) branch otherwise; the
star velocity follows, and the self-similar solution is sampled at
xi = (x - x0)/t through the left wave, the contact and the right wave.  The
reference's own Sod solution (tests/Hydro/Sod/SodAnalyticSolution.py:77-170)
is the special case left rarefaction + right shock with u_L = u_R = 0, which
every two-state tube of the ensemble workload is (P_L > P_R, v = 0).
"""
import math

import numpy as np


def _fk(p, rho, P, c, g):
    """f_K(p) and its derivative (Toro eq. 4.6, 4.7, 4.37)."""
    if p > P:
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * P
        s = math.sqrt(A / (p + B))
        return (p - P) * s, s * (1.0 - 0.5 * (p - P) / (B + p))
    r = p / P
    e = (g - 1.0) / (2.0 * g)
    return 2.0 * c / (g - 1.0) * (r ** e - 1.0), 1.0 / (rho * c) * r ** (-(g + 1.0) / (2.0 * g))


class RiemannSolution:
    def __init__(self, rhoL, uL, PL, rhoR, uR, PR, gamma, x0=0.5):
        g = float(gamma)
        self.g, self.x0 = g, x0
        self.L = (rhoL, uL, PL, math.sqrt(g * PL / rhoL))
        self.R = (rhoR, uR, PR, math.sqrt(g * PR / rhoR))
        cL, cR = self.L[3], self.R[3]
        assert 2.0 * (cL + cR) / (g - 1.0) > uR - uL, "vacuum is generated"
        # two-rarefaction guess (Toro eq. 4.46), then Newton on f(p) = 0
        e = (g - 1.0) / (2.0 * g)
        p = ((cL + cR - 0.5 * (g - 1.0) * (uR - uL)) / (cL / PL ** e + cR / PR ** e)) ** (1.0 / e)
        p = max(p, 1.0e-12)
        for _ in range(100):
            fL, dL = _fk(p, rhoL, PL, cL, g)
            fR, dR = _fk(p, rhoR, PR, cR, g)
            pn = p - (fL + fR + uR - uL) / (dL + dR)
            pn = max(pn, 1.0e-14)
            done = abs(pn - p) <= 1.0e-15 * 0.5 * (pn + p)
            p = pn
            if done:
                break
        fL, _ = _fk(p, rhoL, PL, cL, g)
        fR, _ = _fk(p, rhoR, PR, cR, g)
        self.pstar = p
        self.ustar = 0.5 * (uL + uR) + 0.5 * (fR - fL)

    def _star_rho(self, K, sign):
        rho, u, P, c = K
        g, p = self.g, self.pstar
        if p > P:
            r = p / P
            q = (g - 1.0) / (g + 1.0)
            return rho * (r + q) / (q * r + 1.0)
        return rho * (p / P) ** (1.0 / g)

    def star_densities(self):
        return self._star_rho(self.L, -1), self._star_rho(self.R, +1)

    def shock_speed_right(self):
        rho, u, P, c = self.R
        g = self.g
        return u + c * math.sqrt((g + 1.0) / (2.0 * g) * self.pstar / P + (g - 1.0) / (2.0 * g))

    def sample(self, x, t):
        """rho, u, P at positions x (array) and time t > 0."""
        g, p, us = self.g, self.pstar, self.ustar
        xi = (np.asarray(x, dtype=np.float64) - self.x0) / t
        rho = np.empty_like(xi)
        u = np.empty_like(xi)
        P = np.empty_like(xi)
        for i, s in enumerate(xi):
            if s <= us:   # left of the contact
                rK, uK, PK, cK = self.L
                if p > PK:   # left shock
                    SL = uK - cK * math.sqrt((g + 1.0) / (2.0 * g) * p / PK + (g - 1.0) / (2.0 * g))
                    st = (rK, uK, PK) if s <= SL else (self._star_rho(self.L, -1), us, p)
                else:        # left rarefaction
                    cs = cK * (p / PK) ** ((g - 1.0) / (2.0 * g))
                    head, tail = uK - cK, us - cs
                    if s <= head:
                        st = (rK, uK, PK)
                    elif s >= tail:
                        st = (self._star_rho(self.L, -1), us, p)
                    else:
                        f = 2.0 / (g + 1.0) + (g - 1.0) / ((g + 1.0) * cK) * (uK - s)
                        st = (rK * f ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (cK + 0.5 * (g - 1.0) * uK + s),
                              PK * f ** (2.0 * g / (g - 1.0)))
            else:         # right of the contact
                rK, uK, PK, cK = self.R
                if p > PK:   # right shock
                    SR = uK + cK * math.sqrt((g + 1.0) / (2.0 * g) * p / PK + (g - 1.0) / (2.0 * g))
                    st = (rK, uK, PK) if s >= SR else (self._star_rho(self.R, +1), us, p)
                else:        # right rarefaction
                    cs = cK * (p / PK) ** ((g - 1.0) / (2.0 * g))
                    head, tail = uK + cK, us + cs
                    if s >= head:
                        st = (rK, uK, PK)
                    elif s <= tail:
                        st = (self._star_rho(self.R, +1), us, p)
                    else:
                        f = 2.0 / (g + 1.0) - (g - 1.0) / ((g + 1.0) * cK) * (uK - s)
                        st = (rK * f ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (-cK + 0.5 * (g - 1.0) * uK + s),
                              PK * f ** (2.0 * g / (g - 1.0)))
            rho[i], u[i], P[i] = st
        return rho, u, P
