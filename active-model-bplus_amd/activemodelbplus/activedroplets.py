"""Mean-field coexistence in Active Model B+: bulk binodals, planar interfaces, droplets
and the pseudo-tension (the parameter-selection theory of Tjhung, Nardini and Cates,
Phys. Rev. X 8, 031080 (2018)).

The reference provides this as ``activemodelbplus/activedroplets.py`` (classes
``ActiveModelBPlus``, ``ActiveBinodal``, ``ActiveDroplet``, function
``binodal_densities``; ``activedroplets.py:117-569``) over symbolic expressions in
``activefield.py`` and a Hermite finite-element solver.  This module keeps those names,
arguments and conventions and computes the same quantities from closed forms and
scipy's collocation solver.  It is host-side theory (no GPU), used to choose
parameters and to check the integrator's steady states.

Free energy and activity (``activefield.py:233-247``, ``:126-170``):

    f(phi) = t phi^2 / 2 + u phi^4 / 4,   mu = f'(phi) - K lap phi + lambda |grad phi|^2
    alpha = (zeta - 2 lambda) / K
    pseudodensity   psi(phi) = (exp(alpha phi) - 1) / alpha        (psi = phi at alpha = 0)
    pseudopotential g(phi)   = int_0^phi f'(s) psi'(s) ds,  g(0) = f(0)
    pseudopressure  P(phi)   = psi(phi) f'(phi) - g(phi)

Bulk binodals (``activedroplets.py:519-530``): f'(phi0) = f'(phi1), P(phi0) = P(phi1).

Planar interface: J = 0 in one dimension gives mu = const; q = phi'^2 as a function of
phi obeys the linear equation dq/dphi = (2/K)(f' - mu) - alpha q, so

    q(phi) = (2/K) exp(-alpha phi) [g(phi) - g(phi0) - mu (psi(phi) - psi(phi0))],

which vanishes at both binodals.  The pseudo-tension (``activedroplets.py:284-318``,
Tjhung 2018) is sigma = (1/alpha) int (zeta e^{alpha phi_ref} - 2 lambda e^{alpha phi})
phi'^2 dx with phi_ref the density of the droplet's own phase (clusters: the dense
binodal, bubbles: the dilute one); int (.) phi'^2 dx = int (.) sqrt(q) dphi.  sigma is
linear in lambda at fixed zeta - 2 lambda, so the boundary sigma = 0 of reverse Ostwald
ripening has a closed form per zeta - 2 lambda.

Droplet (``activefield.py:315-390``, ``activedroplets.py:321-470``): radial steady state
in d dimensions with zero flux,

    d/dr [f'(phi) - K (phi'' + (d-1) phi'/r) + (lambda - zeta/2) phi'^2] = zeta (d-1) phi'^2 / r,

with phi'(0) = 0, phi(R) = 0 (the interface location defines R), phi'(L) = 0, L = 5 R by
default; phi0 = phi(0), phi1 = phi(L).  Solved as two segments [0, R] and [R, L] joined at
R (phi, phi' and the local chemical potential continuous) with scipy.integrate.solve_bvp.
The pseudo-pressure drop obeys P(phi0) - P(phi1) = (d-1) sigma_R / R exactly, with
sigma_R the pseudo-tension weighted by R/r (``activedroplets.py:440-468``).
"""
from __future__ import annotations

import numpy as np
from scipy.integrate import solve_bvp
from scipy.optimize import fsolve

_GL_X, _GL_W = np.polynomial.legendre.leggauss(64)


class ActiveModelBPlus:
    """Field theory of Active Model B+ with the phi^4 free energy (``activedroplets.py:471``).

    ``ActiveModelBPlus(zeta, lamb, K=1, t=-0.25, u=0.25, d=3)``; the integrator's
    ``Model(a, b, c, kappa, lamb, zeta, T)`` with b = 0 corresponds to t = a, u = c, K = kappa."""

    def __init__(self, zeta, lamb, K=1.0, t=-0.25, u=0.25, d=3):
        self.zeta, self.lamb, self.K, self.t, self.u, self.d = float(zeta), float(lamb), float(K), float(t), float(u), int(d)
        if self.K <= 0 or self.t >= 0 or self.u <= 0:
            raise ValueError("need K > 0, t < 0 < u (a phase-separating phi^4 free energy)")
        self.alpha = (self.zeta - 2 * self.lamb) / self.K

    def __repr__(self):
        return "<ActiveModelBPlus zeta=%g lamb=%g>" % (self.zeta, self.lamb)

    # -- local functions -------------------------------------------------------------
    @property
    def passive_bulk_interfacial_width(self):
        return float(np.sqrt(-2 * self.K / self.t))

    def free_energy_density(self, phi, derivative=0):
        phi = np.asarray(phi, dtype=float)
        t, u = self.t, self.u
        if derivative == 0:
            return t * phi ** 2 / 2 + u * phi ** 4 / 4
        if derivative == 1:
            return t * phi + u * phi ** 3
        if derivative == 2:
            return t + 3 * u * phi ** 2
        raise ValueError("derivative must be 0, 1 or 2")

    def pseudodensity(self, phi, derivative=0):
        phi = np.asarray(phi, dtype=float)
        a = self.alpha
        if derivative == 1:
            return np.exp(a * phi)
        if derivative != 0:
            raise ValueError("derivative must be 0 or 1")
        if abs(a * np.max(np.abs(phi), initial=0.0)) < 1e-8:
            return phi + a * phi ** 2 / 2  # series; exact to round-off in this range
        return np.expm1(a * phi) / a

    def pseudopotential(self, phi):
        """g(phi) = int_0^phi f'(s) exp(alpha s) ds (64-point Gauss-Legendre; the integrand is
        a cubic times an exponential, integrated to round-off for |alpha phi| of order 10)."""
        phi = np.asarray(phi, dtype=float)
        s = 0.5 * phi[..., None] * (1 + _GL_X)
        vals = self.free_energy_density(s, 1) * np.exp(self.alpha * s)
        return 0.5 * phi * np.sum(_GL_W * vals, axis=-1)

    def bulk_pseudopressure(self, phi, mu=None):
        """P = psi mu - g with mu = f'(phi) (bulk) unless given."""
        mu = self.free_energy_density(phi, 1) if mu is None else mu
        return self.pseudodensity(phi) * mu - self.pseudopotential(phi)

    # -- bulk coexistence -------------------------------------------------------------
    @property
    def binodal_densities(self):
        """(phi_a, phi_b): the coexisting bulk densities, phi_a on the branch through +1 and
        phi_b through -1 (the reference's Newton start [1, -1])."""
        if abs(self.alpha) < 1e-12:
            s = np.sqrt(-self.t / self.u)
            return np.array([s, -s])
        mu = lambda p: self.free_energy_density(p, 1)
        P = self.bulk_pseudopressure

        def residual(x):
            return [float(P(x[0]) - P(x[1])), float(mu(x[0]) - mu(x[1]))]

        s = np.sqrt(-self.t / self.u)
        x, info, ier, msg = fsolve(residual, [s, -s], full_output=True, xtol=1e-13)
        if ier != 1 and max(abs(r) for r in residual(x)) > 1e-10:
            raise RuntimeError("binodal densities did not converge: " + msg)
        if abs(x[0] - x[1]) < 1e-6:
            raise RuntimeError("binodal search collapsed onto one density")
        return np.array(x)

    def binodal(self):
        return ActiveBinodal(self)

    def droplet(self, R, domain_size=None, guess=None, tol=1e-8, nodes=400):
        """Steady droplet of radius R (phi(R) = 0) of the phase through +1 in a sphere of
        radius ``domain_size`` (default 5 R).  ``guess``: a droplet solved at another R of the
        same model (continuation)."""
        return ActiveDroplet(self, R, domain_size=domain_size, guess=guess, tol=tol, nodes=nodes)

    # -- pseudo-tension -----------------------------------------------------------------
    def _tension_integrals(self, phi_ref):
        """(I0, I1) = int sqrt(q) dphi and int (e^{alpha phi_ref} - e^{alpha phi}) sqrt(q) dphi
        across the planar interface."""
        hi, lo = np.sort(self.binodal_densities)[::-1]
        mu = float(self.free_energy_density(hi, 1))
        s = lo + (hi - lo) * 0.5 * (1 + _GL_X)
        w = _GL_W * (hi - lo) * 0.5
        q = self.planar_gradient_squared(s, hi, mu)
        sq = np.sqrt(np.maximum(q, 0.0))
        a = self.alpha
        return float(np.sum(w * sq)), float(np.sum(w * (np.exp(a * phi_ref) - np.exp(a * s)) * sq))

    def planar_gradient_squared(self, phi, phi0=None, mu=None):
        """q(phi) = phi'(x)^2 on the planar interface between the binodals."""
        if phi0 is None:
            phi0 = float(np.max(self.binodal_densities))
        if mu is None:
            mu = float(self.free_energy_density(phi0, 1))
        phi = np.asarray(phi, dtype=float)
        a = self.alpha
        num = self.pseudopotential(phi) - self.pseudopotential(phi0) - mu * (self.pseudodensity(phi) - self.pseudodensity(phi0))
        return (2.0 / self.K) * np.exp(-a * phi) * num

    def surface_tension(self, which="clusters"):
        """Pseudo-tension of the planar interface for droplets of the dense phase
        (``"clusters"``) or of the dilute phase (``"bubbles"``)."""
        hi, lo = np.sort(self.binodal_densities)[::-1]
        ref = hi if which == "clusters" else lo
        if which not in ("clusters", "bubbles"):
            raise ValueError("which must be 'clusters' or 'bubbles'")
        I0, I1 = self._tension_integrals(ref)
        a = self.alpha
        if abs(a) < 1e-8:  # limit alpha -> 0: integrand K + zeta phi_ref - 2 lambda phi
            hi_lo = np.sort(self.binodal_densities)[::-1]
            s = hi_lo[1] + (hi_lo[0] - hi_lo[1]) * 0.5 * (1 + _GL_X)
            w = _GL_W * (hi_lo[0] - hi_lo[1]) * 0.5
            sq = np.sqrt(np.maximum(self.planar_gradient_squared(s), 0.0))
            return float(np.sum(w * (self.K + self.zeta * ref - 2 * self.lamb * s) * sq))
        # zeta e^{a ref} - 2 lambda e^{a phi} = zl e^{a ref} + 2 lambda (e^{a ref} - e^{a phi})
        zl = self.zeta - 2 * self.lamb
        return float((zl * np.exp(a * ref) * I0 + 2 * self.lamb * I1) / a)


def reversal_lambda(zeta_lamb, which="clusters", K=1.0, t=-0.25, u=0.25):
    """lambda at which the pseudo-tension vanishes for fixed zeta - 2 lambda (the boundary of
    reverse Ostwald ripening, Tjhung 2018 Fig. 4; reference notebook
    examples/literature-compare.ipynb cells 24-29).  Returns (zeta, lambda)."""
    m = ActiveModelBPlus(zeta_lamb, 0.0, K, t, u)
    hi, lo = np.sort(m.binodal_densities)[::-1]
    ref = hi if which == "clusters" else lo
    I0, I1 = m._tension_integrals(ref)
    lam = -zeta_lamb * np.exp(m.alpha * ref) * I0 / (2 * I1)
    return zeta_lamb + 2 * lam, lam


def binodal_densities(zeta_lamb, *args, **kwargs):
    """Bulk binodals for each value of zeta - 2 lambda (``activedroplets.py:558-569``):
    returns (phi0, phi1) arrays shaped like ``zeta_lamb``."""
    zl = np.asarray(zeta_lamb, dtype=float)
    out = np.array([ActiveModelBPlus(1.0, 0.5 * (1.0 - z), *args, **kwargs).binodal_densities for z in zl.ravel()])
    return out[:, 0].reshape(zl.shape), out[:, 1].reshape(zl.shape)


class ActiveBinodal:
    """Planar interface between the bulk binodals (``activedroplets.py:117-318``):
    ``phi0`` (x -> -inf) is the dense density, ``phi1`` the dilute one; the profile is
    phi(x) with phi(0) at the mean of the two."""

    def __init__(self, field_theory):
        self.field_theory = field_theory
        self.phi0, self.phi1 = np.sort(field_theory.binodal_densities)[::-1]
        self.mu = float(field_theory.free_energy_density(self.phi0, 1))
        # x(phi) = -int_{phi_mid}^{phi} dphi / sqrt(q): tabulated between the binodals
        eps = 1e-7 * (self.phi0 - self.phi1)
        self._phi = np.linspace(self.phi1 + eps, self.phi0 - eps, 4001)
        sq = np.sqrt(np.maximum(field_theory.planar_gradient_squared(self._phi, self.phi0, self.mu), 1e-300))
        inv = 1.0 / sq
        x = np.concatenate(([0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(self._phi))))
        mid = 0.5 * (self.phi0 + self.phi1)
        self._x = -(x - np.interp(mid, self._phi, x))

    def __repr__(self):
        m = self.field_theory
        return "<ActiveBinodal zeta=%g lamb=%g phi=[%.4f->%.4f]>" % (m.zeta, m.lamb, self.phi0, self.phi1)

    def __call__(self, x):
        """phi(x), dense for x -> -inf."""
        xs, ps = self._x[::-1], self._phi[::-1]
        return np.interp(np.asarray(x, dtype=float), xs, ps, left=self.phi0, right=self.phi1)

    @property
    def pseudopressure0(self):
        return float(self.field_theory.bulk_pseudopressure(self.phi0, self.mu))

    @property
    def pseudopressure1(self):
        return float(self.field_theory.bulk_pseudopressure(self.phi1, self.mu))

    @property
    def surface_tension(self):
        return self.field_theory.surface_tension("clusters")


class ActiveDroplet:
    """Steady droplet (``activedroplets.py:321-470``); see the module docstring."""

    def __init__(self, field_theory, R, domain_size=None, guess=None, tol=1e-8, nodes=400):
        m = field_theory
        self.field_theory, self.R = m, float(R)
        self.L = 5.0 * self.R if domain_size is None else float(domain_size)
        if not (0 < self.R < self.L):
            raise ValueError("need 0 < R < domain_size")
        R, L, d, K, zeta = self.R, self.L, m.d, m.K, m.zeta
        c = m.lamb - zeta / 2
        fp = lambda p: m.free_energy_density(p, 1)

        # y = [phi, phi', m] on [0, R] (r = R s) and on [R, L] (r = R + (L - R) s); m is the
        # local chemical potential f' - K lap phi + (lambda - zeta/2) phi'^2.
        def rhs(s, y):
            out = np.empty_like(y)
            for k, (scale, r) in enumerate(((R, R * s), (L - R, R + (L - R) * s))):
                phi, p, mloc = y[3 * k], y[3 * k + 1], y[3 * k + 2]
                inv_r = np.where(r > 0, 1.0 / np.where(r > 0, r, 1.0), 0.0)
                rest = (fp(phi) + c * p * p - mloc) / K
                # the left segment's (d-1) phi'/r term enters through the singular matrix S
                curv = (d - 1) * p * inv_r if k == 1 else 0.0
                out[3 * k] = scale * p
                out[3 * k + 1] = scale * (rest - curv)
                out[3 * k + 2] = scale * zeta * (d - 1) * p * p * inv_r
            return out

        S = np.zeros((6, 6))
        S[1, 1] = -(d - 1)

        def bc(ya, yb):
            return np.array([ya[1], yb[0], ya[3], yb[1] - ya[4], yb[2] - ya[5], yb[4]])

        s = np.linspace(0.0, 1.0, nodes)
        if guess is not None:
            y0 = guess._sol.sol(s)
        else:
            w = m.passive_bulk_interfacial_width
            phi_a = float(np.max(m.binodal_densities))
            phi_b = float(np.min(m.binodal_densities))
            y0 = np.zeros((6, nodes))
            for k, r in enumerate((R * s, R + (L - R) * s)):
                th = np.tanh((r - R) / w)
                y0[3 * k] = np.where(th < 0, -phi_a * th, -(-phi_b) * th)
                y0[3 * k + 1] = -np.where(th < 0, phi_a, -phi_b) * (1 - th * th) / w
                y0[3 * k + 2] = fp(phi_b)
        sol = solve_bvp(rhs, bc, s, y0, S=S, tol=tol, max_nodes=400000)
        if not sol.success:
            raise RuntimeError("droplet profile did not converge: " + sol.message)
        self._sol = sol
        self.phi0 = float(sol.y[0, 0])
        self.phi1 = float(sol.y[3, -1])
        self.mu = float(sol.y[5, -1])

    def __repr__(self):
        m = self.field_theory
        return "<ActiveDroplet zeta=%g lamb=%g d=%d R=%g phi=[%.4f->%.4f]>" % (m.zeta, m.lamb, m.d, self.R, self.phi0,
                                                                            self.phi1)

    def __call__(self, r, derivative=0):
        """phi(r) (derivative 0) or phi'(r) (derivative 1) for 0 <= r <= L."""
        r = np.asarray(r, dtype=float)
        k = derivative
        inner = self._sol.sol(np.clip(r / self.R, 0, 1))[k]
        outer = self._sol.sol(np.clip((r - self.R) / (self.L - self.R), 0, 1))[3 + k]
        return np.where(r <= self.R, inner, outer)

    @property
    def pseudopressure0(self):
        return float(self.field_theory.bulk_pseudopressure(self.phi0, self.mu))

    @property
    def pseudopressure1(self):
        return float(self.field_theory.bulk_pseudopressure(self.phi1, self.mu))

    @property
    def pseudopressure_drop(self):
        return self.pseudopressure0 - self.pseudopressure1

    @property
    def surface_tension(self):
        """sigma_R = (1/alpha) int (zeta e^{alpha phi0} - 2 lambda e^{alpha phi}) phi'^2 (R/r) dr."""
        m = self.field_theory
        a = m.alpha
        # Gauss-Legendre on each segment, split at the interface where phi' peaks
        xs, ws = np.polynomial.legendre.leggauss(200)
        total = 0.0
        for lo, hi in ((0.0, self.R), (self.R, self.L)):
            for a0, b0 in ((lo, 0.5 * (lo + hi)), (0.5 * (lo + hi), hi)):
                r = a0 + (b0 - a0) * 0.5 * (1 + xs)
                w = ws * (b0 - a0) * 0.5
                phi, p = self(r), self(r, 1)
                if abs(a) < 1e-8:
                    weight = m.K + m.zeta * self.phi0 - 2 * m.lamb * phi
                else:
                    weight = (m.zeta * np.exp(a * self.phi0) - 2 * m.lamb * np.exp(a * phi)) / a
                total += float(np.sum(w * weight * p * p * self.R / r))
        return total

    @property
    def surface_tension_pseudopressure_drop(self):
        return (self.field_theory.d - 1) * self.surface_tension / self.R
