"""Textbook lattice Boltzmann in numpy float64, used to pin the oracle.

This is deliberately a different algorithm from the reference's in-place AA update:
two-array pull streaming with ``np.roll``, the generic equilibrium
``w_i rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u.u)`` formed from the velocity table, and
wall cells that reverse what arrives (full-way bounce-back).  Its state after every step
is the post-collision, pre-streaming population array, which is what the reference's AA
array holds after an even number of steps (``F2[x, i] = C(P(C(P(F0))))[x, i]``).

Semantics it encodes (and that the oracle must reproduce):
* fluid cells: rho = sum f, u = sum c f / rho, u_eq = u + a / omega (velocity-shift
  forcing, test_core.py:96-103), relax towards feq(rho, u_eq);
* fixed fluid cells (flag & 6): populations replaced by feq(rho_fixed, u_fixed);
* TRT: symmetric and antisymmetric halves relax with omega+ and omega-;
* advection-diffusion: rho = sum g, u = the fluid's u_eq of the same step; fixed scalar
  cells (flag & 8) are replaced by geq(C_fixed, u).
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np

C = {
    (2, 9): [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)],
    (2, 5): [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)],
    (3, 7): [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)],
}
# D3Q27 as the tensor product of D2Q9 and D1Q3 in the reference's sorted order is long; it
# is rebuilt from the rule "rest, axes, xy/xz/yz diagonals, corners, each +/- paired".
C[(3, 27)] = [
    (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (0, 1, 1),
    (0, -1, -1), (1, 0, -1), (-1, 0, 1), (0, 1, -1), (0, -1, 1), (1, 1, 1), (-1, -1, -1),
    (1, -1, 1), (-1, 1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, -1), (-1, 1, 1),
]


def weights(d: int, q: int) -> np.ndarray:
    """Weights from the lattice definition: products of 1D weights (D2Q9, D3Q27) or the
    advection-diffusion sets (D2Q5: rest 1/3; D3Q7: rest 1/20)."""
    c = np.array(C[(d, q)])
    if q in (9, 27):
        w1 = {0: Fraction(2, 3), 1: Fraction(1, 6)}
        w = [np.prod([w1[abs(int(x))] for x in row]) for row in c]
    elif (d, q) == (2, 5):
        w = [Fraction(1, 3)] + [Fraction(1, 6)] * 4
    else:
        w = [Fraction(1, 20)] + [Fraction(19, 120)] * 6
    return np.array([float(x) for x in w])


def opposite(d: int, q: int) -> np.ndarray:
    c = [tuple(r) for r in C[(d, q)]]
    return np.array([c.index(tuple(-x for x in r)) for r in c])


def feq(d, q, rho, u):
    c = np.array(C[(d, q)], np.float64)
    w = weights(d, q)
    cu = u @ c.T
    uu = np.sum(u * u, axis=-1, keepdims=True)
    return w * rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)


def stream(F, d, q):
    """Pull streaming: out[x, i] = F[x - c_i, i], periodic."""
    c = C[(d, q)]
    out = np.empty_like(F)
    for i in range(q):
        out[..., i] = np.roll(F[..., i], shift=c[i], axis=tuple(range(d)))
    return out


class TextbookSim:
    """Two-array LBM mirroring the reference's feature set, float64 throughout."""

    def __init__(self, counts, q, omega, flags, rho, vel, acc=None):
        self.counts = tuple(counts)
        self.d = len(counts)
        self.q = q
        self.omega = omega
        self.flags = np.asarray(flags)
        self.rho = np.asarray(rho, np.float64).copy()
        self.vel = np.asarray(vel, np.float64).copy()
        self.acc = None if acc is None else np.asarray(acc, np.float64)
        self.F = feq(self.d, q, self.rho, self.vel)
        self.tracers = []  # [q, omega, G, Cval]

    def add_tracer(self, q, omega, C):
        C = np.asarray(C, np.float64).copy()
        self.tracers.append([q, omega, feq(self.d, q, C, self.vel), C])

    def step(self):
        d, q = self.d, self.q
        wall = (self.flags & 1) != 0
        fixed = ((self.flags & 6) != 0) & ~wall
        fluid = ~wall
        opp = opposite(d, q)
        c = np.array(C[(d, q)], np.float64)
        pre = stream(self.F, d, q)
        rho = pre.sum(-1)
        u = (pre @ c) / rho[..., None]
        kind = self.omega[0]
        wp = self.omega[1]
        wn = self.omega[2] if kind == "trt" else wp
        if self.acc is not None:
            u = u + self.acc / wp
        rho = np.where(fixed, self.rho, rho)
        u = np.where(fixed[..., None], self.vel, u)
        wpf = np.where(fixed, 1.0, wp)[..., None]
        wnf = np.where(fixed, 1.0, wn)[..., None]
        eq = feq(d, q, rho, u)
        if kind == "bgk":
            post = pre + wpf * (eq - pre)
        else:
            pp, pm = 0.5 * (pre + pre[..., opp]), 0.5 * (pre - pre[..., opp])
            ep, em = 0.5 * (eq + eq[..., opp]), 0.5 * (eq - eq[..., opp])
            post = pre + wpf * (ep - pp) + wnf * (em - pm)
        post[wall] = pre[wall][:, opp]
        self.F = post
        self.rho = np.where(fluid, rho, self.rho)
        self.vel = np.where(fluid[..., None], u, self.vel)
        for t in self.tracers:
            tq, tom, G, Cv = t
            topp = opposite(d, tq)
            gpre = stream(G, d, tq)
            r = gpre.sum(-1)
            tfixed = ((self.flags & 8) != 0) & ~wall
            r = np.where(tfixed, Cv, r)
            om = np.where(tfixed, 1.0, tom)[..., None]
            gpost = gpre + om * (feq(d, tq, r, self.vel) - gpre)
            gpost[wall] = gpre[wall][:, topp]
            t[2] = gpost
            t[3] = np.where(fluid, r, Cv)


def reachable_mask(flags, d, q):
    """[..., q] mask of slots the reference's array and the textbook array both define:
    every slot of a non-wall cell, and slot j of a wall cell w when w + c_j is not a wall
    (slots between two walls are never read by a fluid cell and evolve differently in the
    two formulations)."""
    wall = (np.asarray(flags) & 1) != 0
    m = np.repeat(~wall[..., None], q, axis=-1)
    c = C[(d, q)]
    for j in range(q):
        nb_wall = np.roll(wall, shift=tuple(-x for x in c[j]), axis=tuple(range(d)))
        m[..., j] |= wall & ~nb_wall
    return m
