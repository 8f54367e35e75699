"""McMurchie-Davidson ERIs of single primitive Gaussians in mpmath (test
helper).

Every operation runs in mpmath at `DPS` decimal digits: the Hermite expansion
coefficients E_t^{ij} (HJO eqs. 9.5.6-9.5.7), the Hermite Coulomb integrals
R^n_{tuv} (HJO eqs. 9.9.18-9.9.20) with Boys values from mpmath's incomplete
gamma function, and the final sums.  The inputs (FP64 coordinates and
exponents) convert to mpmath exactly, so a result is the exact integral of the
FP64 inputs to far below FP64 rounding.  A different algorithm from the Rys
quadrature of the oracle and the device kernels; used to pin the oracle's
long-double build on ill-conditioned high-l quartets, one Cartesian component
at a time (unnormalised primitives (x-A)^i (y-A)^j (z-A)^k e^{-a|r-A|^2}).
"""
import functools

import mpmath as mp

DPS = 40


def _E_axis(la, lb, a, b, A, B):
    """E[(i, j, t)] of one axis (without the Gaussian product factor)."""
    p = a + b
    P = (a * A + b * B) / p
    xpa, xpb = P - A, P - B
    E = {(0, 0, 0): mp.mpf(1)}

    def g(i, j, t):
        return E.get((i, j, t), mp.mpf(0))

    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = xpa * g(i - 1, j, t) + (t + 1) * g(i - 1, j, t + 1)
                    if t > 0:
                        v += g(i - 1, j, t - 1) / (2 * p)
                else:
                    v = xpb * g(i, j - 1, t) + (t + 1) * g(i, j - 1, t + 1)
                    if t > 0:
                        v += g(i, j - 1, t - 1) / (2 * p)
                E[(i, j, t)] = v
    return E


def cart_comps(l):
    """libcint order: lx from l down, then ly from l-lx down."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


class PrimitiveQuartet:
    """(ab|cd) of four primitive Cartesian shells, component by component."""

    def __init__(self, ls, exps, centers):
        with mp.workdps(DPS):
            self.ls = ls
            a, b, c, d = [mp.mpf(float(e)) for e in exps]
            A, B, C, D = [[mp.mpf(float(v)) for v in r] for r in centers]
            p, q = a + b, c + d
            P = [(a * A[x] + b * B[x]) / p for x in range(3)]
            Q = [(c * C[x] + d * D[x]) / q for x in range(3)]
            self.Eab = [_E_axis(ls[0], ls[1], a, b, A[x], B[x]) for x in range(3)]
            self.Ecd = [_E_axis(ls[2], ls[3], c, d, C[x], D[x]) for x in range(3)]
            rab = sum((A[x] - B[x]) ** 2 for x in range(3))
            rcd = sum((C[x] - D[x]) ** 2 for x in range(3))
            self.alpha = p * q / (p + q)
            self.XPQ = [P[x] - Q[x] for x in range(3)]
            L = sum(ls)
            T = self.alpha * sum(v * v for v in self.XPQ)
            self.F = [self._boys(n, T) for n in range(L + 1)]
            self.pre = (2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) *
                        mp.exp(-a * b / p * rab) * mp.exp(-c * d / q * rcd))
            self._R = functools.lru_cache(maxsize=None)(self._Rn)

    @staticmethod
    def _boys(n, T):
        if T == 0:
            return mp.mpf(1) / (2 * n + 1)
        h = mp.mpf(n) + mp.mpf(1) / 2
        return mp.gammainc(h, 0, T) / (2 * T ** h)

    def _Rn(self, n, t, u, v):
        if t < 0 or u < 0 or v < 0:
            return mp.mpf(0)
        if t == 0 and u == 0 and v == 0:
            return (-2 * self.alpha) ** n * self.F[n]
        X, Y, Z = self.XPQ
        if t > 0:
            return (t - 1) * self._R(n + 1, t - 2, u, v) + X * self._R(n + 1, t - 1, u, v)
        if u > 0:
            return (u - 1) * self._R(n + 1, t, u - 2, v) + Y * self._R(n + 1, t, u - 1, v)
        return (v - 1) * self._R(n + 1, t, u, v - 2) + Z * self._R(n + 1, t, u, v - 1)

    def component(self, ia, ib, ic, id_):
        """The integral for Cartesian component indices (libcint order)."""
        with mp.workdps(DPS):
            ca = cart_comps(self.ls[0])[ia]
            cb = cart_comps(self.ls[1])[ib]
            cc = cart_comps(self.ls[2])[ic]
            cd = cart_comps(self.ls[3])[id_]
            bra = [[(t, self.Eab[x][(ca[x], cb[x], t)]) for t in range(ca[x] + cb[x] + 1)]
                   for x in range(3)]
            ket = [[(t, self.Ecd[x][(cc[x], cd[x], t)] * (-1) ** t)
                    for t in range(cc[x] + cd[x] + 1)] for x in range(3)]
            tot = mp.mpf(0)
            for t, ex in bra[0]:
                for u, ey in bra[1]:
                    for v, ez in bra[2]:
                        e1 = ex * ey * ez
                        if e1 == 0:
                            continue
                        s = mp.mpf(0)
                        for tt, fx in ket[0]:
                            for uu, fy in ket[1]:
                                for vv, fz in ket[2]:
                                    s += fx * fy * fz * self._R(0, t + tt, u + uu, v + vv)
                        tot += e1 * s
            return self.pre * tot
