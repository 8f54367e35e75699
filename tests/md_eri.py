"""Independent reference ERIs for pinning the oracle (test helper).

McMurchie-Davidson Hermite-Gaussian scheme (L. E. McMurchie, E. R. Davidson,
J. Comput. Phys. 26 (1978) 218; Helgaker-Jorgensen-Olsen ch. 9): a different
algorithm from the Rys quadrature of the oracle and of the device kernels,
so agreement pins VRR/HRR/root mistakes.  Boys values come from mpmath at 30
digits.  Spherical transforms are fitted numerically against scipy's complex
spherical harmonics, independent of the closed form used elsewhere.
"""
import functools
import math

import mpmath as mp
import numpy as np
import scipy.special as sps

mp.mp.dps = 30


def cart_comps(l):
    """libcint order: lx from l down, then ly from l-lx down."""
    out = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            out.append((lx, ly, l - lx - ly))
    return out


def boys(nmax, T):
    T = mp.mpf(T)
    if T == 0:
        return [1.0 / (2 * n + 1) for n in range(nmax + 1)]
    return [float(mp.gammainc(n + mp.mpf(1) / 2, 0, T) / (2 * T ** (n + mp.mpf(1) / 2)))
            for n in range(nmax + 1)]


def hermite_E(la, lb, a, b, A, B):
    """E[ia, ib, t] per axis for the product of 1D Gaussians (without K_AB)."""
    p = a + b
    P = (a * A + b * B) / p
    XPA, XPB = P - A, P - B
    E = np.zeros((la + 1, lb + 1, la + lb + 2))
    E[0, 0, 0] = 1.0
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t] + (t + 1) * E[i - 1, j, t + 1]
                    if t > 0:
                        v += E[i - 1, j, t - 1] / (2 * p)
                else:
                    v = XPB * E[i, j - 1, t] + (t + 1) * E[i, j - 1, t + 1]
                    if t > 0:
                        v += E[i, j - 1, t - 1] / (2 * p)
                E[i, j, t] = v
    return E


def hermite_R(L, alpha, RPQ):
    """R[t,u,v] = R^0_{tuv} for t+u+v <= L."""
    X, Y, Z = RPQ
    F = boys(L, alpha * (X * X + Y * Y + Z * Z))
    R = np.zeros((L + 1, L + 1, L + 1, L + 1))  # [n, t, u, v]
    for n in range(L + 1):
        R[n, 0, 0, 0] = (-2 * alpha) ** n * F[n]
    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = X * R[n + 1, t - 1, u, v]
                        if t > 1:
                            val += (t - 1) * R[n + 1, t - 2, u, v]
                    elif u > 0:
                        val = Y * R[n + 1, t, u - 1, v]
                        if u > 1:
                            val += (u - 1) * R[n + 1, t, u - 2, v]
                    else:
                        val = Z * R[n + 1, t, u, v - 1]
                        if v > 1:
                            val += (v - 1) * R[n + 1, t, u, v - 2]
                    R[n, t, u, v] = val
    return R[0]


def prim_eri_cart(ls, exps, centers):
    """Primitive Cartesian (ab|cd), shape (nca, ncb, ncc, ncd), unnormalised
    Gaussians (x-A)^i ... e^{-a r^2}."""
    la, lb, lc, ld = ls
    a, b, c, d = exps
    A, B, C, D = [np.asarray(x, dtype=float) for x in centers]
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    Kab = math.exp(-a * b / p * np.dot(A - B, A - B))
    Kcd = math.exp(-c * d / q * np.dot(C - D, C - D))
    Eab = [hermite_E(la, lb, a, b, A[x], B[x]) for x in range(3)]
    Ecd = [hermite_E(lc, ld, c, d, C[x], D[x]) for x in range(3)]
    L = la + lb + lc + ld
    alpha = p * q / (p + q)
    R = hermite_R(L, alpha, P - Q)
    pre = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kcd
    ca, cb, cc, cd = cart_comps(la), cart_comps(lb), cart_comps(lc), cart_comps(ld)
    nab = la + lb + 1
    ncd = lc + ld + 1
    # bra Hermite coefficients: [ia, ib, (t,u,v)]
    Wab = np.zeros((len(ca), len(cb), nab, nab, nab))
    for i, (ax, ay, az) in enumerate(ca):
        for j, (bx, by, bz) in enumerate(cb):
            Wab[i, j] = np.einsum("t,u,v->tuv", Eab[0][ax, bx, :nab], Eab[1][ay, by, :nab],
                                  Eab[2][az, bz, :nab])
    Wcd = np.zeros((len(cc), len(cd), ncd, ncd, ncd))
    for i, (cx, cy, cz) in enumerate(cc):
        for j, (dx, dy, dz) in enumerate(cd):
            Wcd[i, j] = np.einsum("t,u,v->tuv", Ecd[0][cx, dx, :ncd], Ecd[1][cy, dy, :ncd],
                                  Ecd[2][cz, dz, :ncd])
    # M[tuv, tau nu phi] = (-1)^(tau+nu+phi) R[t+tau, u+nu, v+phi]
    t1 = np.arange(nab)
    t2 = np.arange(ncd)
    idx = t1[:, None] + t2[None, :]
    sign = (-1.0) ** t2
    M = R[idx[:, None, None, :, None, None], idx[None, :, None, None, :, None],
          idx[None, None, :, None, None, :]]
    M = M * sign[None, None, None, :, None, None] * sign[None, None, None, None, :, None] \
        * sign[None, None, None, None, None, :]
    M = M.reshape(nab ** 3, ncd ** 3)
    out = Wab.reshape(len(ca) * len(cb), -1) @ M @ Wcd.reshape(len(cc) * len(cd), -1).T
    return pre * out.reshape(len(ca), len(cb), len(cc), len(cd))


FAC_SP = {0: 0.282094791773878143, 1: 0.488602511902919921}


def shell(atm, bas, env, s):
    b = bas[s]
    l, nprim, nctr = int(b[1]), int(b[2]), int(b[3])
    r = env[atm[b[0], 1]:atm[b[0], 1] + 3]
    e = env[b[5]:b[5] + nprim]
    c = env[b[6]:b[6] + nprim * nctr].reshape(nctr, nprim)
    return l, np.array(r), np.array(e), c


def int2e_cart_md(shls, atm, bas, env):
    """libcint int2e_cart block (F-order dims (nci*nctr_i, ...)), MD route."""
    sh = [shell(atm, bas, env, s) for s in shls]
    ls = [x[0] for x in sh]
    nc = [len(cart_comps(l)) for l in ls]
    ctr = [x[3].shape[0] for x in sh]
    out = np.zeros((ctr[0], nc[0], ctr[1], nc[1], ctr[2], nc[2], ctr[3], nc[3]))
    fac = np.prod([FAC_SP.get(l, 1.0) for l in ls])
    for ia, ea in enumerate(sh[0][2]):
        for ib, eb in enumerate(sh[1][2]):
            for ic, ec in enumerate(sh[2][2]):
                for id_, ed in enumerate(sh[3][2]):
                    blk = prim_eri_cart(ls, (ea, eb, ec, ed), [x[1] for x in sh])
                    cfs = np.einsum("a,b,c,d->abcd", sh[0][3][:, ia], sh[1][3][:, ib],
                                    sh[2][3][:, ic], sh[3][3][:, id_])
                    out += np.einsum("abcd,ijkl->aibjckdl", cfs, blk)
    return fac * out.reshape(ctr[0] * nc[0], ctr[1] * nc[1], ctr[2] * nc[2], ctr[3] * nc[3])


@functools.lru_cache(maxsize=None)
def c2s_fit(l):
    """(2l+1, ncart) matrix mapping libcint Cartesian monomials to r^l Y_lm
    (real, unit-sphere normalised, m = -l..l; p kept as the identity since
    libcint moves its factor into the common factor)."""
    comps = cart_comps(l)
    if l < 2:
        return np.eye(len(comps))
    rng = np.random.default_rng(1234 + l)
    v = rng.normal(size=(400, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    theta = np.arccos(v[:, 2])
    phi = np.arctan2(v[:, 1], v[:, 0])
    mono = np.stack([v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c for a, b, c in comps], axis=1)
    rows = []
    for m in range(-l, l + 1):
        Y = sps.sph_harm_y(l, abs(m), theta, phi)  # Condon-Shortley phase included
        if m > 0:
            y = math.sqrt(2) * (-1) ** m * Y.real
        elif m < 0:
            y = math.sqrt(2) * (-1) ** m * Y.imag
        else:
            y = Y.real
        coef, *_ = np.linalg.lstsq(mono, y, rcond=None)
        rows.append(coef)
    return np.array(rows)


def cart_to_sph(block, ls, ctrs):
    """Apply c2s along each of the four axes of a (ctr*ncart, ...) block."""
    out = block
    for ax, (l, nct) in enumerate(zip(ls, ctrs)):
        T = c2s_fit(l)
        nc = T.shape[1]
        shp = list(out.shape)
        o = np.moveaxis(out, ax, 0).reshape(nct, nc, -1)
        o = np.einsum("sc,kcx->ksx", T, o).reshape(nct * T.shape[0], *np.moveaxis(
            out, ax, 0).shape[1:])
        out = np.moveaxis(o, 0, ax)
    return out


def int2e_sph_md(shls, atm, bas, env):
    cart = int2e_cart_md(shls, atm, bas, env)
    ls = [int(bas[s][1]) for s in shls]
    ctrs = [int(bas[s][3]) for s in shls]
    return cart_to_sph(cart, ls, ctrs)
