"""Seeded synthetic molecules in libcint's atm/bas/env layout.

Five workloads (sizes are the ones the hot path is measured on):

  cfg 1  3 atoms uniform in an 8-bohr cube (min separation 1.5 bohr); one
         heavy atom (Z=8, 3s 2p 1d) and two light atoms (Z=1, 2s 1p).
  cfg 2  64 atoms on a jittered 4x4x4 lattice (2.8 bohr, +-0.3 bohr); 1/3
         heavy (3s2p1d), 2/3 light (2s1p); exponents log-uniform [0.08, 8000].
  cfg 3  256 atoms on a jittered 8x8x4 lattice with the cfg-1 shell recipe;
         2/3 heavy, which gives ~2,800 spherical functions.
  cfg 4  the cfg-2 geometry with the cfg-1 orbital recipe plus an auxiliary
         basis of 6s 5p 4d 3f 1g uncontracted shells per atom, exponents
         geometric in [0.1, 200].
  cfg 5  32 atoms on a jittered 4x4x2 lattice (3.0 bohr), per atom
         2s 2p 2d 1f 1g with 1-4 primitives, exponents log-uniform [0.2, 2000].

Primitive counts are uniform in 1..6 (1..4 for cfg 5), exponents sorted in
descending order (libcint's screening takes the last one as the most
diffuse, libcint-master/src/cint2e.c:123).  Contraction coefficients are
N(0,1) draws times the primitive norm CINTgto_norm (libcint-master/src/
misc.c:70-92), rescaled to unit self-overlap of the contracted radial
function.  Random numbers: splitmix64 -> uniform [0,1) -> Box-Muller.
"""
import math

import numpy as np

PTR_ENV_START = 20
MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64
        self._spare = None

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        if self._spare is not None:
            v, self._spare = self._spare, None
            return v
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self._spare = r * math.sin(2 * math.pi * u2)
        return r * math.cos(2 * math.pi * u2)

    def randint(self, lo, hi):
        """uniform integer in [lo, hi]"""
        return lo + int(self.uniform() * (hi - lo + 1))


def gaussian_int(n, alpha):
    n1 = (n + 1) * 0.5
    return math.exp(math.lgamma(n1)) / (2.0 * alpha ** n1)


def gto_norm(l, a):
    return 1.0 / math.sqrt(gaussian_int(2 * l + 2, 2 * a))


def make_shell(rng, l, nprim, lo, hi):
    exps = sorted((math.exp(math.log(lo) + rng.uniform() * (math.log(hi) - math.log(lo)))
                   for _ in range(nprim)), reverse=True)
    c = [rng.normal() * gto_norm(l, a) for a in exps]
    s = 0.0
    for p in range(nprim):
        for q in range(nprim):
            s += c[p] * c[q] * gaussian_int(2 * l + 2, exps[p] + exps[q])
    scale = 1.0 / math.sqrt(s)
    return exps, [x * scale for x in c]


class Molecule:
    """atm (natm, 6) int32, bas (nbas, 8) int32, env float64 (libcint layout)."""

    def __init__(self):
        self.coords = []
        self.charges = []
        self.shells = []  # (atom, l, exps, coefs[nctr][nprim])
        self.aux_start = None

    def add_atom(self, z, r):
        self.coords.append(tuple(float(v) for v in r))
        self.charges.append(z)
        return len(self.coords) - 1

    def add_shell(self, atom, l, exps, coefs):
        coefs = np.atleast_2d(np.asarray(coefs, dtype=np.float64))
        self.shells.append((atom, l, list(exps), coefs))

    def build(self):
        env = [0.0] * PTR_ENV_START
        atm = np.zeros((len(self.coords), 6), dtype=np.int32)
        for i, (z, r) in enumerate(zip(self.charges, self.coords)):
            atm[i, 0] = z
            atm[i, 1] = len(env)
            env.extend(r)
            env.append(0.0)  # zeta slot
            atm[i, 3] = len(env) - 1
        bas = np.zeros((len(self.shells), 8), dtype=np.int32)
        for s, (atom, l, exps, coefs) in enumerate(self.shells):
            bas[s, 0] = atom
            bas[s, 1] = l
            bas[s, 2] = len(exps)
            bas[s, 3] = coefs.shape[0]
            bas[s, 5] = len(env)
            env.extend(exps)
            bas[s, 6] = len(env)
            env.extend(coefs.reshape(-1).tolist())  # column-major [ctr][prim]
        self.atm, self.bas, self.env = atm, bas, np.asarray(env, dtype=np.float64)
        return self

    @property
    def nbas(self):
        return len(self.shells)

    def nao(self, shells=None):
        idx = range(self.nbas) if shells is None else shells
        return sum((2 * self.shells[s][1] + 1) * self.shells[s][3].shape[0] for s in idx)


HEAVY = [0, 0, 0, 1, 1, 2]  # 3s 2p 1d
LIGHT = [0, 0, 1]           # 2s 1p


def _lattice(rng, nx, ny, nz, spacing, jitter):
    pts = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                base = np.array([i, j, k], dtype=float) * spacing
                d = np.array([rng.uniform() * 2 - 1 for _ in range(3)]) * jitter
                pts.append(base + d)
    return pts


def _heavy_mask(rng, n, frac):
    """exactly round(frac*n) heavy atoms, positions chosen by a seeded shuffle"""
    nh = int(round(frac * n))
    order = list(range(n))
    for i in range(n - 1, 0, -1):
        j = rng.randint(0, i)
        order[i], order[j] = order[j], order[i]
    heavy = set(order[:nh])
    return [i in heavy for i in range(n)]


def _orbital_basis(mol, rng, atoms, heavy, lo, hi, pmax=6, lmax=2):
    for a, h in zip(atoms, heavy):
        for l in (HEAVY if h else LIGHT):
            if l > lmax:
                continue
            e, c = make_shell(rng, l, rng.randint(1, pmax), lo, hi)
            mol.add_shell(a, l, e, [c])


def config(cfg, seed=None):
    """Build synthetic molecule `cfg` (1..5)."""
    rng = SplitMix64(20261202 + cfg if seed is None else seed)
    mol = Molecule()
    if cfg == 1:
        pts = []
        while len(pts) < 3:
            p = np.array([rng.uniform() * 8.0 for _ in range(3)])
            if all(np.linalg.norm(p - q) >= 1.5 for q in pts):
                pts.append(p)
        atoms = [mol.add_atom(8 if i == 0 else 1, p) for i, p in enumerate(pts)]
        _orbital_basis(mol, rng, atoms, [True, False, False], 0.1, 5000.0)
    elif cfg in (2, 4):
        pts = _lattice(rng, 4, 4, 4, 2.8, 0.3)
        heavy = _heavy_mask(rng, len(pts), 1.0 / 3.0)
        atoms = [mol.add_atom(8 if h else 1, p) for p, h in zip(pts, heavy)]
        if cfg == 2:
            _orbital_basis(mol, rng, atoms, heavy, 0.08, 8000.0)
        else:
            _orbital_basis(mol, rng, atoms, heavy, 0.1, 5000.0)
            mol.aux_start = mol.nbas
            ratio = (200.0 / 0.1)
            for a in atoms:
                for l, cnt in ((0, 6), (1, 5), (2, 4), (3, 3), (4, 1)):
                    for k in range(cnt):
                        t = k / (cnt - 1) if cnt > 1 else 0.5
                        e = 0.1 * ratio ** t
                        mol.add_shell(a, l, [e], [[gto_norm(l, e)]])
    elif cfg == 3:
        pts = _lattice(rng, 8, 8, 4, 2.8, 0.3)
        heavy = _heavy_mask(rng, len(pts), 2.0 / 3.0)
        atoms = [mol.add_atom(8 if h else 1, p) for p, h in zip(pts, heavy)]
        _orbital_basis(mol, rng, atoms, heavy, 0.1, 5000.0)
    elif cfg == 5:
        pts = _lattice(rng, 4, 4, 2, 3.0, 0.3)
        atoms = [mol.add_atom(8, p) for p in pts]
        for a in atoms:
            for l, cnt in ((0, 2), (1, 2), (2, 2), (3, 1), (4, 1)):
                for _ in range(cnt):
                    e, c = make_shell(rng, l, rng.randint(1, 4), 0.2, 2000.0)
                    mol.add_shell(a, l, e, [c])
    else:
        raise ValueError(f"unknown config {cfg}")
    return mol.build()


def general_contraction_h2():
    """Two centres, shells s..g with up to 3 contraction columns (the input
    design of libcint-master/testsuite/test_int2e.py:8-36, raw coefficients)."""
    rng = SplitMix64(7)
    mol = Molecule()
    a0 = mol.add_atom(1, (0.0, 0.0, -0.7))
    a1 = mol.add_atom(1, (0.1, 0.2, 0.7))
    for atom in (a0, a1):
        for l, (nprim, nctr) in enumerate([(3, 2), (3, 3), (2, 2), (2, 1), (1, 1)]):
            exps = sorted((0.3 + 4.0 * rng.uniform() for _ in range(nprim)), reverse=True)
            coefs = [[rng.normal() for _ in range(nprim)] for _ in range(nctr)]
            mol.add_shell(atom, l, exps, coefs)
    return mol.build()
