"""Pins of the oracle's integrals (int2e_sph, int2e_cart, int2e_ip1_sph,
int2e_ip2_sph, int3c2e_sph, int3c2e_cart, int3c2e_ip1_sph, int3c2e_ip2_sph,
int2c2e_sph) against checks that do not reuse its algorithm:

* an independent McMurchie-Davidson implementation (tests/md_eri.py) with
  numerically fitted spherical harmonics;
* the (ss|ss) closed form;
* 8-fold permutational symmetry of the full cfg-1 tensor;
* central finite differences in the nuclear coordinate (ip1 = -d/dA_i,
  ip2 = -d/dC_k), translational invariance, and ip2 as the ip1 of the
  bra-ket swapped quartet;
* int3c2e as int2e with a zero-exponent s function on the third centre, and
  int2c2e as int2e with zero-exponent s functions in the j and l slots;
* the (s|s) Coulomb-metric closed form;
* linearity in the contraction coefficients (general contraction layout).
"""
import math

import numpy as np
import pytest

import md_eri
from cintx_amd import molecule as M


def block(O, name, shls, atm, bas, env):
    out, not0 = O.eval_tuple(name, shls, atm, bas, env)
    d = O.tuple_dims(name, bas, shls)
    nc = O.ncomp(name)
    return out.reshape([nc] + d[::-1]).transpose([0] + list(range(len(d), 0, -1))), not0


def assert_close(a, b, atol=1e-12, rtol=1e-11):
    err = np.abs(a - b)
    tol = atol + rtol * np.abs(b)
    assert np.all(err <= tol), (err.max(), np.abs(b).max())


def test_against_mcmurchie_davidson_cfg1(oracle, cfg1):
    rng = np.random.default_rng(11)
    for _ in range(40):
        shls = [int(s) for s in rng.integers(0, cfg1.nbas, 4)]
        ref = md_eri.int2e_sph_md(shls, cfg1.atm, cfg1.bas, cfg1.env)
        got, _ = block(oracle, "int2e_sph", shls, cfg1.atm, cfg1.bas, cfg1.env)
        assert_close(got[0], ref)


def test_int2e_cart_against_mcmurchie_davidson(oracle, cfg1):
    rng = np.random.default_rng(15)
    gc = M.general_contraction_h2()
    cases = [(cfg1, [int(s) for s in rng.integers(0, cfg1.nbas, 4)]) for _ in range(25)]
    cases += [(gc, [int(s) for s in rng.integers(0, gc.nbas, 4)]) for _ in range(6)]
    for mol, shls in cases:
        ref = md_eri.int2e_cart_md(shls, mol.atm, mol.bas, mol.env)
        got, _ = block(oracle, "int2e_cart", shls, mol.atm, mol.bas, mol.env)
        assert_close(got[0], ref, atol=1e-11, rtol=1e-11)


def test_int3c2e_cart_equals_int2e_cart_with_unit_function(oracle):
    mol = M.config(4)
    rng = np.random.default_rng(16)
    for _ in range(10):
        i, j = [int(s) for s in rng.integers(0, mol.aux_start, 2)]
        k = int(rng.integers(mol.aux_start, mol.nbas))
        m2 = M.Molecule()
        m2.coords, m2.charges = list(mol.coords), list(mol.charges)
        m2.shells = list(mol.shells)
        m2.add_shell(mol.shells[k][0], 0, [0.0], [[1.0]])
        m2.build()
        three, _ = block(oracle, "int3c2e_cart", [i, j, k], mol.atm, mol.bas, mol.env)
        four, _ = block(oracle, "int2e_cart", [i, j, k, m2.nbas - 1], m2.atm, m2.bas, m2.env)
        assert_close(three[0], four[0][..., 0] / 0.282094791773878143, atol=1e-13)


def test_against_mcmurchie_davidson_general_contraction(oracle):
    mol = M.general_contraction_h2()
    rng = np.random.default_rng(5)
    tuples = [(0, 5, 1, 6), (2, 7, 3, 8), (4, 9, 0, 5), (3, 3, 8, 9), (9, 4, 4, 2),
              (1, 1, 1, 1), (2, 2, 7, 7)]
    for _ in range(6):
        tuples.append(tuple(int(s) for s in rng.integers(0, mol.nbas, 4)))
    for shls in tuples:
        ref = md_eri.int2e_sph_md(list(shls), mol.atm, mol.bas, mol.env)
        got, _ = block(oracle, "int2e_sph", list(shls), mol.atm, mol.bas, mol.env)
        assert_close(got[0], ref, atol=1e-11, rtol=1e-11)


def test_ssss_closed_form(oracle):
    mol = M.Molecule()
    ra, rb, rc, rd = (0, 0, 0), (0.3, -0.2, 0.9), (1.1, 0.4, -0.5), (-0.6, 1.2, 0.2)
    exps = (1.3, 0.7, 2.2, 0.45)
    for r, e in zip((ra, rb, rc, rd), exps):
        a = mol.add_atom(1, r)
        mol.add_shell(a, 0, [e], [[1.0]])
    mol.build()
    got, _ = block(oracle, "int2e_sph", [0, 1, 2, 3], mol.atm, mol.bas, mol.env)
    a, b, c, d = exps
    A, B, C, D = map(np.array, (ra, rb, rc, rd))
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    T = p * q / (p + q) * np.dot(P - Q, P - Q)
    f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    ref = (2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * math.exp(-a * b / p * np.dot(A - B, A - B))
           * math.exp(-c * d / q * np.dot(C - D, C - D)) * f0)
    ref *= 0.282094791773878143 ** 4  # libcint's s normalisation (g1e.c:566)
    assert abs(got[0, 0, 0, 0, 0] - ref) <= 1e-15 * ref


@pytest.fixture(scope="module")
def cfg1_tensor(oracle, cfg1):
    return oracle.full_tensor_int2e(cfg1.atm, cfg1.bas, cfg1.env)


def test_eightfold_symmetry_cfg1(cfg1_tensor):
    e = cfg1_tensor
    assert e.shape == (24, 24, 24, 24)
    scale = np.abs(e).max()
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.abs(e - e.transpose(perm)).max() <= 1e-12 + 1e-14 * scale


def four_centre_molecule(ls, seed=3):
    rng = M.SplitMix64(seed)
    mol = M.Molecule()
    for k, l in enumerate(ls):
        r = [rng.uniform() * 2.0 - 1.0 for _ in range(3)]
        a = mol.add_atom(1, r)
        e, c = M.make_shell(rng, l, 2, 0.4, 3.0)
        mol.add_shell(a, l, e, [c])
    return mol.build()


def moved(mol, atom, axis, h):
    env = mol.env.copy()
    env[mol.atm[atom, 1] + axis] += h
    return env


@pytest.mark.parametrize("ls", [(0, 0, 0, 0), (1, 0, 1, 0), (2, 1, 0, 1), (1, 2, 2, 0)])
def test_ip1_finite_difference(oracle, ls):
    mol = four_centre_molecule(ls)
    shls = [0, 1, 2, 3]
    ip1, _ = block(oracle, "int2e_ip1_sph", shls, mol.atm, mol.bas, mol.env)
    h = 1e-4
    for ax in range(3):
        ep = moved(mol, 0, ax, h)
        em = moved(mol, 0, ax, -h)
        fp, _ = block(oracle, "int2e_sph", shls, mol.atm, mol.bas, ep)
        fm, _ = block(oracle, "int2e_sph", shls, mol.atm, mol.bas, em)
        fd = -(fp[0] - fm[0]) / (2 * h)  # ip1 = (nabla_r i j|kl) = -d/dA_i
        assert np.abs(ip1[ax] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_ip1_translational_invariance(oracle):
    # d/dA + d/dB + d/dC + d/dD = 0; with 8-fold symmetry every centre's
    # derivative is an ip1 of a permuted tuple.
    mol = four_centre_molecule((1, 2, 0, 1), seed=9)
    tot = 0
    for perm in [(0, 1, 2, 3), (1, 0, 2, 3), (2, 3, 0, 1), (3, 2, 0, 1)]:
        b, _ = block(oracle, "int2e_ip1_sph", list(perm), mol.atm, mol.bas, mol.env)
        # back to (0,1,2,3) axis order: b[c, i0, i1, i2, i3] indexes shells perm
        inv = np.argsort(perm)
        tot = tot + b.transpose([0] + [1 + int(k) for k in inv])
    assert np.abs(tot).max() <= 1e-12 * max(1, np.abs(b).max())


@pytest.mark.parametrize("ls", [(0, 0, 0, 0), (1, 0, 1, 0), (2, 1, 0, 1), (0, 2, 2, 1)])
def test_ip2_finite_difference(oracle, ls):
    # int2e_ip2 = (ij|nabla_r k l) = -d/dC_k with k alone on atom 2
    mol = four_centre_molecule(ls, seed=6)
    shls = [0, 1, 2, 3]
    ip2, _ = block(oracle, "int2e_ip2_sph", shls, mol.atm, mol.bas, mol.env)
    h = 1e-4
    for ax in range(3):
        fp, _ = block(oracle, "int2e_sph", shls, mol.atm, mol.bas, moved(mol, 2, ax, h))
        fm, _ = block(oracle, "int2e_sph", shls, mol.atm, mol.bas, moved(mol, 2, ax, -h))
        fd = -(fp[0] - fm[0]) / (2 * h)
        assert np.abs(ip2[ax] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_ip2_is_ip1_of_the_swapped_quartet(oracle):
    # (ij|nabla k l) = (nabla k l|ij): the nabla_k recurrence against the
    # nabla_i one through bra-ket symmetry, general contractions included
    gc = M.general_contraction_h2()
    rng = np.random.default_rng(8)
    for _ in range(10):
        i, j, k, l = (int(x) for x in rng.integers(0, gc.nbas, 4))
        ip2, _ = block(oracle, "int2e_ip2_sph", [i, j, k, l], gc.atm, gc.bas, gc.env)
        ip1, _ = block(oracle, "int2e_ip1_sph", [k, l, i, j], gc.atm, gc.bas, gc.env)
        ref = ip1.transpose(0, 3, 4, 1, 2)
        assert_close(ip2, ref, atol=1e-11 * max(1.0, np.abs(ref).max()), rtol=1e-10)


@pytest.mark.parametrize("ls", [(0, 0, 1), (1, 2, 3), (2, 0, 4)])
def test_int3c2e_ip2_finite_difference(oracle, ls):
    mol = four_centre_molecule(ls + (0,), seed=5)
    shls = [0, 1, 2]
    ip2, _ = block(oracle, "int3c2e_ip2_sph", shls, mol.atm, mol.bas, mol.env)
    h = 1e-4
    for ax in range(3):
        fp, _ = block(oracle, "int3c2e_sph", shls, mol.atm, mol.bas, moved(mol, 2, ax, h))
        fm, _ = block(oracle, "int3c2e_sph", shls, mol.atm, mol.bas, moved(mol, 2, ax, -h))
        fd = -(fp[0] - fm[0]) / (2 * h)
        assert np.abs(ip2[ax] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    # translational invariance: ip1(i) + ip1(j) + ip2(k) = 0
    ipi, _ = block(oracle, "int3c2e_ip1_sph", [0, 1, 2], mol.atm, mol.bas, mol.env)
    ipj, _ = block(oracle, "int3c2e_ip1_sph", [1, 0, 2], mol.atm, mol.bas, mol.env)
    tot = ipi + ipj.transpose(0, 2, 1, 3) + ip2
    assert np.abs(tot).max() <= 1e-11 * max(1.0, np.abs(ip2).max())


def test_int3c2e_equals_int2e_with_unit_function(oracle):
    mol = M.config(4)
    rng = np.random.default_rng(2)
    # append an s function with exponent 0 and coefficient 1 on each aux centre
    for _ in range(12):
        i, j = [int(s) for s in rng.integers(0, mol.aux_start, 2)]
        k = int(rng.integers(mol.aux_start, mol.nbas))
        m2 = M.Molecule()
        m2.coords, m2.charges = list(mol.coords), list(mol.charges)
        m2.shells = list(mol.shells)
        m2.add_shell(mol.shells[k][0], 0, [0.0], [[1.0]])
        m2.build()
        three, _ = block(oracle, "int3c2e_sph", [i, j, k], mol.atm, mol.bas, mol.env)
        four, _ = block(oracle, "int2e_sph", [i, j, k, m2.nbas - 1], m2.atm, m2.bas, m2.env)
        assert_close(three[0], four[0][..., 0] / 0.282094791773878143, atol=1e-13)


def test_int2c2e_ss_closed_form(oracle):
    mol = M.Molecule()
    ra, rb = (0.2, -0.1, 0.3), (1.0, 0.6, -0.4)
    a0, a1 = mol.add_atom(1, ra), mol.add_atom(1, rb)
    mol.add_shell(a0, 0, [1.7], [[0.8]])
    mol.add_shell(a1, 0, [0.35], [[1.3]])
    mol.build()
    got, _ = block(oracle, "int2c2e_sph", [0, 1], mol.atm, mol.bas, mol.env)
    a, b = 1.7, 0.35
    R2 = float(np.dot(np.subtract(ra, rb), np.subtract(ra, rb)))
    T = a * b / (a + b) * R2
    f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    ref = 0.8 * 1.3 * 2 * math.pi ** 2.5 / (a * b * math.sqrt(a + b)) * f0
    ref *= 0.282094791773878143 ** 2  # s normalisation of the two shells
    assert abs(got[0, 0, 0] - ref) <= 1e-15 * ref
    # symmetric metric: (i|k) = (k|i)
    back, _ = block(oracle, "int2c2e_sph", [1, 0], mol.atm, mol.bas, mol.env)
    assert back[0, 0, 0] == pytest.approx(got[0, 0, 0], rel=1e-15)


def test_int2c2e_equals_int2e_with_unit_functions(oracle):
    mol = M.config(4)
    gc = M.general_contraction_h2()
    rng = np.random.default_rng(12)
    cases = [(mol, int(i), int(k)) for i, k in rng.integers(mol.aux_start, mol.nbas, (14, 2))]
    cases += [(gc, int(i), int(k)) for i, k in rng.integers(0, gc.nbas, (8, 2))]
    for m, i, k in cases:
        m2 = M.Molecule()
        m2.coords, m2.charges = list(m.coords), list(m.charges)
        m2.shells = list(m.shells)
        m2.add_shell(m.shells[i][0], 0, [0.0], [[1.0]])
        m2.add_shell(m.shells[k][0], 0, [0.0], [[1.0]])
        m2.build()
        two, _ = block(oracle, "int2c2e_sph", [i, k], m.atm, m.bas, m.env)
        four, _ = block(oracle, "int2e_sph", [i, m2.nbas - 2, k, m2.nbas - 1], m2.atm, m2.bas,
                        m2.env)
        ref = four[0][:, 0, :, 0] / 0.282094791773878143 ** 2
        assert_close(two[0], ref, atol=1e-13 * max(1.0, np.abs(ref).max()))
        # the metric is symmetric: (i|k) = (k|i)^T
        back, _ = block(oracle, "int2c2e_sph", [k, i], m.atm, m.bas, m.env)
        assert_close(back[0], two[0].T, atol=1e-13 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("ls", [(0, 0, 1), (1, 2, 3), (2, 0, 4)])
def test_int3c2e_ip1_finite_difference(oracle, ls):
    mol = four_centre_molecule(ls + (0,), seed=4)
    shls = [0, 1, 2]
    ip1, _ = block(oracle, "int3c2e_ip1_sph", shls, mol.atm, mol.bas, mol.env)
    h = 1e-4
    for ax in range(3):
        fp, _ = block(oracle, "int3c2e_sph", shls, mol.atm, mol.bas, moved(mol, 0, ax, h))
        fm, _ = block(oracle, "int3c2e_sph", shls, mol.atm, mol.bas, moved(mol, 0, ax, -h))
        fd = -(fp[0] - fm[0]) / (2 * h)
        assert np.abs(ip1[ax] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_general_contraction_is_linear(oracle):
    # a 2-column shell gives the same numbers as two 1-column shells
    mol = M.general_contraction_h2()
    split = M.Molecule()
    split.coords, split.charges = list(mol.coords), list(mol.charges)
    cols = {}
    for s, (a, l, e, c) in enumerate(mol.shells):
        cols[s] = []
        for col in range(c.shape[0]):
            cols[s].append(len(split.shells))
            split.add_shell(a, l, e, [c[col]])
    split.build()
    for shls in [(0, 5, 1, 6), (2, 7, 1, 1), (1, 6, 2, 7)]:
        big, _ = block(oracle, "int2e_sph", list(shls), mol.atm, mol.bas, mol.env)
        ns = [2 * mol.shells[s][1] + 1 for s in shls]
        for combo in np.ndindex(*[len(cols[s]) for s in shls]):
            sub = [cols[s][c] for s, c in zip(shls, combo)]
            small, _ = block(oracle, "int2e_sph", sub, split.atm, split.bas, split.env)
            sl = tuple(slice(c * n, (c + 1) * n) for c, n in zip(combo, ns))
            assert np.array_equal(big[0][sl], small[0])


def test_dims_and_zero_fill(oracle):
    mol = M.Molecule()
    a0 = mol.add_atom(1, (0, 0, 0))
    a1 = mol.add_atom(1, (0, 0, 40.0))
    mol.add_shell(a0, 1, [50.0], [[1.0]])
    mol.add_shell(a1, 2, [60.0], [[1.0]])
    mol.build()
    # far apart tight shells: libcint's screen drops every primitive pair
    dims = [5, 7, 4, 6]
    out = np.full(int(np.prod(dims)), 7.0)
    o, not0 = oracle.eval_tuple("int2e_sph", [0, 1, 0, 1], mol.atm, mol.bas, mol.env,
                                dims=dims, out=out)
    assert not0 == 0
    o = o.reshape(dims[::-1]).transpose(3, 2, 1, 0)
    assert np.all(o[:3, :5, :3, :5] == 0)
    assert np.all(o[3:] == 7.0) and np.all(o[:, 5:] == 7.0)
    # dims larger than the block: values land in the natural corner
    m = M.config(1)
    nat, _ = oracle.eval_tuple("int2e_sph", [5, 3, 4, 8], m.atm, m.bas, m.env)
    d0 = oracle.tuple_dims("int2e_sph", m.bas, [5, 3, 4, 8])
    dims = [d + 2 for d in d0]
    big, _ = oracle.eval_tuple("int2e_sph", [5, 3, 4, 8], m.atm, m.bas, m.env, dims=dims)
    big = big.reshape(dims[::-1]).transpose(3, 2, 1, 0)
    nat = nat.reshape(d0[::-1]).transpose(3, 2, 1, 0)
    assert np.array_equal(big[:d0[0], :d0[1], :d0[2], :d0[3]], nat)


def test_work_model_counts(oracle, cfg1):
    # with nothing screened, surviving primitive tuples = product of nprims
    mol = four_centre_molecule((1, 0, 2, 1))
    c = oracle.work_model("int2e_sph", [0, 1, 2, 3], mol.atm, mol.bas, mol.env)
    assert c[0] == 16 and c[1] == 16 * 3 and c[2] > 0
