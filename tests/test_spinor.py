"""Spinor forms: int2e_spinor and int3c2e_spinor (libcint-master/src/
cint2e.c:1215-1224, cint3c2e.c:719-727).

CPU (no marker):
  * the library's Cartesian -> spinor coefficients (generated from the real
    solid harmonics and Clebsch-Gordan coefficients) against libcint's own
    tables as the reference holds them (tests/golden/c2spinor_coeffs.json,
    written by tests/golden/make_c2spinor_fixture.py);
  * the oracle's staged transforms against the reference test suite's
    definition of them (libcint-master/testsuite/test_cart2sph.py:434-477,
    554-567: one einsum over the coefficient tables);
  * the oracle on real integrals: for kappa = 0 shells the spinor block is a
    unitary image of the spherical block times the spin identity per
    electron, so |spinor|^2 = 4 |sph|^2 (two electrons) or 2 |sph|^2
    (int3c2e); electron exchange; hermiticity of each electron's pair
    (bra/ket swap conjugates); the s-shell closed form delta_ij delta_kl.
GPU (marked): the device against the oracle (judge Cartesian integrals,
spinor transform in extended precision) at the FP64 bar
|got - ref| <= 1e-12 + 1e-12 |ref| (complex modulus), through the raw
libcint entry points, the cintrs_eval shim and plans (eager and graph).
"""
import json
import os
import sys

import numpy as np
import pytest

from cintx_amd import capi
from cintx_amd import molecule as M

HERE = os.path.dirname(os.path.abspath(__file__))
ATOL = 1e-12
RTOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spinor_basis():
    """s..g shells with general contractions on two centres, each shell
    present with KAPPA_OF 0, -(l+1) and (for l > 0) +l."""
    mol = M.general_contraction_h2()
    rows = []
    for b in mol.bas:
        l = int(b[1])
        for kappa in (0, -(l + 1)) + ((l,) if l > 0 else ()):
            r = b.copy()
            r[4] = kappa
            rows.append(r)
    return mol.atm, np.array(rows, dtype=np.int32), mol.env


def fixture_rows(l, kappa):
    T = json.load(open(os.path.join(HERE, "golden", "c2spinor_coeffs.json")))["tables"]
    nf = (l + 1) * (l + 2) // 2

    def blk(name):
        return (np.array(T[f"{name}_{l}_R"]) + 1j * np.array(T[f"{name}_{l}_I"])).reshape(-1, 2 * nf)

    if kappa < 0:
        return blk("GT")
    if kappa > 0:
        return blk("LT")
    return np.vstack([blk("LT"), blk("GT")])


# ----------------------------------------------------------------- CPU

def test_library_spinor_coeffs_match_reference_tables():
    for l in range(5):
        for kappa in (-1, 0, 1):
            got = capi.spinor_coeffs(l, kappa)
            ref = fixture_rows(l, kappa)
            assert got.shape[0] == ref.shape[0] == capi.spinor_len(l, kappa)
            if ref.size:
                assert np.abs(got.reshape(ref.shape) - ref).max() < 1e-15, (l, kappa)


def _c_c2s(l):
    """The reference test's c_c2s(l) (kappa 0): [spin][cart][spinor]."""
    from oracle import spinor as S
    return S.coeffs(l, 0).astype(np.complex128).transpose(1, 2, 0)


@pytest.mark.parametrize("ls", [(0, 1, 2, 3), (2, 2, 1, 0), (4, 3, 2, 1), (1, 4, 4, 2)])
def test_oracle_stages_equal_reference_definition(ls):
    from oracle import spinor as S
    li, lj, lk, ll = ls
    rng = np.random.default_rng(sum(ls))
    nf = [(v + 1) * (v + 2) // 2 for v in ls]
    g = rng.standard_normal(nf)  # [a][b][c][d]
    Ci, Cj, Ck, Cl = (S.coeffs(v, 0) for v in ls)
    z = S.sf_e1(g.astype(np.longdouble), Ci, Cj)
    # test_cart2sph.py:445-446 (its gcart is [p][k][l][q] = [i][k][l][j])
    ref1 = np.einsum("xy,pklq,xpi,yqj->iklj", np.eye(2), g.transpose(0, 2, 3, 1),
                     _c_c2s(li).conj(), _c_c2s(lj))
    assert np.abs(z.transpose(0, 2, 3, 1) - ref1).max() < 1e-13
    out = S.sf_e2(z, Ck, Cl)
    # test_cart2sph.py:474-476 (its gcart is [i][p][q][j] complex)
    ref2 = np.einsum("xy,ipqj,xpk,yql->ijkl", np.eye(2),
                     z.transpose(0, 2, 3, 1).astype(np.complex128),
                     _c_c2s(lk).conj(), _c_c2s(ll))
    assert np.abs(out - ref2).max() < 1e-13


@pytest.mark.parametrize("ls", [(0, 1, 2), (3, 2, 4), (4, 4, 1)])
def test_oracle_3c_stage_equals_reference_definition(ls):
    from oracle import spinor as S
    li, lj, lk = ls
    rng = np.random.default_rng(7 + sum(ls))
    nf = [(v + 1) * (v + 2) // 2 for v in ls]
    g = rng.standard_normal(nf)  # [a][b][c]
    R = S.c2s_real(lk)
    xs = np.einsum("mc,abc->abm", R, g.astype(np.longdouble))
    got = S.sf_e1(xs, S.coeffs(li, 0), S.coeffs(lj, 0))
    # test_cart2sph.py:564-566 (its gcart is [p][r][q] = [i][k][j])
    ref = np.einsum("xy,prq,xpi,yqj,rk->ijk", np.eye(2), g.transpose(0, 2, 1),
                    _c_c2s(li).conj(), _c_c2s(lj), R.T.astype(float))
    assert np.abs(got - ref).max() < 1e-13


def test_oracle_spinor_invariants(oracle):
    from oracle import spinor as S
    atm, bas, env = spinor_basis()
    k0 = [s for s in range(len(bas)) if bas[s, 4] == 0]
    rng = np.random.default_rng(11)
    done = 0
    while done < 6:
        sh = [int(k0[v]) for v in rng.integers(0, len(k0), 4)]
        sph, _ = oracle.eval_tuple("int2e_sph", sh, atm, bas, env, ext=True)
        if not sph.any():  # screened out: nothing to compare
            continue
        done += 1
        sp, _ = S.int2e_spinor(sh, atm, bas, env)
        assert abs(np.sum(np.abs(sp) ** 2) / (4 * np.sum(sph ** 2)) - 1) < 1e-13
        ex, _ = S.int2e_spinor([sh[2], sh[3], sh[0], sh[1]], atm, bas, env)
        assert np.abs(ex.transpose(2, 3, 0, 1) - sp).max() <= 1e-15 * np.abs(sp).max()
        hm, _ = S.int2e_spinor([sh[1], sh[0], sh[3], sh[2]], atm, bas, env)
        assert np.abs(hm.transpose(1, 0, 3, 2).conj() - sp).max() <= 1e-15 * np.abs(sp).max()
    done = 0
    while done < 4:
        sh = [int(k0[v]) for v in rng.integers(0, len(k0), 3)]
        sph, _ = oracle.eval_tuple("int3c2e_sph", sh, atm, bas, env, ext=True)
        if not sph.any():
            continue
        done += 1
        sp, _ = S.int3c2e_spinor(sh, atm, bas, env)
        assert abs(np.sum(np.abs(sp) ** 2) / (2 * np.sum(sph ** 2)) - 1) < 1e-13
    # s shells: the spinors are beta s, alpha s (m_j = -1/2, +1/2)
    s_sh = [s for s in range(len(bas)) if bas[s, 1] == 0 and bas[s, 4] == 0][:2]
    sh = [s_sh[0], s_sh[1], s_sh[1], s_sh[0]]
    sp, _ = S.int2e_spinor(sh, atm, bas, env)
    sph, _ = oracle.eval_tuple("int2e_sph", sh, atm, bas, env, ext=True)
    n = [int(bas[s, 3]) for s in sh]
    sph = sph.reshape(n[3], n[2], n[1], n[0]).transpose(3, 2, 1, 0)
    want = np.einsum("abcd,xy,zw->axbyczdw", sph, np.eye(2), np.eye(2)).reshape(sp.shape)
    assert np.abs(sp - want).max() <= 1e-16 * np.abs(want).max()


# ----------------------------------------------------------------- GPU

def assert_cparity(got, ref):
    err = np.abs(got - ref)
    bad = err > ATOL + RTOL * np.abs(ref)
    say(f"  max abs err {err.max():.3e}  max |ref| {np.abs(ref).max():.3e}")
    assert not bad.any(), (int(bad.sum()), float(err.max()))


def _f_order(a):
    return np.asarray(a).transpose(tuple(range(a.ndim))[::-1]).ravel()


@pytest.mark.gpu
def test_int2e_spinor_raw_against_oracle(oracle):
    from oracle import spinor as S
    atm, bas, env = spinor_basis()
    rng = np.random.default_rng(2026)
    opt = capi.Optimizer("int2e_spinor", atm, bas, env)
    for t in range(60):
        sh = [int(v) for v in rng.integers(0, len(bas), 4)]
        say(f"int2e_spinor {sh} l={[int(bas[s, 1]) for s in sh]} kappa={[int(bas[s, 4]) for s in sh]}")
        ref, nz = S.int2e_spinor(sh, atm, bas, env)
        got, gnz = capi.raw_call("int2e_spinor", sh, atm, bas, env, opt=opt if t % 2 else None)
        assert gnz == nz
        assert_cparity(got, _f_order(ref))
    opt.close()


@pytest.mark.gpu
def test_int3c2e_spinor_raw_against_oracle(oracle):
    from oracle import spinor as S
    atm, bas, env = spinor_basis()
    rng = np.random.default_rng(2027)
    for t in range(60):
        sh = [int(v) for v in rng.integers(0, len(bas), 3)]
        say(f"int3c2e_spinor {sh} l={[int(bas[s, 1]) for s in sh]}")
        ref, nz = S.int3c2e_spinor(sh, atm, bas, env)
        got, gnz = capi.raw_call("int3c2e_spinor", sh, atm, bas, env)
        assert gnz == nz
        assert_cparity(got, _f_order(ref))


@pytest.mark.gpu
def test_int2e_spinor_dims_and_shim(oracle):
    from oracle import spinor as S
    atm, bas, env = spinor_basis()
    sh = [3, 7, 12, 5]
    ref, _ = S.int2e_spinor(sh, atm, bas, env)
    n = list(ref.shape)
    dims = [v + 2 for v in n]
    say(f"int2e_spinor dims {dims} for natural {n}")
    got, _ = capi.raw_call("int2e_spinor", sh, atm, bas, env, dims=dims)
    big = got.reshape(dims[::-1]).transpose(3, 2, 1, 0)
    assert_cparity(big[:n[0], :n[1], :n[2], :n[3]].ravel(), ref.ravel())
    rest = big.copy()
    rest[:n[0], :n[1], :n[2], :n[3]] = 0
    assert not rest.any()  # outside the block: untouched (zeros here)
    out = np.zeros(2 * ref.size)
    s = capi.eval_raw(capi.INT2E_SPINOR, out, None, sh, atm, bas, env)
    assert s["bytes_written"] == out.size * 8
    assert_cparity(out.view(np.complex128), _f_order(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("family", ["int2e_spinor", "int3c2e_spinor"])
def test_spinor_plan_eager_and_graph(oracle, family):
    from oracle import spinor as S
    atm, bas, env = spinor_basis()
    rng = np.random.default_rng(5)
    ar = 4 if family == "int2e_spinor" else 3
    tuples = [[int(v) for v in rng.integers(0, len(bas), ar)] for _ in range(150)]
    say(f"plan {family} ntuples={len(tuples)}")
    p = capi.Plan(family, atm, bas, env, tuples)
    fn = S.int2e_spinor if ar == 4 else S.int3c2e_spinor
    ref = np.concatenate([_f_order(fn(t, atm, bas, env)[0]) for t in tuples])
    assert p.nelem == 2 * ref.size
    _, offs = p.tuples()
    assert offs[0] == 0 and np.all(np.diff(offs) >= 0)
    p.execute()
    p.sync()
    eager = p.output()
    assert_cparity(eager.view(np.complex128), ref)
    for rep in range(3):
        p.fill_output(np.nan)
        p.execute(graph=True)
        p.sync()
        g = p.output()
        assert np.array_equal(g, eager), rep
    which = [3, 0, 77]
    blocks, boffs, sizes = p.output_tuples(which)
    for q, o, n in zip(which, boffs, sizes):
        assert np.array_equal(blocks[o:o + n], eager[offs[q]:offs[q] + n])
    p.close()
