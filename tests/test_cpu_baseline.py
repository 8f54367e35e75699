"""The CPU baseline (oracle/cpu_simd.cpp, cintx's f64x4 CPU kernel restated)
computes the same integrals: every int2e_sph block of the 3-atom molecule,
and sampled blocks of the s..g molecule, within the reference's element rule
of the long-double judge; general contractions are refused."""
import numpy as np
import pytest

from cintx_amd import molecule as M


def test_cpu_baseline_cfg1_all_unique_quartets(oracle):
    mol = M.config(1)
    n = mol.nbas
    t = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(i + 1)
         for l in range(k + 1)]
    got, _, _ = oracle.cpu_baseline_int2e(t, mol.atm, mol.bas, mol.env, nthreads=4, keep=True)
    ref, _, _ = oracle.eval_list("int2e_sph", t, mol.atm, mol.bas, mol.env, ext=True)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-12 + 1e-12 * np.abs(ref)), err.max()


def test_cpu_baseline_high_l(oracle):
    mol = M.config(5)
    rng = np.random.default_rng(12)
    t = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(300)]
    got, _, _ = oracle.cpu_baseline_int2e(t, mol.atm, mol.bas, mol.env, nthreads=4, keep=True)
    ref, _, _ = oracle.eval_list("int2e_sph", t, mol.atm, mol.bas, mol.env, ext=True)
    # libcint's transfer direction (kept by this restatement) loses up to
    # ~1e-10 on the worst (g.|g g) quartets of this basis (test_oracle_judge)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-10 + 1e-10 * np.abs(ref)), err.max()
    assert np.quantile(err, 0.999) <= 1e-12


def test_cpu_baseline_refuses_general_contractions(oracle):
    mol = M.general_contraction_h2()
    with pytest.raises(ValueError):
        oracle.cpu_baseline_int2e([(0, 1, 2, 3)], mol.atm, mol.bas, mol.env, keep=True)


def test_cpu_jk_equals_the_full_tensor_contraction(oracle):
    """The CPU J/K from unique quartets (each distinct 8-fold image added)
    equals the contraction of the judge's full cfg-1 tensor with D."""
    mol = M.config(1)
    n = mol.nbas
    t = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(i + 1)
         for l in range(k + 1) if i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l]
    eri = oracle.full_tensor_int2e(mol.atm, mol.bas, mol.env, ext=True)
    nao = eri.shape[0]
    rng = np.random.default_rng(11)
    D = rng.normal(size=(nao, nao)) / nao
    D = D + D.T
    Jr = np.einsum("pqrs,rs->pq", eri, D)
    Kr = np.einsum("pqrs,qs->pr", eri, D)
    J, K = oracle.cpu_baseline_jk(t, mol.atm, mol.bas, mol.env, D, nthreads=4)
    assert np.abs(J - Jr).max() <= 1e-11 * np.abs(Jr).max()
    assert np.abs(K - Kr).max() <= 1e-11 * np.abs(Kr).max()
