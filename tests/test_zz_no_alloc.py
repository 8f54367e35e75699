"""The allocation-free J/K entry point (cintx_plan_jk_noalloc) against the
allocating one (cintx_plan_jk) through the C ABI: two calls in a row on one
GPU, the second with a smaller basis, each equal to the allocating build
within the reordering of FP64 atomic sums (1e-13 of the largest element),
and the scratch limit reported as an error without touching the outputs."""
import numpy as np
import pytest

from cintx_amd import capi
from cintx_amd import molecule as M

pytestmark = pytest.mark.gpu


def _subset(mol, natom):
    """The first `natom` atoms of a molecule with their shells."""
    sub = M.Molecule()
    keep = {}
    for a in range(natom):
        keep[a] = sub.add_atom(mol.charges[a], mol.coords[a])
    for atom, l, exps, coefs in mol.shells:
        if atom in keep:
            sub.add_shell(keep[atom], l, exps, coefs)
    sub.build()
    return sub


def _jk_both(mol, seed, stride):
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-12, stride))
    nao = mol.nao()
    rng = np.random.default_rng(seed)
    D = rng.normal(size=(nao, nao)) / nao  # not symmetric: both take (D + D^T)/2
    J0, K0 = p.jk(D)
    Dd, Jd, Kd = (capi.DeviceArray(nao * nao) for _ in range(3))
    Dd.upload(D)
    # outputs start as NaN: every element must be written
    Jd.upload(np.full(nao * nao, np.nan))
    Kd.upload(np.full(nao * nao, np.nan))
    p.jk_noalloc(Dd.address, Jd.address, Kd.address)
    p.sync()
    J1 = Jd.download().reshape(nao, nao)
    K1 = Kd.download().reshape(nao, nao)
    p.close()
    return (J0, K0), (J1, K1)


def test_jk_noalloc_equals_jk_twice():
    big = M.config(2)
    small = _subset(M.config(1), 2)
    # the 64-atom lattice (every 37th bra pair), then the first two atoms of
    # the 3-atom molecule (every unique quartet)
    for mol, seed, stride in ((big, 1, 37), (small, 2, 1)):
        (J0, K0), (J1, K1) = _jk_both(mol, seed, stride)
        for a, b in ((J0, J1), (K0, K1)):
            assert not np.isnan(b).any()
            scale = max(1.0, np.abs(a).max())
            assert np.abs(a - b).max() <= 1e-13 * scale, np.abs(a - b).max()


def test_jk_noalloc_rejects_a_density_beyond_the_scratch():
    """1000 atoms with one p shell each (3000 functions, a 72 MB density
    against the 64 MiB scratch): refused before anything is enqueued, and the
    output keeps its contents; a null output is refused too."""
    mol = M.Molecule()
    for a in range(1000):
        at = mol.add_atom(1, (1.7 * (a % 10), 1.7 * ((a // 10) % 10), 1.7 * (a // 100)))
        mol.add_shell(at, 1, [0.8], [[1.0]])
    mol.build()
    assert mol.nao() == 3000
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, [(0, 0, 0, 0)])
    buf = capi.DeviceArray(16)
    buf.upload(np.full(16, 7.0))
    with pytest.raises(capi.CintxError) as e:
        p.jk_noalloc(buf.address, buf.address, buf.address)
    assert e.value.status == 5  # MemoryLimitExceeded
    p.sync()
    assert np.all(buf.download() == 7.0)
    small = M.config(1)
    q = capi.Plan("int2e_sph", small.atm, small.bas, small.env, screened=(1e-300, 1))
    with pytest.raises(capi.CintxError):
        q.jk_noalloc(buf.address, buf.address, 0)
    q.close()
    p.close()
