"""GPU parity: libcintx_hip.so (HIP kernels on the MI355X) against the oracle
on the same seeded inputs.  FP64 bar: the reference's own element criterion
|got - ref| <= 1e-12 + 1e-12 |ref| (UNIFIED_ATOL/UNIFIED_RTOL and the pass rule
of crates/cintx-oracle/src/compare.rs:21-22, 256-259) for every family,
derivatives included, with `ref` from the oracle's long-double build (the
judge: the exact integrals of the FP64 inputs, pinned in
test_oracle_judge.py).  The oracle's FP64 build restates libcint's rounding
as well as its algorithm, and that rounding alone breaks the criterion on
ill-conditioned quartets (up to 1e-10 on (g d|g g) of the s..g molecule, ~1e-11
on cfg-1 nabla integrals), so it is not the reference value here."""
import ctypes
import sys

import numpy as np
import pytest

from cintx_amd import capi
from cintx_amd import molecule as M

pytestmark = pytest.mark.gpu
ATOL = 1e-12
RTOL = 1e-12


def assert_parity(got, ref):
    err = np.abs(got - ref)
    tol = ATOL + RTOL * np.abs(ref)
    bad = err > tol
    say(f"  max abs err {err.max():.3e}  max |ref| {np.abs(ref).max():.3e}  "
        f"max rel(>1e-6) {np.max(err[np.abs(ref) > 1e-6] / np.abs(ref[np.abs(ref) > 1e-6]), initial=0):.3e}")
    assert not bad.any(), (int(bad.sum()), float(err.max()))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_blocks(O, name, tuples, mol):
    """The judge's blocks (long-double oracle) of the tuples, libcint layout."""
    out, offs, sizes = O.eval_list(name, tuples, mol.atm, mol.bas, mol.env, ext=True)
    return out, offs, sizes


def check_plan(O, name, tuples, mol, variant=None):
    say(f"plan {name} ntuples={len(tuples)} variant={variant}")
    p = capi.Plan(name, mol.atm, mol.bas, mol.env, tuples)
    if variant is not None:
        p.set_variant(variant)
    p.execute()
    p.sync()
    got = p.output()
    ref, offs, sizes = oracle_blocks(O, name, tuples, mol)
    t, o = p.tuples()
    assert np.array_equal(o, offs)
    assert_parity(got, ref)
    p.close()
    return got


def test_device_rys_rule_matches_oracle(oracle):
    """The device rule against the oracle's, and against the host evaluation
    of the same table (which test_rys_table_moment_equations pins to the
    moment equations with mpmath alone, independently of the oracle)."""
    L = capi.lib()
    L.cintx_rys_eval_device.argtypes = [ctypes.c_int, capi.DP, ctypes.c_int64, capi.DP, capi.DP]
    L.cintx_rys_eval_host.argtypes = [ctypes.c_int, capi.DP, ctypes.c_int64, capi.DP, capi.DP]
    for n in range(1, 15):
        xs = np.concatenate([np.linspace(0.0, 40.0 + 5 * n, 301), [1e-12, 3e-7, 1e4]])
        u = np.zeros(len(xs) * n)
        w = np.zeros(len(xs) * n)
        assert L.cintx_rys_eval_device(n, capi._dp(xs), len(xs), capi._dp(u), capi._dp(w)) == 0
        uh = np.zeros(len(xs) * n)
        wh = np.zeros(len(xs) * n)
        assert L.cintx_rys_eval_host(n, capi._dp(xs), len(xs), capi._dp(uh), capi._dp(wh)) == 0
        assert np.max(np.abs(u - uh) / uh) < 4e-14 and np.max(np.abs(w - wh) / wh) < 4e-14, n
        u, w = u.reshape(-1, n), w.reshape(-1, n)
        for k, x in enumerate(xs):
            uo, wo = oracle.rys_roots(n, x)
            tol = 1e-13 if x >= 35.0 + 5 * n else 2e-14
            assert np.max(np.abs(u[k] - uo) / uo) < tol
            assert np.max(np.abs(w[k] - wo) / wo) < tol


def test_int2e_sph_cfg1_full_tensor(oracle):
    mol = M.config(1)
    n = mol.nbas
    tuples = [(i, j, k, l) for l in range(n) for k in range(n) for j in range(n) for i in range(n)]
    check_plan(oracle, "int2e_sph", tuples, mol)


def test_int2e_sph_cfg2_sampled(oracle):
    mol = M.config(2)
    rng = np.random.default_rng(2)
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(3000)]
    check_plan(oracle, "int2e_sph", tuples, mol)


def test_int2e_ip1_sph_cfg1(oracle):
    mol = M.config(1)
    rng = np.random.default_rng(3)
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(2000)]
    check_plan(oracle, "int2e_ip1_sph", tuples, mol)


def test_int3c2e_sph_cfg4_sampled(oracle):
    mol = M.config(4)
    rng = np.random.default_rng(4)
    tuples = [(int(rng.integers(0, mol.aux_start)), int(rng.integers(0, mol.aux_start)),
               int(rng.integers(mol.aux_start, mol.nbas))) for _ in range(3000)]
    check_plan(oracle, "int3c2e_sph", tuples, mol)


def test_int3c2e_ip1_sph_cfg4_sampled(oracle):
    mol = M.config(4)
    rng = np.random.default_rng(5)
    tuples = [(int(rng.integers(0, mol.aux_start)), int(rng.integers(0, mol.aux_start)),
               int(rng.integers(mol.aux_start, mol.nbas))) for _ in range(1500)]
    check_plan(oracle, "int3c2e_ip1_sph", tuples, mol)


def test_int2e_ip2_sph_cfg1(oracle):
    mol = M.config(1)
    rng = np.random.default_rng(13)
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(2000)]
    check_plan(oracle, "int2e_ip2_sph", tuples, mol)


def test_int3c2e_ip2_sph_cfg4_sampled(oracle):
    mol = M.config(4)
    rng = np.random.default_rng(14)
    tuples = [(int(rng.integers(0, mol.aux_start)), int(rng.integers(0, mol.aux_start)),
               int(rng.integers(mol.aux_start, mol.nbas))) for _ in range(1500)]
    check_plan(oracle, "int3c2e_ip2_sph", tuples, mol)


def test_cart_families(oracle):
    mol = M.config(1)
    n = mol.nbas
    full = [(i, j, k, l) for l in range(n) for k in range(n) for j in range(n) for i in range(n)]
    check_plan(oracle, "int2e_cart", full, mol)
    aux = M.config(4)
    rng = np.random.default_rng(17)
    t3 = [(int(rng.integers(0, aux.aux_start)), int(rng.integers(0, aux.aux_start)),
           int(rng.integers(aux.aux_start, aux.nbas))) for _ in range(3000)]
    check_plan(oracle, "int3c2e_cart", t3, aux)
    gc = M.general_contraction_h2()
    low = [s for s in range(gc.nbas) if gc.bas[s][1] <= 2]
    for _ in range(10):
        sh = [int(low[i]) for i in rng.integers(0, len(low), 4)]
        dims = [d + 1 for d in capi.natural_dims("int2e_cart", sh, gc.bas)]
        got, not0 = capi.raw_call("int2e_cart", sh, gc.atm, gc.bas, gc.env, dims)
        ref, rnot0 = oracle.eval_tuple("int2e_cart", sh, gc.atm, gc.bas, gc.env, dims=dims, ext=True)
        assert not0 == rnot0
        assert_parity(got, ref)
        got, not0 = capi.raw_call("int3c2e_cart", sh[:3], gc.atm, gc.bas, gc.env)
        ref, rnot0 = oracle.eval_tuple("int3c2e_cart", sh[:3], gc.atm, gc.bas, gc.env, ext=True)
        assert not0 == rnot0
        assert_parity(got, ref)
    with pytest.raises(capi.CintxError):
        capi.Plan("int2e_cart", mol.atm, mol.bas, mol.env, full[:10], precision=capi.PREC_MIXED,
                  mixed_threshold=1.0)


def test_int2c2e_sph_cfg4_aux_metric(oracle):
    # every (P|Q) of the cfg-4 auxiliary basis (s..g on 64 centres)
    mol = M.config(4)
    aux = np.arange(mol.aux_start, mol.nbas)
    rng = np.random.default_rng(9)
    tuples = [(int(i), int(k)) for i in aux for k in aux]
    tuples = [tuples[j] for j in rng.permutation(len(tuples))[:20000]]
    check_plan(oracle, "int2c2e_sph", tuples, mol)
    gc = M.general_contraction_h2()
    check_plan(oracle, "int2c2e_sph", [(i, k) for i in range(gc.nbas) for k in range(gc.nbas)],
               gc)


def test_raw_signature_dims_and_general_contraction(oracle):
    mol = M.general_contraction_h2()
    low = [s for s in range(mol.nbas) if mol.bas[s][1] <= 2]
    rng = np.random.default_rng(6)
    for _ in range(25):
        shls = [int(low[i]) for i in rng.integers(0, len(low), 4)]
        got, not0 = capi.int2e_sph(shls, mol.atm, mol.bas, mol.env)
        ref, rnot0 = oracle.eval_tuple("int2e_sph", shls, mol.atm, mol.bas, mol.env, ext=True)
        assert not0 == rnot0
        assert_parity(got, ref)
        d0 = capi.natural_dims("int2e_sph", shls, mol.bas)
        dims = [d + 1 for d in d0]
        gd, _ = capi.int2e_sph(shls, mol.atm, mol.bas, mol.env, dims=dims)
        rd, _ = oracle.eval_tuple("int2e_sph", shls, mol.atm, mol.bas, mol.env, dims=dims,
                                  ext=True)
        assert_parity(gd, rd)
    shls = [int(low[0]), int(low[1]), int(low[2])]
    for name in ("int2e_ip1_sph", "int3c2e_sph", "int3c2e_ip1_sph", "int2c2e_sph",
                 "int2e_ip2_sph", "int3c2e_ip2_sph"):
        sh = shls + [int(low[3])] if name.startswith("int2e") else shls
        sh = sh[:capi.arity(name)]
        got, not0 = capi.raw_call(name, sh, mol.atm, mol.bas, mol.env)
        ref, rnot0 = oracle.eval_tuple(name, sh, mol.atm, mol.bas, mol.env, ext=True)
        assert not0 == rnot0
        assert_parity(got, ref)
    for _ in range(10):
        sh = [int(low[i]) for i in rng.integers(0, len(low), 2)]
        dims = [d + 2 for d in capi.natural_dims("int2c2e_sph", sh, mol.bas)]
        gd, not0 = capi.int2c2e_sph(sh, mol.atm, mol.bas, mol.env, dims=dims)
        rd, rnot0 = oracle.eval_tuple("int2c2e_sph", sh, mol.atm, mol.bas, mol.env, dims=dims,
                                       ext=True)
        assert not0 == rnot0
        assert_parity(gd, rd)


def test_raw_calls_with_optimizers_from_four_threads(oracle):
    """Four host threads call the libcint-signature entry points at once,
    sharing one device-resident optimizer per family (the role of CINTOpt's
    pair data, libcint-master/src/optimizer.c:344-425), with some calls
    passing none; the calls take no lock and every block equals the judge."""
    import threading
    m1 = M.config(1)
    gc = M.general_contraction_h2()
    n = m1.nbas
    rng = np.random.default_rng(31)
    uniq = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(i + 1)
            for l in range(k + 1)]
    opts = {("int2e_sph", 1): capi.Optimizer("int2e_sph", m1.atm, m1.bas, m1.env),
            ("int2e_ip1_sph", 1): capi.Optimizer("int2e_ip1_sph", m1.atm, m1.bas, m1.env),
            ("int2e_sph", 0): capi.Optimizer("int2e_sph", gc.atm, gc.bas, gc.env),
            ("int3c2e_sph", 0): capi.Optimizer("int3c2e_sph", gc.atm, gc.bas, gc.env)}
    mols = {1: m1, 0: gc}
    jobs = [("int2e_sph", 1, t, True) for t in uniq]
    jobs += [("int2e_ip1_sph", 1, tuple(int(v) for v in rng.integers(0, n, 4)), True)
             for _ in range(800)]
    jobs += [("int2e_sph", 1, tuple(int(v) for v in rng.integers(0, n, 4)), False)
             for _ in range(400)]
    jobs += [("int2e_sph", 0, tuple(int(v) for v in rng.integers(0, gc.nbas, 4)), True)
             for _ in range(150)]
    jobs += [("int3c2e_sph", 0, tuple(int(v) for v in rng.integers(0, gc.nbas, 3)), True)
             for _ in range(150)]
    order = rng.permutation(len(jobs))
    jobs = [jobs[i] for i in order]
    results = [None] * len(jobs)
    errors = []

    def worker(t):
        try:
            for i in range(t, len(jobs), 4):
                fam, mk, sh, use = jobs[i]
                mol = mols[mk]
                results[i] = capi.raw_call(fam, sh, mol.atm, mol.bas, mol.env,
                                           opt=opts[(fam, mk)] if use else None)
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:3]
    say(f"  {len(jobs)} raw calls from 4 threads")
    for fam in ("int2e_sph", "int2e_ip1_sph", "int3c2e_sph"):
        for mk, mol in mols.items():
            idx = [i for i, j in enumerate(jobs) if j[0] == fam and j[1] == mk]
            if not idx:
                continue
            ref, offs, sizes = oracle.eval_list(fam, [jobs[i][2] for i in idx], mol.atm, mol.bas,
                                                mol.env, ext=True)
            got = np.concatenate([results[i][0] for i in idx])
            say(f"  {fam} on {'cfg1' if mk else 'gc basis'}: {len(idx)} calls")
            assert_parity(got, ref)
    for o in opts.values():
        o.close()


def test_cintrs_eval_on_device(oracle):
    mol = M.config(1)
    shls = [5, 3, 4, 8]
    req = int(np.prod(capi.natural_dims("int2e_sph", shls, mol.bas)))
    out = np.zeros(req)
    s = capi.eval_raw(capi.INT2E_SPH, out, None, shls, mol.atm, mol.bas, mol.env)
    ref, not0 = oracle.eval_tuple("int2e_sph", shls, mol.atm, mol.bas, mol.env, ext=True)
    assert s["not0"] == not0 and s["bytes_written"] == req * 8
    assert_parity(out, ref)
    # api 9 (int2e_cart): Cartesian extents
    req = int(np.prod(capi.natural_dims("int2e_cart", shls, mol.bas)))
    out = np.zeros(req)
    s = capi.eval_raw(capi.INT2E_CART, out, None, shls, mol.atm, mol.bas, mol.env)
    ref, not0 = oracle.eval_tuple("int2e_cart", shls, mol.atm, mol.bas, mol.env, ext=True)
    assert s["not0"] == not0 and s["bytes_written"] == req * 8
    assert_parity(out, ref)


def test_graph_replay_equals_eager(oracle):
    mol = M.config(1)
    n = mol.nbas
    tuples = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(n)
              for l in range(k + 1)]
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, tuples)
    p.execute()
    p.sync()
    eager = p.output()
    assert not np.isnan(eager).any()
    for rep in range(4):
        # poison the buffer first: a bin the graph left out would stay NaN
        p.fill_output(np.nan)
        p.execute(graph=True)
        p.sync()
        again = p.output()
        assert np.array_equal(eager, again), rep
        p.fill_output(np.nan)
        p.execute()
        p.sync()
        assert np.array_equal(eager, p.output()), rep
    p.close()


def test_graph_chains_with_runtime_class_bins(oracle):
    """The captured graph runs the bins as concurrent chains and gives the
    runtime-class bins of different chains their own workspace slices: many
    f/g classes replayed several times must equal the one-bin-at-a-time
    execution bit for bit, and the oracle."""
    mol = M.config(5)
    rng = np.random.default_rng(21)
    hi = [s for s in range(mol.nbas) if mol.bas[s][1] >= 3]
    tuples = [(int(hi[i]), int(hi[j]), int(rng.integers(0, mol.nbas)), int(rng.integers(0, mol.nbas)))
              for i, j in rng.integers(0, len(hi), (120, 2))]
    say(f"plan int2e_sph cfg5 ntuples={len(tuples)} (graph chains)")
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, tuples)
    assert sum(b["generic"] for b in p.bin_stats()) >= 8
    p.execute()
    p.sync()
    eager = p.output()
    assert not np.isnan(eager).any()
    for rep in range(3):
        p.fill_output(np.nan)
        p.execute(graph=True)
        p.sync()
        assert np.array_equal(eager, p.output()), rep
        p.fill_output(np.nan)
        p.execute()
        p.sync()
    ref, offs, sizes = oracle_blocks(oracle, "int2e_sph", tuples, mol)
    assert_parity(eager, ref)
    p.close()


def test_screened_plan_cfg2(oracle):
    mol = M.config(2)
    say("screened plan cfg2")
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-12, 97))
    p.execute()
    p.sync()
    got = p.output()
    t, offs = p.tuples()
    say(f"  tuples {p.ntuples}")
    assert p.ntuples > 1000
    rng = np.random.default_rng(7)
    pick = rng.choice(p.ntuples, size=min(2000, p.ntuples), replace=False)
    sub = [tuple(int(v) for v in t[q]) for q in pick]
    ref, roffs, sizes = oracle.eval_list("int2e_sph", sub, mol.atm, mol.bas, mol.env, ext=True)
    blk = np.concatenate([got[offs[q]:offs[q] + sizes[k]] for k, q in enumerate(pick)])
    assert_parity(blk, ref)
    p.close()


def test_screened_enumeration_complete_and_unique(oracle):
    """The screened plan holds each unordered quartet of shells at most once
    (i >= j, k >= l, ij >= kl in the pair order), and exactly the quartets
    whose Schwarz product sqrt(max|(ij|ij)|) sqrt(max|(kl|kl)|) reaches the
    threshold, with bra pairs restricted to every stride-th pair: the set is
    recomputed here from the judge's diagonal integrals (pairs within 1e-6
    relative of the threshold, where the device's and the judge's rounding
    may decide differently, are left out of the comparison)."""
    mol = M.config(2)
    thr, stride = 1e-12, 37
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(thr, stride))
    t, _ = p.tuples()
    p.close()
    n = mol.nbas
    ls = mol.bas[:, 1]
    pid = {}
    diag = []
    for i in range(n):
        for j in range(i + 1):
            pid[(i, j)] = len(diag)
            diag.append((i, j, i, j))
    ref, offs, sizes = oracle.eval_list("int2e_sph", diag, mol.atm, mol.bas, mol.env, ext=True)
    q = np.sqrt([np.abs(ref[a:a + k]).max() for a, k in zip(offs, sizes)])
    qmax = q.max()

    def key(a, b):
        return pid[(max(a, b), min(a, b))]

    got = {}
    for a, b, c, d in t:
        ka, kb = key(a, b), key(c, d)
        assert ls[a] >= ls[b] and ls[c] >= ls[d]
        assert ka >= kb
        k = (ka, kb)
        assert k not in got, k
        got[k] = True
    say(f"  screened cfg2 stride {stride}: {len(got)} unique quartets")
    live = np.where(q * qmax >= thr)[0]
    border = 0
    want = set()
    for ka in live:
        if ka % stride:
            continue
        for kb in live[live <= ka]:
            v = q[ka] * q[kb]
            if abs(v - thr) <= 1e-6 * thr:
                border += 1
                got.pop((int(ka), int(kb)), None)
                continue
            if v >= thr:
                want.add((int(ka), int(kb)))
    assert want == set(got), (len(want - set(got)), len(set(got) - want), border)


def test_high_l_against_judge_sg_sample(oracle):
    """The f/g classes of the 32-atom s..g molecule (runtime-class kernel):
    the benchmark's own sample (300 tuples, seed 5, of the screened plan with
    bra stride 64, and of its every-8th-tuple int2e_ip1_sph plan), the
    (g d|g g) tuple on which libcint's FP64 route loses 1e-10, and 1500
    random tuples with an f or g shell, all at the element rule against the
    judge."""
    mol = M.config(5)
    p5 = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-12, 64))
    t5, _ = p5.tuples()
    p5.execute()
    p5.sync()
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(t5), size=min(300, len(t5)), replace=False))
    got, _, _ = p5.output_tuples(pick)
    ref, _, _ = oracle.eval_list("int2e_sph", [tuple(int(v) for v in t5[q]) for q in pick],
                                 mol.atm, mol.bas, mol.env, ext=True)
    say(f"plan int2e_sph s..g screened sample {len(pick)} of {len(t5)}")
    assert_parity(got, ref)
    p5.close()
    t8 = np.array(t5[::8], dtype=np.int32)
    pi = capi.Plan("int2e_ip1_sph", mol.atm, mol.bas, mol.env, t8)
    pi.execute()
    pi.sync()
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(t8), size=min(300, len(t8)), replace=False))
    got, _, _ = pi.output_tuples(pick)
    ref, _, _ = oracle.eval_list("int2e_ip1_sph", [tuple(int(v) for v in t8[q]) for q in pick],
                                 mol.atm, mol.bas, mol.env, ext=True)
    say(f"plan int2e_ip1_sph s..g sample {len(pick)}")
    assert_parity(got, ref)
    pi.close()
    ls = mol.bas[:, 1]
    rng = np.random.default_rng(23)
    hl = [(215, 140, 79, 71)]
    while len(hl) < 1500:
        sh = tuple(int(s) for s in rng.integers(0, mol.nbas, 4))
        if max(ls[list(sh)]) >= 3:
            hl.append(sh)
    check_plan(oracle, "int2e_sph", hl, mol)
    check_plan(oracle, "int2e_ip1_sph", hl[:600], mol)


def test_high_l_generic_kernel_cfg5(oracle):
    # (ff|..) and (gg|gg) classes run the runtime-class kernel
    mol = M.config(5)
    rng = np.random.default_rng(8)
    ls = mol.bas[:, 1]
    hi = [s for s in range(mol.nbas) if ls[s] >= 3]
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(150)]
    tuples += [(int(hi[i]), int(hi[j]), int(hi[k]), int(hi[l])) for i, j, k, l in
               rng.integers(0, len(hi), (40, 4))]
    check_plan(oracle, "int2e_sph", tuples, mol)


def test_high_l_ip1_cfg5(oracle):
    mol = M.config(5)
    rng = np.random.default_rng(9)
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(120)]
    check_plan(oracle, "int2e_ip1_sph", tuples, mol)


@pytest.mark.parametrize("variant", range(capi.NVARIANTS))
def test_high_l_template_classes_every_variant(oracle, variant):
    """The int2e_sph classes with an f or g shell that run as template
    kernels (at most 3500 Cartesian elements and 135 values per chain; cintx_amd.build.high_l_classes):
    on the s..g molecule, eight tuples of every such class, shells taken from
    two neighbouring atoms (so the blocks are not screened to zero) and put
    in a random one of the eight orders of the quartet's symmetry, in every
    compiled launch shape, against the judge."""
    from cintx_amd import build as B
    mol = M.config(5)
    ls, at = mol.bas[:, 1], mol.bas[:, 0]
    natm = int(at.max()) + 1
    rng = np.random.default_rng(40 + variant)
    tuples = []
    for cls in B.high_l_classes():
        for _ in range(8):
            a = int(rng.integers(0, natm))
            near = (at == a) | (at == (a + 1) % natm)
            sh = [int(rng.choice(np.where(near & (ls == l))[0])) for l in cls]
            i, j, k, l = sh
            orders = [(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                      (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)]
            tuples.append(orders[int(rng.integers(0, 8))])
    check_plan(oracle, "int2e_sph", tuples, mol, variant)


@pytest.mark.parametrize("variant", range(capi.NVARIANTS))
def test_high_l_nabla_templates_every_variant(oracle, variant):
    """The nabla_i template classes with an f or g shell i and an s shell j
    (cintx_amd.build.high_l_nabla_classes): int2e_ip1_sph (f s|..) and
    (g s|..) tuples of neighbouring atoms on the s..g molecule, and
    int3c2e_ip2_sph triples whose aux shell is f or g (nabla on it), in
    every compiled launch shape, against the judge."""
    from cintx_amd import build as B
    rng = np.random.default_rng(60 + variant)
    mol = M.config(5)
    ls, at = mol.bas[:, 1], mol.bas[:, 0]
    natm = int(at.max()) + 1
    tuples = []
    for cls in B.high_l_nabla_classes():
        for _ in range(10):
            a = int(rng.integers(0, natm))
            near = (at == a) | (at == (a + 1) % natm)
            i, j, k, l = [int(rng.choice(np.where(near & (ls == v))[0])) for v in cls]
            tuples.append((i, j, k, l) if rng.integers(0, 2) else (i, j, l, k))
    check_plan(oracle, "int2e_ip1_sph", tuples, mol, variant)
    aux = M.config(4)
    ls, at = aux.bas[:, 1], aux.bas[:, 0]
    natm = int(at.max()) + 1
    orb = np.arange(aux.nbas) < aux.aux_start
    triples = []
    for lk in (3, 4):
        for li in range(3):
            for lj in range(3):
                for _ in range(6):
                    a = int(rng.integers(0, natm))
                    near = (at == a) | (at == (a + 1) % natm)
                    pick = [np.where(near & orb & (ls == li))[0],
                            np.where(near & orb & (ls == lj))[0],
                            np.where(near & ~orb & (ls == lk))[0]]
                    if all(len(c) for c in pick):
                        triples.append(tuple(int(rng.choice(c)) for c in pick))
    check_plan(oracle, "int3c2e_ip2_sph", triples, aux, variant)


def test_high_l_ip1_neighbouring_atoms(oracle):
    """int2e_ip1_sph f/g classes on tuples of neighbouring atoms of the s..g
    molecule, where the integrals are large and libcint's FP64 route loses
    most (measured against the judge: up to 1e4 times the 1e-12 bar on
    (f g|d s), 2e3 on (g g|s s), 110 on (s p|g g)): the device, whose
    transfers start from the centre of the larger exponent per primitive
    pair, stays within the bar."""
    mol = M.config(5)
    ls, at = mol.bas[:, 1], mol.bas[:, 0]
    natm = int(at.max()) + 1
    rng = np.random.default_rng(7)
    tuples = []
    for c in ("3030", "4130", "1131", "3032", "1300", "3420", "0144", "2311", "4400",
              "3300", "1311"):
        for _ in range(40):
            a = int(rng.integers(0, natm))
            near = (at == a) | (at == (a + 1) % natm) | (at == (a + 8) % natm)
            tuples.append(tuple(int(rng.choice(np.where(near & (ls == int(l)))[0])) for l in c))
    check_plan(oracle, "int2e_ip1_sph", tuples, mol)


def test_general_contraction_all_classes(oracle):
    mol = M.general_contraction_h2()
    rng = np.random.default_rng(10)
    tuples = [tuple(int(s) for s in rng.integers(0, mol.nbas, 4)) for _ in range(200)]
    check_plan(oracle, "int2e_sph", tuples, mol)
    check_plan(oracle, "int3c2e_sph", [t[:3] for t in tuples[:100]], mol)


def test_jk_build_cfg1(oracle):
    # J/K from the screened unique-quartet plan (screen off) against an
    # einsum over the oracle's full cfg-1 tensor
    mol = M.config(1)
    eri = oracle.full_tensor_int2e(mol.atm, mol.bas, mol.env, ext=True)
    nao = eri.shape[0]
    rng = np.random.default_rng(11)
    D = rng.normal(size=(nao, nao)) / nao
    D = D + D.T
    Jr = np.einsum("pqrs,rs->pq", eri, D)
    Kr = np.einsum("pqrs,qs->pr", eri, D)
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-300, 1))
    n = mol.nbas
    # every unique quartet except those with a zero Schwarz factor (a pair
    # whose primitives libcint's screen drops entirely)
    assert 0 < p.ntuples <= (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2
    J, K = p.jk(D)
    say(f"  J err {np.abs(J - Jr).max():.3e}  K err {np.abs(K - Kr).max():.3e}  "
        f"|J| {np.abs(Jr).max():.3e}")
    assert np.abs(J - Jr).max() <= 1e-11 * max(1.0, np.abs(Jr).max())
    assert np.abs(K - Kr).max() <= 1e-11 * max(1.0, np.abs(Kr).max())
    p.close()


def _jk_reference(eri, nao, seed):
    rng = np.random.default_rng(seed)
    D = rng.normal(size=(nao, nao)) / nao
    D = D + D.T
    return D, np.einsum("pqrs,rs->pq", eri, D), np.einsum("pqrs,qs->pr", eri, D)


def test_jk_fused_every_launch_variant(oracle):
    """The fused J/K epilogues (one-lane kernels: blocks staged in LDS per
    chunk and digested by their lane; wide kernels: the staged spherical
    block digested by the tuple's lanes) in every compiled launch shape,
    against an einsum over the oracle's full cfg-1 tensor."""
    mol = M.config(1)
    eri = oracle.full_tensor_int2e(mol.atm, mol.bas, mol.env, ext=True)
    D, Jr, Kr = _jk_reference(eri, eri.shape[0], 12)
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-300, 1))
    for v in range(capi.NVARIANTS):
        p.set_variant(v)
        J, K = p.jk(D)
        say(f"variant {v}: J err {np.abs(J - Jr).max():.3e}  K err {np.abs(K - Kr).max():.3e}")
        assert np.abs(J - Jr).max() <= 1e-11 * max(1.0, np.abs(Jr).max())
        assert np.abs(K - Kr).max() <= 1e-11 * max(1.0, np.abs(Kr).max())
    p.close()


def test_jk_with_runtime_class_bins(oracle):
    """J/K over s..g shells on two centres: the f and g classes run the
    runtime-class kernel (blocks written, then the digest kernel), the rest
    the fused class kernels; against an einsum over the oracle tensor."""
    mol = M.Molecule()
    a0 = mol.add_atom(8, (0.0, 0.0, 0.0))
    a1 = mol.add_atom(8, (0.3, -0.4, 1.9))
    rng = np.random.default_rng(21)
    for atom in (a0, a1):
        for l in range(5):
            e = sorted(rng.uniform(0.3, 6.0, 2), reverse=True)
            mol.add_shell(atom, l, e, [[M.gto_norm(l, x) * c for x, c in zip(e, (0.6, 0.5))]])
    mol.build()
    eri = oracle.full_tensor_int2e(mol.atm, mol.bas, mol.env)
    D, Jr, Kr = _jk_reference(eri, eri.shape[0], 13)
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-300, 1))
    J, K = p.jk(D)
    say(f"s..g: J err {np.abs(J - Jr).max():.3e}  K err {np.abs(K - Kr).max():.3e}  "
        f"|J| {np.abs(Jr).max():.3e}")
    assert np.abs(J - Jr).max() <= 1e-11 * max(1.0, np.abs(Jr).max())
    assert np.abs(K - Kr).max() <= 1e-11 * max(1.0, np.abs(Kr).max())
    p.close()


@pytest.mark.parametrize("variant", range(capi.NVARIANTS))
def test_every_launch_variant(oracle, variant):
    """Every compiled launch shape (lanes per tuple, primitive tuples per
    step) of every class meets the same parity bar; classes with fewer
    variants fall back to their default."""
    mol = M.config(1)
    n = mol.nbas
    rng = np.random.default_rng(20 + variant)
    full = [(i, j, k, l) for l in range(n) for k in range(n) for j in range(n) for i in range(n)]
    check_plan(oracle, "int2e_sph", full, mol, variant)
    check_plan(oracle, "int2e_ip1_sph",
               [tuple(int(s) for s in rng.integers(0, n, 4)) for _ in range(1500)], mol, variant)
    aux = M.config(4)
    t3 = [(int(rng.integers(0, aux.aux_start)), int(rng.integers(0, aux.aux_start)),
           int(rng.integers(aux.aux_start, aux.nbas))) for _ in range(1500)]
    check_plan(oracle, "int3c2e_sph", t3, aux, variant)
    check_plan(oracle, "int3c2e_ip1_sph", t3[:700], aux, variant)
    check_plan(oracle, "int2e_cart", full[::7], mol, variant)
    check_plan(oracle, "int3c2e_cart", t3[:700], aux, variant)
    # FP32 instantiations: every tuple of a mixed plan with a huge threshold
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, full, precision=capi.PREC_MIXED,
                  mixed_threshold=1e300)
    assert p.ntuples_fp32 == p.ntuples
    p.set_variant(variant)
    p.execute()
    p.sync()
    got = p.output()
    ref, offs, sizes = oracle_blocks(oracle, "int2e_sph", full, mol)
    # FP32 rounding is relative to the terms summed, not to the (possibly
    # cancelled) result: bound it by the tuple's Schwarz factor
    # sqrt(max|(ij|ij)|) sqrt(max|(kl|kl)|)
    diag = [(i, j, i, j) for i in range(n) for j in range(n)]
    dref, doffs, dsizes = oracle_blocks(oracle, "int2e_sph", diag, mol)
    qf = np.sqrt([np.abs(dref[a:a + k]).max() for a, k in zip(doffs, dsizes)]).reshape(n, n)
    scale = np.repeat([qf[i, j] * qf[k, l] for i, j, k, l in full], sizes)
    err = np.abs(got - ref)
    ratio = err[scale > 0] / scale[scale > 0]
    # a pair whose (ij|ij) libcint's screen drops entirely (e^-cce beyond
    # the cutoff twice) can still meet a compact ket pair: its integrals are
    # then below e^-30 or so
    assert np.all(err[scale == 0] <= 1e-11)
    worst = np.argsort(err / np.maximum(scale, 1e-300))[-3:]
    say(f"  fp32 variant {variant}: max err {err.max():.3e}, err/Schwarz bound "
        f"max {ratio.max():.3e} p99.9 {np.quantile(ratio, 0.999):.3e}; worst "
        + ", ".join(f"{got[w]:.6e} vs {ref[w]:.6e} (bound {scale[w]:.2e})" for w in worst))
    # FP32 recurrences: rounding relative to the bound, amplified by the HRR
    # cancellation of distant pairs (measured up to ~3e-4 of the bound on this
    # basis); an indexing error would be O(1) here
    assert ratio.max() <= 1e-2
    assert np.quantile(ratio, 0.999) <= 1e-4
    p.close()


def test_mixed_plan_error_bound_cfg2(oracle):
    """The mixed path (FP32 recurrences for tuples whose Schwarz bound is
    below 1e-9) stays within 1e-10 absolute of the oracle."""
    mol = M.config(2)
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-12, 97),
                  precision=capi.PREC_MIXED, mixed_threshold=1e-9)
    assert p.ntuples_fp32 > 0
    p.execute()
    p.sync()
    t, _ = p.tuples()
    rng = np.random.default_rng(11)
    first32 = p.ntuples - p.ntuples_fp32
    pick = np.sort(first32 + rng.choice(p.ntuples_fp32, size=min(3000, p.ntuples_fp32),
                                        replace=False))
    got, _, _ = p.output_tuples(pick)
    ref, _, _ = oracle.eval_list("int2e_sph", [tuple(int(v) for v in t[q]) for q in pick],
                                 mol.atm, mol.bas, mol.env, ext=True)
    err = float(np.abs(got - ref).max())
    say(f"  mixed cfg2: {p.ntuples_fp32}/{p.ntuples} tuples in FP32, max abs err {err:.3e}")
    assert err <= 1e-10
    p.close()


def test_mixed_plan_cfg3_256_atoms(oracle):
    """The 256-atom lattice (the headline molecule) in the mixed mode at the
    bench's thresholds, every 16384th bra pair: tuples sampled from the FP32
    and FP64 bins within 1e-10 of the judge, and the plan's J/K against the CPU
    J/K of the same unique quartets (every distinct 8-fold image added, an
    independent digestion)."""
    mol = M.config(3)
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, screened=(1e-12, 16384),
                  precision=capi.PREC_MIXED, mixed_threshold=1e-9)
    say(f"  cfg3 mixed: {p.ntuples} quartets, {p.ntuples_fp32} in FP32")
    assert p.ntuples_fp32 > 0
    p.execute()
    p.sync()
    t, _ = p.tuples()
    rng = np.random.default_rng(12)
    first32 = p.ntuples - p.ntuples_fp32
    pick = np.sort(np.concatenate([rng.choice(first32, size=min(1000, first32), replace=False),
                                   first32 + rng.choice(p.ntuples_fp32,
                                                        size=min(1000, p.ntuples_fp32),
                                                        replace=False)]))
    got, _, _ = p.output_tuples(pick)
    ref, _, _ = oracle.eval_list("int2e_sph", [tuple(int(v) for v in t[q]) for q in pick],
                                 mol.atm, mol.bas, mol.env, ext=True)
    err = float(np.abs(got - ref).max())
    say(f"  sampled max abs err vs judge {err:.3e}")
    assert err <= 1e-10
    nao = mol.nao()
    D = rng.normal(size=(nao, nao)) / nao
    D = D + D.T
    J, K = p.jk(D)
    Jc, Kc = oracle.cpu_baseline_jk(t, mol.atm, mol.bas, mol.env, D, nthreads=16)
    ej, ek = float(np.abs(J - Jc).max()), float(np.abs(K - Kc).max())
    say(f"  J/K vs CPU: {ej:.3e} {ek:.3e} (|J| {np.abs(Jc).max():.3e})")
    assert ej <= 1e-10 and ek <= 1e-10
    p.close()


def test_int3c2e_mixed_cfg4(oracle):
    """int3c2e_sph in the mixed mode on the 64-atom lattice with the
    6s5p4d3f1g auxiliary basis: a triple takes the FP32 recurrences when
    sqrt((ij|ij)) sqrt((P|P)) over the absolute coefficients is below 1e-9;
    every triple within 1e-10 of the judge."""
    mol = M.config(4)
    rng = np.random.default_rng(41)
    a0 = mol.aux_start
    tuples = [(int(rng.integers(0, a0)), int(rng.integers(0, a0)),
               int(rng.integers(a0, mol.nbas))) for _ in range(4000)]
    p = capi.Plan("int3c2e_sph", mol.atm, mol.bas, mol.env, tuples, precision=capi.PREC_MIXED,
                  mixed_threshold=1e-9)
    say(f"  int3c2e mixed: {p.ntuples_fp32} of {p.ntuples} triples in FP32")
    assert 0 < p.ntuples_fp32 < p.ntuples
    p.execute()
    p.sync()
    got = p.output()
    ref, _, _ = oracle.eval_list("int3c2e_sph", tuples, mol.atm, mol.bas, mol.env, ext=True)
    err = float(np.abs(got - ref).max())
    say(f"  max abs err vs judge {err:.3e}")
    assert err <= 1e-10
    p.close()


def test_autotune_keeps_results(oracle):
    mol = M.config(1)
    n = mol.nbas
    tuples = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(n)
              for l in range(k + 1)]
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, tuples)
    p.execute()
    p.sync()
    before = p.output()
    p.autotune(reps=2)
    info = p.bin_variants()
    say("  variants " + " ".join(f"{b['variant']}/{b['nvariants']}:T{b['lanes']}P{b['prims']}"
                                 for b in info))
    assert all(0 <= b["variant"] < b["nvariants"] for b in info)
    p.fill_output(np.nan)
    p.execute(graph=True)
    p.sync()
    after = p.output()
    scale = np.abs(before).max()
    assert np.abs(after - before).max() <= 1e-13 * scale
    p.close()


@pytest.mark.parametrize("variant", range(capi.NVARIANTS))
def test_work_counters_match_oracle_model(oracle, variant):
    """The device work counters (surviving primitive tuples after libcint's
    screen, algorithmic FLOPs) equal the oracle's count for every launch
    shape."""
    mol = M.config(1)
    n = mol.nbas
    tuples = [(i, j, k, l) for i in range(n) for j in range(i + 1) for k in range(n)
              for l in range(k + 1)]
    p = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env, tuples)
    p.set_variant(variant)
    prim, roots, flops = p.work()
    ref = np.sum([oracle.work_model("int2e_sph", t, mol.atm, mol.bas, mol.env) for t in tuples],
                 axis=0)
    say(f"  variant {variant}: prim {prim:.0f} vs {ref[0]:.0f}, flops {flops:.4e} vs {ref[2]:.4e}")
    assert prim == ref[0]
    assert roots == ref[1]
    # the device transforms cart->sph in the canonical shell order of its
    # class, the oracle in the caller's: the c2s FLOPs differ by the axis order
    assert abs(flops - ref[2]) <= 1e-3 * ref[2]
    p.close()
