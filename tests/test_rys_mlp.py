"""Learned Rys initializer: gradient checks (CPU back-propagation against
finite differences, device kernels against the CPU back-propagation),
training on the device, and what the network gains over Newton started from
the same rule without it (the asymptotic Hermite rule)."""
import numpy as np
import pytest

from cintx_amd import rys_mlp as R


def batch(seed, B=96):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(1e-6), np.log(1e3), B))
    n = rng.integers(1, 14, B)
    z = R.features(x, n)
    tgt = rng.normal(size=(B, R.OUT)) * R.mask(n)
    return z, n, tgt


def test_backprop_matches_finite_differences():
    p = R.init_params(3).astype(np.float64)
    z, n, tgt = batch(1)
    g = R.grad_np(p, z, n, tgt)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(np.arange(s, e), 6, replace=False) for s, e in
                          [(R.OW1, R.OB1), (R.OB1, R.OW2), (R.OW2, R.OB2), (R.OB2, R.OW3),
                           (R.OW3, R.OB3), (R.OB3, R.NPARAM)]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_np(pp, z, n, tgt).mean() - R.loss_np(pm, z, n, tgt).mean()) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(g[k]), (k, fd, g[k])


@pytest.mark.gpu
def test_device_gradient_matches_cpu_backprop():
    p = R.init_params(5)
    z, n, tgt = batch(2, B=1000)
    g_dev, loss_dev = R.device_grad(p, z, n, tgt)
    g_ref = R.grad_np(p.astype(np.float64), z.astype(np.float32).astype(np.float64), n,
                      tgt.astype(np.float32).astype(np.float64))
    loss_ref = R.loss_np(p.astype(np.float64), z.astype(np.float32).astype(np.float64), n,
                         tgt.astype(np.float32).astype(np.float64))
    assert np.allclose(loss_dev, loss_ref, rtol=2e-5, atol=1e-6)
    scale = np.abs(g_ref).max()
    assert np.max(np.abs(g_dev - g_ref)) <= 2e-5 * scale


@pytest.mark.gpu
def test_learned_initializer_ablation(oracle):
    """Ablation of the learned initializer.  Two FP64 Newton steps on the
    moment equations, started from (a) the trained network, (b) the
    asymptotic (large-x) rule, (c) the table rule's own roots; the table
    rule (the method without the network) judges each at 1e-14 relative and
    is returned when the polished rule misses.  Findings the asserts hold:
    the network starts Newton far closer than the asymptotic rule, but the
    polish itself cannot keep 1e-14 for n >= 4 even from exact roots (the
    moment system's conditioning), so the table rule, exact and faster, is
    what the ERI kernels use."""
    p, loss = R.train(nsamples=1 << 20, epochs=8, batch=1024, lr=3e-3)
    print("training loss", loss)
    assert loss < 0.05
    rng = np.random.default_rng(4)
    x = np.exp(rng.uniform(np.log(1e-3), np.log(60.0), 400))
    report = {}
    for n in (1, 2, 3, 4, 5):
        u, w, st, ms_net, err = R.evaluate(p, n, x, newton_steps=2, init_mode=0, with_error=True)
        _, _, st0, _, err0 = R.evaluate(p, n, x, newton_steps=2, init_mode=1, with_error=True)
        _, _, st3, _, err3 = R.evaluate(p, n, x, newton_steps=2, init_mode=3, with_error=True)
        _, _, _, ms_tab = R.evaluate(p, n, x, newton_steps=0, init_mode=2)
        med, med0, med3 = np.median(err), np.median(err0), np.median(err3)
        report[n] = (float(st.mean()), float(st0.mean()), float(st3.mean()), float(med),
                     float(med0), float(med3), ms_net, ms_tab)
        # whatever the path, the returned rule is the oracle's to 1e-12
        for k in range(0, len(x), 37):
            uo, wo = oracle.rys_roots(n, x[k])
            assert np.max(np.abs(u[k] - uo) / uo) < 1e-12
            assert np.max(np.abs(w[k] - wo) / wo) < 1e-12
        # the network's start is never worse than the asymptotic rule's, and
        # from two roots on it is orders of magnitude closer after the polish
        assert st.mean() >= st0.mean() - 0.02, (n, st.mean(), st0.mean())
        if n >= 2:
            assert med < 1e-4 and med < 1e-2 * med0, (n, med, med0)
        # no start settles more often than the exact roots do
        assert st.mean() <= st3.mean() + 0.02, (n, st.mean(), st3.mean())
    # the polish floor: exact roots stay exact for one root and drift past
    # the 1e-14 gate from four roots on
    assert report[1][5] <= 1e-15
    assert report[4][5] > 1e-14 and report[5][5] > 1e-14, report
    print("n: settled (network, asymptotic, exact start), median polish error (same), "
          "ms (network path, table rule):", report)
