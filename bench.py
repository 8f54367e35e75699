"""int2e_sph throughput on one MI355X (quartets/s), mixed vs plain FP64.

Workload: the 256-atom synthetic molecule (8x8x4 jittered lattice, cfg-1 shell
recipe, 2,819 spherical functions), unique 8-fold shell quartets that pass a
Schwarz screen of 1e-12, bra pairs restricted to every `--bra-stride`-th
significant pair so one step (one pass over that quartet set) stays within
seconds.  Inputs (pair data, quartet lists) are resident in HBM before timing;
outputs are written to HBM.

Prints one JSON line: metric/value/unit, ms_per_step, the roofline of the
dominant class kernel (FP64 VALU bound) of each plan, the FP64 path's rate
beside the mixed path's, the max abs error of both against the long-double
judge on a seeded sample, and a CPU baseline: oracle/cpu_simd.cpp (cintx's
f64x4 CPU kernel restated) on the host's cores, median of 5 runs on a bounded
sample of the same quartets.

`--calibrate` runs a plain device copy of known size instead (torch's copy
kernel, nothing of this library), the reference for the FETCH_SIZE /
WRITE_SIZE counter runs that the roofline traffic figure is scaled by.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bra-stride", type=int, default=256)
    ap.add_argument("--threshold", type=float, default=1e-12)
    ap.add_argument("--mixed-threshold", type=float, default=1e-9)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sample", type=int, default=4000)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip training/evaluating the learned Rys initializer")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run every class with its default launch shape")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the int3c2e (aux-basis lattice) and int2e/ip1 (s..g basis) lines")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the class kernels directly instead of replaying a hipGraph")
    ap.add_argument("--secondary-only", action="store_true",
                    help="run only the int3c2e and s..g-basis lines and print them as JSON "
                         "(for profiling those kernels; not the benchmark line)")
    ap.add_argument("--calibrate", type=int, default=0, metavar="NELEM",
                    help="copy NELEM doubles on the device (3 times) and exit: the "
                         "known-size reference for the byte counters")
    args = ap.parse_args()
    if args.calibrate:
        calibrate_copy(args.calibrate)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would give each rank its own bra pairs and its own output "
              "region, with one all-reduce of J and K at the end)", file=sys.stderr)
        sys.exit(2)

    if args.secondary_only:
        print(json.dumps({"secondary": secondary_configs(), "raw_path": raw_path_rate()}),
              flush=True)
        return

    from cintx_amd import capi, molecule
    mol = molecule.config(3)
    log(f"molecule: {len(mol.coords)} atoms, {mol.nbas} shells, {mol.nao()} functions")

    t0 = time.time()
    p64 = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env,
                    screened=(args.threshold, args.bra_stride), precision=capi.PREC_FP64)
    log(f"fp64 plan: {p64.ntuples} quartets, {p64.nelem} doubles, {p64.nbins} bins "
        f"({time.time() - t0:.1f}s)")
    t0 = time.time()
    pmx = capi.Plan("int2e_sph", mol.atm, mol.bas, mol.env,
                    screened=(args.threshold, args.bra_stride), precision=capi.PREC_MIXED,
                    mixed_threshold=args.mixed_threshold)
    log(f"mixed plan: {pmx.ntuples} quartets, {pmx.ntuples_fp32} in FP32, {pmx.nbins} bins "
        f"({time.time() - t0:.1f}s)")

    tuned = {}
    if not args.no_autotune:
        for plan, label in ((p64, "fp64"), (pmx, "mixed")):
            t0 = time.time()
            plan.autotune(reps=2)
            tuned[label] = ["{}{}{}{}{}:T{}P{}{}".format(
                *b["class"], "df"[b["precision"]], v["lanes"], v["prims"],
                "(fp64 kernel)" if v["variant"] >= capi.NVARIANTS else "")
                for b, v in zip(plan.bin_stats(), plan.bin_variants())]
            log(f"autotune {label}: {time.time() - t0:.1f}s")

    def timed(plan, label):
        for _ in range(args.warmup):
            plan.execute(graph=not args.no_graph)
            plan.sync()
        ms = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            plan.execute(graph=not args.no_graph)
            plan.sync()
            ms.append(plan.last_ms())
        wall = (time.perf_counter() - t_start) * 1e3 / args.steps
        log(f"{label}: events {np.mean(ms):.2f} ms/step, wall {wall:.2f} ms/step")
        return float(np.mean(ms)), wall

    ms64, wall64 = timed(p64, "fp64")
    msmx, wallmx = timed(pmx, "mixed")

    # accuracy on a seeded sample of tuples: both GPU paths against the judge
    err64, errmx = sample_errors(mol, p64, pmx, args.sample)
    log(f"sampled max abs err vs the judge: fp64 {err64:.3e}, mixed {errmx:.3e}")

    # J/K against a random symmetric density (N(0,1)/nbf): the FP64 plan's
    # quartets digested with their 8-fold images inside the class kernels
    # (fused: no ERI block is written)
    nao = mol.nao()
    rngd = np.random.default_rng(2)
    dens = rngd.normal(size=(nao, nao)) / nao
    dens = dens + dens.T
    p64.jk(dens)  # warm-up
    t0 = time.perf_counter()
    J64, K64 = p64.jk(dens)
    jk_ms = (time.perf_counter() - t0) * 1e3
    log(f"J/K build (host arrays, copies included): {jk_ms:.1f} ms")
    # the same on device-resident D, J, K (the allocation-free entry point):
    # the J/K step itself
    from cintx_amd import capi as C
    dD, dJ, dK = C.DeviceArray(nao * nao), C.DeviceArray(nao * nao), C.DeviceArray(nao * nao)
    dD.upload(dens)
    jk_dev = []
    for _ in range(4):
        t0 = time.perf_counter()
        p64.jk_noalloc(dD.address, dJ.address, dK.address)
        p64.sync()
        jk_dev.append((time.perf_counter() - t0) * 1e3)
    jk_dev_ms = float(np.median(jk_dev[1:]))
    assert np.abs(dJ.download().reshape(nao, nao) - J64).max() <= 1e-12 * max(1.0, np.abs(J64).max())
    for a in (dD, dJ, dK):
        a.free()
    log(f"J/K step on device-resident matrices: {jk_dev_ms:.1f} ms "
        f"({p64.ntuples / (jk_dev_ms / 1e3):.3e} quartets/s with their J/K)")
    # the mixed plan's J/K against the FP64 plan's (whose ERIs match the
    # oracle to the sampled error above)
    Jm, Km = pmx.jk(dens)
    jk_err = {"J_max_abs_err_mixed_vs_fp64": float(np.abs(Jm - J64).max()),
              "K_max_abs_err_mixed_vs_fp64": float(np.abs(Km - K64).max()),
              "J_max_abs": float(np.abs(J64).max()), "K_max_abs": float(np.abs(K64).max())}
    log(f"J/K mixed vs FP64: J err {jk_err['J_max_abs_err_mixed_vs_fp64']:.3e}, "
        f"K err {jk_err['K_max_abs_err_mixed_vs_fp64']:.3e}")

    prim64, roots64, flops64 = p64.work()
    log(f"fp64 work: {prim64:.3e} primitive quartets, {flops64:.3e} FLOP")
    step_s = ms64 / 1e3
    achieved_step = flops64 / step_s / 1e12
    bins, dom, achieved, kname, dom_grid = dominant(p64, args.steps, "double")
    _, dommx, achievedmx, knamemx, _ = dominant(pmx, args.steps, None)
    traffic = profiled_traffic(kname, dom["tuples"])
    value = pmx.ntuples / (msmx / 1e3)
    fp64_rate = p64.ntuples / step_s
    out_bytes = p64.nelem * 8

    learned = None
    if not args.no_learned:
        learned = learned_roots_report()
        log(f"learned roots: {learned['summary']}")

    secondary = None
    if not args.no_secondary:
        secondary = secondary_configs()
        for k, v in secondary.items():
            log(f"{k}: {v['per_s']:.3e} {v['unit']}, {v['ms']:.1f} ms, "
                f"{v['tflops']:.2f} TFLOP/s, max err {v['max_abs_err_vs_judge']:.2e}")

    raw = None
    if not args.no_secondary:
        raw = raw_path_rate()
        log(f"raw int2e_sph calls: {raw['calls_per_s_with_optimizer']:.0f}/s with an optimizer, "
            f"{raw['calls_per_s_without_optimizer']:.0f}/s without")

    cpu = None
    if not args.no_cpu:
        t64, _ = p64.tuples_cached()
        cpu = cpu_baseline(mol, t64, args.cpu_seconds)
        log(f"cpu baseline: {cpu['value']:.3e} quartets/s on {cpu['cores']} cores")

    res = {
        "metric": "int2e_sph quartets/s: mixed vs FP64 GPU at <=1e-10 abs err, 1 MI355X",
        "value": value,
        "unit": "quartets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": msmx,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": "int2e_sph, 256-atom 8x8x4 lattice (cfg-1 shells, 2819 sph functions), "
                        f"unique quartets, Schwarz 1e-12, bra stride {args.bra_stride}",
            "quartets_per_step": pmx.ntuples,
            "mixed_threshold": args.mixed_threshold,
            "fp32_quartets": pmx.ntuples_fp32,
        },
        "fp64": {"quartets_per_s": fp64_rate, "ms_per_step": ms64,
                 "wall_ms_per_step": wall64, "gflops": achieved_step * 1e3,
                 "prim_quartets": prim64, "root_evals": roots64},
        "mixed_over_fp64": value / fp64_rate,
        "jk": {"ms": jk_ms, "device_ms": jk_dev_ms,
               "quartets_per_s_with_jk": p64.ntuples / (jk_dev_ms / 1e3), "nbf": nao, **jk_err,
               "note": "J and K from the FP64 plan's unique quartets and a random symmetric "
                       "density, digested inside the class kernels (no ERI block written); "
                       "ms: host arrays (64 MB copies each way), device_ms: device-resident "
                       "D/J/K through cintx_plan_jk_noalloc, host-timed around a sync"},
        "fp64_max_abs_err_vs_judge": err64,
        "mixed_max_abs_err_vs_judge": errmx,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "traffic_note": "HBM bytes per launch of the dominant kernel: calibrated "
                                     "FETCH_SIZE + WRITE_SIZE per tuple from the committed "
                                     "rocprofv3 counter runs x this launch's tuples (null when "
                                     "the dominant kernel differs from the profiled one)",
                     "kernel": kname, "kernel_ms_per_launch": dom["ms"],
                     "kernel_grid_workitems": dom_grid,
                     "kernel_flops_per_launch": dom["flops"],
                     "kernel_tuples_per_launch": dom["tuples"],
                     "note": "FP64 vector-ALU bound class kernel of the FP64 plan (the only "
                             "matrix-core work is the cart->sph of the wide variants, 2-5% of "
                             "FLOPs); achieved = algorithmic FLOPs of the dominant class launch "
                             "(libcint loop trip counts) / its HIP-event duration; peak = "
                             "MI355X FP64 vector 78.6 TFLOP/s",
                     "step_achieved_tflops": achieved_step,
                     "output_GBps": out_bytes / step_s / 1e9},
        "roofline_mixed": {"bound": "valu", "achieved": achievedmx,
                           "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                           "frac": achievedmx / PEAK_FP64_TFLOPS, "kernel": knamemx,
                           "kernel_ms_per_launch": dommx["ms"],
                           "note": "dominant class launch of the mixed plan (the `value` "
                                   "plan); FP32 without packed math issues at the FP64 "
                                   "rate on CDNA4, so the same peak applies"},
        "fp64_bins": [{"class": "".join(map(str, b["class"])), "ms": round(b["ms"], 3),
                       "tflops": round(b["flops"] / max(b["ms"], 1e-9) / 1e9, 3),
                       "tuples": b["tuples"]} for b in bins],
        "cpu_baseline": cpu,
        "launch_shapes": tuned,
        "secondary": secondary,
        "raw_path": raw,
        "learned_roots": learned,
    }
    print(json.dumps(res), flush=True)


def secondary_configs():
    """Throughput of the other families on their own synthetic inputs (FP64,
    autotuned, graph replay): int3c2e_sph and int3c2e_sph mixed on the 64-atom
    lattice with the 6s5p4d3f1g aux basis (every orbital pair i >= j with
    every aux shell, 3.97e7 triples), int2e_sph (unique quartets, Schwarz 1e-12, bra stride 64)
    and int2e_ip1_sph (every 8th of those quartets) on the 32-atom s..g
    basis (runtime-class kernel for f and g classes).  Max abs error against
    the judge (the oracle's long-double build) on a 300-tuple sample."""
    from cintx_amd import capi, molecule
    from oracle import oracle as O
    out = {}

    def run(key, plan, family, mol, unit, reps=3, sample=300):
        log(f"secondary {key}: {plan.ntuples} tuples")
        plan.autotune(reps=1)
        plan.execute(graph=True)
        plan.sync()
        ms = []
        for _ in range(reps):
            plan.execute(graph=True)
            plan.sync()
            ms.append(plan.last_ms())
        _, _, flops = plan.work()
        t, _ = plan.tuples_cached()
        rng = np.random.default_rng(5)
        pick = np.sort(rng.choice(len(t), size=min(sample, len(t)), replace=False))
        got, _, _ = plan.output_tuples(pick)
        arity = {"int2c2e": 2, "int3c2e": 3}.get(family.split("_")[0], 4)
        sh = [tuple(int(v) for v in t[q][:arity]) for q in pick]
        if family.endswith("_spinor"):
            # the spinor oracle (extended precision on the judge's Cartesian
            # blocks), libcint's F-order block per tuple
            from oracle import spinor as S
            fn = S.int2e_spinor if arity == 4 else S.int3c2e_spinor
            got = got.view(np.complex128)
            ref = np.concatenate([np.asarray(fn(q, mol.atm, mol.bas, mol.env)[0]).transpose(
                tuple(range(arity))[::-1]).ravel() for q in sh])
        else:
            ref, _, _ = O.eval_list(family, sh, mol.atm, mol.bas, mol.env, ext=True)
        m = float(np.mean(ms))
        out[key] = {"family": family, "tuples": plan.ntuples, "ms": m,
                    "per_s": plan.ntuples / (m / 1e3), "unit": unit,
                    "tflops": flops / (m / 1e3) / 1e12,
                    "fp32_tuples": plan.ntuples_fp32,
                    "max_abs_err_vs_judge": float(np.abs(got - ref).max()),
                    # relative to each element's scale, |ref| floored at 1
                    "max_err_over_max_1_ref": float(
                        (np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)).max())}
        plan.close()

    log("secondary: building the aux-basis lattice plans")
    m4 = molecule.config(4)
    ii, jj = np.tril_indices(m4.aux_start)
    aux = np.arange(m4.aux_start, m4.nbas, dtype=np.int32)
    trip = np.zeros((len(ii) * len(aux), 4), dtype=np.int32)
    trip[:, 0] = np.repeat(ii, len(aux))
    trip[:, 1] = np.repeat(jj, len(aux))
    trip[:, 2] = np.tile(aux, len(ii))
    run("int3c2e_sph_aux_lattice", capi.Plan("int3c2e_sph", m4.atm, m4.bas, m4.env, trip),
        "int3c2e_sph", m4, "triples/s")
    # mixed: a triple takes the FP32 recurrences when its Schwarz bound over
    # the absolute coefficients is below 1e-9
    run("int3c2e_sph_aux_lattice_mixed",
        capi.Plan("int3c2e_sph", m4.atm, m4.bas, m4.env, trip, precision=capi.PREC_MIXED,
                  mixed_threshold=1e-9),
        "int3c2e_sph", m4, "triples/s")
    m5 = molecule.config(5)
    p5 = capi.Plan("int2e_sph", m5.atm, m5.bas, m5.env, screened=(1e-12, 64))
    t5, _ = p5.tuples_cached()
    t5 = np.array(t5[::8], dtype=np.int32)
    run("int2e_sph_sg_basis", p5, "int2e_sph", m5, "quartets/s")
    run("int2e_ip1_sph_sg_basis", capi.Plan("int2e_ip1_sph", m5.atm, m5.bas, m5.env, t5),
        "int2e_ip1_sph", m5, "quartets/s")
    # the other forms, each on a share of the same inputs: the unique
    # quartets of the 64-atom lattice (Schwarz 1e-12, bra stride 32) and the
    # aux-lattice triples
    m2 = molecule.config(2)
    p2 = capi.Plan("int2e_sph", m2.atm, m2.bas, m2.env, screened=(1e-12, 32))
    t2, _ = p2.tuples_cached()
    t2 = np.array(t2, dtype=np.int32)
    p2.close()
    run("int2e_cart_lattice64", capi.Plan("int2e_cart", m2.atm, m2.bas, m2.env, t2),
        "int2e_cart", m2, "quartets/s")
    run("int2e_ip2_sph_lattice64", capi.Plan("int2e_ip2_sph", m2.atm, m2.bas, m2.env, t2[::8]),
        "int2e_ip2_sph", m2, "quartets/s")
    run("int2e_spinor_lattice64",
        capi.Plan("int2e_spinor", m2.atm, m2.bas, m2.env, t2[::64]),
        "int2e_spinor", m2, "quartets/s", sample=40)
    run("int3c2e_cart_aux_lattice", capi.Plan("int3c2e_cart", m4.atm, m4.bas, m4.env, trip[::8]),
        "int3c2e_cart", m4, "triples/s")
    run("int3c2e_ip2_sph_aux_lattice",
        capi.Plan("int3c2e_ip2_sph", m4.atm, m4.bas, m4.env, trip[::32]),
        "int3c2e_ip2_sph", m4, "triples/s")
    run("int3c2e_spinor_aux_lattice",
        capi.Plan("int3c2e_spinor", m4.atm, m4.bas, m4.env, trip[::256]),
        "int3c2e_spinor", m4, "triples/s", sample=40)
    # the density-fitting metric: every (P|Q), P >= Q, of the aux basis
    pp, qq = np.tril_indices(len(aux))
    pq = np.zeros((len(pp), 4), dtype=np.int32)
    pq[:, 0] = aux[pp]
    pq[:, 1] = aux[qq]
    run("int2c2e_sph_aux_metric", capi.Plan("int2c2e_sph", m4.atm, m4.bas, m4.env, pq),
        "int2c2e_sph", m4, "pairs/s")
    return out


def raw_path_rate(ncalls=4000):
    """The libcint raw signature from a host caller: one int2e_sph call per
    quartet of the 64-atom lattice (host buffers: each call uploads its
    tuple, launches, waits and copies its block back), with a device
    CINTOpt and without (the call then builds its pair data itself); calls
    from Python through ctypes, one thread."""
    from cintx_amd import capi, molecule
    m2 = molecule.config(2)
    rng = np.random.default_rng(11)
    sh = rng.integers(0, m2.nbas, (ncalls, 4)).astype(np.int32)
    opt = capi.Optimizer("int2e_sph", m2.atm, m2.bas, m2.env)
    for q in sh[:50]:
        capi.raw_call("int2e_sph", q, m2.atm, m2.bas, m2.env, opt=opt)
    t0 = time.perf_counter()
    for q in sh:
        capi.raw_call("int2e_sph", q, m2.atm, m2.bas, m2.env, opt=opt)
    with_opt = ncalls / (time.perf_counter() - t0)
    n0 = ncalls // 8
    t0 = time.perf_counter()
    for q in sh[:n0]:
        capi.raw_call("int2e_sph", q, m2.atm, m2.bas, m2.env)
    without = n0 / (time.perf_counter() - t0)
    opt.close()
    return {"family": "int2e_sph", "calls": ncalls, "calls_per_s_with_optimizer": with_opt,
            "calls_per_s_without_optimizer": without,
            "note": "PCIe-inclusive host-buffer boundary, one quartet per call, Python ctypes "
                    "caller, one thread; random shell quartets of the 64-atom lattice"}


def profiled_traffic(kname, tuples):
    """HBM bytes per launch of `kname` from the counter summary in profiles/
    (calibrated FETCH_SIZE + WRITE_SIZE per tuple), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "int2e_sph_256atom_hbm_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_tuple "):
                name, per = line[len("traffic_bytes_per_tuple "):].rsplit(" ", 1)
                if name.strip() == kname:
                    return float(per) * tuples
    except OSError:
        pass
    return None


def learned_roots_report(nsamples=1 << 22, epochs=8):
    """Train the MLP initializer on the device (2^22 samples, x log-uniform in
    [1e-6, 1e3], n = 1..13) and compare, per n, two Newton steps started from
    the network with the same steps started from the asymptotic rule and
    from the table rule itself (the floor of the Newton polish: FP64 Newton
    on the moment equations cannot hold 1e-14 for n >= 4 even from exact
    roots), and time the table rule alone (the method without the network);
    the table rule judges every polished rule at 1e-14 relative."""
    from cintx_amd import rys_mlp
    from oracle import oracle as O
    t0 = time.perf_counter()
    params, loss = rys_mlp.train(nsamples=nsamples, epochs=epochs)
    train_s = time.perf_counter() - t0
    rng = np.random.default_rng(3)
    x = np.exp(rng.uniform(np.log(1e-6), np.log(1e3), 1 << 16))
    per_n = {}
    for n in range(1, 10):
        u, w, st, ms_mlp, e = rys_mlp.evaluate(params, n, x, newton_steps=2, init_mode=0,
                                               with_error=True)
        _, _, st0, _, e0 = rys_mlp.evaluate(params, n, x, newton_steps=2, init_mode=1,
                                            with_error=True)
        _, _, st3, _, e3 = rys_mlp.evaluate(params, n, x, newton_steps=2, init_mode=3,
                                            with_error=True)
        _, _, _, ms_tab = rys_mlp.evaluate(params, n, x, newton_steps=0, init_mode=2)
        err = 0.0
        for k in range(0, len(x), 2048):
            uo, wo = O.rys_roots(n, x[k])
            err = max(err, float(np.max(np.abs(u[k] - uo) / uo)),
                      float(np.max(np.abs(w[k] - wo) / wo)))
        per_n[n] = {"settled": float(st.mean()), "settled_without_network": float(st0.mean()),
                    "median_polish_err": float(np.median(e)),
                    "median_polish_err_without_network": float(np.median(e0)),
                    "settled_from_exact_start": float(st3.mean()),
                    "median_polish_err_from_exact_start": float(np.median(e3)),
                    "fallback_rate": float(1 - st.mean()), "max_rel_err_vs_oracle": err,
                    "ms_mlp_newton_judge": ms_mlp, "ms_table": ms_tab,
                    "table_speedup_over_mlp": ms_mlp / ms_tab if ms_tab > 0 else None}
    summ = ", ".join(f"n={n}: {v['settled']:.2f}/{v['settled_without_network']:.2f}/"
                     f"{v['settled_from_exact_start']:.2f} "
                     f"({v['median_polish_err']:.1e}/{v['median_polish_err_without_network']:.1e}/"
                     f"{v['median_polish_err_from_exact_start']:.1e})"
                     for n, v in per_n.items())
    return {"train_samples": nsamples, "epochs": epochs, "final_loss": loss,
            "train_s": train_s, "eval_points": len(x), "per_nroots": per_n,
            "summary": "settled fraction (median polish error) from network / asymptotic "
                       "rule / exact table roots: " + summ,
            "conclusion": "not wired into the ERI kernels: the table rule is exact and faster "
                          "than network + Newton + judge, and the moment-equation Newton "
                          "polish misses 1e-14 for n >= 4 even from the exact roots"}


def dominant(plan, steps, real):
    """Per-bin HIP-event times of `plan` (execute_timed, averaged over
    `steps`), the slowest bin, its algorithmic TFLOP/s and kernel name."""
    per = {}
    shapes = plan.bin_variants()
    for _ in range(steps):
        plan.execute_timed()
        for b, v in zip(plan.bin_stats(), shapes):
            per.setdefault((b["class"], b["precision"]), []).append(dict(b, **v))
    bins = []
    for (cls, prec), lst in per.items():
        bins.append({"class": cls, "precision": prec,
                     "ms": float(np.mean([b["ms"] for b in lst])),
                     "flops": lst[0]["flops"], "tuples": lst[0]["tuples"],
                     "prim_tuples": lst[0]["prim_tuples"], "generic": lst[0]["generic"],
                     "lanes": lst[0]["lanes"], "prims": lst[0]["prims"],
                     "variant": lst[0]["variant"],
                     "tuples_per_block": lst[0]["tuples_per_block"]})
    bins.sort(key=lambda b: -b["ms"])
    dom = bins[0]
    if dom["flops"] == 0:  # work counters not taken for this plan yet
        plan.work()
        st = {(b["class"], b["precision"]): b["flops"] for b in plan.bin_stats()}
        for b in bins:
            b["flops"] = st[(b["class"], b["precision"])]
    achieved = dom["flops"] / (dom["ms"] / 1e3) / 1e12
    from cintx_amd import capi
    r = real or ("float" if dom["precision"] == 1 and dom["variant"] < capi.NVARIANTS
                 else "double")
    kname = "eri_class_kernel<{}, {}, {}, {}, 0, {}, {}, {}>".format(*dom["class"], r,
                                                                     dom["lanes"], dom["prims"])
    q = dom["tuples_per_block"]
    grid = -(-dom["tuples"] // q) * q * dom["lanes"]  # work-items of the launch
    log(f"dominant kernel {kname}: {dom['ms']:.2f} ms/launch, {achieved:.3f} TFLOP/s, "
        f"grid {grid} work-items")
    return bins, dom, achieved, kname, grid


def calibrate_copy(nelem):
    """A plain device copy of nelem doubles, 3 times (torch's copy kernel):
    the known-size reference for FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.ones(nelem, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 3
    mb = nelem * 8 / 1e6
    print(f"copy of {nelem} doubles ({mb:.1f} MB read + {mb:.1f} MB written): {ms:.3f} ms, "
          f"{2 * nelem * 8 / ms / 1e6:.1f} GB/s", flush=True)


def sample_errors(mol, p64, pmx, n):
    """max abs error of each GPU path against the judge (the oracle's
    long-double build) on n seeded tuples (for the mixed plan, drawn from its
    FP32 bins when it has any)."""
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    out = []
    for p in (p64, pmx):
        t, _ = p.tuples_cached()
        pick = np.sort(rng.choice(len(t), size=min(n, len(t)), replace=False))
        if p is pmx and p.ntuples_fp32 > 0:
            # mixed plans keep the FP32 tuples after the FP64 ones: take half of
            # the sample from each part
            first32 = len(t) - p.ntuples_fp32
            a = rng.choice(first32, size=min(n // 2, first32), replace=False) if first32 else []
            b = first32 + rng.choice(p.ntuples_fp32, size=min(n // 2, p.ntuples_fp32),
                                     replace=False)
            pick = np.sort(np.concatenate([np.asarray(a, dtype=np.int64), b]))
        got, offs, sizes = p.output_tuples(pick)
        ref, roffs, rsizes = O.eval_list("int2e_sph", [tuple(int(v) for v in t[q]) for q in pick],
                                         mol.atm, mol.bas, mol.env, ext=True)
        out.append(float(np.abs(got - ref).max()))
    return out[0], out[1]


def cpu_baseline(mol, tuples, seconds):
    """oracle/cpu_simd.cpp (cintx's f64x4 CPU kernel restated, OpenMP
    schedule(dynamic, 2) over bra rows) on a bounded random sample of the same
    quartet list, sized so that 5 runs take about `seconds`; median of the 5."""
    from oracle import oracle as O
    ncores = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(0)
    pick = np.sort(rng.permutation(len(tuples))[:200000])
    sub = np.ascontiguousarray(np.asarray(tuples)[pick], dtype=np.int32)
    # size the sample on a probe run
    n = min(len(sub), 4000)
    t0 = time.perf_counter()
    O.cpu_baseline_int2e(sub[:n], mol.atm, mol.bas, mol.env, nthreads=ncores)
    per = (time.perf_counter() - t0) / n
    n = int(min(len(sub), max(2000, seconds / 5 / max(per, 1e-9))))
    sample = sub[:n]
    rates = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.cpu_baseline_int2e(sample, mol.atm, mol.bas, mol.env, nthreads=ncores)
        rates.append(n / (time.perf_counter() - t0))
    return {"value": float(np.median(rates)), "unit": "quartets/s", "cores": ncores,
            "kind": "port",
            "sample": f"{n} random quartets of the same screened set (bra-row order kept), "
                      "median of 5 runs; oracle/cpu_simd.cpp: cintx-simd's f64x4 two-electron "
                      "kernel restated (libcint recurrences, primitives unscreened, table Rys "
                      "rule, cart->sph), OpenMP schedule(dynamic,2) over bra rows",
            "runs_quartets_per_s": [float(r) for r in rates]}


if __name__ == "__main__":
    main()
