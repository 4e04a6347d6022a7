#!/usr/bin/env python3
"""Benchmark: FP64 Fock build (J+K) of the 16-atom random cluster on one MI355X.

A "step" is one full J+K build with the density, J and K resident in HBM.  The workload is the
seeded 16-atom random cluster: heavy atoms 3s2p1d and H 2s1p, each shell a 3-primitive
contraction.  Odd electron counts are evened by removing one H, which leaves 15 atoms here.

``value`` is ERI quartets/s: unique contracted shell quartets (8-fold symmetry) computed per
second, counted over the timed steps.  The JSON line also carries:

* ``roofline``: the dominant class kernel.  Its algorithmic FLOPs per launch are divided by its
  average launch time, taken from HIP events on the engine stream over the timed region.
* ``cpu_baseline``: the CPU oracle's direct J/K, timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
METRIC = "Fock-build s & ERI quartets/s; SCF time-to-1e-6 Eh: mixed+learned vs FP64 GPU"
VARIANT_KERNEL = {0: "light", 1: "heavy", 2: "heavy2", 3: "lightc", 4: "heavy2", 5: "heavy2", 6: "heavy2", 7: "heavy2",
                  8: "heavy2", 9: "heavy2", 10: "heavy2"}


def cpu_baseline(mol, D, budget_s=15.0, threads=None):
    """Oracle direct J/K (OpenMP) on the first bra shell pairs of the same workload, scaled to quartets/s."""
    import oracle

    threads = threads or min(16, os.cpu_count() or 1)
    ob = oracle.Basis(mol["shells"])
    # calibrate on a small prefix, then size the sample to ~budget_s
    limit = 64
    t0 = time.perf_counter()
    _, _, cnt = ob.jk_direct(D, nthreads=threads, pair_limit=limit)
    dt = time.perf_counter() - t0
    ns = len(mol["shells"])
    npairs = ns * (ns + 1) // 2
    rate_pairs = max(dt, 1e-3) / limit
    # quartets grow ~ pair index, so scale the prefix length by sqrt of the time ratio
    limit2 = int(min(npairs, limit * max(1.0, (budget_s / max(dt, 1e-3)) ** 0.5)))
    t0 = time.perf_counter()
    _, _, cnt = ob.jk_direct(D, nthreads=threads, pair_limit=limit2)
    dt = time.perf_counter() - t0
    del rate_pairs
    return {"value": cnt / dt, "unit": "quartets/s", "cores": threads, "kind": "port",
            "sample": "oracle direct J+K (8-fold, FP64, OpenMP) over the first %d of %d bra shell pairs "
                      "(%d contracted quartets, %.1f s)" % (limit2, npairs, cnt, dt)}


def profiled_traffic(kernel_prefix):
    """HBM bytes per dispatch of the kernel (FETCH_SIZE + WRITE_SIZE, copy-calibrated) from the
    committed rocprofv3 summary of this workload, or None when it has no entry for the kernel."""
    path = os.path.join(ROOT, "profiles", "jk_build_random_cluster_fp64.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("TRAFFIC_JSON:"):
            table = json.loads(line.split(":", 1)[1])
            hits = [v for k, v in table.items() if k.startswith(kernel_prefix)]
            return hits[0] if hits else None
    return None


def time_to(e_hist, t_hist, e_ref, tol=1e-6):
    for e, t in zip(e_hist, t_hist):
        if abs(e - e_ref) < tol:
            return float(t)
    return None


def scf_gpu(qa, mol, **kw):
    """GPU RHF (C++ driver through the HF faces) converged tightly; time-to-1e-6 against itself."""
    B = qa.Basis(mol["shells"])
    res = qa.rhf(B, mol["Z"], mol["R"], e_tol=1e-10, fd_tol=1e-7, **kw)
    return {"energy": res["energy"], "iterations": res["iterations"], "converged": res["converged"],
            "time_s": res["t_total"], "jk_s": res["t_jk"], "diag_s": res["t_diag"],
            "time_to_1e-6_s": time_to(res["energy_history"], res["time_history"], res["energy"])}


def scf_cpu_reference(mol, e_conv):
    """The reference algorithm on the CPU: ERI4 materialised once, MatMul + DIIS per iteration."""
    import oracle
    import oracle.scf as oscf

    t0 = time.perf_counter()
    ob = oracle.Basis(mol["shells"])
    e4 = oracle.ERI4(ob)
    t_eri4 = time.perf_counter() - t0
    res = oscf.rhf(ob, mol["Z"], mol["R"], e4.jk, e_tol=1e-10, fd_tol=1e-7)
    t1 = time_to(res["history"], res["times"], e_conv)
    return {"energy": res["energy"], "iterations": res["iterations"], "cores": 1,
            "eri4_build_s": t_eri4, "time_to_1e-6_s": None if t1 is None else t_eri4 + t1,
            "delta_e_vs_gpu": res["energy"] - e_conv}


def scf_time_to_solution(qa, S, geometries=tuple(range(100000, 100008)), mixed_tol=1e-7):
    """Held-out C3H6O-like geometries: time to |E - E_conv| < 1e-6 Eh for the plain FP64 GPU path
    (SAD seed) and for mixed precision + the learned seed; E_conv from the FP64 run."""
    from qchem_amd import guess as G
    if not os.path.exists(G.DEFAULT_MODEL):
        return {"skipped": "no trained model at %s" % os.path.relpath(G.DEFAULT_MODEL, ROOT)}
    model = G.LearnedGuess.load()
    rows = G.evaluate(model, list(geometries), log=lambda *_: None, mixed_tol=mixed_tol)
    summ = G.summarize(rows)
    summ["geometries"] = list(geometries)
    summ["first_energy_error_learned_eh"] = [r["learned_fp64"]["first_energy_error"] for r in rows]
    return summ


def mixed_precision_chain(qa, S, mixed_tol=1e-7):
    """C16H32-like chain (784 functions): FP64 vs mixed J/K for one density, errors and build times."""
    import numpy as np
    import torch
    m = S.alkane_chain()
    B = qa.Basis(m["shells"])
    eng = qa.JKEngine(B, mixed_tol=mixed_tol)
    from qchem_amd import tune
    tune.apply(eng, workload="random-cluster-16")
    n = B.nbf
    rng = np.random.default_rng(11)
    C = rng.normal(size=(n, sum(m["Z"]) // 2)) / np.sqrt(n)
    D = torch.from_numpy(2.0 * C @ C.T).cuda()
    out = {"workload": m["name"], "n_basis": n, "mixed_tol": mixed_tol}
    res = {}
    for name, fl in (("fp64", qa.QCJK_J | qa.QCJK_K), ("mixed", qa.QCJK_J | qa.QCJK_K | qa.QCJK_MIXED)):
        J, K = torch.empty_like(D), torch.empty_like(D)
        st = eng.build_device(D, J, K, fl, stats=True)
        res[name] = (J, K, st)
        out[name + "_build_s"] = st.device_ms / 1e3
        out[name + "_quartets"] = st.quartets_computed
    J0, K0, _ = res["fp64"]
    J1, K1, st1 = res["mixed"]
    out["fp32_fraction"] = st1.quartets_fp32 / max(1, st1.quartets_computed)
    out["J_max_rel_err"] = float((J1 - J0).abs().max() / J0.abs().max())
    out["K_max_rel_err"] = float((K1 - K0).abs().max() / K0.abs().max())
    out["delta_E_two_electron_eh"] = float((0.5 * (D * (J1 - J0)).sum() - 0.25 * (D * (K1 - K0)).sum()).item())
    return out


def batched_scf(qa, S, nmol=512, max_iter=60):
    """Config 4's SCF: 512 molecules, batched RHF from the SAD seed, one graph-replayed build per
    iteration for all of them.  Converged molecules stop updating; the loop runs until all converge
    or max_iter.  Progress goes to stderr, one line per batched iteration."""
    mols = S.batch(nmol)
    bases = [qa.Basis(m["shells"]) for m in mols]
    t0 = time.perf_counter()
    out, summ = qa.rhf_batch(bases, mols, e_tol=1e-8, fd_tol=1e-5, max_iter=max_iter, guess=qa.GUESS_SAD,
                             verbose=True)
    wall = time.perf_counter() - t0
    its = sorted(r["iterations"] for r in out if r["converged"])
    return {"molecules": nmol, "converged": summ["nconverged"], "batched_iterations": summ["iterations"],
            "wall_s": wall, "per_molecule_s": wall / nmol, "jk_s": summ["t_jk"], "host_s": summ["t_host"],
            "setup_s": summ["t_setup"], "quartets": summ["quartets"],
            "median_iterations_converged": its[len(its) // 2] if its else None,
            "quartets_per_s_in_scf": summ["quartets"] / max(summ["t_jk"], 1e-9)}


def batched_builds(qa, S, nmol=512, reps=3):
    """512 molecules of 2-12 atoms in one handle: one launch per class per build, graph-replayed."""
    import numpy as np
    import torch
    mols = S.batch(nmol)
    bases = [qa.Basis(m["shells"]) for m in mols]
    t0 = time.perf_counter()
    eng = qa.JKEngine(bases)
    t_create = time.perf_counter() - t0
    rng = np.random.default_rng(5)
    Ds = []
    for b, m in zip(bases, mols):
        C = rng.normal(size=(b.nbf, max(1, sum(m["Z"]) // 2))) / np.sqrt(b.nbf)
        Ds.append(2.0 * C @ C.T)
    D = torch.from_numpy(eng.pack(Ds)).cuda()
    J, K = torch.empty_like(D), torch.empty_like(D)
    st = eng.build_device(D, J, K, stats=True)
    eng.graph_build(D, J, K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.graph_build(D, J, K)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"molecules": nmol, "n_basis_range": [min(eng.nbfs), max(eng.nbfs)], "n_basis_total": sum(eng.nbfs),
            "create_s": t_create, "quartets_per_build": st.quartets_computed, "build_s": dt,
            "quartets_per_s": st.quartets_computed / dt, "per_molecule_build_ms": 1e3 * dt / nmol}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF, mixed-precision and batched sections")
    ap.add_argument("--no-tune", action="store_true", help="use the untuned default kernel variant per class")
    ap.add_argument("--retune", action="store_true", help="re-run the per-class kernel-variant autotuner")
    ap.add_argument("--tune-out", default=os.path.join(ROOT, "tuning_retuned.json"),
                    help="where --retune writes the new per-class table")
    ap.add_argument("--batch-scf", action="store_true", help="also run the 512-molecule batched RHF (minutes)")
    ap.add_argument("--streams", type=int, default=0,
                    help="concurrent class streams per build (1-4; 0 = the engine's default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a copy kernel of known size (HBM counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 64 * 1024 * 1024  # 512 MiB of FP64 read, 512 MiB written per launch
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy": {"bytes_read": n * 8, "bytes_written": n * 8, "launches": 3}}))
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; the multi-GPU path is design-only (--gpus %d requested)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    torch.zeros(1, device="cuda:0")  # torch's HIP runtime up first: it is the device-memory plumbing below
    import qchem_amd as qa
    from qchem_amd import systems as S

    mol = S.random_cluster()
    B = qa.Basis(mol["shells"])
    eng = qa.JKEngine(B, device=0)
    n = B.nbf
    rng = np.random.default_rng(7)
    C = rng.normal(size=(n, sum(mol["Z"]) // 2)) / np.sqrt(n)
    Dh = 2.0 * C @ C.T
    dev = torch.device("cuda:0")
    D = torch.from_numpy(Dh).to(dev)
    J = torch.empty_like(D)
    K = torch.empty_like(D)
    stream = torch.cuda.current_stream(dev)

    flags = qa.QCJK_J | qa.QCJK_K
    tuning = None
    from qchem_amd import tune
    workload_key = "random-cluster-16"
    if args.retune:  # untimed setup: time every kernel variant per class on this workload
        rep = tune.tune(eng, D, reps=2, cache=args.tune_out, workload=workload_key)
        tuning = {"source": "autotuned now", "classes": {"%d%d%d%d" % (r["la"], r["lb"], r["lc"], r["ld"]): [
            r["wavefront_per_quartet"], round(r["device_ms"], 3), round(r["alt_ms"], 3)] for r in rep}}
    elif not args.no_tune:  # the committed per-class choices for this workload on gfx950
        table = tune.apply(eng, workload=workload_key)
        tuning = {"source": "qchem_amd/tuning/gfx950.json", "classes": table}
    if args.streams:
        eng.set_streams(args.streams)
    st = eng.build_device(D, J, K, flags, stream=stream, stats=True)  # counts + first build
    for _ in range(args.warmup):
        eng.build_device(D, J, K, flags, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.build_device(D, J, K, flags, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms = 1e3 * elapsed / args.steps
    # Per-class kernel durations: the same K builds again with the classes one after another on the
    # build stream (under the default overlap a class's events would include the time it shares
    # the device with its neighbours), HIP events around every class launch.
    nstreams = eng.streams()
    eng.set_streams(1)
    eng.profile(True)
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for _ in range(args.steps):
        eng.build_device(D, J, K, flags, stream=stream)
    torch.cuda.synchronize()
    ms_serial = 1e3 * (time.perf_counter() - s0) / args.steps
    rep = eng.class_report()
    eng.profile(False)
    eng.set_streams(nstreams)
    quartets = st.quartets_computed
    value = quartets * args.steps / elapsed

    dom = max(rep, key=lambda r: r["device_ms"])
    launches = max(dom["launches"], 1)
    per_launch_flops = dom["flops"] / launches
    per_launch_s = dom["device_ms"] / launches / 1e3
    achieved = per_launch_flops / per_launch_s / 1e12
    total_flops = sum(r["flops"] for r in rep) / max(1, args.steps)
    kname = "k_jk_%s<%d, %d, %d, %d," % (VARIANT_KERNEL[dom["wavefront_per_quartet"]], dom["la"], dom["lb"],
                                          dom["lc"], dom["ld"])
    roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": profiled_traffic(kname),
            "kernel": kname.rstrip(",") + {4: ", factorised bra>", 5: ", small LDS>",
                                           6: ", factorised bra, small LDS>", 7: ", factorised bra and ket>",
                                           8: ", factorised bra and ket, small LDS>", 9: ", lane R>",
                                           10: ", lane R, narrow batches>"}.get(
                                               dom["wavefront_per_quartet"], ">"),
            "kernel_ms_per_launch": per_launch_s * 1e3, "kernel_flops_per_launch": per_launch_flops,
            "build_tflops": total_flops / (ms / 1e3) / 1e12,
            "measured": "HIP events around each class launch, %d builds with the classes in sequence on one "
                        "stream (%.2f ms per build)" % (args.steps, ms_serial)}

    cpu = None if args.no_cpu else cpu_baseline(mol, Dh)
    scf = mixed = batched = None
    if not args.no_scf:
        n2 = S.n2_even_tempered()
        scf = {"n2_even_tempered_fp64_gpu_core_guess": scf_gpu(qa, n2),
               "c3h6o_time_to_1e-6": scf_time_to_solution(qa, S)}
        if cpu is not None:
            cpu["scf_n2_reference_algorithm"] = scf_cpu_reference(
                n2, scf["n2_even_tempered_fp64_gpu_core_guess"]["energy"])
        mixed = mixed_precision_chain(qa, S)
        batched = batched_builds(qa, S)
        if args.batch_scf:
            batched["scf"] = batched_scf(qa, S)
    out = {"metric": METRIC, "value": value, "unit": "quartets/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded random cluster, random PSD density)",
           "config": {"workload": "FP64 J+K Fock build, %s, n=%d basis functions, %d shell quartets (8-fold)"
                                  % (mol["name"], n, quartets),
                      "n_basis": n, "quartets_per_build": quartets, "fock_build_s": ms / 1e3,
                      "class_streams": nstreams, "parallelism": "single GPU"},
           "roofline": roof, "cpu_baseline": cpu, "scf": scf, "mixed_precision": mixed, "batched": batched,
           "tuning": tuning,
           "classes": [{k: r[k] for k in ("la", "lb", "lc", "ld", "wavefront_per_quartet", "quartets")} |
                       {"ms": r["device_ms"] / max(1, r["launches"]),
                        "tflops": r["flops"] / max(1e-12, r["device_ms"] / 1e3) / 1e12} for r in rep]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
