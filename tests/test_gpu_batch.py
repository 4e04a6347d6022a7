"""GPU: batched uneven Fock builds (qcjk_create_batch) and the captured-graph replay.

A batch handle holds many molecules; every build runs one launch per shell-quartet class over all
of them.  Checked against per-molecule handles (same kernels, 1e-12 relative), against the CPU
oracle for small members (the FP64 bar, 1e-10 relative per element, floor 1e-4 max|ref|), and the
hipGraph replay against eager builds on several replays with eager work between them.
All calls go through the C ABI (libqcjk.so) via qchem_amd; the oracle is the checker only.
"""
import numpy as np
import pytest

import oracle
import qchem_amd as qa
from qchem_amd import systems as S

from test_gpu_jk import TOL, rand_density, rel_err

pytestmark = pytest.mark.gpu


def _batch(nmol, seed):
    mols = S.batch(nmol, seed=seed, lo=2, hi=6)
    return mols, [qa.Basis(m["shells"]) for m in mols]


def test_batch_matches_single_and_oracle(require_gpu):
    mols, bases = _batch(12, 901)
    eng = qa.JKEngine(bases)
    assert eng.nbfs == [b.nbf for b in bases]
    Ds = [rand_density(b.nbf, 900 + i) for i, b in enumerate(bases)]
    Js, Ks, st = eng.build(Ds)
    total = 0
    for i, (m, b) in enumerate(zip(mols, bases)):
        single = qa.JKEngine(b)
        J1, K1, s1 = single.build(Ds[i])
        total += s1.quartets_computed
        assert rel_err(Js[i], J1) < 1e-12 and rel_err(Ks[i], K1) < 1e-12, i
        if i < 4:
            Jr, Kr, _ = oracle.Basis(m["shells"]).jk_direct(Ds[i])
            assert rel_err(Js[i], Jr) < TOL and rel_err(Ks[i], Kr) < TOL, i
    assert st.quartets_computed == total
    print("batch of %d: %d quartets, device %.3f ms" % (len(mols), st.quartets_computed, st.device_ms))


def test_batch_one_electron(require_gpu):
    mols, bases = _batch(5, 902)
    eng = qa.JKEngine(bases)
    for i, (m, b) in enumerate(zip(mols, bases)):
        S1, T1, V1 = eng.one_electron(m["Z"], m["R"], mol=i)
        S0, T0, V0 = qa.JKEngine(b).one_electron(m["Z"], m["R"])
        assert np.abs(S1 - S0).max() < 1e-14 and np.abs(T1 - T0).max() < 1e-12 and np.abs(V1 - V0).max() < 1e-12


def test_batch_mixed(require_gpu):
    mols, bases = _batch(10, 903)
    Ds = [rand_density(b.nbf, 910 + i) for i, b in enumerate(bases)]
    eng = qa.JKEngine(bases, mixed_tol=1e-7)
    J0, K0, _ = eng.build(Ds)
    J1, K1, st = eng.build(Ds, flags=qa.QCJK_J | qa.QCJK_K | qa.QCJK_MIXED)
    assert st.quartets_fp32 > 0
    for a, b in zip(J0 + K0, J1 + K1):
        assert np.abs(a - b).max() < 1e-9 * np.abs(a).max()


def test_graph_replay_matches_eager(require_gpu):
    import torch
    mols, bases = _batch(16, 904)
    eng = qa.JKEngine(bases)
    tune_table = {"2 1 1 0": 3, "2 2 0 0": 2}
    eng.apply_tuning(tune_table)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    D = torch.zeros(eng.nelem, dtype=torch.float64, device=dev)
    J = torch.zeros_like(D)
    K = torch.zeros_like(D)
    Je, Ke = torch.zeros_like(D), torch.zeros_like(D)
    for rep in range(4):
        Ds = [rand_density(b.nbf, 1000 * rep + i) for i, b in enumerate(bases)]
        D.copy_(torch.from_numpy(eng.pack(Ds)))
        eng.graph_build(D, J, K, stream=stream)
        eng.build_device(D, Je, Ke, stream=stream)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(torch.isfinite(J), torch.ones_like(J, dtype=torch.bool))
        dj = float((J - Je).abs().max() / Je.abs().max())
        dk = float((K - Ke).abs().max() / Ke.abs().max())
        print("replay %d: J %.2e K %.2e" % (rep, dj, dk))
        assert dj < 1e-12 and dk < 1e-12
    # a variant change re-captures
    eng.set_variant(2, 1, 1, 0, 0)
    eng.graph_build(D, J, K, stream=stream)
    torch.cuda.synchronize()
    assert float((J - Je).abs().max() / Je.abs().max()) < 1e-12


def test_batched_rhf_matches_single(require_gpu):
    """Batched RHF (one graph-replayed build per iteration for all molecules) gives each molecule
    the energy and iteration count of its own single-molecule RHF."""
    mols = S.batch(6, seed=905, lo=2, hi=4)
    bases = [qa.Basis(m["shells"]) for m in mols]
    out, summ = qa.rhf_batch(bases, mols, e_tol=1e-10, fd_tol=1e-7, max_iter=60)
    print("batched RHF: %d/%d converged, %d builds, %.3f s" % (summ["nconverged"], len(mols), summ["iterations"],
                                                              summ["t_total"]))
    both = 0
    for m, b, r in zip(mols, bases, out):
        single = qa.rhf(b, m["Z"], m["R"], e_tol=1e-10, fd_tol=1e-7, max_iter=60)
        if r["converged"] and single["converged"]:
            both += 1
            assert abs(r["energy"] - single["energy"]) < 1e-8
            assert r["iterations"] == single["iterations"]
        else:
            assert r["converged"] == single["converged"]
    assert both >= 3
