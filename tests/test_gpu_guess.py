"""GPU: the committed learned seed (qchem_amd/guess_data/c3h6o_mlp.npz) on held-out C3H6O-like
geometries, against the method without it (the SAD seed).

* The seed is a valid closed-shell density: symmetric, Tr(DS) = N, idempotent in the S metric.
* Its first-iteration energy is within 0.05 Eh of the converged energy (SAD: ~1.7 Eh off).
* The SCF from it converges to the same energy as from SAD (1e-8 Eh) in fewer iterations to 1e-6 Eh.
The network's forward pass is the library's own kernels (libqcml.so); the converged FP64 SCF judges it.
"""
import numpy as np
import pytest

import qchem_amd as qa
from qchem_amd import guess as G
from qchem_amd import systems as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("index", [100000, 100007])
def test_learned_seed_gains_over_sad(require_gpu, index):
    import torch
    torch.zeros(1, device="cuda")
    model = G.LearnedGuess.load()
    m = S.acetone_like(index)
    B = qa.Basis(m["shells"])
    D0 = model.density(m)
    _, Smat = G.overlap(m)
    nel = sum(m["Z"])
    assert np.abs(D0 - D0.T).max() < 1e-12
    assert abs(np.trace(D0 @ Smat) - nel) < 1e-8
    assert np.abs(D0 @ Smat @ D0 - 2.0 * D0).max() < 1e-8
    tight = dict(e_tol=1e-10, fd_tol=1e-7, max_iter=150)
    sad = qa.rhf(B, m["Z"], m["R"], guess=qa.GUESS_SAD, **tight)
    lrn = qa.rhf(B, m["Z"], m["R"], D0=D0, **tight)
    e = sad["energy"]
    it_sad, it_lrn = G.iters_to(sad["energy_history"], e), G.iters_to(lrn["energy_history"], e)
    print("geometry %d: first-iteration error SAD %.3e learned %.3e; iterations to 1e-6 Eh SAD %d learned %d" % (
        index, sad["energy_history"][0] - e, lrn["energy_history"][0] - e, it_sad, it_lrn))
    assert sad["converged"] and lrn["converged"]
    assert abs(lrn["energy"] - e) < 1e-8
    assert abs(lrn["energy_history"][0] - e) < 0.05
    assert it_lrn < it_sad
