"""CPU: the C-ABI libraries load, export every declared symbol, and the host conventions hold.

No compute call reaches the device here; on a machine without a GPU the engine must refuse to
start (QCJK_ERR_NODEVICE) rather than fall back to the CPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import qchem_amd as qa
from qchem_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(qc(?:jk|host|scf|ml)_[a-z_]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    lib = ctypes.CDLL(qa.lib_paths()[0])
    names = _declared("qcjk.h")
    assert names and set(names) <= set(qa.QCJK_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n


def test_host_exports_every_declared_symbol():
    qa.lib()
    lib = ctypes.CDLL(qa.lib_paths()[1])
    names = _declared("qchem_host.h")
    assert names and set(names) <= set(qa.HOST_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n


def test_network_exports_every_declared_symbol():
    from qchem_amd import mlp
    lib = ctypes.CDLL(mlp.lib_path())
    names = _declared("qcml.h")
    assert names and set(names) <= set(mlp.QCML_SYMBOLS)
    for n in names:
        assert hasattr(lib, n), n


def test_network_needs_a_device():
    from conftest import gpu_available
    from qchem_amd import mlp
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(mlp.NetError) as e:
        mlp.Net(4, 8, 2, max_batch=4)
    assert e.value.code == 5


def test_strerror_codes():
    L = qa.lib()
    assert L.qcjk_strerror(0) == b"ok"
    assert L.qcjk_strerror(5) == b"no usable HIP device"


def test_host_basis_conventions_match_oracle():
    for mol in (S.n2_even_tempered(), S.random_cluster(6, seed=99), S.acetone_like(0)):
        b = qa.Basis(mol["shells"])
        ob = oracle.Basis(mol["shells"])
        assert b.nbf == ob.nbf
        assert np.allclose(b.ns, ob.ns(), rtol=1e-13, atol=0)
        assert np.allclose(b.folded, ob.folded_coefs(), rtol=1e-14, atol=0)


def test_component_order_is_reference_order():
    # m outer, l inner, n = L-m-l  (PG_Cart/Imp/IrrepBasisSet.C:46-54): d = xx xy yy xz yz zz
    assert [oracle.cart_comp(2, k) for k in range(6)] == [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)]


def test_workload_sizes():
    assert S.nbf(S.n2_even_tempered()["shells"]) == 70
    assert S.nbf(S.alkane_chain()["shells"]) == 784
    assert S.nbf(S.acetone_like(3)["shells"]) == 182
    c = S.random_cluster()
    assert sum(c["Z"]) % 2 == 0
    assert 150 <= S.nbf(c["shells"]) <= 260
    sizes = [S.nbf(m["shells"]) for m in S.batch(64)]
    assert min(sizes) >= 5 and max(sizes) <= 200


def test_no_cpu_fallback_without_gpu():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    b = qa.Basis(S.n2_even_tempered()["shells"])
    with pytest.raises(qa.EngineError) as e:
        qa.JKEngine(b)
    assert e.value.code == 5
    with pytest.raises(qa.EngineError):
        qa.rhf(b, [7, 7], [[0, 0, -1.037], [0, 0, 1.037]])


def test_rejects_unsupported_shells():
    bad = [dict(L=3, exps=[1.0], coefs=[1.0], center=[0, 0, 0], contracted=False)]
    b = qa.Basis(bad)
    with pytest.raises(qa.EngineError) as e:
        qa.JKEngine(b)
    assert e.value.code in (2, 5)
