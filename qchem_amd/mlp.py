"""Python binding of the learned-guess network (include/qcml.h, libqcml.so).

The network's forward, backward and Adam kernels are the library's own HIP code; torch tensors are
used only as device memory.  There is no CPU path: without the built library or a GPU every call
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

QCML_SYMBOLS = ["qcml_create", "qcml_nparams", "qcml_get_params", "qcml_set_params", "qcml_forward",
                "qcml_train_step", "qcml_gradients", "qcml_gemm", "qcml_destroy"]


def lib_path():
    return os.path.join(_HERE, "libqcml.so")


def lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("qchem_amd: %s is not built (run __graft_entry__.build())" % path)
        L = ctypes.CDLL(path)
        vp, f32p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)
        L.qcml_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                  ctypes.c_int32, ctypes.POINTER(vp)]
        L.qcml_nparams.argtypes = [vp]
        L.qcml_nparams.restype = ctypes.c_int64
        L.qcml_get_params.argtypes = [vp, f32p]
        L.qcml_set_params.argtypes = [vp, f32p]
        L.qcml_forward.argtypes = [vp, vp, ctypes.c_int32, vp, vp]
        L.qcml_train_step.argtypes = [vp, vp, vp, ctypes.c_int32, ctypes.c_float, f32p, vp]
        L.qcml_gradients.argtypes = [vp, vp, vp, ctypes.c_int32, f32p, f32p]
        L.qcml_gemm.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, ctypes.c_int32, ctypes.c_int64,
                                ctypes.c_int64, vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, vp, ctypes.c_int64,
                                ctypes.c_float, ctypes.c_float, vp]
        L.qcml_destroy.argtypes = [vp]
        _LIB = L
    return _LIB


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ptr(x):
    return x if isinstance(x, int) else x.data_ptr()


def _stream(stream):
    """The HIP stream to run on: the given torch stream, else torch's current stream, so the
    kernels are ordered after the torch work that produced their inputs."""
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return stream if isinstance(stream, int) else stream.cuda_stream


class NetError(RuntimeError):
    def __init__(self, where, code):
        super().__init__("%s failed with status %d" % (where, code))
        self.code = code


class Net:
    """3-layer perceptron with BF16 weights / FP32 accumulation on the GPU (qcml_*)."""

    def __init__(self, nin, width, nout, max_batch, seed=1, device=0):
        self.dims = (nin, width, nout)
        self.max_batch = max_batch
        h = ctypes.c_void_p()
        rc = lib().qcml_create(nin, width, nout, max_batch, seed, device, ctypes.byref(h))
        if rc != 0:
            raise NetError("qcml_create", rc)
        self.h = h
        self.nparams = int(lib().qcml_nparams(h))

    def close(self):
        if getattr(self, "h", None):
            lib().qcml_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        rc = lib().qcml_get_params(self.h, _f32p(p))
        if rc != 0:
            raise NetError("qcml_get_params", rc)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        rc = lib().qcml_set_params(self.h, _f32p(p))
        if rc != 0:
            raise NetError("qcml_set_params", rc)

    def forward(self, X, Y, stream=None):
        """X (B x nin) and Y (B x nout): float32 CUDA tensors."""
        rc = lib().qcml_forward(self.h, _ptr(X), X.shape[0], _ptr(Y), _stream(stream))
        if rc != 0:
            raise NetError("qcml_forward", rc)

    def train_step(self, X, T, lr, want_loss=False, stream=None):
        loss = ctypes.c_float(0.0)
        rc = lib().qcml_train_step(self.h, _ptr(X), _ptr(T), X.shape[0], float(lr),
                                   ctypes.byref(loss) if want_loss else None, _stream(stream))
        if rc != 0:
            raise NetError("qcml_train_step", rc)
        return loss.value if want_loss else None

    def gradients(self, X, T):
        import torch
        torch.cuda.synchronize()  # inputs complete before the default-stream gradient pass reads them
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        rc = lib().qcml_gradients(self.h, _ptr(X), _ptr(T), X.shape[0], _f32p(g), ctypes.byref(loss))
        if rc != 0:
            raise NetError("qcml_gradients", rc)
        return loss.value, g


def gemm(A, B, C, transA=False, transB=False, alpha=1.0, beta=0.0, stream=None):
    """C = alpha op(A) op(B) + beta C on the library's MFMA kernel.  A, B: float32 or int16 (raw BF16)
    CUDA tensors, 2-D contiguous; C float32."""
    M, K = (A.shape[1], A.shape[0]) if transA else A.shape
    Kb, N = (B.shape[1], B.shape[0]) if transB else B.shape
    assert K == Kb and tuple(C.shape) == (M, N)
    lda, ldb = A.shape[1], B.shape[1]
    sam, sak = (1, lda) if transA else (lda, 1)
    sbk, sbn = (1, ldb) if transB else (ldb, 1)
    abf = 1 if A.element_size() == 2 else 0
    bbf = 1 if B.element_size() == 2 else 0
    rc = lib().qcml_gemm(M, N, K, _ptr(A), abf, sam, sak, _ptr(B), bbf, sbk, sbn, _ptr(C), N, float(alpha),
                         float(beta), _stream(stream))
    if rc != 0:
        raise NetError("qcml_gemm", rc)
