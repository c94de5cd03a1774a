"""ctypes driver for the MEX gateways (sssi_amd/mex/*.c) compiled against the
test double of the MEX API in this folder.

call(name, *args, nargout=2) converts numpy arrays (column-major doubles),
Python numbers and dicts (1x1 structs) to mock mxArrays, runs the gateway's
mexFunction and returns the outputs as numpy arrays; a gateway error raises
MexError(identifier, message)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(HERE, "_build")
GATEWAYS = ["fwdTimeCpmlFor2dAw_mex", "rvsTimeCpmlFor2dAw_mex", "fwdTimeCpmlFor3dAw_mex",
            "fwdTimeSpongeFor2dAw_mex", "rvsTimeSpongeFor2dAw_mex", "eikonal2d_mex",
            "migrate_mex", "freqCpmlFor2dAw_mex"]


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident
        self.msg = msg


def build():
    """Compile every gateway (make, a few seconds; needs sssi_amd/libsssi.so)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def _lib(name):
    if name not in _libs:
        path = os.path.join(BUILD, name + ".so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.mock_double.restype = ctypes.c_void_p
        L.mock_double.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), ctypes.c_void_p]
        L.mock_int32.restype = ctypes.c_void_p
        L.mock_struct.restype = ctypes.c_void_p
        L.mock_struct.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                  ctypes.POINTER(ctypes.c_void_p)]
        L.mock_free.argtypes = [ctypes.c_void_p]
        L.mock_invoke.restype = ctypes.c_int
        L.mock_invoke.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p,
                                  ctypes.c_char_p]
        L.mock_complex.restype = ctypes.c_void_p
        L.mock_complex.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                   ctypes.c_void_p]
        L.mock_nnz.restype = ctypes.c_longlong
        L.mock_nnz.argtypes = [ctypes.c_void_p]
        L.mock_kind.restype = ctypes.c_int
        L.mock_kind.argtypes = [ctypes.c_void_p]
        L.mock_sparse.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_double)]
        L.mock_data.restype = ctypes.POINTER(ctypes.c_double)
        L.mock_data.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_longlong)]
        _libs[name] = L
    return _libs[name]


class Int32Scalar:
    """Stands for a non-double argument (class check)."""


def _to_mx(L, x):
    if isinstance(x, dict):
        names = (ctypes.c_char_p * len(x))(*[k.encode() for k in x])
        vals = (ctypes.c_void_p * len(x))(*[_to_mx(L, v) for v in x.values()])
        return L.mock_struct(len(x), names, vals)
    if isinstance(x, Int32Scalar):
        return L.mock_int32()
    if np.iscomplexobj(x):
        c = np.asarray(x, dtype=np.complex128)
        c = c.reshape(-1, 1) if c.ndim == 1 else c
        c = np.atleast_2d(c)
        flat = np.ascontiguousarray(c.ravel(order="F"))
        dims = (ctypes.c_longlong * c.ndim)(*c.shape)
        return L.mock_complex(c.ndim, dims, flat.ctypes.data)
    a = np.asfortranarray(np.atleast_2d(np.asarray(x, dtype=np.float64))
                          if np.ndim(x) < 2 else np.asarray(x, dtype=np.float64))
    if np.ndim(x) == 1:
        a = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(-1, 1))
    dims = (ctypes.c_longlong * a.ndim)(*a.shape)
    return L.mock_double(a.ndim, dims, a.ctypes.data)


def _from_mx(L, p):
    nd = ctypes.c_int()
    dims = (ctypes.c_longlong * 8)()
    pr = L.mock_data(p, ctypes.byref(nd), dims)
    shape = tuple(dims[k] for k in range(nd.value))
    kind = L.mock_kind(p)
    cplx = bool(kind & 1)
    if kind & 2:  # sparse: compressed columns -> scipy CSC
        import scipy.sparse as sp
        m, ncol = shape
        nnz = int(L.mock_nnz(p))
        jc = np.zeros(ncol + 1, dtype=np.int64)
        ir = np.zeros(max(1, nnz), dtype=np.int64)
        val = np.zeros(max(1, 2 * nnz if cplx else nnz))
        L.mock_sparse(p, jc.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                      ir.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                      val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        data = val[:2 * nnz].view(np.complex128) if cplx else val[:nnz]
        return sp.csc_matrix((data.copy(), ir[:nnz].copy(), jc), shape=(m, ncol))
    n = int(np.prod(shape))
    if cplx:
        flat = (np.ctypeslib.as_array(pr, shape=(2 * n,)).copy().view(np.complex128)
                if n else np.zeros(0, dtype=np.complex128))
    else:
        flat = np.ctypeslib.as_array(pr, shape=(n,)).copy() if n else np.zeros(0)
    return flat.reshape(shape, order="F")


def call(name, *args, nargout=2):
    L = _lib(name)
    prhs = [_to_mx(L, a) for a in args]
    parr = (ctypes.c_void_p * max(len(prhs), 1))(*prhs)
    plhs = (ctypes.c_void_p * max(nargout, 1))()
    eid = ctypes.create_string_buffer(128)
    emsg = ctypes.create_string_buffer(512)
    try:
        rc = L.mock_invoke(nargout, plhs, len(prhs), parr, eid, emsg)
        outs = [_from_mx(L, plhs[k]) if plhs[k] else None for k in range(nargout)]
    finally:
        for k in range(nargout):
            if plhs[k]:
                L.mock_free(plhs[k])
        for p in prhs:
            L.mock_free(p)
    if rc:
        raise MexError(eid.value.decode(), emsg.value.decode())
    return outs
