"""ctypes binding of the CPU oracle (oracle/_build/liboracle_h2eri.so).

Test infrastructure only: the product library never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle_h2eri.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    lib = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    lib.orc_load.restype = C.c_void_p
    lib.orc_load.argtypes = [C.c_char_p]
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_size.restype = C.c_int
    lib.orc_size.argtypes = [C.c_void_p]
    lib.orc_nbf.restype = C.c_int
    lib.orc_nbf.argtypes = [C.c_void_p]
    lib.orc_nshell.restype = C.c_int
    lib.orc_nshell.argtypes = [C.c_void_p]
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_matvec.argtypes = [C.c_void_p, dp, dp]
    lib.orc_build_coulomb.argtypes = [C.c_void_p, dp, dp]
    lib.orc_dense_matrix.argtypes = [C.c_void_p, dp]
    if hasattr(lib, "orc_build_exchange"):
        lib.orc_build_exchange.argtypes = [C.c_void_p, dp, dp]
        lib.orc_dlist.restype = C.c_int
        lib.orc_dlist.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, rep_path):
        self.lib = load()
        self.r = self.lib.orc_load(os.fsencode(rep_path))
        if not self.r:
            raise RuntimeError(f"oracle could not read {rep_path}")
        self.n = self.lib.orc_size(self.r)
        self.nbf = self.lib.orc_nbf(self.r)
        self.nshell = self.lib.orc_nshell(self.r)

    def close(self):
        if self.r:
            self.lib.orc_free(self.r)
            self.r = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        self.lib.orc_matvec(self.r, _p(x), _p(y))
        return y

    def dense(self):
        A = np.empty((self.n, self.n))
        self.lib.orc_dense_matrix(self.r, _p(A))
        return A

    def coulomb(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        J = np.empty_like(D)
        self.lib.orc_build_coulomb(self.r, _p(D), _p(J))
        return J

    def dlist(self, D):
        """Density screening list: (sidx, list) as the reference's dlist_sidx / dlist."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        ns = self.nshell
        sidx = np.zeros(ns + 1, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        n = self.lib.orc_dlist(self.r, _p(D), sidx.ctypes.data_as(ip), None)
        lst = np.zeros(max(n, 1), dtype=np.int32)
        self.lib.orc_dlist(self.r, _p(D), sidx.ctypes.data_as(ip), lst.ctypes.data_as(ip))
        return sidx, lst[:n]

    def exchange(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        K = np.empty_like(D)
        self.lib.orc_build_exchange(self.r, _p(D), _p(K))
        return K
