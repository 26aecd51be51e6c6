"""Builds the MEX gateway (mnad-modeling-rtm_amd/matlab/mnad_mex.c) against
the mock mx API of tests/mex_mock and calls it the way MATLAB would."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "mex_mock")
OUT = os.path.join(MOCK, "_build", "libmnad_mex_mock.so")
LIBDIR = os.path.join(ROOT, "mnad-modeling-rtm_amd", "_build")


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__(f"{ident}: {msg}")
        self.ident = ident


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    src = os.path.join(ROOT, "mnad-modeling-rtm_amd", "matlab", "mnad_mex.c")
    cmd = ["gcc", "-O1", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fPIC", "-shared",
           f"-I{MOCK}", f"-I{os.path.join(ROOT, 'include')}", src,
           os.path.join(MOCK, "mock_mx.c"), f"-L{LIBDIR}", "-lmnad", f"-Wl,-rpath,{LIBDIR}",
           "-lm", "-o", OUT]
    subprocess.run(cmd, check=True)
    return OUT


class Mex:
    def __init__(self):
        L = C.CDLL(build())
        vp = C.c_void_p
        L.mock_double.restype = vp
        L.mock_double.argtypes = [C.c_size_t, C.c_size_t, C.POINTER(C.c_double)]
        L.mock_string.restype = vp
        L.mock_string.argtypes = [C.c_char_p]
        L.mock_data.restype = C.POINTER(C.c_double)
        L.mock_data.argtypes = [vp]
        L.mock_str.restype = C.c_char_p
        L.mock_str.argtypes = [vp]
        L.mock_free.argtypes = [vp]
        L.mock_call.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p,
                                C.c_char_p]
        L.mxGetM.restype = C.c_size_t
        L.mxGetM.argtypes = [vp]
        L.mxGetN.restype = C.c_size_t
        L.mxGetN.argtypes = [vp]
        self.L = L

    def _arg(self, a):
        if isinstance(a, str):
            return self.L.mock_string(a.encode())
        arr = np.asarray(a, dtype=np.float64)
        if arr.ndim == 0:
            arr = arr.reshape(1, 1)
        elif arr.ndim == 1:
            arr = arr.reshape(1, -1)
        arr = np.asfortranarray(arr)
        return self.L.mock_double(arr.shape[0], arr.shape[1],
                                  arr.ctypes.data_as(C.POINTER(C.c_double)))

    def __call__(self, *args, nlhs=1):
        L = self.L
        prhs = (C.c_void_p * len(args))(*[self._arg(a) for a in args])
        plhs = (C.c_void_p * max(nlhs, 1))()
        ident, msg = C.create_string_buffer(128), C.create_string_buffer(1024)
        rc = L.mock_call(nlhs, plhs, len(args), prhs, ident, msg)
        for p in prhs:
            L.mock_free(p)
        if rc:
            raise MexError(ident.value.decode(), msg.value.decode())
        outs = []
        for p in plhs[:nlhs]:
            if not p:
                outs.append(None)
                continue
            s = L.mock_str(p)
            if s is not None:
                outs.append(s.decode())
            else:
                m, n = L.mxGetM(p), L.mxGetN(p)
                d = np.ctypeslib.as_array(L.mock_data(p), shape=(max(m * n, 1),))[:m * n]
                outs.append(d.reshape(m, n, order="F").copy())
            L.mock_free(p)
        return outs[0] if nlhs == 1 else outs
