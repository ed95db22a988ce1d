"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes bindings to liboracle.so (the C++ FP64 restatement of the reference's
algorithms, see oracle_*.cpp) plus numpy restatements of the host-side SCF pieces.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker; the product (iqcp_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False):
    so = _HERE / "liboracle.so"
    if force or not so.exists():
        subprocess.run(["make", "-C", str(_HERE), "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = _HERE / "liboracle.so"
        if not so.exists():
            build()
        _LIB = ctypes.CDLL(str(so))
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _basis_args(b):
    return (b.nshell, _i(b.shell_l), _i(b.shell_nprim), _i(b.shell_prim_off), _d(b.exps), _d(b.coefs),
            _d(np.ascontiguousarray(b.shell_center)), _i(b.shell_bf_off), b.nbf)


def boys_all(m: int, t: float) -> np.ndarray:
    out = np.zeros(m + 1)
    rc = lib().or_boys_all(ctypes.c_int(m), ctypes.c_double(t), _d(out))
    if rc:
        raise ValueError("boys: invalid argument")
    return out


def rys_roots(n: int, t: float):
    r = np.zeros(n)
    w = np.zeros(n)
    rc = lib().or_rys_roots(ctypes.c_int(n), ctypes.c_double(t), _d(r), _d(w))
    if rc:
        raise ValueError(f"rys_roots failed rc={rc}")
    return r, w


def shell_eri(b, si, sj, sk, sl) -> np.ndarray:
    n = [(b.shell_l[s] + 1) * (b.shell_l[s] + 2) // 2 for s in (si, sj, sk, sl)]
    out = np.zeros(int(np.prod(n)))
    rc = lib().or_shell_eri(*_basis_args(b), si, sj, sk, sl, _d(out))
    if rc:
        raise ValueError("shell_eri failed")
    return out.reshape(n)


def schwarz(b) -> np.ndarray:
    q = np.zeros((b.nshell, b.nshell))
    lib().or_schwarz(*_basis_args(b), _d(q))
    return q


def eri_compressed(b):
    npair = b.nbf * (b.nbf + 1) // 2
    out = np.zeros(npair * (npair + 1) // 2)
    cnt = ctypes.c_int64(0)
    lib().or_eri_compressed(*_basis_args(b), _d(out), ctypes.byref(cnt))
    return out, cnt.value


def build_fock(h, d, eri, nbf):
    f = np.zeros((nbf, nbf))
    lib().or_build_fock(_d(np.ascontiguousarray(h)), _d(np.ascontiguousarray(d)), _d(eri), nbf, _d(f))
    return f


def build_jk(d, eri, nbf):
    j = np.zeros((nbf, nbf))
    k = np.zeros((nbf, nbf))
    lib().or_build_jk(_d(np.ascontiguousarray(d)), _d(eri), nbf, _d(j), _d(k))
    return j, k


def pair_index(i, j):
    i, j = max(i, j), min(i, j)
    return i * (i + 1) // 2 + j


def eri_index(i, j, k, l):
    a, b = pair_index(i, j), pair_index(k, l)
    a, b = max(a, b), min(a, b)
    return a * (a + 1) // 2 + b


def expand_eri(eri, nbf) -> np.ndarray:
    """Dense (nbf^4) tensor from the 8-fold compressed layout (small nbf only)."""
    idx = np.arange(nbf)
    p = np.maximum.outer(idx, idx) * (np.maximum.outer(idx, idx) + 1) // 2 + np.minimum.outer(idx, idx)
    P = p[:, :, None, None]
    Q = p[None, None, :, :]
    a = np.maximum(P, Q)
    bq = np.minimum(P, Q)
    return eri[a * (a + 1) // 2 + bq]
