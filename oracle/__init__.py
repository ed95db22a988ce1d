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


def one_electron(b):
    """S, T, V (nbf x nbf) of the basis b with the nuclei b.Z at b.xyz."""
    n = b.nbf
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    Z = np.ascontiguousarray(b.Z, dtype=np.int32)
    xyz = np.ascontiguousarray(b.xyz, dtype=np.float64)
    rc = lib().or_one_electron(*_basis_args(b), len(Z), _i(Z), _d(xyz), _d(S), _d(T), _d(V))
    if rc:
        raise ValueError("one_electron failed")
    return S, T, V


def build_fock(h, d, eri, nbf):
    f = np.zeros((nbf, nbf))
    lib().or_build_fock(_d(np.ascontiguousarray(h)), _d(np.ascontiguousarray(d)), _d(eri), nbf, _d(f))
    return f


def build_fock_omp(h, d, eri, nbf, nthreads):
    f = np.zeros((nbf, nbf))
    lib().or_build_fock_omp(_d(np.ascontiguousarray(h)), _d(np.ascontiguousarray(d)), _d(eri), nbf, _d(f),
                            int(nthreads))
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


# ------------------------------------------------------------------ DFT (oracle_dft.cpp)
def becke_grid(Z, xyz, n_radial=75, quality=0, pruning=1):
    """build_becke_grid (dft/grid.rs:103): (points (G,3), weights (G,), atom index (G,))."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    cap = len(Z) * n_radial * (302 if quality == 0 else 590)
    pts = np.zeros((cap, 3))
    w = np.zeros(cap)
    at = np.zeros(cap, dtype=np.int32)
    L = lib()
    L.or_becke_grid.restype = ctypes.c_int64
    n = L.or_becke_grid(len(Z), _i(Z), _d(xyz), n_radial, quality, pruning, ctypes.c_int64(cap), _d(pts), _d(w),
                        _i(at))
    if n < 0:
        raise ValueError("grid construction failed")
    return pts[:n].copy(), w[:n].copy(), at[:n].copy()


def b3lyp(rho, sigma):
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    exc, vr, vs = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    lib().or_b3lyp(ctypes.c_int64(rho.size), _d(rho), _d(sigma), _d(exc), _d(vr), _d(vs))
    return exc, vr, vs


def basis_on_grid(b, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    G = len(pts)
    chi = np.zeros((G, b.nbf))
    grad = np.zeros((G, b.nbf, 3))
    lib().or_basis_on_grid(*_basis_args(b), ctypes.c_int64(G), _d(pts), _d(chi), _d(grad))
    return chi, grad


def basis_hessian_on_grid(b, pts):
    """evaluate_basis_hessian_on_grid (scf/gradient.rs:1938): (G, nbf, 6) xx xy xz yy yz zz."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    G = len(pts)
    hess = np.zeros((G, b.nbf, 6))
    lib().or_basis_hessian_on_grid(*_basis_args(b), ctypes.c_int64(G), _d(pts), _d(hess))
    return hess


def xc_gradient(b, pts, w, D):
    """xc_gradient_analytical (scf/gradient.rs:1687), B3LYP, no grid-weight response: (natom, 3)."""
    chi, grad = basis_on_grid(b, pts)
    hess = basis_hessian_on_grid(b, pts)
    g = np.zeros((b.natom, 3))
    w = np.ascontiguousarray(w, dtype=np.float64)
    lib().or_xc_gradient(b.nshell, _i(b.shell_l), _i(b.shell_bf_off), _i(np.ascontiguousarray(b.shell_atom, dtype=np.int32)),
                         b.nbf, ctypes.c_int64(len(w)), _d(w), _d(chi), _d(grad), _d(hess),
                         _d(np.ascontiguousarray(D, dtype=np.float64)), _d(g))
    return g


def vxc(chi, grad, w, D):
    """build_vxc_gga_matmul (dft/ks_scf.rs:1150) incl. the 512-point block screening: (V, Exc, rho)."""
    G, nbf = chi.shape
    V = np.zeros((nbf, nbf))
    rho = np.zeros(G)
    L = lib()
    L.or_vxc.restype = ctypes.c_double
    exc = L.or_vxc(nbf, ctypes.c_int64(G), _d(np.ascontiguousarray(chi)), _d(np.ascontiguousarray(grad)),
                   _d(np.ascontiguousarray(w)), _d(np.ascontiguousarray(D)), _d(V), _d(rho))
    return V, exc, rho


def xc_component(which, rho, sigma):
    """which: 'slater' | 'vwn5' | 'b88' | 'lyp'."""
    k = {"slater": 0, "vwn5": 1, "b88": 2, "lyp": 3}[which]
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    exc, vr, vs = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    lib().or_xc_component(k, ctypes.c_int64(rho.size), _d(rho), _d(sigma), _d(exc), _d(vr), _d(vs))
    return exc, vr, vs


def jk_direct(b, D, thr=1e-12, nthreads=None):
    """J, K from the unique Schwarz-surviving shell quartets without the stored tensor
    (or_jk_direct); returns (J, K, kept unique shell quartets)."""
    n = b.nbf
    J, K = np.zeros((n, n)), np.zeros((n, n))
    L = lib()
    L.or_jk_direct.restype = ctypes.c_int64
    nthr = int(nthreads or min(16, os.cpu_count() or 1))
    kept = L.or_jk_direct(*_basis_args(b), _d(np.ascontiguousarray(D, dtype=np.float64)), ctypes.c_double(thr),
                          nthr, _d(J), _d(K))
    return J, K, int(kept)


def screen_counts(b, thr=1e-12, cut=None):
    """Unique shell quartets (pair-of-pairs, 8-fold) kept by the reference's Schwarz predicate
    Q_ij * Q_kl >= thr (integrals/eri/mod.rs:1008-1023), evaluated as that exact product, and
    (cut given) how many of those have Q_ij * Q_kl < cut."""
    Q = schwarz(b)
    q = Q[np.tril_indices(b.nshell)]
    kept = low = 0
    for i in range(len(q)):
        prod = q[i] * q[: i + 1]
        keep = prod >= thr
        kept += int(np.count_nonzero(keep))
        if cut is not None:
            low += int(np.count_nonzero(keep & (prod < cut)))
    return kept if cut is None else (kept, low)
