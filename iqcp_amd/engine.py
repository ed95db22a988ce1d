"""Python binding of libqcg.so (the C ABI in include/qcg.h).

Mirrors the qc-core entry points the GPU path replaces (paths relative to crates/qc-core/src):
  boys_eval_all (boys/mod.rs:327), rys_roots (rys/mod.rs:253), compute_schwarz_bounds /
  eri_compressed (integrals/eri/mod.rs:1263 / :992), build_fock (scf/mod.rs:825),
  build_jk (dft/ks_scf.rs:1369), rhf_scf (scf/mod.rs:1363), ks_scf (dft/ks_scf.rs:1567),
  rhf_gradient (scf/gradient.rs:900).
There is no CPU fallback: if the HIP library or a GPU is missing, every call raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .basis import BasisSet

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libqcg.so"
_LIB = None

QCG_PREC_FP64 = 0
QCG_PREC_MIXED = 1
STATUS = {0: "OK", 1: "INVALID_ARG", 2: "UNSUPPORTED_L", 3: "NUMERICAL", 4: "DEVICE", 5: "NOT_CONVERGED",
          6: "NO_BASIS", 7: "NO_GRID"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# Every exported symbol of include/qcg.h (tests check that the library exports all of them).
EXPORTS = [
    "qcg_create", "qcg_destroy", "qcg_last_error", "qcg_set_basis", "qcg_set_options", "qcg_set_charge",
    "qcg_nbf", "qcg_boys_eval_all", "qcg_rys_roots", "qcg_schwarz_bounds", "qcg_screen_counts",
    "qcg_eri_compressed", "qcg_build_jk", "qcg_build_fock", "qcg_build_jk_device", "qcg_one_electron",
    "qcg_rhf_scf", "qcg_build_fock_device", "qcg_set_profiling", "qcg_class_stats",
    "qcg_build_grid", "qcg_set_grid", "qcg_get_grid", "qcg_xc_eval", "qcg_xc", "qcg_xc_device", "qcg_ks_scf",
    "qcg_eri_gradient", "qcg_one_electron_gradient", "qcg_rhf_gradient",
    "qcg_set_basis_batch", "qcg_batch_build_jk", "qcg_batch_rhf_scf", "qcg_autotune",
    "qcg_set_tiles", "qcg_get_tiles", "qcg_xc_gradient", "qcg_ks_gradient", "qcg_set_graph",
    "qcg_set_prim_split", "qcg_set_xc_mode", "qcg_build_fock_stored", "qcg_rhf_scf_stored",
    "qcg_autotune_refine", "qcg_error_bound", "qcg_scf_history", "qcg_build_jk_device_noalloc",
    "qcg_set_class_cost", "qcg_get_class_cost",
]


class QcgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qcg error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class ScfConfig(ctypes.Structure):
    """qcg_scf_config = ScfConfig + ConvergenceProfile (scf/mod.rs:123-262)."""
    _fields_ = [("energy_tol", ctypes.c_double), ("density_tol", ctypes.c_double),
                ("max_iterations", ctypes.c_int), ("use_diis", ctypes.c_int), ("diis_size", ctypes.c_int),
                ("diis_start", ctypes.c_int), ("damp", ctypes.c_double), ("damp_start", ctypes.c_int),
                ("level_shift", ctypes.c_double)]

    @classmethod
    def tight(cls, use_diis=True, **kw):
        c = cls(1e-10, 1e-8, 100, int(use_diis), 8, 1, 0.0, 5, 0.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    @classmethod
    def medium(cls, use_diis=True, **kw):
        c = cls(1e-8, 1e-6, 100, int(use_diis), 8, 1, 0.0, 5, 0.0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class ClassStat(ctypes.Structure):
    """qcg_class_stat: per quartet-class counts, flop model and last profiled kernel times."""
    _fields_ = [("la", ctypes.c_int), ("lb", ctypes.c_int), ("lc", ctypes.c_int), ("ld", ctypes.c_int),
                ("quartets_fp64", ctypes.c_int64), ("quartets_fp32", ctypes.c_int64),
                ("prim_quartets_fp64", ctypes.c_int64), ("prim_quartets_fp32", ctypes.c_int64),
                ("flops_fp64", ctypes.c_double), ("flops_fp32", ctypes.c_double),
                ("ms_fp64", ctypes.c_double), ("ms_fp32", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScfIteration(ctypes.Structure):
    """qcg_scf_iteration = ScfIteration (scf/mod.rs:627-650) plus the mixed-precision fields."""
    _fields_ = [("iteration", ctypes.c_int), ("converged", ctypes.c_int), ("diis_applied", ctypes.c_int),
                ("fp32_build", ctypes.c_int), ("energy_total", ctypes.c_double),
                ("energy_electronic", ctypes.c_double), ("delta_e", ctypes.c_double),
                ("rms_density_change", ctypes.c_double), ("error_bound", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScfResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int), ("energy_total", ctypes.c_double),
                ("energy_electronic", ctypes.c_double), ("energy_nuclear", ctypes.c_double),
                ("last_delta_e", ctypes.c_double), ("last_rms_d", ctypes.c_double),
                ("energy_xc", ctypes.c_double), ("energy_j", ctypes.c_double), ("energy_k", ctypes.c_double),
                ("energy_1e", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load libqcg.so; raise if it was not built (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.qcg_last_error.restype = ctypes.c_char_p
        L.qcg_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
        L.qcg_destroy.argtypes = [ctypes.c_void_p]
        L.qcg_last_error.argtypes = [ctypes.c_void_p]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


@dataclass
class JK:
    J: np.ndarray
    K: np.ndarray


class Engine:
    """One qcg context bound to a GPU and a basis set."""

    def __init__(self, basis: BasisSet, device: int = 0, precision: str = "fp64",
                 fp32_cutoff: float = 1e-6, schwarz_threshold: float = 1e-12, charge: int = 0):
        self.L = lib()
        self.ctx = ctypes.c_void_p()
        rc = self.L.qcg_create(ctypes.byref(self.ctx), device)
        if rc:
            raise QcgError(rc, "qcg_create failed (no GPU?)")
        self.basis = basis
        self.nbf = basis.nbf
        self.set_options(precision, fp32_cutoff, schwarz_threshold)
        b = basis
        self._check(self.L.qcg_set_basis(
            self.ctx, b.nshell, _i(b.shell_l), _i(b.shell_nprim), _i(b.shell_prim_off), _d(b.exps), _d(b.coefs),
            _d(np.ascontiguousarray(b.shell_center)), _i(b.shell_atom), b.natom, _i(b.Z),
            _d(np.ascontiguousarray(b.xyz))))
        self._check(self.L.qcg_set_charge(self.ctx, int(sum(b.Z)) - b.nelec))

    def _check(self, rc):
        if rc:
            raise QcgError(rc, self.L.qcg_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.L.qcg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, precision="fp64", fp32_cutoff=1e-6, schwarz_threshold=1e-12):
        p = {"fp64": QCG_PREC_FP64, "mixed": QCG_PREC_MIXED}[precision]
        self.precision = precision
        self._check(self.L.qcg_set_options(self.ctx, p, ctypes.c_double(fp32_cutoff),
                                           ctypes.c_double(schwarz_threshold)))

    # ---- special functions
    def boys_eval_all(self, m_max: int, T) -> np.ndarray:
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        F = np.zeros((len(T), m_max + 1))
        self._check(self.L.qcg_boys_eval_all(self.ctx, m_max, len(T), _d(T), _d(F)))
        return F

    def rys_roots(self, nroots: int, T):
        T = np.ascontiguousarray(T, dtype=np.float64).ravel()
        x = np.zeros((len(T), nroots))
        w = np.zeros((len(T), nroots))
        self._check(self.L.qcg_rys_roots(self.ctx, nroots, len(T), _d(T), _d(x), _d(w)))
        return x, w

    # ---- two-electron
    def schwarz_bounds(self) -> np.ndarray:
        Q = np.zeros((self.basis.nshell, self.basis.nshell))
        self._check(self.L.qcg_schwarz_bounds(self.ctx, _d(Q)))
        return Q

    def screen_counts(self):
        kept = ctypes.c_int64()
        f32 = ctypes.c_int64()
        self._check(self.L.qcg_screen_counts(self.ctx, ctypes.byref(kept), ctypes.byref(f32)))
        return kept.value, f32.value

    def error_bound(self, D, hf_frac: float = 1.0) -> float:
        """Bound on |E2(mixed) - E2(FP64)| of a J/K build at D (0 in FP64 mode)."""
        b = ctypes.c_double()
        Dc = np.ascontiguousarray(D, dtype=np.float64)
        self._check(self.L.qcg_error_bound(self.ctx, _d(Dc), ctypes.c_double(hf_frac), ctypes.byref(b)))
        return b.value

    def scf_history(self):
        """Per-iteration trace of the last rhf_scf / ks_scf (list of dicts)."""
        n = ctypes.c_int()
        self._check(self.L.qcg_scf_history(self.ctx, None, 0, ctypes.byref(n)))
        arr = (ScfIteration * max(1, n.value))()
        self._check(self.L.qcg_scf_history(self.ctx, arr, n.value, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def eri_compressed(self) -> np.ndarray:
        npair = self.nbf * (self.nbf + 1) // 2
        out = np.zeros(npair * (npair + 1) // 2)
        self._check(self.L.qcg_eri_compressed(self.ctx, _d(out)))
        return out

    def build_jk(self, D, want_j=True, want_k=True) -> JK:
        D = np.ascontiguousarray(D, dtype=np.float64)
        J = np.zeros_like(D) if want_j else None
        K = np.zeros_like(D) if want_k else None
        self._check(self.L.qcg_build_jk(self.ctx, _d(D), _d(J) if want_j else None, _d(K) if want_k else None))
        return JK(J, K)

    def build_fock(self, H, D) -> np.ndarray:
        H = np.ascontiguousarray(H, dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64)
        F = np.zeros_like(D)
        self._check(self.L.qcg_build_fock(self.ctx, _d(H), _d(D), _d(F)))
        return F

    def build_jk_device(self, dD_ptr: int, dJ_ptr: int, dK_ptr: int, stream_ptr: int = 0):
        """Device-pointer J/K build (e.g. torch tensor .data_ptr()) on a HIP stream."""
        self._check(self.L.qcg_build_jk_device(self.ctx, ctypes.c_void_p(dD_ptr), ctypes.c_void_p(dJ_ptr),
                                               ctypes.c_void_p(dK_ptr), ctypes.c_void_p(stream_ptr)))

    def build_jk_device_noalloc(self, dD_ptr: int, dJ_ptr: int, dK_ptr: int, stream_ptr: int = 0):
        """qcg_build_jk_device with the library's static scratch (allocates nothing; the plan must
        exist: one build_jk_device call after the last change of options or tiles)."""
        self._check(self.L.qcg_build_jk_device_noalloc(self.ctx, ctypes.c_void_p(dD_ptr), ctypes.c_void_p(dJ_ptr),
                                                       ctypes.c_void_p(dK_ptr), ctypes.c_void_p(stream_ptr)))

    def build_fock_device(self, dH_ptr: int, dD_ptr: int, dF_ptr: int, k_factor: float = 0.5,
                          stream_ptr: int = 0):
        self._check(self.L.qcg_build_fock_device(self.ctx, ctypes.c_void_p(dH_ptr), ctypes.c_void_p(dD_ptr),
                                                 ctypes.c_double(k_factor), ctypes.c_void_p(dF_ptr),
                                                 ctypes.c_void_p(stream_ptr)))

    def autotune(self, dD_ptr: int, iters: int = 2):
        """Pick the fastest tile per quartet class and precision part on a device density.
        Returns [[(block, kets) FP64 part, (block, kets) FP32 part] per class]."""
        ch = (ctypes.c_int * 84)()
        self._check(self.L.qcg_autotune(self.ctx, ctypes.c_void_p(dD_ptr), int(iters), ch))
        return [[(ch[4 * i], ch[4 * i + 1]), (ch[4 * i + 2], ch[4 * i + 3])] for i in range(21)]

    def autotune_refine(self, dD_ptr: int, builds: int = 30):
        """Refine the autotuned tiles on whole builds (all classes on the four streams at once).
        Returns the tiles in the form of autotune()."""
        ch = (ctypes.c_int * 84)()
        self._check(self.L.qcg_autotune_refine(self.ctx, ctypes.c_void_p(dD_ptr), int(builds), ch))
        return [[(ch[4 * i], ch[4 * i + 1]), (ch[4 * i + 2], ch[4 * i + 3])] for i in range(21)]

    def set_tiles(self, tiles):
        """Tiles per quartet class: either (block, kets) for both precision parts or a pair
        [(block, kets) FP64, (block, kets) FP32].  Block 0 selects the wave-cooperative kernel."""
        t = list(tiles)
        if len(t) != 21:
            raise ValueError("21 tiles expected")
        flat = []
        for x in t:
            parts = [tuple(x), tuple(x)] if isinstance(x[0], int) else [tuple(x[0]), tuple(x[1])]
            for bk in parts:
                flat += [int(bk[0]), int(bk[1])]
        arr = (ctypes.c_int * 84)(*flat)
        self._check(self.L.qcg_set_tiles(self.ctx, arr))

    def get_tiles(self):
        arr = (ctypes.c_int * 84)()
        self._check(self.L.qcg_get_tiles(self.ctx, arr))
        return [[(arr[4 * i], arr[4 * i + 1]), (arr[4 * i + 2], arr[4 * i + 3])] for i in range(21)]

    def set_class_cost(self, cost):
        """Per-class costs (21, tile order) that order the class launches over the streams;
        None or all zeros: the work plan's model flops."""
        if cost is None:
            self._check(self.L.qcg_set_class_cost(self.ctx, None))
            return
        c = [float(x) for x in cost]
        if len(c) != 21:
            raise ValueError("21 class costs expected")
        self._check(self.L.qcg_set_class_cost(self.ctx, (ctypes.c_double * 21)(*c)))

    def get_class_cost(self):
        """(costs in use, True if measured by autotune or set, False for the plan's estimate)."""
        arr = (ctypes.c_double * 21)()
        m = ctypes.c_int()
        self._check(self.L.qcg_get_class_cost(self.ctx, arr, ctypes.byref(m)))
        return list(arr), bool(m.value)

    def set_graph(self, on: bool):
        """Replay device Fock builds from a captured HIP graph."""
        self._check(self.L.qcg_set_graph(self.ctx, int(bool(on))))

    def set_xc_mode(self, mode: int):
        """XC kernels: 0 = MFMA kernels, FP32 products for mixed and FP64 products for FP64 up to
        352 basis functions, library DGEMMs for FP64 above (default); 1 = library DGEMMs for both;
        2 = MFMA kernels for both."""
        self._check(self.L.qcg_set_xc_mode(self.ctx, int(mode)))

    def set_prim_split(self, split: int):
        """Cut bra pairs with more than `split` primitive pairs into primitive ranges (0 = off)."""
        self._check(self.L.qcg_set_prim_split(self.ctx, int(split)))

    def set_profiling(self, on: bool):
        self._check(self.L.qcg_set_profiling(self.ctx, int(bool(on))))

    def class_stats(self):
        arr = (ClassStat * 21)()
        n = ctypes.c_int()
        self._check(self.L.qcg_class_stats(self.ctx, arr, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    # ---- exchange-correlation
    def build_grid(self, n_radial=75, quality="standard", pruning=True) -> int:
        """build_becke_grid (dft/grid.rs:103) + Phi / grad Phi on the device; returns the point count."""
        q = {"standard": 0, "fine": 1}[quality]
        n = ctypes.c_int64()
        self._check(self.L.qcg_build_grid(self.ctx, int(n_radial), q, int(bool(pruning)), ctypes.byref(n)))
        self.npts = n.value
        return n.value

    def set_grid(self, xyz, w):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._check(self.L.qcg_set_grid(self.ctx, ctypes.c_int64(len(w)), _d(xyz), _d(w)))
        self.npts = len(w)

    def get_grid(self):
        xyz = np.zeros((self.npts, 3))
        w = np.zeros(self.npts)
        at = np.zeros(self.npts, dtype=np.int32)
        self._check(self.L.qcg_get_grid(self.ctx, _d(xyz), _d(w), _i(at)))
        return xyz, w, at

    def xc_eval(self, rho, sigma):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        exc, vr, vs = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
        self._check(self.L.qcg_xc_eval(self.ctx, ctypes.c_int64(rho.size), _d(rho), _d(sigma), _d(exc), _d(vr),
                                       _d(vs)))
        return exc, vr, vs

    def xc(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        V = np.zeros_like(D)
        exc = ctypes.c_double()
        self._check(self.L.qcg_xc(self.ctx, _d(D), _d(V), ctypes.byref(exc)))
        return V, exc.value

    def xc_device(self, dD_ptr, dV_ptr, dExc_ptr, stream_ptr=0):
        self._check(self.L.qcg_xc_device(self.ctx, ctypes.c_void_p(dD_ptr), ctypes.c_void_p(dV_ptr),
                                         ctypes.c_void_p(dExc_ptr), ctypes.c_void_p(stream_ptr)))

    def ks_scf(self, config: ScfConfig | None = None, D_init=None):
        """B3LYP KS-SCF on the current grid (ks_scf, dft/ks_scf.rs:1567)."""
        cfg = config or ScfConfig.tight()
        n = self.nbf
        res = ScfResult()
        D, C, eps, F = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros((n, n))
        Di = None if D_init is None else np.ascontiguousarray(D_init, dtype=np.float64)
        self._check(self.L.qcg_ks_scf(self.ctx, ctypes.byref(cfg), _d(Di) if Di is not None else None,
                                      ctypes.byref(res), _d(D), _d(C), _d(eps), _d(F)))
        out = res.as_dict()
        out.update(D=D, C=C, eps=eps, F=F)
        return out

    # ---- gradients
    def eri_gradient(self, D, hf_frac=1.0):
        D = np.ascontiguousarray(D, dtype=np.float64)
        g = np.zeros((self.basis.natom, 3))
        self._check(self.L.qcg_eri_gradient(self.ctx, _d(D), ctypes.c_double(hf_frac), _d(g)))
        return g

    def one_electron_gradient(self, D, W):
        D = np.ascontiguousarray(D, dtype=np.float64)
        W = np.ascontiguousarray(W, dtype=np.float64)
        g = np.zeros((self.basis.natom, 3))
        self._check(self.L.qcg_one_electron_gradient(self.ctx, _d(D), _d(W), _d(g)))
        return g

    def rhf_gradient(self, D, C, eps, n_occ):
        """rhf_gradient (scf/gradient.rs:900): (natom, 3) in Ha/bohr."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        C = np.ascontiguousarray(C, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        g = np.zeros((self.basis.natom, 3))
        self._check(self.L.qcg_rhf_gradient(self.ctx, _d(D), _d(C), _d(eps), int(n_occ), _d(g)))
        return g

    def xc_gradient(self, D):
        """B3LYP XC nuclear gradient on the current grid (xc_gradient_analytical,
        scf/gradient.rs:1687; no grid-weight response): (natom, 3)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        g = np.zeros((self.basis.natom, 3))
        self._check(self.L.qcg_xc_gradient(self.ctx, _d(D), _d(g)))
        return g

    def ks_gradient(self, D, C, eps, n_occ):
        """B3LYP nuclear gradient (ks_dft_gradient_with_guess, scf/gradient.rs:1214, no D3): (natom, 3)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        C = np.ascontiguousarray(C, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        g = np.zeros((self.basis.natom, 3))
        self._check(self.L.qcg_ks_gradient(self.ctx, _d(D), _d(C), _d(eps), int(n_occ), _d(g)))
        return g

    # ---- one-electron + SCF
    def one_electron(self):
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        self._check(self.L.qcg_one_electron(self.ctx, _d(S), _d(T), _d(V)))
        return S, T, V

    def rhf_scf(self, config: ScfConfig | None = None, D_init=None):
        cfg = config or ScfConfig.tight()
        n = self.nbf
        res = ScfResult()
        D, C, eps, F = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros((n, n))
        Di = None if D_init is None else np.ascontiguousarray(D_init, dtype=np.float64)
        self._check(self.L.qcg_rhf_scf(self.ctx, ctypes.byref(cfg), _d(Di) if Di is not None else None,
                                       ctypes.byref(res), _d(D), _d(C), _d(eps), _d(F)))
        out = res.as_dict()
        out.update(D=D, C=C, eps=eps, F=F)
        return out


class BatchEngine:
    """Many molecules in one context (qcg_set_basis_batch): batched J/K and lockstep SCF."""

    def __init__(self, bases, device: int = 0, precision: str = "fp64", fp32_cutoff: float = 1e-6):
        self.L = lib()
        self.ctx = ctypes.c_void_p()
        rc = self.L.qcg_create(ctypes.byref(self.ctx), device)
        if rc:
            raise QcgError(rc, "qcg_create failed (no GPU?)")
        self.bases = bases
        p = {"fp64": QCG_PREC_FP64, "mixed": QCG_PREC_MIXED}[precision]
        self._check(self.L.qcg_set_options(self.ctx, p, ctypes.c_double(fp32_cutoff), ctypes.c_double(1e-12)))
        ns = np.array([b.nshell for b in bases], np.int32)
        na = np.array([b.natom for b in bases], np.int32)
        l = np.concatenate([b.shell_l for b in bases]).astype(np.int32)
        npr = np.concatenate([b.shell_nprim for b in bases]).astype(np.int32)
        po, off = [], 0
        for b in bases:
            po.append(b.shell_prim_off + off)
            off += len(b.exps)
        po = np.concatenate(po).astype(np.int32)
        ex = np.concatenate([b.exps for b in bases])
        co = np.concatenate([b.coefs for b in bases])
        ce = np.ascontiguousarray(np.concatenate([b.shell_center for b in bases]))
        sa = np.concatenate([b.shell_atom for b in bases]).astype(np.int32)
        Z = np.concatenate([b.Z for b in bases]).astype(np.int32)
        xyz = np.ascontiguousarray(np.concatenate([b.xyz for b in bases]))
        self._check(self.L.qcg_set_basis_batch(self.ctx, len(bases), _i(ns), _i(na), _i(l), _i(npr), _i(po),
                                               _d(ex), _d(co), _d(ce), _i(sa), _i(Z), _d(xyz)))
        self.nbf = [b.nbf for b in bases]
        self.offsets = np.cumsum([0] + [n * n for n in self.nbf])

    def set_tiles(self, tiles):
        """Tiles per quartet class, as Engine.set_tiles."""
        Engine.set_tiles(self, tiles)

    def get_tiles(self):
        return Engine.get_tiles(self)

    def set_class_cost(self, cost):
        """Per-class launch costs, as Engine.set_class_cost."""
        Engine.set_class_cost(self, cost)

    def get_class_cost(self):
        return Engine.get_class_cost(self)

    def class_stats(self):
        """Per quartet-class counts and flop model over all molecules of the batch."""
        arr = (ClassStat * 21)()
        n = ctypes.c_int()
        self._check(self.L.qcg_class_stats(self.ctx, arr, ctypes.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def autotune(self, dD_ptr: int, iters: int = 2):
        """Per-class tile choice on a packed device density (all molecules), as Engine.autotune."""
        return Engine.autotune(self, dD_ptr, iters)

    def _check(self, rc):
        if rc:
            raise QcgError(rc, self.L.qcg_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            self.L.qcg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, mats):
        return np.concatenate([np.ascontiguousarray(m, dtype=np.float64).ravel() for m in mats])

    def unpack(self, flat):
        return [flat[self.offsets[i]:self.offsets[i + 1]].reshape(n, n) for i, n in enumerate(self.nbf)]

    def build_jk(self, Ds):
        D = self.pack(Ds)
        J, K = np.zeros_like(D), np.zeros_like(D)
        self._check(self.L.qcg_batch_build_jk(self.ctx, _d(D), _d(J), _d(K)))
        return self.unpack(J), self.unpack(K)

    def rhf_scf(self, config: ScfConfig | None = None, use_graph: bool = True):
        cfg = config or ScfConfig.tight()
        res = (ScfResult * len(self.bases))()
        D = np.zeros(int(self.offsets[-1]))
        ms = ctypes.c_double()
        self._check(self.L.qcg_batch_rhf_scf(self.ctx, ctypes.byref(cfg), int(bool(use_graph)), res, _d(D),
                                             ctypes.byref(ms)))
        return [res[i].as_dict() for i in range(len(self.bases))], self.unpack(D), ms.value
