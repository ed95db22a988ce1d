"""iqcp_amd — MI355X-native Fock-build engine (HIP kernels behind the C ABI in include/qcg.h)."""
from .basis import BasisSet, build_basis  # noqa: F401
from .engine import Engine, ScfConfig  # noqa: F401
