"""bem_amd: MI355X-native frequency-sweep hot path of the boundary-element
heat-transfer code (RWG/PMCHWT impedance assembly, dense complex solve,
spectral-power trace), behind the C ABI of include/bem_hip.h.

The Python layer is a thin ctypes binding of libbem_hip.so used by the tests
and the benchmark; the compute path is the HIP library, and it fails loudly
when the library is missing or no GPU is present.
"""
from .runtime import (  # noqa: F401
    BEM_E_INVALID,
    BEM_E_NOMEM,
    BEM_E_SINGULAR,
    BEM_PENDING,
    BEM_PREC_FP64,
    BEM_PREC_MIXED,
    BEM_SOLVE_LOWRANK,
    BEM_LOWRANK_ACA,
    BEM_SOLVE_NORECOVER,
    BemError,
    Body,
    Context,
    Request,
    lib_path,
    load_library,
    omega_tilde,
    rwg_tree,
    theta,
    lorentz_eps,
)
from . import mesh  # noqa: F401
