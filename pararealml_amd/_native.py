"""ctypes binding of ``libpml.so`` (the C ABI declared in ``include/pml.h``).

The device path has no fallback: if the library is missing, or no GPU is
visible, every call raises.
"""

import ctypes
import threading
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpml.so"

PML_OK = 0
PML_E_INVALID = 1
PML_E_UNSUPPORTED = 2
PML_E_HIP = 3

PML_F32 = 0
PML_F64 = 1
PML_F32C = 2

PML_EQ = {
    "diffusion": 0,
    "cahn_hilliard": 1,
    "shallow_water": 2,
    "convection_diffusion": 3,
    "wave": 4,
    "burgers": 5,
    "navier_stokes": 6,
}

PML_REDUCE_SLOTS = 2048
PML_FDM_SCRATCH_BYTES = 4 << 20
PML_JACOBI_STATE_BYTES = 32 + 8 * PML_REDUCE_SLOTS

PML_FLAG_GRAPH = 1
PML_FLAG_CONSTRAIN_INPUT = 2
PML_FLAG_PREPARE = 4

PML_INT_FE = 0
PML_INT_MIDPOINT = 1
PML_INT_RK4 = 2

PML_INTERP_LINEAR = 0
PML_INTERP_NEAREST = 1
PML_LAYOUT_CHANNEL_LAST = 0
PML_LAYOUT_FIELD_MAJOR = 1
PML_LAYOUT_TIME_INNER = 2


class PmlGrid(ctypes.Structure):
    """``pml_grid`` of include/pml.h."""

    _fields_ = [
        ("n0", c_int32),
        ("n1", c_int32),
        ("dx0", c_double),
        ("dx1", c_double),
        ("state_offset", c_int64),
        ("bc_offset", c_int64),
    ]


# Every symbol include/pml.h declares, with (restype, argtypes).
SIGNATURES = {
    "pml_init": (c_int, [c_int]),
    "pml_last_error": (c_char_p, []),
    "pml_version": (c_int, []),
    "pml_device_sync": (c_int, [c_void_p]),
    "pml_fdm_solve": (
        c_int,
        [
            c_int, POINTER(c_double), c_int, c_int, c_int, c_int,
            POINTER(PmlGrid), c_int, c_void_p,
            c_void_p, c_void_p, c_void_p, c_void_p,
            c_int64, c_double, c_int64,
            c_void_p, c_void_p, c_int64, c_int, c_int, c_void_p,
        ],
    ),
    "pml_graph_cache_clear": (c_int, []),
    "pml_fdm_solve_noalloc": (
        c_int,
        [
            c_int, POINTER(c_double), c_int, c_int, c_int, c_int,
            POINTER(PmlGrid), c_int, c_void_p,
            c_void_p, c_void_p, c_void_p, c_void_p,
            c_int64, c_double, c_int64,
            c_void_p, c_void_p, c_int64, c_int, c_int, c_void_p,
        ],
    ),
    "pml_program_create": (
        c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int, POINTER(c_void_p)]
    ),
    "pml_program_destroy": (c_int, [c_void_p]),
    "pml_expr_eval": (
        c_int,
        [c_int, c_void_p, POINTER(c_void_p), POINTER(c_int64), POINTER(c_void_p),
         POINTER(c_int64), c_double, c_int64, c_void_p],
    ),
    "pml_ode_solve": (
        c_int,
        [c_int, c_void_p, c_int, c_int64, c_void_p, c_void_p, c_double, c_int64, c_void_p,
         c_void_p, c_int64, c_void_p],
    ),
    "pml_snapshots_time_inner": (
        c_int, [c_int, c_int, c_void_p, c_int64, c_int64, c_int, c_void_p, c_void_p]
    ),
    "pml_interp_regular": (
        c_int,
        [c_int, c_int, c_int, c_void_p, c_int64, c_int, POINTER(c_int64), c_int, c_void_p,
         c_void_p, c_int64, c_void_p, c_void_p],
    ),
    "pml_snapshot_diff": (
        c_int,
        [c_int, c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p,
         c_int64, c_int64, c_int, c_void_p],
    ),
    "pml_supervised_rows": (
        c_int,
        [c_void_p, c_int64, c_void_p, c_int64, c_int64, c_int64, c_int, c_void_p, c_void_p,
         c_int, c_void_p, c_void_p, c_void_p],
    ),
    "pml_fdm_configs": (c_int, [c_int, c_int, c_int]),
    "pml_fdm_config_name": (c_char_p, [c_int, c_int, c_int, c_int]),
    "pml_stencil_first": (
        c_int,
        [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_double,
         c_void_p, c_void_p, c_int, c_void_p],
    ),
    "pml_stencil_second": (
        c_int,
        [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_double, c_double,
         c_void_p, c_void_p, c_int, c_void_p],
    ),
    "pml_weighted_add": (
        c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p]
    ),
    "pml_set_values": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "pml_secant_step": (
        c_int,
        [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_int64, c_double, c_void_p, c_void_p],
    ),
    "pml_jacobi_sweep": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, POINTER(c_int64), c_int,
         POINTER(c_double), POINTER(c_void_p), POINTER(c_void_p), c_void_p,
         c_void_p, c_void_p, c_void_p],
    ),
    "pml_jacobi_solve": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, POINTER(c_int64), c_int,
         POINTER(c_double), POINTER(c_void_p), POINTER(c_void_p), c_void_p,
         c_double, c_int64, c_int64, c_void_p, POINTER(c_int64), POINTER(c_double),
         c_void_p],
    ),
    "pml_jacobi_solve_weighted": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, POINTER(c_int64), c_int,
         POINTER(c_double), POINTER(c_void_p), POINTER(c_void_p), c_void_p,
         c_double, c_int64, c_int64, c_void_p, POINTER(c_int64), POINTER(c_double),
         POINTER(c_void_p), POINTER(c_void_p), c_void_p, c_void_p],
    ),
    "pml_parareal_correction": (
        c_int,
        [c_int, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64,
         c_void_p],
    ),
    "pml_parareal_update": (
        c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]
    ),
    "pml_from_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "pml_to_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "pml_slice_guard": (
        c_int, [c_void_p, c_void_p, c_int64, c_double, c_void_p, c_void_p, c_void_p]
    ),
    "pml_restrict2x": (
        c_int, [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p]
    ),
    "pml_prolong2x": (
        c_int, [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p]
    ),
    "pml_prolong2x_cubic": (
        c_int, [c_int, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p]
    ),
    "pml_apply_dirichlet": (
        c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int64, c_void_p]
    ),
    "pml_conv64": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
         c_int, c_void_p],
    ),
    "pml_wgrad64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                            c_void_p, c_int64, c_void_p]),
    "pml_grad_partials_floats": (c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "pml_conv_small": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
         c_int, c_int, c_int, c_int, c_void_p],
    ),
    "pml_conv_small_wgrad": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
         c_int64, c_void_p],
    ),
    "pml_bias_grad64": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "pml_adam": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float, c_float, c_float, c_float,
         c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p],
    ),
    "pml_surrogate_normalize": (
        c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int64, c_void_p]
    ),
    "pml_surrogate_residual": (
        c_int,
        [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int64, c_void_p],
    ),
    "pml_surrogate_loss": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int64,
         c_void_p],
    ),
}

_lock = threading.Lock()
_lib = None
_initialised = set()


def load_library() -> ctypes.CDLL:
    """Loads libpml.so and declares its signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m pararealml_amd.build`"
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        return _lib


def check(rc: int) -> None:
    """Maps a status code to the reference's exception types."""
    if rc == PML_OK:
        return
    msg = load_library().pml_last_error().decode("utf-8", "replace")
    if rc == PML_E_INVALID:
        raise ValueError(msg)
    if rc == PML_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def device_lib():
    """The library, initialised for torch's current GPU.  Raises when no GPU
    is visible: there is no host fallback."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError(
            "pararealml_amd needs a gfx950 GPU: torch.cuda.is_available() is False"
        )
    lib = load_library()
    dev = torch.cuda.current_device()
    if dev not in _initialised:
        check(lib.pml_init(dev))
        _initialised.add(dev)
    return lib
