"""ctypes binding of libdfcsr.so (the gfx950 C ABI declared in include/dfcsr.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or does not load, every entry point raises DfcsrError.
Device arrays are torch tensors on a HIP device; they cross the boundary as
raw pointers (``data_ptr()``) together with torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os

import torch

LIB_NAME = "libdfcsr.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

c_dp = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f64 = ctypes.c_double

ABI_VERSION = 8   # DFCSR_ABI_VERSION
DFCSR_PREC_FP64 = 0
DFCSR_PREC_MIXED = 1
MOMENTS_COUNT = 10
# beam statistics (DFCSR_STAT_*): means, covariance triangle, Twiss x / y, energy
STAT_MEAN, STAT_COV, STAT_TWISS_X, STAT_TWISS_Y = 0, 6, 27, 34
STAT_MEAN_ENERGY, STAT_SIGMA_ENERGY, STATS_COUNT = 41, 42, 43
TWISS_KEYS = ("alpha", "beta", "gamma", "emit", "eta", "etap", "norm_emit")
HIST_REC64 = 20   # corner-packed records: field f, corner c at [4 f + c]
HIST_REC32 = 20
HIST_CHUNKS32 = 5
CELL = 5   # values per cell of the cell-major rings: the five fields
# sgolay2d's derivative argument: the reference's None, 'col', 'row', 'both'
SG2D_DERIVATIVE = {None: 0, "col": 1, "row": 2, "both": 3}
SG2D_MAX_WINDOW = 13
# rings: FP64 (Xi, Zi, cap, 20); FP32 chunked (Xi, cap, 5, Zi, 4); FP32
# time-innermost (Xi, Zi, cap, 20); cell-major (cap, Xi, Zi, 5) float64 / float32
FIELDS = ("density", "density_x", "density_z", "vx", "vx_x")   # DFCSR_F_* order


class DfcsrError(RuntimeError):
    pass


class RegridJob(ctypes.Structure):
    _fields_ = [("src_slot", c_i32), ("dst_slot", c_i32), ("nbx", c_i32), ("nbz", c_i32),
                ("x0", c_f64), ("x1", c_f64), ("z0", c_f64), ("z1", c_f64)]


class History(ctypes.Structure):
    _fields_ = [("data", c_dp), ("data32", c_dp), ("data32t", c_dp), ("cells", c_dp),
                ("cells32", c_dp), ("cap", c_i32), ("head", c_i32), ("nt", c_i32), ("Xi", c_i32),
                ("Zi", c_i32), ("min_t", c_f64), ("min_x", c_f64), ("min_z", c_f64),
                ("dt", c_f64), ("dx", c_f64), ("dz", c_f64), ("dcells32", c_dp),
                ("zspan", c_dp)]


class Lattice(ctypes.Structure):
    _fields_ = [("table", c_dp), ("ns", c_i32), ("nelem", c_i32), ("min_s", c_f64),
                ("ds", c_f64), ("distance", c_dp), ("rho", c_dp)]


DFCSR_ELE_DRIFT, DFCSR_ELE_SBEND, DFCSR_ELE_QUAD, DFCSR_ELE_SEXTUPOLE = 0, 1, 2, 3


class Element(ctypes.Structure):
    _fields_ = [("type", c_i32), ("num_steps", c_i32), ("L", c_f64), ("g", c_f64),
                ("e1", c_f64), ("e2", c_f64), ("fint", c_f64), ("hgap", c_f64),
                ("fintx", c_f64), ("hgapx", c_f64), ("k1", c_f64), ("k2", c_f64),
                ("x_offset", c_f64), ("y_offset", c_f64), ("tilt", c_f64)]


class WakeBatch(ctypes.Structure):
    _fields_ = [("tasks", c_i32), ("blocks_far", c_i32), ("blocks_near", c_i32),
                ("observers", c_i32), ("partial_bytes", c_i64), ("precision", c_i32)]


class WakeParams(ctypes.Structure):
    _fields_ = [("t", c_f64), ("sigma_x", c_f64), ("sigma_z", c_f64), ("slope0", c_f64),
                ("slope1", c_f64), ("mean_x", c_f64), ("formation_length", c_f64),
                ("n_formation_length", c_f64), ("csr_scaling", c_f64), ("Nx", c_i32),
                ("Nz", c_i32), ("mesh_xbins", c_i32), ("mesh_zbins", c_i32), ("xr0", c_f64),
                ("xr1", c_f64), ("zr0", c_f64), ("zr1", c_f64), ("observers", c_dp),
                ("points_evaluated", c_dp)]


# (name, restype, argtypes) for every exported symbol of include/dfcsr.h
_SIGNATURES = [
    ("dfcsr_abi_version", c_i32, []),
    ("dfcsr_last_error", ctypes.c_char_p, []),
    ("dfcsr_moments_workspace_bytes", c_i64, [c_i64]),
    ("dfcsr_beam_moments", c_i32, [c_dp, c_dp, c_i64, c_dp, c_i64, c_dp, c_dp]),
    ("dfcsr_beam_statistics_workspace_bytes", c_i64, [c_i64]),
    ("dfcsr_beam_statistics", c_i32, [c_dp, c_dp, c_dp, c_dp, c_dp, c_dp, c_i64, c_f64, c_f64,
                                      c_dp, c_i64, c_dp, c_dp]),
    ("dfcsr_histogram_workspace_bytes", c_i64, [c_i32, c_i32]),
    ("dfcsr_histogram_cic_2d", c_i32, [c_dp, c_dp, c_dp, c_i64, c_i32, c_f64, c_f64, c_i32,
                                       c_f64, c_f64, c_dp, c_dp, c_dp, c_dp, c_i64, c_dp]),
    ("dfcsr_deposit_density_vx", c_i32, [c_dp, c_dp, c_dp, c_i64, c_i32, c_f64, c_f64, c_i32,
                                         c_f64, c_f64, c_dp, c_dp, c_dp, c_dp, c_dp, c_i64,
                                         c_dp]),
    ("dfcsr_df_workspace_bytes", c_i64, [c_i32, c_i32]),
    ("dfcsr_df_process", c_i32, [c_dp, c_dp, c_dp, c_dp, c_dp, c_i32, c_f64, c_f64, c_i32,
                                 c_f64, c_f64, c_f64, c_i32, c_i32, c_dp, c_dp, c_i64, c_dp]),
    ("dfcsr_savgol_operator", c_i32, [c_i32, c_i32, c_dp, c_dp]),
    ("dfcsr_sgolay2d", c_i32, [c_dp, c_i32, c_i32, c_i32, c_i32, c_i32, c_dp, c_dp, c_dp]),
    ("dfcsr_sgolay2d_kernels", c_i32, [c_i32, c_i32, c_i32, c_dp, c_dp]),
    ("dfcsr_regrid_workspace_bytes", c_i64, [c_i32]),
    ("dfcsr_regrid", c_i32, [c_dp, c_i64, ctypes.POINTER(RegridJob), c_i32, c_dp, c_dp, c_dp,
                             c_dp, c_dp, c_i32, c_i32, c_i32, c_f64, c_f64, c_f64, c_f64, c_dp,
                             c_i64, c_dp]),
    ("dfcsr_cells32_tdiff", c_i32, [c_dp, c_dp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, c_dp]),
    ("dfcsr_history_zspan", c_i32, [c_dp, c_dp, c_dp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32,
                                    c_dp]),
    ("dfcsr_interpolate3d", c_i32, [ctypes.POINTER(History), c_i32, c_dp, c_i64, c_dp, c_dp]),
    ("dfcsr_interpolate1d", c_i32, [ctypes.POINTER(Lattice), c_i32, c_dp, c_i64, c_dp, c_dp]),
    ("dfcsr_csr_integrand", c_i32, [ctypes.POINTER(History), ctypes.POINTER(Lattice), c_f64,
                                    c_f64, c_f64, c_dp, c_i32, c_dp, c_i32, c_dp, c_dp, c_dp]),
    ("dfcsr_wake_workspace_bytes", c_i64, [ctypes.POINTER(WakeParams)]),
    ("dfcsr_wake_mesh", c_i32, [ctypes.POINTER(History), ctypes.POINTER(Lattice),
                                ctypes.POINTER(WakeParams), c_i32, c_dp, c_dp, c_dp, c_dp, c_dp,
                                c_i64, c_dp]),
    ("dfcsr_wake_mesh_noalloc", c_i32, [ctypes.POINTER(History), ctypes.POINTER(Lattice),
                                        ctypes.POINTER(WakeParams), c_i32, c_dp, c_dp, c_dp,
                                        c_dp, c_dp]),
    ("dfcsr_wake_batched_workspace_bytes", c_i64, [c_i32, ctypes.POINTER(WakeParams)]),
    ("dfcsr_wake_task_table_bytes", c_i64, [c_i32]),
    ("dfcsr_wake_mesh_batched", c_i32, [c_i32, ctypes.POINTER(History), ctypes.POINTER(Lattice),
                                        ctypes.POINTER(WakeParams), c_i32,
                                        ctypes.POINTER(c_dp), ctypes.POINTER(c_dp), c_dp, c_i64,
                                        c_dp]),
    ("dfcsr_wake_tasks_pack", c_i32, [c_i32, ctypes.POINTER(History), ctypes.POINTER(Lattice),
                                      ctypes.POINTER(WakeParams), c_i32, ctypes.POINTER(c_dp),
                                      ctypes.POINTER(c_dp), c_dp, c_i64,
                                      ctypes.POINTER(WakeBatch)]),
    ("dfcsr_wake_batched_launch", c_i32, [c_dp, ctypes.POINTER(WakeBatch), c_i32, c_dp, c_i64,
                                          c_dp]),
    ("dfcsr_track_element", c_i32, [c_dp, c_dp, c_dp, c_dp, c_dp, c_dp, c_i64,
                                    ctypes.POINTER(Element), c_f64, c_f64, c_dp]),
    ("dfcsr_apply_wakes", c_i32, [c_dp, c_dp, c_dp, c_dp, c_i64, c_dp, c_dp, c_i32, c_i32, c_f64,
                                  c_f64, c_f64, c_f64, c_f64, c_f64, c_f64, c_f64, c_i32, c_dp]),
]

EXPORTED = tuple(name for name, _, _ in _SIGNATURES)

_lib = None


def load(path: str | None = None):
    """Loads and binds the library (once).  Raises DfcsrError if absent.
    DFCSR_LIB overrides the path (e.g. an alternative build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or os.environ.get("DFCSR_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise DfcsrError(f"{path} not found: build it first (python -c 'import __graft_entry__ as g; "
                         "g.build()' or make -C pydfcsr_amd/csrc)")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise DfcsrError(f"cannot load {path}: {e}") from e
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.dfcsr_abi_version() != ABI_VERSION:
        raise DfcsrError(f"ABI version mismatch: {lib.dfcsr_abi_version()}")
    if _lib is None:
        _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().dfcsr_last_error().decode(errors="replace")
        raise DfcsrError(f"{what} failed ({rc}): {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_device_tensor(t: torch.Tensor, dtype=torch.float64, name="tensor"):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise DfcsrError(f"{name} must be a HIP device tensor")
    if t.dtype != dtype:
        raise DfcsrError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise DfcsrError(f"{name} must be contiguous")
    return t


class Workspace:
    """Caller-owned scratch buffers, grown on demand and reused across calls."""

    def __init__(self, device):
        self.device = torch.device(device)
        self._bufs = {}

    def get(self, name: str, nbytes: int) -> torch.Tensor:
        nbytes = max(int(nbytes), 256)
        b = self._bufs.get(name)
        if b is None or b.numel() < nbytes:
            b = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._bufs[name] = b
        return b
