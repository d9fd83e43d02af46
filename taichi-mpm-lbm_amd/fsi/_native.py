"""ctypes binding of the gfx950 kernel library (include/fsi_hip.h).

The library is loaded after torch so that both share torch's HIP runtime.
There is no fallback: a missing library or a missing GPU raises RuntimeError
at solver construction, before any state is allocated.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libfsi_hip.so"
# include/fsi_hip.h FSI_ABI_VERSION the structures below follow
ABI_VERSION = 5
_lib = None

POP_F32, POP_F16, POP_F64, POP_I16 = 0, 1, 2, 3
PHASE_STREAM_COLLIDE, PHASE_COLLIDE, PHASE_STREAM_MACRO = 0, 1, 2
REAL_F32, REAL_F64 = 0, 1
BC_MODES = {"periodic": 0, "pressure": 1, "velocity": 2}
STORAGE_CODES = {"fp32": POP_F32, "fp16": POP_F16, "fp64": POP_F64, "i16": POP_I16}
# floats of the I16 storage control block (FSI_STORE_CTRL_FLOATS), and the
# layout the host reads it with: four slots (three write epochs, one
# measuring pass) of 65 lines of 32 floats, the scale at the start of a slot,
# the 64 partial maxima one per following line
STORE_CTRL_FLOATS, STORE_CTRL_SLOT, STORE_CTRL_LINE = 8320, 2080, 32

c_double3 = ctypes.c_double * 3
_vp = ctypes.c_void_p


class LbmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("rho0", ctypes.c_double),
        ("initial_velocity", c_double3),
        ("s_v", ctypes.c_double), ("s_o", ctypes.c_double),
        ("left_mode", ctypes.c_int32), ("right_mode", ctypes.c_int32),
        ("bc_rho", ctypes.c_double * 2),
        ("bc_vel", c_double3 * 2),
        ("pop_scale", ctypes.c_double),
        ("store_epoch", ctypes.c_int32),
        ("store_pass", ctypes.c_int32),
    ]


class LbmBuffers(ctypes.Structure):
    _fields_ = [("solid", _vp), ("nbmask", _vp), ("rho", _vp), ("vel", _vp), ("force", _vp),
                ("face_vel", _vp), ("store_ctrl", _vp), ("solid_bits", _vp)]


def solid_bits_words(nx: int, ny: int, nz: int) -> int:
    """Words of LbmBuffers.solid_bits (FSI_SOLID_BITS_WORDS)."""
    return nx * ny * (nz // 32 + 2)


class LbmForce(ctypes.Structure):
    _fields_ = [("field", _vp), ("uniform", c_double3)]


class MpmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("real", ctypes.c_int32), ("num_particles", ctypes.c_int32), ("capacity", ctypes.c_int32),
        ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("two_mu", ctypes.c_double), ("lam", ctypes.c_double),
        ("gravity", c_double3),
        ("boundary_width", ctypes.c_int32),
        ("boundary_damping", ctypes.c_double),
        ("apic", ctypes.c_int32),
        ("clamp_lower", ctypes.c_double), ("clamp_upper", c_double3),
        ("num_chunks", ctypes.c_int32), ("split_max_chunks", ctypes.c_int32),
        ("wide_max_workgroups", ctypes.c_int32),
    ]


class MpmBuffers(ctypes.Structure):
    _fields_ = [("x", _vp), ("v", _vp), ("C", _vp), ("F", _vp), ("mass", _vp), ("volume", _vp),
                ("active", _vp), ("pforce", _vp), ("grid", _vp), ("grid_alt", _vp), ("chunks", _vp),
                ("zero_boxes", _vp), ("zero_boxes_cap", ctypes.c_int32), ("grid_alt2", _vp),
                ("material", _vp)]


class CouplingParams(ctypes.Structure):
    _fields_ = [
        ("enabled", ctypes.c_int32), ("build_fraction", ctypes.c_int32),
        ("gamma_eff", ctypes.c_double), ("dt_ratio", ctypes.c_double),
        ("cell_volume", ctypes.c_double), ("min_valid_weight", ctypes.c_double),
        ("use_relative_velocity_limit", ctypes.c_int32), ("relative_velocity_limit", ctypes.c_double),
        ("use_force_limit", ctypes.c_int32), ("force_limit", ctypes.c_double),
        ("contact_enabled", ctypes.c_int32), ("contact_threshold", ctypes.c_double),
        ("contact_damping", ctypes.c_double), ("refine", ctypes.c_int32),
    ]


class CouplingBuffers(ctypes.Structure):
    _fields_ = [("solid", _vp), ("lbm_vel", _vp), ("cforce", _vp), ("valid_weight", _vp),
                ("coupling_mask", _vp), ("contact_mask", _vp), ("counters", _vp), ("closure_force", _vp),
                ("counter_rows", _vp), ("counter_rows_cap", ctypes.c_int32), ("force_mean", _vp)]


class ForceAssembly(ctypes.Structure):
    _fields_ = [
        ("coupling", ctypes.c_int32), ("clamp_fraction", ctypes.c_int32),
        ("ib_enabled", ctypes.c_int32),
        ("ib_drag", ctypes.c_double), ("ib_threshold", ctypes.c_double),
        ("ib_max_force", ctypes.c_double), ("ib_use_max", ctypes.c_int32),
        ("base_force", c_double3),
    ]


# name -> argument types; every entry point returns int
_SIGNATURES = {
    "fsi_lbm_init": [_vp, _vp, _vp, _vp, _vp],
    "fsi_lbm_step": [_vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp],
    "fsi_lbm_force_op": [_vp, _vp, ctypes.c_int32, _vp, _vp, _vp, _vp],
    "fsi_lbm_force_import": [_vp, _vp, _vp, _vp],
    "fsi_lbm_export_populations": [_vp, _vp, _vp, _vp, _vp],
    "fsi_soa_to_aos": [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp],
    "fsi_aos_to_soa": [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp],
    "fsi_mpm_clear_grid": [_vp, _vp, _vp],
    "fsi_mpm_p2g": [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp],
    "fsi_mpm_grid_g2p": [_vp, _vp, _vp, _vp, _vp],
    "fsi_mpm_substeps": [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp],
    "fsi_mpm_substeps_noalloc": [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp],
    "fsi_contact": [_vp, _vp, _vp, _vp, _vp],
    "fsi_polar_rotation": [ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp],
    "fsi_force_assemble": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp],
    "fsi_lbm_collide_assembled": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "fsi_mpm_sort": [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_int32, _vp, _vp],
    "fsi_reduce_lattice": [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp],
    "fsi_reduce_particles": [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp],
    "fsi_abi_version": [],
    "fsi_fill_zero": [ctypes.c_int64, _vp, _vp],
    "fsi_closure_features": [_vp, _vp, _vp, _vp, _vp],
    "fsi_mlp_train_step": [ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp],
    "fsi_mlp_adam": [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                     ctypes.c_float, _vp],
    "fsi_mlp_forward": [ctypes.c_int32, _vp, _vp, _vp, _vp],
    "fsi_mlp_infer": [ctypes.c_int32, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp],
}

# learned-closure MLP shapes (FSI_MLP_* of the header)
MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS, MLP_NORM = 18, 64, 3, 7312, 84


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> ctypes.CDLL:
    """Load the kernel library (no GPU needed to load it)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"fsi kernel library not built: {_LIB_PATH} is missing "
                "(run `python taichi-mpm-lbm_amd/build.py`)"
            )
        lib = ctypes.CDLL(str(_LIB_PATH))
        for name, args in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.fsi_mpm_sort_workspace.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
        lib.fsi_mpm_sort_workspace.restype = ctypes.c_int
        lib.fsi_mlp_train_workspace.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
        lib.fsi_mlp_train_workspace.restype = ctypes.c_int
        lib.fsi_last_error.argtypes = []
        lib.fsi_last_error.restype = ctypes.c_char_p
        if lib.fsi_abi_version() != ABI_VERSION:
            raise RuntimeError(
                f"{_LIB_PATH} implements ABI version {lib.fsi_abi_version()}, the bindings expect "
                f"{ABI_VERSION}: rebuild it (`python taichi-mpm-lbm_amd/build.py`)"
            )
        _lib = lib
    return _lib


def device() -> torch.device:
    """The GPU the solvers run on; raises when there is none."""
    if not torch.cuda.is_available():
        raise RuntimeError("fsi solvers need a ROCm GPU; torch.cuda.is_available() is False")
    load_library()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def ref(s: ctypes.Structure) -> ctypes.c_void_p:
    """Address of a parameter block (the caller keeps `s` alive for the call)."""
    return ctypes.c_void_p(ctypes.addressof(s))


def call(name: str, *args) -> None:
    """Call an entry point on the current stream; map failures to RuntimeError."""
    lib = load_library()
    rc = getattr(lib, name)(*args, stream_handle())
    if rc != 0:
        msg = lib.fsi_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")


# doubles per workgroup row of a reduction workspace (FSI_SUM_SLOTS) and the
# workspace size the solvers allocate: 1024 rows
SUM_SLOTS = 16
SUM_WORKSPACE = 1024 * SUM_SLOTS


def reduction_buffers(dev: torch.device) -> tuple[torch.Tensor, torch.Tensor]:
    """(workspace, out) of the FP64 reductions: 128 KiB and 16 doubles."""
    return (torch.empty(SUM_WORKSPACE, dtype=torch.float64, device=dev),
            torch.zeros(SUM_SLOTS, dtype=torch.float64, device=dev))


def fill_zero(t: torch.Tensor) -> None:
    call("fsi_fill_zero", ctypes.c_int64(t.numel() * t.element_size()), ptr(t))
