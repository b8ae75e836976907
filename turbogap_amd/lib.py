"""ctypes binding of libtgx.so (the C ABI declared in include/tgx.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
turbogap_amd/csrc``) into ``turbogap_amd/_lib/libtgx.so``. Loading fails loudly
when it is missing: there is no CPU fallback for any compute entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "_lib" / "libtgx.so"

TGX_OK = 0
TGX_PATH_FP64_PLAIN = 0
TGX_PATH_FP64_ADJOINT = 1
TGX_PATH_MIXED = 2
TGX_PATH_SURROGATE = 3

PATHS = {
    "fp64_plain": TGX_PATH_FP64_PLAIN,
    "fp64_adjoint": TGX_PATH_FP64_ADJOINT,
    "mixed": TGX_PATH_MIXED,
    "surrogate": TGX_PATH_SURROGATE,
}


class TgxError(RuntimeError):
    """A libtgx call returned a non-zero status."""


class SoapHypersC(C.Structure):
    """tgx_soap_hypers (include/tgx.h)."""

    _fields_ = [
        ("n_species", C.c_int),
        ("l_max", C.c_int),
        ("basis", C.c_int),
        ("radial_enhancement", C.c_int),
        ("n_quad_soft", C.c_int),
        ("n_quad_buffer", C.c_int),
        ("alpha_max", C.POINTER(C.c_int)),
        ("rcut_hard", C.POINTER(C.c_double)),
        ("rcut_soft", C.POINTER(C.c_double)),
        ("nf", C.POINTER(C.c_double)),
        ("global_scaling", C.POINTER(C.c_double)),
        ("atom_sigma_r", C.POINTER(C.c_double)),
        ("atom_sigma_r_scaling", C.POINTER(C.c_double)),
        ("atom_sigma_t", C.POINTER(C.c_double)),
        ("atom_sigma_t_scaling", C.POINTER(C.c_double)),
        ("amplitude_scaling", C.POINTER(C.c_double)),
        ("central_weight", C.POINTER(C.c_double)),
    ]


_lib = None

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_I64 = C.c_int64
_U64 = C.c_uint64
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PF = C.POINTER(C.c_float)

_SIGS = {
    "tgx_init": [_I, C.POINTER(_P)],
    "tgx_finalize": [_P],
    "tgx_memory_budget": [_P, _PD, _PD],
    "tgx_device_name": [_P, C.c_char_p, _I],
    "tgx_set_system": [_P, _I, _I, _PD, _PI, _PD],
    "tgx_set_positions": [_P, _PD],
    "tgx_set_positions_device": [_P, _P],
    "tgx_build_neighbors_list": [_P, _D, _PI, _PI, C.POINTER(_I64)],
    "tgx_refresh_neighbors_list": [_P],
    "tgx_get_neighbors_list": [_P, _PI, _PI, _PD, _PD, _PD, _PD, _PI],
    "tgx_soap_model_create": [_P, C.POINTER(SoapHypersC), _PI, _I, _D, _D, _D, _I, _I, _PD, _PD, _PI],
    "tgx_soap_model_create_compressed": [_P, C.POINTER(SoapHypersC), _PI, _I, _D, _D, _D, _I, _I, _PD, _PD, _I, _PI,
                                         _PI, _PD, _PI],
    "tgx_soap_backend_begin": [_P, _I],
    "tgx_soap_backend_end": [_P],
    "tgx_soap_model_dim": [_P, _I, _PI],
    "tgx_get_gap_soap": [_P, _I, _I, _PD, _PD, _PD, _PI, C.POINTER(_I64)],
    "tgx_soap_block_size": [_P, _PI, C.POINTER(_I64)],
    "tgx_local_property_predict": [_P, _I, _PD, _PD, _PI, C.POINTER(_I64)],
    "tgx_get_soap": [_P, _PD, _PD],
    "tgx_set_memory_budget": [_P, _D],
    "tgx_get_soap_list": [_P, _I, _PI, _I64, _PI, _PD, _PD, _I, _PD, _PD],
    "tgx_device_count": [_PI],
    "tgx_get_gap_soap_sites": [_P, _I, _I, _I, _I, _PD, _PD, _PD, _PI, C.POINTER(_I64)],
    "tgx_get_soap_energy_and_forces": [_P, _I, _I, _I64, _PD, _PD, _D, _D, _D, _PI, _PI, _PD, _I, _I, _PD, _PD, _PD],
    "tgx_get_soap_energy_and_forces_noalloc": [_P, _I, _I, _I64, _P, _P, _D, _D, _D, _P, _P, _P, _I, _I, _P, _P,
                                               _P],
    "tgx_compute": [_P, _I, _PD, _PD],
    "tgx_get_results": [_P, _PD, _PD],
    "tgx_results_device": [_P, C.POINTER(_P), C.POINTER(_P)],
    "tgx_mlp_create": [_P, _I, _I, _I, _U64, C.POINTER(_P)],
    "tgx_mlp_destroy": [_P],
    "tgx_mlp_add_samples": [_P, _I, _I, _I, _PI],
    "tgx_mlp_dataset_stats": [_P, _I, _PD, _PD],
    "tgx_mlp_set_normalization": [_P, _D, _D],
    "tgx_mlp_get_normalization": [_P, _PD, _PD],
    "tgx_mlp_train": [_P, _I, _I, _I, C.c_float, C.c_float, C.c_float, _U64, _PD, _PD],
    "tgx_mlp_evaluate": [_P, _I, _I, _PD, _PD, _PD],
    "tgx_mlp_forward_host": [_P, _I, _PF, _PF, _PF],
    "tgx_mlp_gradients_host": [_P, _I, _PF, _PF, _PF, C.c_float, C.c_float, _PF, _PF, _PF, _PF, _PF, _PF, _PD, _PD],
    "tgx_mlp_force_gradients": [_P, _I, C.POINTER(_P), _PI, _D, _D, _PD, _PD],
    "tgx_mlp_get_gradients": [_P, _PF, _PF, _PF, _PF, _PF, _PF],
    "tgx_mlp_adam_step": [_P, C.c_float],
    "tgx_mlp_get_weights": [_P, _PF, _PF, _PF, _PF, _PF, _PF],
    "tgx_mlp_set_weights": [_P, _PF, _PF, _PF, _PF, _PF, _PF],
    "tgx_surrogate_install": [_P, _I, _I, _P],
    "tgx_surrogate_gate": [_P, _I, _D],
    "tgx_surrogate_stats": [_P, _PD, _PD],
    "tgx_md_set_state": [_P, _PD, _PD],
    "tgx_md_get_state": [_P, _PD, _PD],
    "tgx_velocity_verlet": [_P, _D, _I],
    "tgx_kinetic": [_P, _PD, _PD],
    "tgx_berendsen": [_P, _D, _D, _D, _D],
    "tgx_remove_cm_vel": [_P],
    "tgx_wrap_pbc": [_P],
    "tgx_md_displacement": [_P, _D, _I, _PI],
    "tgx_md_set_fixed": [_P, _PI],
    "tgx_md_advance": [_P, _D, _I, _I, _D, _D, _D, _PD, _PD, _PI],
    "tgx_box_scaling": [_P, _PD, _I, _I, _PD],
    "tgx_bussi_thermostat": [_P, _D, _D, _D, _D, _D, _PD],
    "tgx_gradient_descent": [_P, _D, _D, _I, _PD, _PI],
    "tgx_variable_time_step": [_P, _I, _D, _D, _D, _PD],
    "tgx_adaptive_time_step": [_P, _I, _D, _D, _D, _D, _D, _PD],
    "tgx_md_pressure": [_P, _PD, _D, _PD, _PD],
    "tgx_berendsen_barostat": [_P, _D, _PD, C.c_char_p, _D, _D, _D, _PD],
    "tgx_distance_2b_create": [_P, _I, _I, _D, _D, _D, _D, _D, _I, _PD, _PD, _PD, _PI],
    "tgx_core_pot_create": [_P, _I, _I, _I, _PD, _PD, _D, _D, _D, _D, _PI],
    "tgx_angle_3b_create": [_P, _I, _I, _I, _D, _D, _D, _PD, _D, C.c_char_p, _I, _PD, _PD, _PD, _PI],
    "tgx_pair_energy_and_forces": [_P, _I, _PD, _PD, _PD],
    "tgx_set_profiling": [_P, _I],
    "tgx_set_block_streams": [_P, _I],
    "tgx_kernel_time": [_P, C.c_char_p, _PI, _PD],
    "tgx_reset_profile": [_P],
    "tgx_synchronize": [_P],
    "tgx_hbm_probe": [_P, C.c_size_t, _I, _PD],
    "tgx_number_of_unit_cells": [_PD, _D, _PI],
    "tgx_build_supercell": [_I, _PD, _PI, _PD, _PI, _PD, _PI, _PD],
    "tgx_generate_rsa": [_I, _PD, _D, _U64, _PD],
    "tgx_generate_permutation": [_I, _U64, _PI],
    "tgx_generate_normal": [_I, _D, _U64, _PD],
}


def load():
    """Load libtgx.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("TGX_LIB", str(LIB_PATH))
    if not os.path.exists(path):
        raise ImportError(
            f"libtgx.so not found at {path}: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C turbogap_amd/csrc`"
        )
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.tgx_last_error.argtypes = []
    lib.tgx_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != TGX_OK:
        msg = load().tgx_last_error().decode(errors="replace")
        raise TgxError(f"{what or 'libtgx'} failed with status {status}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def dptr(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_PI) if a is not None else None


def fptr(a):
    return a.ctypes.data_as(_PF) if a is not None else None
