"""ctypes bindings of include/cgfd_core.h and include/cgfd_setup.h.

The product path is the HIP library ``_build/libcgfd.so``; there is no CPU
fallback: loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .build import LIB_PATH

MAX_FAULTS = 5
NCMP = 9
NFAULT_OUT = 11
NPG = 15
COMPAT_FIX_F2W_SWAP = 1
COMPAT_FIX_STATION_INTERP = 2

ERR = {0: "OK", -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_NOMEM", -4: "ERR_STATE", -5: "ERR_NAN", -6: "ERR_NODEV"}

VP = C.c_void_p
I32 = C.c_int
F32 = C.c_float


class FaultDesc(C.Structure):
    _fields_ = [
        ("fault_x_index", I32),
        ("T0x", VP), ("T0y", VP), ("T0z", VP), ("mu_s", VP), ("mu_d", VP), ("Dc", VP), ("C0", VP),
        ("united", VP), ("faultgrid", VP), ("rup_index_y", VP), ("rup_index_z", VP),
        ("Tn0", VP), ("Ts10", VP), ("Ts20", VP),
        ("rho_f", VP), ("vec_s1", VP), ("vec_s2", VP),
        ("D", VP * 12), ("matPlus2Min", VP * 5), ("matMin2Plus", VP * 5), ("matM2", VP * 6), ("matF", VP * 14),
    ]


class Desc(C.Structure):
    _fields_ = [
        ("ni", I32), ("nj", I32), ("nk", I32),
        ("metric", VP), ("lambda_", VP), ("mu", VP), ("rho", VP),
        ("free_top", I32),
        ("pml_is_sides", (I32 * 2) * 3),
        ("pml_ni1", (I32 * 2) * 3), ("pml_ni2", (I32 * 2) * 3),
        ("pml_nj1", (I32 * 2) * 3), ("pml_nj2", (I32 * 2) * 3),
        ("pml_nk1", (I32 * 2) * 3), ("pml_nk2", (I32 * 2) * 3),
        ("pml_A", (VP * 2) * 3), ("pml_B", (VP * 2) * 3), ("pml_D", (VP * 2) * 3),
        ("imethod", I32), ("number_fault", I32),
        ("fault", FaultDesc * MAX_FAULTS),
        ("dt", F32), ("t0", F32), ("nt_total", I32), ("device_id", I32),
        ("qc_check_nan_every", I32), ("compat_flags", I32),
    ]


class SetupParams(C.Structure):
    _fields_ = [
        ("ni", I32), ("nj", I32), ("nk", I32),
        ("x", VP), ("y", VP), ("z", VP),
        ("lambda_", VP), ("mu", VP), ("rho", VP),
        ("free_top", I32),
        ("pml_layers", (I32 * 2) * 3),
        ("pml_alpha_max", (F32 * 2) * 3), ("pml_beta_max", (F32 * 2) * 3), ("pml_velocity", (F32 * 2) * 3),
        ("imethod", I32), ("number_fault", I32),
        ("fault_x_index", I32 * MAX_FAULTS), ("fault_grid", I32 * (4 * MAX_FAULTS)),
        ("init_stress", (VP * 7) * MAX_FAULTS),
        ("united_margin", I32),
        ("dt", F32), ("t0", F32), ("nt_total", I32), ("device_id", I32),
        ("metric", VP), ("device_setup", I32),
    ]


STATION_NAME_MAX = 256


class Station(C.Structure):
    """cgfd_station (include/cgfd_io.h)"""
    _fields_ = [("name", C.c_char * STATION_NAME_MAX), ("i", I32), ("j", I32), ("k", I32),
                ("di", F32), ("dj", F32), ("dk", F32), ("x", F32), ("y", F32), ("z", F32)]


class FaultStation(C.Structure):
    """cgfd_fault_station (include/cgfd_io.h)"""
    _fields_ = [("name", C.c_char * STATION_NAME_MAX), ("fault_id", I32), ("j", I32), ("k", I32),
                ("dj", F32), ("dk", F32), ("x", F32), ("y", F32), ("z", F32)]


_LIB = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libcgfd.so (built by cgfd.build.build_core)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() "
                           "(the HIP extension is required; there is no CPU fallback)")
    lib = C.CDLL(str(p))
    sig = {
        "cgfd_last_error": (C.c_char_p, []),
        "cgfd_version": (C.c_char_p, []),
        "cgfd_core_create": (I32, [C.POINTER(Desc), C.POINTER(VP)]),
        "cgfd_core_destroy": (I32, [VP]),
        "cgfd_core_run": (I32, [VP, I32]),
        "cgfd_core_step_index": (I32, [VP]),
        "cgfd_core_sync": (I32, [VP]),
        "cgfd_core_read_wavefield": (I32, [VP, VP]),
        "cgfd_core_write_wavefield": (I32, [VP, VP]),
        "cgfd_core_read_fault": (I32, [VP, I32, VP]),
        "cgfd_core_read_fault_state": (I32, [VP, I32, VP]),
        "cgfd_core_write_fault_state": (I32, [VP, I32, VP, VP]),
        "cgfd_core_read_pg": (I32, [VP, VP]),
        "cgfd_core_set_stations": (I32, [VP, I32, VP, I32]),
        "cgfd_core_read_stations": (I32, [VP, VP, C.POINTER(I32)]),
        "cgfd_core_run_timed": (I32, [VP, I32, C.POINTER(F32), C.POINTER(F32), C.POINTER(I32)]),
        "cgfd_core_set_variant": (I32, [VP, I32]),
        "cgfd_core_set_zchunk": (I32, [VP, I32]),
        "cgfd_core_set_vertical_columns": (I32, [VP, I32]),
        "cgfd_core_get_vertical_columns": (I32, [VP, C.POINTER(I32), C.POINTER(I32)]),
        "cgfd_core_set_graph": (I32, [VP, I32]),
        "cgfd_core_set_overlap": (I32, [VP, I32]),
        "cgfd_core_state_bytes": (I32, [VP, C.POINTER(C.c_size_t)]),
        "cgfd_adj_backward": (I32, [VP, VP, I32, VP]),
        "cgfd_adj_forward": (I32, [VP, I32]),
        "cgfd_batch_adj_backward": (I32, [VP, I32, VP, I32, VP, VP]),
        "cgfd_adj_backward_fault": (I32, [VP, VP, VP, I32, VP, VP]),
        "cgfd_core_set_fault_receivers": (I32, [VP, I32, VP, I32]),
        "cgfd_core_read_slice": (I32, [VP, I32, I32, VP]),
        "cgfd_io_fault_create": (I32, [VP, I32, C.c_char_p, C.c_char_p, VP]),
        "cgfd_io_fault_put": (I32, [VP, VP, I32, C.c_float]),
        "cgfd_io_fault_end": (I32, [VP, VP]),
        "cgfd_io_slice_create": (I32, [VP, I32, I32, C.c_char_p, C.c_char_p, VP]),
        "cgfd_io_slice_put": (I32, [VP, VP, I32, C.c_float]),
        "cgfd_io_close": (I32, [VP]),
        "cgfd_io_snap_create": (I32, [VP, C.c_char_p, VP, I32, C.c_char_p, C.c_char_p, VP]),
        "cgfd_io_snap_put": (I32, [VP, VP, I32, C.c_float]),
        "cgfd_io_pg_write": (I32, [VP, C.c_char_p, C.c_char_p]),
        "cgfd_io_stations_sac": (I32, [VP, C.c_char_p, VP, VP, I32]),
        "cgfd_io_read_init_stress": (I32, [C.c_char_p, I32, I32, VP]),
        "cgfd_io_read_fault_coord": (I32, [C.c_char_p, I32, I32, VP]),
        "cgfd_core_read_snapshot": (I32, [VP, VP, I32, VP]),
        "cgfd_batch_read_snapshot": (I32, [VP, I32, VP, I32, VP]),
        "cgfd_core_read_fault_receivers": (I32, [VP, VP, VP]),
        "cgfd_core_save_state": (I32, [VP, VP, I32]),
        "cgfd_core_load_state": (I32, [VP, VP, I32]),
        "cgfd_core_autotune": (I32, [VP, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(F32)]),
        "cgfd_calib_stream": (I32, [I32, I32, I32, I32, I32, I32, I32, C.POINTER(F32)]),
        "cgfd_core_set_stations_interp": (I32, [VP, I32, VP, VP, I32]),
        "cgfd_core_set_fault_receivers_interp": (I32, [VP, I32, VP, VP, I32]),
        "cgfd_io_read_stations": (I32, [C.c_char_p, VP, VP, I32, C.POINTER(I32), C.POINTER(I32)]),
        "cgfd_io_read_fault_stations": (I32, [C.c_char_p, VP, VP, I32, C.POINTER(I32)]),
        "cgfd_io_line_stations": (I32, [VP, C.c_char_p, VP, VP, I32, VP, C.POINTER(I32)]),
        "cgfd_core_get_variant": (I32, [VP, C.POINTER(I32), C.POINTER(I32)]),
        "cgfd_core_read_snapshot_noalloc": (I32, [VP, VP, I32, VP]),
        "cgfd_drv_rk_curv_col_allstep": (I32, [C.POINTER(Desc), VP, VP, VP]),
        "cgfd_batch_create": (I32, [C.POINTER(Desc), I32, C.POINTER(VP)]),
        "cgfd_batch_active": (I32, [VP]),
        "cgfd_batch_steps": (I32, [VP, I32]),
        "cgfd_batch_read_wavefield": (I32, [VP, I32, VP]),
        "cgfd_batch_write_wavefield": (I32, [VP, I32, VP]),
        "cgfd_batch_read_fault": (I32, [VP, I32, I32, VP]),
        "cgfd_batch_read_fault_state": (I32, [VP, I32, I32, VP]),
        "cgfd_batch_write_fault_state": (I32, [VP, I32, I32, VP, VP]),
        "cgfd_batch_read_pg": (I32, [VP, I32, VP]),
        "cgfd_batch_read_stations": (I32, [VP, I32, VP, C.POINTER(I32)]),
        "cgfd_batch_read_fault_receivers": (I32, [VP, I32, VP, C.POINTER(I32)]),
        "cgfd_setup_create": (I32, [C.POINTER(SetupParams), C.POINTER(VP)]),
        "cgfd_setup_destroy": (I32, [VP]),
        "cgfd_setup_get_desc": (I32, [VP, C.POINTER(Desc)]),
        "cgfd_setup_scenario_desc": (I32, [VP, C.POINTER(VP), I32, C.POINTER(Desc)]),
        "cgfd_setup_coords": (C.POINTER(F32), [VP]),
        "cgfd_setup_vec_n": (C.POINTER(F32), [VP, I32]),
        "cgfd_setup_metric": (I32, [I32, I32, I32, VP, VP, VP, VP]),
        "cgfd_setup_dt_max": (I32, [VP, C.POINTER(F32)]),
        "cgfd_pml_cal_R": (F32, [F32]),
        "cgfd_pml_cal_dmax": (F32, [F32, F32, F32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


class CgfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR.get(code, code)}: {msg}")
        self.code = code


def check(lib: C.CDLL, rc: int) -> None:
    if rc != 0:
        raise CgfdError(rc, lib.cgfd_last_error().decode())


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(VP)
