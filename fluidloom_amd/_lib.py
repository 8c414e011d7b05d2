"""ctypes binding of libfluidloom_hip.so (the product library).

The binding fails loudly when the library has not been built: there is no
CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libfluidloom_hip.so"

FL_OK = 0
FL_E_INVALID = 1
FL_E_OOM = 2
FL_E_HIP = 3
FL_E_MESH = 4
FL_E_CAPACITY = 5

FL_BC_WALL = 0
FL_BC_OUTFLOW = 1
FL_BC_PERIODIC = 2
FL_STORE_F32 = 0
FL_STORE_F16 = 1
FL_STORE_F64 = 2
FL_STORE_F16C = 3

vp = C.c_void_p
sz = C.c_size_t
i32 = C.c_int32
u64 = C.c_uint64


class FluidLoomError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"fluidloom status {status}: {msg}")
        self.status = status


class Cells(C.Structure):
    _fields_ = [("x", vp), ("y", vp), ("z", vp), ("level", vp), ("state", vp), ("flag", vp),
                ("material", vp)]


class RegridConfig(C.Structure):
    _fields_ = [("balance_mode", C.c_int), ("anchor_level", C.c_int), ("max_split_level", C.c_int),
                ("max_balance_iters", C.c_int), ("root", i32 * 3),
                ("merge_balance_check", C.c_int), ("ncomp", C.c_int), ("avg_rule", C.c_int)]


class RegridResult(C.Structure):
    _fields_ = [("num_cells_out", u64), ("num_split", u64), ("num_children", u64),
                ("num_groups", u64), ("balance_iterations", u64), ("balance_violations", u64),
                ("balance_marked", u64), ("balance_converged", i32), ("pad", i32)]


class LbmParams(C.Structure):
    _fields_ = [("nx0", i32), ("ny0", i32), ("nz0", i32), ("max_level", i32), ("tau0", C.c_double),
                ("bc", i32 * 6), ("bc_u", (C.c_double * 3) * 6), ("rho0", C.c_double),
                ("storage", i32), ("pad", i32), ("rho_init", C.c_double),
                ("u_init", C.c_double * 3)]


class LbmStats(C.Structure):
    _fields_ = [("leaves", u64 * 24), ("bricks", u64 * 24), ("active_bricks", u64 * 24),
                ("c2f", u64 * 24), ("f2c", u64 * 24), ("capture", u64 * 24), ("nlevels", i32),
                ("pad", i32), ("updates_per_step", u64), ("device_bytes", u64),
                ("variant", i32 * 24), ("ninst", i32), ("graph", i32), ("graph_launches", u64)]


class LbmLayout(C.Structure):
    _fields_ = [(nm, C.c_void_p) for nm in ("nbr", "bcoord", "binst", "bflags", "cflags", "c2f_gchild", "c2f_gcap",
                                            "c2f_gmask", "cap_slots", "f2c_ghost", "f2c_child", "force_links",
                                            "abeg", "lbeg")] + \
               [(nm, C.c_uint64) for nm in ("nb", "nactive", "n_groups", "n_cap", "n_f2c", "n_links")]


_SIGS = {
    "fl_last_error_string": (C.c_char_p, []),
    "fl_version": (C.c_char_p, []),
    "fl_init": (C.c_int, [C.c_int]),
    "fl_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "fl_device_info": (C.c_int, [C.c_int, C.c_char_p, sz, C.POINTER(sz), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
    "fl_malloc": (C.c_int, [C.POINTER(vp), sz]),
    "fl_free": (C.c_int, [vp]),
    "fl_memcpy_h2d": (C.c_int, [vp, vp, sz, vp]),
    "fl_memcpy_d2h": (C.c_int, [vp, vp, sz, vp]),
    "fl_memcpy_d2d": (C.c_int, [vp, vp, sz, vp]),
    "fl_memset": (C.c_int, [vp, C.c_int, sz, vp]),
    "fl_stream_create": (C.c_int, [C.POINTER(vp)]),
    "fl_stream_destroy": (C.c_int, [vp]),
    "fl_stream_sync": (C.c_int, [vp]),
    "fl_device_sync": (C.c_int, []),
    "fl_release_cached_memory": (C.c_int, []),
    "fl_cached_memory_bytes": (C.c_int, [C.POINTER(C.c_uint64)]),
    "fl_event_create": (C.c_int, [C.POINTER(vp)]),
    "fl_event_destroy": (C.c_int, [vp]),
    "fl_event_record": (C.c_int, [vp, vp]),
    "fl_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "fl_mem_info": (C.c_int, [C.POINTER(sz), C.POINTER(sz)]),
    "fl_copy_kernel": (C.c_int, [vp, vp, sz, vp]),
    "fl_hilbert_encode": (C.c_int, [vp, vp, vp, vp, C.c_int, vp, sz, vp]),
    "fl_hilbert_decode": (C.c_int, [vp, vp, C.c_int, vp, vp, vp, sz, vp]),
    "fl_hilbert_encode_aos4": (C.c_int, [vp, C.c_int, vp, sz, vp]),
    "fl_hilbert_decode_aos4": (C.c_int, [vp, C.c_int, vp, sz, vp]),
    "fl_hilbert_key63": (C.c_int, [vp, vp, vp, vp, C.c_int, vp, sz, vp]),
    "fl_cell_hkey": (C.c_int, [vp, vp, vp, vp, vp, sz, vp]),
    "fl_rotdir_encode": (C.c_int, [vp, vp, vp, C.c_int, vp, sz, vp]),
    "fl_sort_temp_bytes": (sz, [sz]),
    "fl_sort_pairs_u64u32": (C.c_int, [vp, vp, vp, vp, sz, C.c_int, C.c_int, vp, sz, vp]),
    "fl_iota_u32": (C.c_int, [vp, sz, vp]),
    "fl_scan_temp_bytes": (sz, [sz]),
    "fl_scan_exclusive_u32": (C.c_int, [vp, vp, vp, sz, vp, sz, vp]),
    "fl_scan_noalloc_max_elements": (sz, []),
    "fl_scan_exclusive_u32_noalloc": (C.c_int, [vp, vp, vp, sz, vp]),
    "fl_hash_capacity": (u64, [sz]),
    "fl_hash_clear": (C.c_int, [vp, vp, u64, vp]),
    "fl_hash_build": (C.c_int, [vp, vp, u64, vp, vp, sz, vp, vp]),
    "fl_hash_query": (C.c_int, [vp, vp, u64, vp, vp, sz, vp, vp]),
    "fl_face_neighbors": (C.c_int, [C.POINTER(Cells), sz, C.POINTER(i32), vp, vp, vp, u64, vp, vp]),
    "fl_regrid_temp_bytes": (sz, [sz]),
    "fl_regrid": (C.c_int, [C.POINTER(RegridConfig), C.POINTER(Cells), sz, C.POINTER(Cells), sz, vp,
                            vp, vp, sz, C.POINTER(RegridResult), vp]),
    "fl_balance": (C.c_int, [C.POINTER(RegridConfig), C.POINTER(Cells), sz, vp, sz,
                             C.POINTER(RegridResult), vp]),
    "fl_split": (C.c_int, [C.POINTER(Cells), sz, C.c_int, vp, C.POINTER(Cells), sz, vp, sz,
                           C.POINTER(u64), vp]),
    "fl_merge": (C.c_int, [C.POINTER(Cells), sz, vp, vp, C.POINTER(Cells), sz, vp, sz,
                           C.POINTER(u64), vp]),
    "fl_gen_cells": (C.c_int, [C.POINTER(Cells), sz, u64, C.c_int, vp]),
    "fl_gen_flags": (C.c_int, [C.POINTER(Cells), sz, u64, C.c_double, C.c_double, vp, sz, vp]),
    "fl_mark_ball": (C.c_int, [C.POINTER(Cells), sz, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_int, C.POINTER(u64), vp]),
    "fl_mark_solid_ball": (C.c_int, [C.POINTER(Cells), sz, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.POINTER(u64), vp]),
    "fl_pack_fields": (C.c_int, [vp, vp, sz, vp, sz, vp]),
    "fl_unpack_fields": (C.c_int, [vp, sz, vp, sz, vp, vp]),
    "fl_ghost_fill_c2f": (C.c_int, [vp, vp, sz, C.c_int, vp, vp]),
    "fl_ghost_fill_f2c": (C.c_int, [vp, vp, sz, C.c_int, C.c_int, vp, vp]),
    "fl_lbm_create": (C.c_int, [C.POINTER(LbmParams), vp, vp, vp, vp, vp, sz, vp, C.POINTER(vp)]),
    "fl_lbm_destroy": (C.c_int, [vp]),
    "fl_lbm_step": (C.c_int, [vp, C.c_int]),
    "fl_lbm_stats_get": (C.c_int, [vp, C.POINTER(LbmStats)]),
    "fl_lbm_get_f": (C.c_int, [vp, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_set_f": (C.c_int, [vp, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_moments": (C.c_int, [vp, vp, vp, vp, vp, sz, vp, vp]),
    "fl_lbm_vorticity": (C.c_int, [vp, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_regrid": (C.c_int, [vp, vp, vp, vp, vp, vp, sz]),
    "fl_lbm_totals": (C.c_int, [vp, vp, vp]),
    "fl_lbm_force": (C.c_int, [vp, vp]),
    "fl_lbm_stream": (C.c_int, [vp, C.POINTER(vp)]),
    "fl_lbm_time_collide": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "fl_lbm_create_batch": (C.c_int, [C.POINTER(LbmParams), C.c_int, vp, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "fl_lbm_plan_bytes": (C.c_int, [C.POINTER(LbmParams), C.c_int, vp, vp, vp, vp, vp, vp, C.POINTER(u64)]),
    "fl_lbm_num_instances": (C.c_int, [vp, C.POINTER(i32)]),
    "fl_lbm_inst_leaves": (C.c_int, [vp, C.c_int, vp, C.POINTER(u64)]),
    "fl_lbm_get_f_inst": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_set_f_inst": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_moments_inst": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, sz, vp, vp]),
    "fl_lbm_vorticity_inst": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, sz, vp]),
    "fl_lbm_regrid_batch": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "fl_lbm_regrid_batch_device": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "fl_lbm_create_batch_device": (C.c_int, [C.POINTER(LbmParams), C.c_int, vp, vp, vp, vp, vp, vp, vp,
                                             C.POINTER(vp)]),
    "fl_lbm_layout_get": (C.c_int, [vp, C.c_int, C.POINTER(LbmLayout)]),
    "fl_lbm_layout_on_host": (C.c_int, [C.c_int]),
    "fl_lbm_totals_batch": (C.c_int, [vp, vp, vp]),
    "fl_lbm_force_batch": (C.c_int, [vp, vp]),
    "fl_lbm_velocity_change": (C.c_int, [vp, vp]),
    "fl_lbm_velocity_sample": (C.c_int, [vp]),
    "fl_lbm_set_live": (C.c_int, [vp, C.c_int, C.c_int]),
    "fl_lbm_set_graph": (C.c_int, [vp, C.c_int]),
    "fl_lbm_num_variants": (C.c_int, []),
    "fl_lbm_variant_name": (C.c_char_p, [C.c_int]),
    "fl_lbm_set_variant": (C.c_int, [vp, C.c_int, C.c_int]),
    "fl_lbm_autotune": (C.c_int, [vp, C.c_int, vp]),
    "fl_lbm_time_collide_variant": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
}

_lib = None


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib() -> C.CDLL:
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m fluidloom_amd.build` (no CPU fallback exists)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != FL_OK:
        raise FluidLoomError(status, lib().fl_last_error_string().decode())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().fl_device_count(C.byref(n)))
    return n.value


class DeviceArray:
    """A device allocation from fl_malloc with numpy-style dtype/shape."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = (shape,) if isinstance(shape, (int, np.integer)) else tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = vp()
        if self.nbytes:
            check(lib().fl_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().fl_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes, stream))
        return d

    def to_host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().fl_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def zero(self, stream=None):
        if self.nbytes:
            check(lib().fl_memset(self.ptr, 0, self.nbytes, stream))
        return self

    @property
    def p(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            lib().fl_free(self.ptr)
            self.ptr = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
