"""Loads csrc/libswl.so (the HIP library) and declares its C-ABI (include/swl.h).

There is no fallback: if the library is missing or fails to load, every session raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "csrc" / "libswl.so"
_lib = None

vp = C.c_void_p
P = C.POINTER


class DeviceInfo(C.Structure):
    _fields_ = [("index", C.c_int), ("name", C.c_char * 128), ("arch", C.c_char * 64), ("total_bytes", C.c_uint64),
                ("free_bytes", C.c_uint64), ("compute_units", C.c_int), ("available", C.c_int)]


class Run2D(C.Structure):
    _fields_ = [("accepted", C.c_int), ("faulted", C.c_int), ("peak_energy", C.c_double)]


class Run3D(C.Structure):
    _fields_ = [("accepted", C.c_int), ("faulted", C.c_int), ("nonfinite", C.c_int), ("peak_energy", C.c_double),
                ("source_peak", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "swl_last_error": (C.c_char_p, []),
    "swl_device_count": (C.c_int, []),
    "swl_device_query": (C.c_int, [C.c_int, P(DeviceInfo)]),
    "swl_engine2d_create": (C.c_int, [vp, vp, C.c_int, C.c_int, P(vp)]),
    "swl_engine2d_destroy": (None, [vp]),
    "swl_engine2d_step": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "swl_engine2d_rollback": (C.c_int, [vp]),
    "swl_engine2d_reset": (C.c_int, [vp]),
    "swl_engine2d_diagnostics": (C.c_int, [vp, vp]),
    "swl_engine2d_diagnostics_noalloc": (C.c_int, [vp, vp]),
    "swl_engine2d_sample": (C.c_int, [vp, vp]),
    "swl_engine2d_frame": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int]),
    "swl_engine2d_snapshot": (C.c_int, [vp, C.c_int, vp]),
    "swl_engine2d_pml": (C.c_int, [vp, vp]),
    "swl_engine2d_probes": (C.c_int, [vp, vp, C.c_int, vp]),
    "swl_engine2d_resident_bytes": (C.c_uint64, [vp]),
    "swl_engine2d_run": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, vp, vp,
                                   P(Run2D)]),
    "swl_engine2d_run_enqueue": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int]),
    "swl_engine2d_run_collect": (C.c_int, [vp, C.c_int, vp, vp, P(Run2D)]),
    "swl_engine2d_profile": (C.c_int, [vp, C.c_int]),
    "swl_engine2d_kernel_timing": (C.c_int, [vp, vp]),
    "swl_engine2d_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_stream_probe": (C.c_int, [C.c_int, C.c_uint64, C.c_int, P(C.c_double)]),
    "swl_sim2d_profile": (C.c_int, [vp, C.c_int]),
    "swl_sim2d_kernel_timing": (C.c_int, [vp, vp]),
    "swl_sim2d_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_sim2d_create": (C.c_int, [vp, C.c_int, C.c_int, P(vp)]),
    "swl_sim2d_plan": (C.c_int, [vp, P(vp)]),
    "swl_sim2d_destroy": (None, [vp]),
    "swl_sim2d_advance": (C.c_int, [vp, C.c_int]),
    "swl_sim2d_reset": (C.c_int, [vp]),
    "swl_sim2d_step": (C.c_int, [vp]),
    "swl_sim2d_override_peak_energy": (C.c_int, [vp, C.c_double]),
    "swl_memory_budget2d": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim2d_parameters": (C.c_int, [vp, vp]),
    "swl_async3d_create": (C.c_int, [P(vp)]),
    "swl_async3d_destroy": (None, [vp]),
    "swl_async3d_initialize": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int]),
    "swl_async3d_queue": (C.c_int, [vp, C.c_int, C.c_double]),
    "swl_async3d_poll": (C.c_int, [vp]),
    "swl_async3d_pause": (None, [vp]),
    "swl_async3d_synchronize": (C.c_int, [vp]),
    "swl_async3d_busy": (C.c_int, [vp]),
    "swl_async3d_timing": (C.c_int, [vp, vp]),
    "swl_async3d_advance": (C.c_int, [vp, C.c_int]),
    "swl_async3d_reset": (C.c_int, [vp]),
    "swl_async3d_state": (C.c_int, [vp, vp]),
    "swl_async3d_dt": (C.c_double, [vp]),
    "swl_async3d_diagnostics": (C.c_int, [vp, vp]),
    "swl_async3d_receivers": (C.c_int, [vp, vp, vp, vp]),
    "swl_async3d_read_display_frame": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
    "swl_async3d_session": (vp, [vp]),
    "swl_memory_estimate3d": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim2d_total_steps": (C.c_int, [vp]),
    "swl_sim2d_dt": (C.c_double, [vp]),
    "swl_sim2d_record_dt": (C.c_double, [vp]),
    "swl_sim2d_layout": (C.c_int, [vp, vp]),
    "swl_sim2d_diagnostics": (C.c_int, [vp, vp]),
    "swl_sim2d_receivers": (C.c_int, [vp, vp, vp, vp, vp]),
    "swl_sim2d_probe_indices": (C.c_int, [vp, vp]),
    "swl_sim2d_snapshot": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim2d_pml": (C.c_int, [vp, vp]),
    "swl_sim2d_model": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim2d_read_frame": (C.c_int, [vp, C.c_int, vp, C.c_int]),
    "swl_sim2d_read_probes": (C.c_int, [vp, vp, C.c_int, vp]),
    "swl_sim2d_write_receivers": (C.c_int, [vp, C.c_char_p]),
    "swl_segy_write": (C.c_int, [C.c_char_p, vp, C.c_int, C.c_int64, C.c_double, C.c_char_p, C.c_char_p, C.c_int,
                                 C.c_double, C.c_double]),
    "swl_segy_read": (C.c_int, [C.c_char_p, P(vp)]),
    "swl_segy_records_destroy": (None, [vp]),
    "swl_segy_records_info": (C.c_int, [vp, P(C.c_int), P(C.c_double), C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
    "swl_segy_records_trace": (C.c_int, [vp, C.c_int, vp, vp, P(C.c_double), P(C.c_int64)]),
    "swl_segy_records_samples": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim2d_write_segy": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double]),
    "swl_sim2d_write_model_vts": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int]),
    "swl_model_vts_read": (C.c_int, [C.c_char_p, P(vp)]),
    "swl_model_doc2d_destroy": (None, [vp]),
    "swl_model_doc2d_params": (C.c_int, [vp, vp]),
    "swl_model_doc2d_info": (C.c_int, [vp, vp]),
    "swl_sim2d_write_snapshot": (C.c_int, [vp, C.c_char_p, vp, C.c_int]),
    "swl_sim3d_write_receivers_csv": (C.c_int, [vp, C.c_char_p, vp]),
    "swl_sim3d_write_receivers_vz_csv": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_sim3d_write_vts": (C.c_int, [vp, C.c_char_p, vp, C.c_int, C.c_int]),
    "swl_async2d_create": (C.c_int, [P(vp)]),
    "swl_async2d_destroy": (None, [vp]),
    "swl_async2d_initialize": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "swl_async2d_queue": (C.c_int, [vp, C.c_int, C.c_double]),
    "swl_async2d_poll": (C.c_int, [vp]),
    "swl_async2d_pause": (None, [vp]),
    "swl_async2d_synchronize": (C.c_int, [vp]),
    "swl_async2d_busy": (C.c_int, [vp]),
    "swl_async2d_timing": (C.c_int, [vp, vp]),
    "swl_async2d_advance": (C.c_int, [vp, C.c_int]),
    "swl_async2d_reset": (C.c_int, [vp]),
    "swl_async2d_state": (C.c_int, [vp, vp]),
    "swl_async2d_times": (C.c_int, [vp, vp]),
    "swl_async2d_diagnostics": (C.c_int, [vp, vp]),
    "swl_async2d_receivers": (C.c_int, [vp, vp, vp]),
    "swl_async2d_read_frame": (C.c_int, [vp, C.c_int, vp, C.c_int, P(C.c_int)]),
    "swl_async2d_session": (vp, [vp]),
    "swl_pml_slot2d": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int]),
    "swl_params2d_size": (C.c_size_t, []),
    "swl_params3d_size": (C.c_size_t, []),
    "swl_batch2d_runner_create": (C.c_int, [C.c_int, P(vp)]),
    "swl_batch2d_runner_destroy": (None, [vp]),
    "swl_batch2d_runner_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_batch2d_run": (C.c_int, [vp, vp, C.c_int]),
    "swl_batch2d_runner_stats": (C.c_int, [vp, vp]),
    "swl_batch2d_create": (C.c_int, [C.c_int, P(vp)]),
    "swl_batch2d_destroy": (None, [vp]),
    "swl_batch2d_add": (C.c_int, [vp, vp]),
    "swl_batch2d_remove": (C.c_int, [vp, vp]),
    "swl_batch2d_size": (C.c_int, [vp]),
    "swl_batch2d_advance": (C.c_int, [vp, C.c_int]),
    "swl_batch2d_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_batch2d_stats": (C.c_int, [vp, vp]),
    "swl_sim2d_fault": (C.c_char_p, [vp]),
    "swl_engine3d_create": (C.c_int, [vp, vp, C.c_int, C.c_int, P(vp)]),
    "swl_engine3d_destroy": (None, [vp]),
    "swl_engine3d_step": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "swl_engine3d_rollback": (C.c_int, [vp]),
    "swl_engine3d_reset": (C.c_int, [vp]),
    "swl_engine3d_diagnostics": (C.c_int, [vp, vp]),
    "swl_engine3d_sample": (C.c_int, [vp, vp]),
    "swl_engine3d_slices": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "swl_engine3d_download": (C.c_int, [vp, C.c_int, vp]),
    "swl_engine3d_resident_bytes": (C.c_uint64, [vp]),
    "swl_engine3d_run": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, vp, vp,
                                   P(Run3D)]),
    "swl_engine3d_profile": (C.c_int, [vp, C.c_int]),
    "swl_engine3d_kernel_timing": (C.c_int, [vp, vp]),
    "swl_sim3d_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(vp)]),
    "swl_sim3d_plan": (C.c_int, [vp, P(vp)]),
    "swl_sim3d_destroy": (None, [vp]),
    "swl_sim3d_advance": (C.c_int, [vp, C.c_int]),
    "swl_sim3d_reset": (C.c_int, [vp]),
    "swl_sim3d_step": (C.c_int, [vp]),
    "swl_sim3d_override_source_peak": (C.c_int, [vp, C.c_double]),
    "swl_sim3d_total_steps": (C.c_int, [vp]),
    "swl_sim3d_dt": (C.c_double, [vp]),
    "swl_sim3d_layout": (C.c_int, [vp, vp]),
    "swl_sim3d_diagnostics": (C.c_int, [vp, vp]),
    "swl_sim3d_receivers": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "swl_sim3d_snapshot": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim3d_model": (C.c_int, [vp, C.c_int, vp]),
    "swl_sim3d_read_display_frame": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "swl_sim3d_profile": (C.c_int, [vp, C.c_int]),
    "swl_sim3d_kernel_timing": (C.c_int, [vp, vp]),
    "swl_sim3d_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "swl_engine3d_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
}


class NativeError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(handle, name):
                continue
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().swl_last_error().decode()


def call(name, *args):
    """Calls an int-status entry point and raises NativeError with swl_last_error() on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise NativeError(last_error())
    return rc


def device_count() -> int:
    return lib().swl_device_count()


def stream_probe(nbytes: int, repeats: int = 20, device: int = 0) -> float:
    """Mean ms per plain copy of `nbytes` (read + write) on the device."""
    ms = C.c_double()
    call("swl_stream_probe", device, int(nbytes), int(repeats), C.byref(ms))
    return ms.value


def device_query(index: int = 0) -> DeviceInfo:
    info = DeviceInfo()
    call("swl_device_query", index, C.byref(info))
    return info
