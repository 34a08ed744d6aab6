"""Loader of the HIP library libmdffgpu.so and ctypes prototypes of the C ABI
declared in include/mdffgpu.h.

There is no fallback: when the library has not been built, every entry point
raises ``LibraryMissing`` (build it with ``python -m mdff_amd.build``).
"""
from __future__ import annotations

import ctypes
import importlib.util
import os

from .params import MdffBatchConfig, MdffEfgOpts, MdffEnergies, MdffMdParams, MdffMdStep, MdffParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmdffgpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); the exported C ABI
PROTOTYPES = {
    "mdffgpu_create": (ctypes.c_int, [ctypes.POINTER(MdffParams), _ip, _dp, ctypes.POINTER(_vp)]),
    "mdffgpu_destroy": (None, [_vp]),
    "mdffgpu_set_cell": (ctypes.c_int, [_vp, _dp]),
    "mdffgpu_engforce": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int32, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _dp, ctypes.POINTER(MdffEnergies)]),
    "mdffgpu_engforce_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                               ctypes.POINTER(MdffEnergies), _vp]),
    "mdffgpu_engforce_noalloc": (ctypes.c_int, [ctypes.POINTER(MdffParams), _ip, _dp, _vp,
                                                 ctypes.c_int32, _vp, _vp, _vp, _vp, _vp,
                                                 ctypes.POINTER(MdffEnergies), _vp]),
    "mdffgpu_md_run": (ctypes.c_int, [_vp, ctypes.POINTER(MdffMdParams), ctypes.c_int32,
                                      ctypes.c_int32, _vp, _vp, _vp, _dp, _dp,
                                      ctypes.POINTER(MdffMdStep), _vp]),
    "mdffgpu_list_info": (ctypes.c_int, [_vp, _ip, _ip]),
    "mdffgpu_efg": (ctypes.c_int, [_vp, ctypes.POINTER(MdffEfgOpts), _dp, _dp, _dp, _dp, _dp,
                                   ctypes.POINTER(MdffEnergies)]),
    "mdffgpu_efg_nmr": (ctypes.c_int, [ctypes.c_int32, _dp, _dp, _dp, _dp, _ip]),
    "mdffgpu_get_dipole_history": (ctypes.c_int, [_vp, _dp]),
    "mdffgpu_set_dipole_history": (ctypes.c_int, [_vp, _dp]),
    "mdffgpu_num_kvectors": (ctypes.c_int32, [_vp]),
    "mdffgpu_last_error": (ctypes.c_char_p, [_vp]),
    "mdffgpu_device_count": (ctypes.c_int32, []),
    "mdffgpu_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mdffgpu_profile_only": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mdffgpu_set_mixed_tuning": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "mdffgpu_profile_read": (ctypes.c_int, [_vp, ctypes.c_int32, _dp,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "mdffgpu_kernel_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "mdffgpu_kernel_count": (ctypes.c_int32, []),
    "mdffgpu_batch_create": (ctypes.c_int, [ctypes.POINTER(MdffParams), ctypes.c_int32,
                                            ctypes.POINTER(MdffBatchConfig), _ip,
                                            ctypes.POINTER(_vp)]),
    "mdffgpu_batch_destroy": (None, [_vp]),
    "mdffgpu_batch_set_model": (ctypes.c_int, [_vp, ctypes.POINTER(MdffParams)]),
    "mdffgpu_batch_set_positions": (ctypes.c_int, [_vp, _vp, _vp]),
    "mdffgpu_batch_forces": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _ip, _dp, _vp]),
    "mdffgpu_batch_num_params": (ctypes.c_int32, [_vp]),
    "mdffgpu_batch_num_atoms": (ctypes.c_int64, [_vp]),
    "mdffgpu_batch_loss_grad": (ctypes.c_int, [_vp, _vp, _vp, _dp, _dp, ctypes.c_int32, _ip,
                                               _dp, _vp]),
    "mdffgpu_batch_graph": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mdffgpu_batch_set_tuning": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "mdffgpu_batch_set_mixed_tuning": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "mdffgpu_batch_last_error": (ctypes.c_char_p, [_vp]),
}


class LibraryMissing(RuntimeError):
    pass


class MdffError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"mdffgpu status {status}: {message}")
        self.status = status


_lib = None


def _adopt_torch_hip_runtime() -> None:
    """When PyTorch is installed, load its HIP runtime before this library.

    torch ships its own libamdhip64 with the same soname as /opt/rocm's; if
    the library bound to /opt/rocm's copy first, a later torch import would
    load a second runtime into the process.  Loading torch's file first
    makes the dynamic linker resolve this library's dependency to it, so the
    process has one HIP runtime and device pointers are shared."""
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def lib():
    """The loaded library; raises LibraryMissing if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} not found: build the HIP library with `python -m mdff_amd.build`")
        _adopt_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check_batch(status: int, b=None) -> None:
    if status != 0:
        msg = lib().mdffgpu_batch_last_error(b)
        raise MdffError(status, msg.decode() if msg else "")


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().mdffgpu_last_error(ctx)
        raise MdffError(status, msg.decode() if msg else "")
