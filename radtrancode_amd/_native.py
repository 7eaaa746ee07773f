"""Loader for the gfx950 library ``lib/libradg.so``.

The library is the product: there is no CPU fallback.  ``lib()`` raises if
the shared object is missing or does not load, and every wrapped call raises
:class:`RadgError` with the library's own message on a non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libradg.so")

RADG_OK = 0
RADG_ECANCELLED = 8
RADG_ETIMEOUT = 9
MODE_FP64 = 0
MODE_MIXED = 1
MODE_BF16 = 2
MODE_F16 = 3

# Every symbol the C header (include/radg.h) declares.
EXPORTS = [
    "radg_set_mode", "radg_get_mode", "radg_last_error", "radg_register_ktables",
    "radg_read_klist", "radg_register_cia", "radg_register_dust", "radg_set_nlte",
    "radg_set_iform", "radg_register_parah2", "radg_register_e3", "radg_register_n2nir",
    "radg_register_solar", "radg_register_gascon",
    "cirsradg_wave_",
    "radg_cirsradg_status",
    "radg_spectrum_r4", "radg_spectrum_r8", "radg_batch_create", "radg_batch_bytes",
    "radg_batch_launch", "radg_batch_sync", "radg_batch_read", "radg_batch_timings",
    "radg_batch_info", "radg_batch_destroy", "radg_batch_timing_stats",
    "radg_server_create", "radg_server_submit", "radg_server_cancel", "radg_server_wait",
    "radg_server_stats", "radg_server_destroy", "radg_batch_set_fetch", "radg_ktable_codes",
    "radg_batch_set_statemaps", "radg_batch_read_mapped", "radg_batch_device_outputs",
    "radg_oe_create", "radg_oe_cost", "radg_oe_propose", "radg_oe_destroy",
    "radg_spectrum_noalloc", "radg_cirsconv", "radg_read_cia",
    "radg_register_lbltables", "radg_read_klbllist", "radg_set_ilbl", "radg_get_ilbl",
]


class RadgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"radg status {code}: {msg}")
        self.code = code


class Request(C.Structure):
    """``radg_request`` (include/radg.h)."""
    _fields_ = [("nlayer", C.c_int), ("npath", C.c_int), ("ngas", C.c_int),
                ("nsw", C.c_int), ("nwave", C.c_int), ("nem", C.c_int),
                ("inormal", C.c_int), ("iray", C.c_int), ("ispace", C.c_int),
                ("itype", C.c_int), ("flagh2p", C.c_int), ("tsurf", C.c_double),
                ("press", C.c_void_p), ("temp", C.c_void_p), ("delh", C.c_void_p),
                ("pp", C.c_void_p), ("amount", C.c_void_p), ("idgas", C.c_void_p),
                ("isogas", C.c_void_p), ("nlayin", C.c_void_p), ("imod", C.c_void_p),
                ("layinc", C.c_void_p), ("scale", C.c_void_p), ("emtemp", C.c_void_p),
                ("isw", C.c_void_p), ("vwave", C.c_void_p), ("vem", C.c_void_p),
                ("emissivity", C.c_void_p), ("path_weight", C.c_void_p),
                ("ncont", C.c_int), ("nsec", C.c_int), ("cont", C.c_void_p),
                ("vsec", C.c_void_p), ("xsec", C.c_void_p),
                ("inlte", C.c_int), ("nlte_k1", C.c_double), ("nlte_a", C.c_double),
                ("iform", C.c_int), ("radius1", C.c_double), ("radextra", C.c_double),
                ("hfp", C.c_void_p), ("dist", C.c_double)]


_lib = None


def lib():
    """Load ``libradg.so`` (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m radtrancode_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip = C.c_void_p, C.c_int
    L.radg_last_error.restype = C.c_char_p
    for name in ["radg_set_mode", "radg_get_mode", "radg_register_ktables", "radg_read_klist",
                 "radg_register_cia", "radg_cirsradg_status", "radg_spectrum_r4",
                 "radg_spectrum_r8", "radg_batch_create", "radg_batch_bytes",
                 "radg_batch_launch", "radg_batch_sync", "radg_batch_read",
                 "radg_batch_timings", "radg_batch_info", "radg_batch_destroy",
                 "radg_batch_timing_stats", "radg_server_create", "radg_server_submit",
                 "radg_server_cancel", "radg_server_wait", "radg_server_stats",
                 "radg_server_destroy", "radg_ktable_codes"]:
        getattr(L, name).restype = ip
    L.radg_ktable_codes.argtypes = [ip, ip, vp, vp]
    L.radg_register_cia.argtypes = [C.c_double, vp, vp]
    L.radg_register_dust.argtypes = [ip, ip, vp, vp]
    L.radg_set_nlte.argtypes = [ip, C.c_double, C.c_double]
    L.radg_set_iform.argtypes = [ip, ip]
    L.radg_register_parah2.argtypes = [C.c_double, vp, ip, vp, vp, vp, vp]
    L.radg_register_e3.argtypes = [ip, vp, vp]
    L.radg_register_n2nir.argtypes = [ip, vp, vp, ip, vp, vp]
    L.radg_register_solar.argtypes = [ip, vp, vp]
    L.radg_read_klist.argtypes = [C.c_char_p, ip]
    for name in ["radg_register_lbltables", "radg_read_klbllist", "radg_set_ilbl", "radg_get_ilbl"]:
        getattr(L, name).restype = ip
    L.radg_register_lbltables.argtypes = [ip, vp, vp, ip, C.c_double, C.c_double, ip, ip, vp, vp, vp]
    L.radg_read_klbllist.argtypes = [C.c_char_p, ip]
    L.radg_set_ilbl.argtypes = [ip]
    L.radg_get_ilbl.argtypes = []
    L.radg_batch_create.argtypes = [ip, vp, ip, C.c_size_t, C.POINTER(vp)]
    L.radg_batch_bytes.argtypes = [ip, vp, ip, C.POINTER(C.c_size_t)]
    L.radg_batch_launch.argtypes = [vp, vp]
    L.radg_batch_sync.argtypes = [vp]
    L.radg_batch_read.argtypes = [vp, ip, vp, vp, vp, vp, vp, vp]
    L.radg_batch_timings.argtypes = [vp, vp]
    L.radg_batch_info.argtypes = [vp, vp, C.POINTER(ip), C.POINTER(ip)]
    L.radg_batch_destroy.argtypes = [vp]
    L.radg_batch_timing_stats.argtypes = [vp, C.POINTER(ip), vp]
    ll = C.c_longlong
    L.radg_server_create.argtypes = [ip, C.c_size_t, ip, ip, ip, C.POINTER(vp)]
    L.radg_batch_set_fetch.restype = ip
    L.radg_batch_set_fetch.argtypes = [vp, ip]
    L.radg_server_submit.argtypes = [vp, vp, C.POINTER(ll)]
    L.radg_server_cancel.argtypes = [vp, ll]
    L.radg_server_wait.argtypes = [vp, ll, C.c_double, vp, vp, vp, vp, vp, vp,
                                   C.POINTER(ip), C.POINTER(C.c_double)]
    L.radg_server_stats.argtypes = [vp, vp]
    L.radg_server_destroy.argtypes = [vp]
    for name in ["radg_batch_set_statemaps", "radg_batch_read_mapped",
                 "radg_batch_device_outputs", "radg_oe_create", "radg_oe_cost",
                 "radg_oe_propose", "radg_oe_destroy"]:
        getattr(L, name).restype = ip
    L.radg_batch_set_statemaps.argtypes = [vp, vp, ip]
    L.radg_batch_read_mapped.argtypes = [vp, ip, vp, vp]
    L.radg_batch_device_outputs.argtypes = [vp, ip, C.POINTER(vp), C.POINTER(vp)]
    L.radg_oe_create.argtypes = [ip, ip, ip, vp, vp, vp, vp, ip, C.POINTER(vp)]
    L.radg_oe_cost.argtypes = [vp, vp, vp, vp, vp]
    L.radg_oe_propose.argtypes = [vp, vp, vp, vp, vp, vp]
    L.radg_oe_destroy.argtypes = [vp]
    L.radg_spectrum_noalloc.restype = ip
    L.radg_spectrum_noalloc.argtypes = [vp, ip, vp, vp, vp, vp]
    L.radg_read_cia.restype = ip
    L.radg_read_cia.argtypes = [C.c_char_p, C.c_double]
    L.radg_cirsconv.restype = ip
    L.radg_cirsconv.argtypes = [ip, vp, ip, vp, vp, ip, vp, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != RADG_OK:
        raise RadgError(rc, lib().radg_last_error().decode())
    return rc
