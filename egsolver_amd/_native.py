"""ctypes bindings of the package's native libraries.

``libegs.so`` (HIP, gfx950) carries every kernel; ``libegs_host.so`` (C) carries the
.mpg reader, the owner partition and the seeded generator.  Both are built in-tree by
``make -C egsolver_amd`` (``__graft_entry__.build()`` runs it).  A missing library raises
at load time: the package has no CPU code path to fall back to.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


class NativeMissing(RuntimeError):
    pass


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeMissing(f"{path} is not built; run `make -C {HERE}` (or __graft_entry__.build())")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


class EgsOptions(ctypes.Structure):
    _fields_ = [
        ("group", ctypes.c_int32),
        ("big_deg", ctypes.c_int32),
        ("huge_deg", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("batch_rounds", ctypes.c_int32),
        ("dense_pct", ctypes.c_int32),
        ("certify_every", ctypes.c_int32),
        ("cert_max_passes", ctypes.c_int32),
        ("max_phases", ctypes.c_int64),
        ("stop_flag", ctypes.c_void_p),
        ("slices", ctypes.c_int32),
        ("phase_wgs_per_cu", ctypes.c_int32),
        ("graphs", ctypes.c_int32),
    ]


class EgsStats(ctypes.Structure):
    _fields_ = [
        ("phases", ctypes.c_int64),
        ("processed", ctypes.c_int64),
        ("edges_relaxed", ctypes.c_int64),
        ("lifted", ctypes.c_int64),
        ("pred_visits", ctypes.c_int64),
        ("converged", ctypes.c_int32),
        ("initial_active", ctypes.c_int32),
        ("ms_prepare", ctypes.c_double),
        ("ms_iterate", ctypes.c_double),
        ("cert_attempts", ctypes.c_int64),
        ("certificates", ctypes.c_int64),
        ("cert_passes", ctypes.c_int64),
        ("top_marked", ctypes.c_int64),
        ("cert_pruned", ctypes.c_int64),
        ("csc_built", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EgsDeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("device_count", ctypes.c_int32),
        ("warp_size", ctypes.c_int32),
        ("max_threads_per_block", ctypes.c_int32),
        ("multiprocessors", ctypes.c_int32),
        ("total_mem", ctypes.c_int64),
        ("shared_mem_per_block", ctypes.c_int64),
        ("clock_khz", ctypes.c_int32),
        ("ecc", ctypes.c_int32),
        ("driver_version", ctypes.c_int32),
        ("runtime_version", ctypes.c_int32),
        ("cap_major", ctypes.c_int32),
        ("cap_minor", ctypes.c_int32),
    ]


class EgsGenParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("n", ctypes.c_int32),
        ("deg_kind", ctypes.c_int32),
        ("deg_lo", ctypes.c_int32),
        ("deg_hi", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("wmax", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class EgsGenReadOrder(ctypes.Structure):
    _fields_ = [("owner", _u8p), ("ptr", _i32p), ("succ", _i32p), ("w", _i32p)]


class EgsCsr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("n0", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("mg", ctypes.c_int32),
        ("max_abs_w", ctypes.c_int32),
        ("ptr", _i32p),
        ("succ", _i32p),
        ("w", _i32p),
        ("ext_of_int", _i32p),
    ]


class EgsMpg(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("ext_of_read", _i32p),
        ("read_of_ext", _i32p),
        ("owner", _u8p),
        ("rptr", _i32p),
        ("rsucc", _i32p),
        ("rw", _i32p),
        ("max_abs_w", ctypes.c_int32),
        ("min_outdeg", ctypes.c_int32),
        ("max_outdeg", ctypes.c_int32),
        ("low_outdeg", ctypes.c_int32 * 4),
        ("names", ctypes.POINTER(ctypes.c_char_p)),
        ("chunks", ctypes.c_int32),
    ]


# symbol -> (restype, argtypes); also the list of exported entry points that the
# CPU test suite checks against include/*.h
EGS_SIGNATURES = {
    "egs_last_error": (ctypes.c_char_p, []),
    "egs_device_query": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(EgsDeviceInfo)]),
    "egs_device_reset": (ctypes.c_int, [ctypes.c_int]),
    "egs_solver_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "egs_solver_destroy": (ctypes.c_int, [_vp]),
    "egs_solver_load": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_int32, _vp, _vp, _vp]),
    "egs_solver_load_packed": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                              _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "egs_solver_prepare": (ctypes.c_int, [_vp, ctypes.POINTER(EgsOptions)]),
    "egs_solver_iterate": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(EgsStats)]),
    "egs_solver_stats": (ctypes.c_int, [_vp, ctypes.POINTER(EgsStats)]),
    "egs_solver_read_values": (ctypes.c_int, [_vp, _vp]),
    "egs_solver_read_winning": (ctypes.c_int, [_vp, _vp]),
    "egs_solver_device_values": (_vp, [_vp]),
    "egs_solver_read_csc": (ctypes.c_int, [_vp, _vp, _vp]),
    "egs_solver_stream": (_vp, [_vp]),
    "egs_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                 _vp, _vp, _vp, _vp, ctypes.POINTER(EgsOptions), ctypes.POINTER(EgsStats)]),
    "egs_solver_count_violations": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "egs_solver_set_values": (ctypes.c_int, [_vp, _vp]),
    "egs_solver_load_read_order": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp,
                                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _vp]),
    "egs_solver_read_game": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "egs_solver_top_certificate": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, ctypes.POINTER(ctypes.c_int32)]),
    "egs_solve_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _vp, _vp,
                                         _vp, _vp, ctypes.POINTER(EgsOptions), ctypes.POINTER(EgsStats), _vp]),
    "egs_host_alloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "egs_host_free": (ctypes.c_int, [_vp]),
    "egs_runtime_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "egs_copy_bandwidth": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
}

EGS_HOST_SIGNATURES = {
    "egs_mpg_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(EgsMpg), ctypes.c_char_p, ctypes.c_int]),
    "egs_mpg_read_mem": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(EgsMpg), ctypes.c_char_p, ctypes.c_int]),
    "egs_mpg_read_mem_par": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(EgsMpg), ctypes.c_char_p, ctypes.c_int,
                                            ctypes.c_int]),
    "egs_mpg_free": (None, [ctypes.POINTER(EgsMpg)]),
    "egs_mpg_internalise": (ctypes.c_int, [ctypes.POINTER(EgsMpg), _vp, _vp, _vp, _vp]),
    "egs_internalise": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "egs_compute_mg": (ctypes.c_int, [ctypes.c_int, _vp, _vp, ctypes.POINTER(ctypes.c_int32)]),
    "egs_max_loop": (ctypes.c_long, [ctypes.c_long] * 4),
    "egs_gen_game": (ctypes.c_int64, [ctypes.POINTER(EgsGenParams), ctypes.c_uint64,
                                      ctypes.POINTER(EgsGenReadOrder), ctypes.c_int]),
    "egs_gen_count_edges": (ctypes.c_int64, [ctypes.POINTER(EgsGenParams), ctypes.c_uint64]),
    "egs_gen_family_params": (None, [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(EgsGenParams)]),
    "egs_rng_test_stream": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _vp]),
    "egs_write_mpg": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "egs_csr_save": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                    _vp, _vp, _vp, _vp, ctypes.c_char_p, ctypes.c_int]),
    "egs_csr_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(EgsCsr), ctypes.c_char_p, ctypes.c_int]),
    "egs_csr_free": (None, [ctypes.POINTER(EgsCsr)]),
    "egs_check_fixpoint": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp,
                                            ctypes.POINTER(ctypes.c_int64)]),
    "egs_check_top_certificate": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp,
                                                 ctypes.POINTER(ctypes.c_int64)]),
}

_egs = None
_host = None


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def egs():
    """libegs.so (GPU library)."""
    global _egs
    if _egs is None:
        _egs = _bind(_load("libegs.so"), EGS_SIGNATURES)
        _bind_extra(_egs)
    return _egs


def host():
    """libegs_host.so (reader + generator)."""
    global _host
    if _host is None:
        _host = _bind(_load("libegs_host.so"), EGS_HOST_SIGNATURES)
    return _host


_EXTRA = []


def register_signatures(sigs):
    """Later modules (batch, serve) add their entry points here."""
    _EXTRA.append(sigs)
    EGS_SIGNATURES.update(sigs)
    if _egs is not None:
        _bind(_egs, sigs)


def _bind_extra(lib):
    for sigs in _EXTRA:
        _bind(lib, sigs)


class EgsError(RuntimeError):
    pass


def check(rc, what="egs call"):
    if rc < 0:
        msg = egs().egs_last_error()
        raise EgsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc
