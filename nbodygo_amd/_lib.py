"""ctypes binding of libnbgpu.so (include/nbody_gpu.h).

The library is the only compute path: if it is missing, importing this module raises.  There is
no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnbgpu.so")

NB_OK = 0
NB_E_INVALID = -1
NB_E_HIP = -2
NB_E_NOMEM = -3
NB_E_PAIR_OVERFLOW = -4
NB_E_STATE = -5
NB_E_UNSUPPORTED = -6

NB_F_EXISTS = 1
NB_F_FRAGMENTING = 2
NB_F_SUN = 4
NB_F_PINNED = 8

NB_CB_NONE = 0
NB_CB_SUBSUME = 1
NB_CB_ELASTIC = 2
NB_CB_FRAGMENT = 3

NB_PREC_FP32 = 0
NB_PREC_FP64 = 1


class NbParams(C.Structure):
    _fields_ = [
        ("G", C.c_double),
        ("eps", C.c_double),
        ("overlap_cutoff", C.c_int32),
        ("collisions", C.c_int32),
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("pair_capacity", C.c_int64),
        ("ipl", C.c_int32),
        ("tile_j", C.c_int32),
        ("jsplit", C.c_int32),
        ("block", C.c_int32),
        ("variant", C.c_int32),
        ("events", C.c_int32),
        ("resolve", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class NbStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("n", C.c_int64),
        ("n_active", C.c_int64),
        ("interactions", C.c_double),
        ("pairs_total", C.c_int64),
        ("pairs_last", C.c_int64),
        ("resolved_total", C.c_int64),
        ("fragment_requests", C.c_int64),
        ("removed_total", C.c_int64),
        ("fixup_tiles", C.c_int64),
        ("pair_overflow", C.c_int32),
        ("pad0", C.c_int32),
        ("wall_ms", C.c_double),
        ("force_ms", C.c_double),
        ("force_launches", C.c_int64),
        ("resolve_ms", C.c_double),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad0"}


class NbDiag(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double),
        ("potential", C.c_double),
        ("momentum", C.c_double * 3),
        ("mass", C.c_double),
        ("n_active", C.c_int64),
    ]


class NbSurParams(C.Structure):
    _fields_ = [
        ("grid", C.c_int32),
        ("shell", C.c_int32),
        ("r0", C.c_double),
        ("use_mlp", C.c_int32),
        ("check_every", C.c_int32),
        ("check_samples", C.c_int32),
        ("pad0", C.c_int32),
        ("tol", C.c_double),
        ("seed", C.c_uint64),
        ("tol_body", C.c_double),
        ("body_quantile", C.c_double),
        ("reserved", C.c_int32 * 4),
    ]


class NbSurStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("checks", C.c_int64),
        ("fallbacks", C.c_int64),
        ("last_error", C.c_double),
        ("near_pairs_total", C.c_double),
        ("nonempty_cells", C.c_int64),
        ("train_rows", C.c_int64),
        ("train_loss", C.c_double),
        ("r0", C.c_double),
        ("trained", C.c_int32),
        ("pad1", C.c_int32),
        ("last_body_error", C.c_double),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad1"}


class NbBhParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("softening", C.c_double), ("max_depth", C.c_int32),
                ("group", C.c_int32), ("mode", C.c_int32), ("term", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class NbBhStats(C.Structure):
    _fields_ = [("builds", C.c_int64), ("bodies", C.c_int64), ("nodes", C.c_int64),
                ("buckets", C.c_int64), ("interactions", C.c_int64), ("depth_max", C.c_int32),
                ("overflow", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class NbEventStats(C.Structure):
    _fields_ = [("subsumed_total", C.c_int64), ("fragmentations", C.c_int64),
                ("fragments_emitted", C.c_int64), ("fragments_pending", C.c_int64),
                ("fragmenting", C.c_int64), ("queue_overflow", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


class NbBody(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("x", "y", "z", "vx", "vy", "vz", "m", "r", "rest",
                                           "frag", "frag_step")] + [
        ("beh", C.c_int32),
        ("flags", C.c_uint8),
        ("pad", C.c_uint8 * 3),
    ]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_ctx = C.c_void_p

# (name, restype, argtypes): every symbol declared in include/nbody_gpu.h
SIGNATURES = [
    ("nb_params_init", C.c_int, [C.POINTER(NbParams)]),
    ("nb_create", C.c_int, [C.POINTER(NbParams), C.POINTER(_ctx)]),
    ("nb_destroy", None, [_ctx]),
    ("nb_last_error", C.c_char_p, [_ctx]),
    ("nb_load", C.c_int,
     [_ctx, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _u8p, _dp, _dp]),
    ("nb_step", C.c_int, [_ctx, C.c_double, C.c_double, C.c_int32, C.POINTER(NbStats)]),
    ("nb_compute_forces", C.c_int, [_ctx, _dp, _dp, _dp]),
    ("nb_accel_fp64", C.c_int, [_ctx, _i64p, C.c_int64, _dp, _dp, _dp]),
    ("nb_get_pairs", C.c_int, [_ctx, _i32p, C.c_int64, _i64p]),
    ("nb_store", C.c_int, [_ctx, _dp, _dp, _dp, _dp, _dp, _dp, _u8p, _dp]),
    ("nb_store_bodies", C.c_int,
     [_ctx, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _u8p, _dp, _dp, _dp]),
    ("nb_store_render", C.c_int, [_ctx, _fp, _u8p]),
    ("nb_diagnostics", C.c_int, [_ctx, C.POINTER(NbDiag)]),
    ("nb_set_timing", C.c_int, [_ctx, C.c_int32]),
    ("nb_get_config", C.c_int, [_ctx, _i32p]),
    ("nb_autotune", C.c_int, [_ctx, C.c_int32, _i32p]),
    ("nb_add_bodies", C.c_int,
     [_ctx, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i32p, _u8p, _dp, _dp]),
    ("nb_cycle", C.c_int, [_ctx, C.c_double, _i64p]),
    ("nb_remove_bodies", C.c_int, [_ctx, C.c_int64, _i64p]),
    ("nb_get_body", C.c_int, [_ctx, C.c_int64, C.POINTER(NbBody)]),
    ("nb_set_body", C.c_int, [_ctx, C.c_int64, C.POINTER(NbBody)]),
    ("nb_set_frag_steps", C.c_int, [_ctx, _dp]),
    ("nb_get_event_stats", C.c_int, [_ctx, C.POINTER(NbEventStats)]),
    ("nb_bh_params_init", C.c_int, [C.POINTER(NbBhParams)]),
    ("nb_bh_enable", C.c_int, [_ctx, C.POINTER(NbBhParams)]),
    ("nb_bh_disable", C.c_int, [_ctx]),
    ("nb_bh_get_stats", C.c_int, [_ctx, C.POINTER(NbBhStats)]),
    ("nb_bh_get_depths", C.c_int, [_ctx, _i32p]),
    ("nb_batch_create", C.c_int, [C.POINTER(NbParams), C.POINTER(_ctx)]),
    ("nb_batch_destroy", None, [_ctx]),
    ("nb_batch_last_error", C.c_char_p, [_ctx]),
    ("nb_batch_load", C.c_int,
     [_ctx, C.c_int32, _i64p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _u8p]),
    ("nb_batch_step", C.c_int, [_ctx, C.c_double, C.c_double, C.c_int32, C.POINTER(NbStats)]),
    ("nb_batch_set_graph", C.c_int, [_ctx, C.c_int32]),
    ("nb_batch_forces", C.c_int, [_ctx, _dp, _dp, _dp]),
    ("nb_batch_store", C.c_int, [_ctx, _dp, _dp, _dp, _dp, _dp, _dp, _u8p]),
    ("nb_batch_get_shapes", C.c_int, [_ctx, _i32p]),
    ("nb_batch_set_shapes", C.c_int, [_ctx, _i32p]),
    ("nb_batch_autotune", C.c_int, [_ctx, C.c_int32, _i32p]),
    ("nb_batch_get_tuning", C.c_int, [_ctx, _i32p]),
    ("nb_batch_set_tuning", C.c_int, [_ctx, _i32p]),
    ("nb_mlp_create", C.c_int,
     [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_ctx)]),
    ("nb_mlp_destroy", None, [_ctx]),
    ("nb_mlp_last_error", C.c_char_p, [_ctx]),
    ("nb_mlp_param_count", C.c_int, [_ctx, _i64p]),
    ("nb_mlp_set_params", C.c_int, [_ctx, _fp, C.c_int64]),
    ("nb_mlp_get_params", C.c_int, [_ctx, _fp, C.c_int64]),
    ("nb_mlp_forward", C.c_int, [_ctx, _fp, C.c_int64, _fp]),
    ("nb_mlp_infer", C.c_int, [_ctx, _fp, C.c_int64, _fp]),
    ("nb_mlp_grad", C.c_int, [_ctx, _fp, _fp, C.c_int64, _fp, _fp]),
    ("nb_mlp_train", C.c_int,
     [_ctx, _fp, _fp, C.c_int64, C.c_int32, C.c_int32, C.c_float, _fp]),
    ("nb_sur_params_init", C.c_int, [C.POINTER(NbSurParams)]),
    ("nb_sur_enable", C.c_int, [_ctx, C.POINTER(NbSurParams)]),
    ("nb_sur_disable", C.c_int, [_ctx]),
    ("nb_sur_collect", C.c_int, [_ctx, C.c_int64, C.c_uint64]),
    ("nb_sur_fit", C.c_int, [_ctx, C.c_int32, C.c_int32, C.c_float, _fp]),
    ("nb_sur_error", C.c_int, [_ctx, C.c_int32, C.c_int32, _dp]),
    ("nb_sur_get_stats", C.c_int, [_ctx, C.POINTER(NbSurStats)]),
    ("nb_sur_save", C.c_int, [_ctx, C.c_char_p]),
    ("nb_sur_load", C.c_int, [_ctx, C.c_char_p]),
    ("nb_forces_noalloc", C.c_int,
     [C.c_int32, C.c_int64] + [C.c_void_p] * 5 + [C.c_double, C.c_double] + [C.c_void_p] * 4),
]

NB_NOALLOC_MAX_BODIES = 262144


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            "libnbgpu.so not built at %s: run __graft_entry__.build() (there is no CPU fallback)"
            % path)
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load_library()


class NBodyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nb error %d: %s" % (code, msg))
        self.code = code


def check(rc, ctx=None):
    if rc != NB_OK:
        msg = lib.nb_last_error(ctx)
        raise NBodyError(rc, msg.decode() if msg else "")
    return rc
