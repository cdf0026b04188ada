"""ctypes binding of libtt.so (include/tt.h).

The library is built in-tree by ``turbulence_tracing_amd.build``.  There is no
CPU fallback: ``load()`` raises if the library is missing, and every wrapped
call raises ``TTError`` with the library's message when it returns non-zero.
Device pointers come from torch tensors (``tensor.data_ptr()``) and calls run
on torch's current HIP stream.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtt.so")

TT_MAX_OPS = 24
TT_MAX_DETECTORS = 4

OP_DISTANCE, OP_LENS, OP_APERTURE, OP_STOP, OP_KNIFE, OP_RECT, OP_ANNULUS = range(1, 8)


class TTError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class Spectrum(C.Structure):
    _fields_ = [("beta", C.c_double), ("aniso", C.c_double), ("kmin", C.c_double),
                ("kmax", C.c_double), ("amp", C.c_double), ("seed", C.c_uint64)]


class TraceOpts(C.Structure):
    _fields_ = [("h", C.c_double), ("cT", C.c_double), ("variant", C.c_int32), ("forms", C.c_void_p)]


class Optic(C.Structure):
    _fields_ = [("op", C.c_int32), ("axis", C.c_int32), ("a", C.c_double), ("b", C.c_double)]


class Train(C.Structure):
    _fields_ = [("n_ops", C.c_int32), ("to_mm", C.c_int32), ("ops", Optic * TT_MAX_OPS)]


class Hist(C.Structure):
    _fields_ = [("nbx", C.c_int32), ("nby", C.c_int32), ("xmin", C.c_double), ("xmax", C.c_double),
                ("ymin", C.c_double), ("ymax", C.c_double)]


class Detector(C.Structure):
    _fields_ = [("train", Train), ("hist", Hist), ("mode", C.c_int32), ("tilt_k", C.c_double),
                ("counts", C.c_void_p), ("wsum", C.c_void_p)]


class GemmArgs(C.Structure):
    _fields_ = [("A", C.c_void_p), ("sam", C.c_int64), ("sak", C.c_int64),
                ("B", C.c_void_p), ("sbk", C.c_int64), ("sbn", C.c_int64),
                ("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32),
                ("bias", C.c_void_p), ("gelu", C.c_int32), ("Cf", C.c_void_p), ("Cb", C.c_void_p),
                ("ldc", C.c_int64), ("Zpre", C.c_void_p), ("ksplit", C.c_int32)]


class BatchItem(C.Structure):
    _fields_ = [("field4", C.c_void_p), ("grid", Grid), ("s0", C.c_void_p), ("ld_s0", C.c_int64),
                ("n", C.c_int64), ("order", C.c_void_p), ("rf", C.c_void_p), ("ld_rf", C.c_int64),
                ("phase", C.c_void_p), ("status", C.c_void_p), ("cT", C.c_double)]


class BatchInfo(C.Structure):
    _fields_ = [("nitems", C.c_int32), ("nchunks", C.c_int32), ("chunk", C.c_int32),
                ("steps_per_cell", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("variant", C.c_int32)]


P = C.c_void_p
I32, I64, U64, D = C.c_int32, C.c_int64, C.c_uint64, C.c_double
SIZE_P = C.POINTER(C.c_size_t)

# name -> argtypes (all return int status)
_SIGNATURES = {
    "tt_fft_plan_create": [I32, I32, I32, C.POINTER(P)],
    "tt_fft_plan_destroy": [P],
    "tt_fft_workspace_bytes": [P, SIZE_P],
    "tt_field_generate": [P, C.POINTER(Spectrum), P, P, P, P],
    "tt_field_from_noise": [P, P, P, P, P, P, P, P],
    "tt_power_spectrum": [P, P, I32, I32, D, D, P, P, P, P],
    "tt_grad_stencil": [P, I32, I64, I64, I64, C.POINTER(Grid), D, D, D, D, P, I64, P, P],
    "tt_plane_forms": [P, C.POINTER(Grid), P, P],
    "tt_trace_rk4": [P, C.POINTER(Grid), C.POINTER(TraceOpts), P, I64, I64, P, P, I64, P, P, I64, P, P],
    "tt_trace_time": [P, C.POINTER(Grid), C.POINTER(TraceOpts), P, I64, I64, P, P, I64, P, P, I64, P, I32, P],
    "tt_trace_rk4_counted": [P, C.POINTER(Grid), C.POINTER(TraceOpts), P, I64, I64, P, P, P, I64, P, P, P],
    "tt_sample_field": [P, C.POINTER(Grid), P, I64, I64, P, I64, P],
    "tt_beam_sample": [I64, D, D, D, I32, U64, U64, P, I64, P],
    "tt_ray_sort_scratch_bytes": [C.POINTER(Grid), I32, SIZE_P],
    "tt_ray_sort": [P, I64, I64, C.POINTER(Grid), I32, P, P, P],
    "tt_ray_sort_noalloc": [P, I64, I64, C.POINTER(Grid), I32, P, P],
    "tt_ray_sort_rays": [P, I64, I64, C.POINTER(Grid), I32, P, P, P, P, I64, P],
    "tt_optics_apply": [C.POINTER(Train), P, I64, I64, P, I64, P],
    "tt_detect": [C.POINTER(Detector), I32, P, I64, P, I64, P],
    "tt_field_blocks": [P, C.POINTER(Grid), I32, P, P],
    "tt_path_features": [P, P, I32, C.POINTER(Grid), P, I64, I64, P, P, I64, P, I64, P],
    "tt_surrogate_apply": [P, I64, P, I64, I64, P, P, P, I64, P, P, P, P, I64, P],
    "tt_gemm_bf16": [C.POINTER(GemmArgs), P],
    "tt_gelu_backward": [P, P, P, P, I64, P],
    "tt_bias_grad": [P, I64, I32, I64, P, P],
    "tt_surrogate_loss": [P, I64, P, I64, P, I64, I64, I32, P, I32, P, P, D, I32, P, P],
    "tt_adam": [P, P, P, P, P, I64, D, D, D, D, I64, D, P],
    "tt_batch_workspace_bytes": [C.POINTER(BatchItem), I32, I32, C.POINTER(C.c_size_t)],
    "tt_batch_prepare": [C.POINTER(BatchItem), I32, D, I32, P, C.c_size_t, C.POINTER(BatchInfo)],
    "tt_batch_trace": [P, C.POINTER(BatchInfo), C.POINTER(BatchTuning), P],
    "tt_mlp_infer": [P, I64, I64, I64, I32, P, P, P, P, P, I64, I32, P],
    "tt_standardize_bf16": [P, I64, I64, I64, I32, P, P, P, I64, P],
}

# every symbol include/tt.h declares
EXPORTS = ["tt_version", "tt_last_error"] + list(_SIGNATURES)


class _Lib:
    def __init__(self, path):
        self.path = path
        self.dll = C.CDLL(path)
        self.dll.tt_version.restype = C.c_int
        self.dll.tt_version.argtypes = []
        self.dll.tt_last_error.restype = C.c_char_p
        self.dll.tt_last_error.argtypes = []
        for name, args in _SIGNATURES.items():
            fn = getattr(self.dll, name)
            fn.restype = C.c_int
            fn.argtypes = args

    def version(self):
        return self.dll.tt_version()

    def call(self, name, *args):
        rc = getattr(self.dll, name)(*args)
        if rc != 0:
            msg = self.dll.tt_last_error().decode(errors="replace")
            raise TTError(f"{name} failed ({rc}): {msg}")
        return rc

    def __getattr__(self, name):
        if name.startswith("tt_"):
            return lambda *a: self.call(name, *a)
        raise AttributeError(name)


_LIB = None


def load(path=None):
    """Load libtt.so (once).  Raises when it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise TTError(f"{p} is missing: build it with `python -m turbulence_tracing_amd.build`")
    lib = _Lib(p)
    if path is None:
        _LIB = lib
    return lib


def stream_ptr(stream=None):
    """The raw hipStream_t of a torch stream (default: the current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())
