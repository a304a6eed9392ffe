"""ctypes binding of libducc_amd.so (the C ABI in include/ducc_amd.h).

The library is built in-tree by ducc_amd/csrc/build.py.  There is no CPU
fallback: a missing library or a missing GPU raises.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libducc_amd.so")

DMX_F32, DMX_F64, DMX_C64, DMX_C128 = 3, 7, 67, 71
DMX_U8, DMX_U32 = 32, 35

_NP_TC = {np.dtype(np.float32): DMX_F32, np.dtype(np.float64): DMX_F64,
          np.dtype(np.complex64): DMX_C64, np.dtype(np.complex128): DMX_C128,
          np.dtype(np.uint8): DMX_U8, np.dtype(np.uint32): DMX_U32}


class DmxArray(ct.Structure):
    _fields_ = [("data", ct.c_void_p), ("ndim", ct.c_int64),
                ("shape", ct.c_int64 * 10), ("stride", ct.c_int64 * 10),
                ("typecode", ct.c_int32), ("device", ct.c_int32)]


class DmxOptions(ct.Structure):
    _fields_ = [("precision", ct.c_int32), ("chooser", ct.c_int32),
                ("log2tile", ct.c_int32), ("kernel_idx", ct.c_int32),
                ("fp64_below", ct.c_double), ("stream", ct.c_void_p),
                ("sigma", ct.c_double), ("spot_check", ct.c_int32),
                ("flags", ct.c_int32), ("deterministic", ct.c_int32),
                ("reserved", ct.c_int32 * 5)]

STREAM_GIVEN, ASYNC, POINT_RECORDS, CHECK_OFFSETS = 1, 2, 4, 8
SPREAD_FP64_ACC = 16


class DmxPlanInfo(ct.Structure):
    _fields_ = [("ndim", ct.c_int32), ("W", ct.c_int32),
                ("kernel_idx", ct.c_int32), ("fp64", ct.c_int32),
                ("log2tile", ct.c_int32), ("chooser_used", ct.c_int32),
                ("nuni", ct.c_int64 * 3), ("nover", ct.c_int64 * 3),
                ("ofactor", ct.c_double), ("eps_tab", ct.c_double),
                ("epsilon", ct.c_double), ("npoints", ct.c_int64)]

    def as_dict(self):
        nd = self.ndim
        return dict(ndim=nd, W=self.W, kernel_idx=self.kernel_idx,
                    fp64=bool(self.fp64), log2tile=self.log2tile,
                    chooser_used=bool(self.chooser_used),
                    nuni=tuple(self.nuni[:nd]), nover=tuple(self.nover[:nd]),
                    ofactor=self.ofactor, eps_tab=self.eps_tab,
                    epsilon=self.epsilon, npoints=self.npoints)


# exported names, each with (restype, argtypes)
_P = ct.c_void_p
_AP = ct.POINTER(DmxArray)
_SIG = {
    "dmx_last_error": (ct.c_char_p, []),
    "dmx_default_options": (None, [ct.POINTER(DmxOptions)]),
    "dmx_best_epsilon": (ct.c_double, [ct.c_size_t, ct.c_int, ct.c_double,
                                       ct.c_double]),
    "dmx_u2nu": (ct.c_int, [_AP, _AP, ct.c_int, ct.c_double, ct.c_size_t, _AP,
                            ct.c_size_t, ct.c_double, ct.c_double, _P,
                            ct.c_int, ct.POINTER(DmxOptions),
                            ct.POINTER(DmxPlanInfo)]),
    "dmx_nu2u": (ct.c_int, [_AP, _AP, ct.c_int, ct.c_double, ct.c_size_t, _AP,
                            ct.c_size_t, ct.c_double, ct.c_double, _P,
                            ct.c_int, ct.POINTER(DmxOptions),
                            ct.POINTER(DmxPlanInfo)]),
    "dmx_plan_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, _AP, _P,
                                   ct.c_size_t, ct.c_double, ct.c_size_t,
                                   ct.c_double, ct.c_double, _P, ct.c_int,
                                   ct.POINTER(DmxOptions)]),
    "dmx_plan_nu2u": (ct.c_int, [_P, ct.c_int, ct.c_size_t, _AP, _AP,
                                 ct.POINTER(DmxOptions)]),
    "dmx_plan_u2nu": (ct.c_int, [_P, ct.c_int, ct.c_size_t, _AP, _AP,
                                 ct.POINTER(DmxOptions)]),
    "dmx_plan_get_info": (ct.c_int, [_P, ct.POINTER(DmxPlanInfo)]),
    "dmx_plan_destroy": (ct.c_int, [_P]),
    "dmx_u2nu_coord_grad": (ct.c_int, [_AP, _AP, ct.c_int, ct.c_double, _AP,
                                       ct.c_double, ct.c_double, _P, ct.c_int,
                                       ct.POINTER(DmxOptions)]),
    "dmx_nudft": (ct.c_int, [ct.c_int, _AP, _AP, _P, ct.c_size_t, ct.c_int,
                             _P, ct.c_int, _P, ct.c_size_t, _AP]),
    "dmx_sort_points": (ct.c_int, [_AP, _P, ct.c_size_t, ct.c_int, ct.c_int,
                                   _P, _AP, _AP]),
    "dmx_sort_records": (ct.c_int, [_AP, _P, ct.c_size_t, ct.c_int, ct.c_int,
                                    _P, _AP, _AP, _AP]),
    "dmx_mlp_nparam": (ct.c_size_t, []),
    "dmx_mlp_forward": (ct.c_int, [_P, _P, ct.c_int, _P]),
    "dmx_mlp_grad": (ct.c_int, [_P, _P, _P, _P, ct.c_int, _P, _P]),
    "dmx_mlp_train": (ct.c_int, [_P, _P, _P, _P, ct.c_int, ct.c_int,
                                 ct.c_float, ct.c_float, _P]),
    "dmx_chooser_set": (ct.c_int, [_P, _P]),
    "dmx_chooser_features": (ct.c_int, [ct.c_size_t, _P, ct.c_size_t,
                                        ct.c_double, ct.c_int, ct.c_int, _P]),
    "dmx_choose": (ct.c_int, [ct.c_size_t, _P, ct.c_size_t, ct.c_double,
                              ct.c_int, ct.c_int, ct.c_double, ct.c_double,
                              ct.POINTER(DmxOptions),
                              ct.POINTER(DmxPlanInfo)]),
    "dmx_profile_enable": (ct.c_int, [ct.c_int]),
    "dmx_profile_reset": (ct.c_int, []),
    "dmx_profile_read": (ct.c_int, [ct.c_char_p, ct.c_size_t]),
    "dmx_c2c": (ct.c_int, [_AP, _AP, _P, ct.c_size_t, ct.c_int, ct.c_double,
                           ct.c_size_t, ct.POINTER(DmxOptions)]),
    "dmx_nufft_noalloc": (ct.c_int, [ct.c_int, _AP, _AP, _AP, ct.c_int,
                                     ct.c_double, ct.c_double, ct.c_double,
                                     _P, ct.c_int, ct.POINTER(DmxOptions),
                                     ct.POINTER(DmxPlanInfo)]),
    "dmx_sht_synthesis_general": (
        ct.c_int, [_AP, _AP, ct.c_size_t, ct.c_size_t, _P, ct.c_size_t,
                   ct.c_int64, _AP, ct.c_double, ct.c_double, ct.c_double,
                   ct.c_size_t, ct.c_int, ct.c_size_t,
                   ct.POINTER(DmxOptions)]),
    "dmx_sht_adjoint_synthesis_general": (
        ct.c_int, [_AP, _AP, ct.c_size_t, ct.c_size_t, _P, ct.c_size_t,
                   ct.c_int64, _AP, ct.c_double, ct.c_double, ct.c_double,
                   ct.c_size_t, ct.c_int, ct.c_size_t,
                   ct.POINTER(DmxOptions)]),
    "dmx_sht_pseudo_analysis_general": (
        ct.c_int, [_AP, _AP, ct.c_size_t, ct.c_size_t, _P, ct.c_size_t,
                   ct.c_int64, _AP, ct.c_double, ct.c_double, ct.c_size_t,
                   ct.c_size_t, ct.c_double, ct.c_size_t,
                   ct.POINTER(ct.c_size_t), ct.POINTER(ct.c_size_t),
                   ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                   ct.POINTER(DmxOptions)]),
    "dmx_vis2dirty": (ct.c_int, [_AP, _AP, _AP, _AP, _AP, ct.c_double,
                                 ct.c_double, ct.c_double, ct.c_int,
                                 ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                 ct.c_double, ct.c_double, ct.c_double,
                                 ct.c_double, ct.c_int, ct.c_size_t,
                                 ct.c_size_t, _AP, ct.POINTER(DmxOptions)]),
    "dmx_dirty2vis": (ct.c_int, [_AP, _AP, _AP, _AP, _AP, ct.c_double,
                                 ct.c_double, ct.c_double, ct.c_int,
                                 ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                 ct.c_double, ct.c_double, ct.c_double,
                                 ct.c_double, ct.c_int, ct.c_size_t,
                                 ct.c_size_t, _AP, ct.POINTER(DmxOptions)]),
    "dmx_device_count": (ct.c_int, []),
}

_lib = None


def lib():
    """Load the library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with "
                "`python ducc_amd/csrc/build.py` (there is no CPU fallback)")
        # torch ships its own libamdhip64.so.7.  Loaded after ours, it would
        # be a second HIP runtime in the process and torch would find no
        # device; loaded first, our library binds to the same runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ct.CDLL(LIBPATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_names():
    return list(_SIG)


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().dmx_last_error().decode())


def require_gpu():
    if lib().dmx_device_count() < 1:
        raise RuntimeError("ducc_amd needs a HIP device (none visible); "
                           "there is no CPU fallback")


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def describe(x):
    """dmx_array of a numpy array (host) or a torch tensor (device or host).
    Returns (descriptor, keepalive)."""
    a = DmxArray()
    if _is_torch(x):
        import torch
        tmap = {torch.float32: DMX_F32, torch.float64: DMX_F64,
                torch.complex64: DMX_C64, torch.complex128: DMX_C128,
                torch.uint8: DMX_U8}
        if x.dtype not in tmap:
            raise RuntimeError("not yet supported")
        a.data = x.data_ptr() if x.numel() else None
        a.ndim = x.dim()
        for i in range(x.dim()):
            a.shape[i] = x.shape[i]
            a.stride[i] = x.stride(i)
        a.typecode = tmap[x.dtype]
        a.device = 1 if x.is_cuda else 0
        return a, x
    x = np.asarray(x)
    if x.dtype not in _NP_TC:
        raise RuntimeError("not yet supported")
    isz = x.dtype.itemsize
    for s in x.strides:
        if s % isz:
            x = np.ascontiguousarray(x)
            break
    a.data = x.ctypes.data if x.size else None
    a.ndim = x.ndim
    for i in range(x.ndim):
        a.shape[i] = x.shape[i]
        a.stride[i] = x.strides[i] // isz
    a.typecode = _NP_TC[x.dtype]
    a.device = 0
    return a, x


def profile_enable(on=True):
    check(lib().dmx_profile_enable(int(bool(on))))


def profile_reset():
    check(lib().dmx_profile_reset())


def profile_read():
    """{phase: (total_ms, count)} accumulated since the last reset."""
    buf = ct.create_string_buffer(1 << 16)
    check(lib().dmx_profile_read(buf, len(buf)))
    res = {}
    for ln in buf.value.decode().splitlines():
        name, ms, cnt = ln.split("\t")
        res[name] = (float(ms), int(cnt))
    return res


def options(precision="auto", chooser="analytic", log2tile=0, kernel_idx=-1,
            fp64_below=1e-6, stream=None, sigma=0.0, spot_check=True,
            deterministic=False, point_records=False, check_offsets=False,
            fp64_accumulation=False):
    o = DmxOptions()
    lib().dmx_default_options(ct.byref(o))
    o.spot_check = 2 if spot_check == "fail" else int(bool(spot_check))
    if chooser == "learned":
        from . import chooser as _ch
        _ch.ensure_loaded()
    o.precision = {"auto": 0, "fp32": 1, "fp64": 2}[precision]
    o.chooser = {"analytic": 0, "learned": 1}[chooser]
    o.log2tile = int(log2tile)
    o.kernel_idx = int(kernel_idx)
    o.fp64_below = float(fp64_below)
    # a stream (torch tensors on the GPU: torch's current stream, NULL for
    # the default stream) means: run on it and return without waiting when
    # all arrays are on the device, as torch operations do
    if stream is not None:
        o.stream = stream
        o.flags = STREAM_GIVEN | ASYNC
    o.sigma = float(sigma)
    o.deterministic = int(bool(deterministic))
    if point_records:
        o.flags |= POINT_RECORDS
    if check_offsets:
        o.flags |= CHECK_OFFSETS
    if fp64_accumulation:
        o.flags |= SPREAD_FP64_ACC
    return o
