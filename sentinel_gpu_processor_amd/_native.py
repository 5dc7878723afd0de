"""ctypes binding of libs1ifg.so (the C ABI declared in include/s1ifg.h).

Device buffers are torch tensors on a ROCm device (plumbing only); every
pixel operation runs in the HIP kernels of the library.  There is no CPU
fallback: if the library is missing or no GPU is visible the calls raise.
"""

import ctypes
import os

import numpy as np

from . import _build

S1IFG_OK = 0
S1IFG_EINVAL = -1
S1IFG_ERANGE = -4
INTERP_BILINEAR = 0
INTERP_SINC = 1
GOLD_FP32 = 0
GOLD_FP16_SPLIT = 1
DEM_F32 = 0
DEM_I16 = 1
TILE_STATS = 6  # counters of s1ifg_ctx_set_tile_stats
COH_FILL_CLAMP = 0
COH_FILL_NAN = 1
COH_FILL_HASH = 2

INTERP = {"bilinear": INTERP_BILINEAR, "sinc": INTERP_SINC}
GOLD_MODE = {"fp32": GOLD_FP32, "fp16_split": GOLD_FP16_SPLIT}


class S1IfgError(RuntimeError):
    """A libs1ifg call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libs1ifg error {code}: {msg}")
        self.code = code


class Orbit(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("t", ctypes.POINTER(ctypes.c_double)),
                ("pos", ctypes.POINTER(ctypes.c_double)),
                ("vel", ctypes.POINTER(ctypes.c_double))]


class Burst(ctypes.Structure):
    _fields_ = [("lines", ctypes.c_int32), ("samples", ctypes.c_int32),
                ("first_valid_line", ctypes.c_int32), ("last_valid_line", ctypes.c_int32),
                ("first_valid_sample", ctypes.c_int32), ("last_valid_sample", ctypes.c_int32),
                ("t_start", ctypes.c_double), ("prf", ctypes.c_double),
                ("azimuth_time_interval", ctypes.c_double),
                ("slant_range_time", ctypes.c_double),
                ("range_sampling_rate", ctypes.c_double), ("wavelength", ctypes.c_double),
                ("radar_frequency", ctypes.c_double),
                ("azimuth_steering_rate", ctypes.c_double),
                ("n_fm", ctypes.c_int32), ("n_dc", ctypes.c_int32),
                ("fm", ctypes.POINTER(ctypes.c_double)),
                ("dc", ctypes.POINTER(ctypes.c_double))]


class Grid(ctypes.Structure):
    _fields_ = [("nrow", ctypes.c_int32), ("ncol", ctypes.c_int32),
                ("lat0", ctypes.c_double), ("lon0", ctypes.c_double),
                ("dlat", ctypes.c_double), ("dlon", ctypes.c_double)]


class Dem(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int32),
                ("nrow", ctypes.c_int32), ("ncol", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("lat0", ctypes.c_double), ("lon0", ctypes.c_double),
                ("dlat", ctypes.c_double), ("dlon", ctypes.c_double)]


class PairTask(ctypes.Structure):
    _fields_ = [("g_ref", ctypes.c_void_p), ("g_sec", ctypes.c_void_p),
                ("grid", Grid), ("dem", Dem), ("dh", ctypes.c_void_p),
                ("offpoly_sec", ctypes.c_double * 6), ("use_offpoly", ctypes.c_int32),
                ("pad_", ctypes.c_int32),
                ("ref_slc", ctypes.c_void_p), ("sec_slc", ctypes.c_void_p),
                ("ref_deramped", ctypes.c_void_p), ("sec_deramped", ctypes.c_void_p),
                ("ifg_ml", ctypes.c_void_p), ("pow_ref", ctypes.c_void_p),
                ("pow_sec", ctypes.c_void_p), ("filt", ctypes.c_void_p),
                ("coh", ctypes.c_void_p)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("interp", ctypes.c_int32), ("rl", ctypes.c_int32), ("cl", ctypes.c_int32),
                ("alpha", ctypes.c_float), ("gold_mode", ctypes.c_int32),
                ("coh_win", ctypes.c_int32), ("coh_fill", ctypes.c_int32),
                ("coh_seed", ctypes.c_uint32), ("pair_block_x", ctypes.c_int32),
                ("pair_blocks_per_cu", ctypes.c_int32), ("gold_blocks_per_cu", ctypes.c_int32),
                ("stream_blocks_per_cu", ctypes.c_int32), ("raw", ctypes.c_int32)]


class StitchBurst(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("left", ctypes.c_int32), ("top", ctypes.c_int32),
                ("right", ctypes.c_int32), ("bottom", ctypes.c_int32),
                ("estimate", ctypes.c_int32), ("pad_", ctypes.c_int32)]


# a *_blocks_per_cu value: one block per work item instead of persistent blocks
BLOCK_PER_ITEM = -1

LAUNCH_FIELDS = ("pair_block_x", "pair_blocks_per_cu", "gold_blocks_per_cu",
                 "stream_blocks_per_cu")

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_F = ctypes.c_float
_U32 = ctypes.c_uint32

# name -> argtypes (all return int unless listed in _VOID / _STR)
SIGNATURES = {
    "s1ifg_abi_version": [],
    "s1ifg_last_error": [],
    "s1ifg_ctx_create": [_I, ctypes.POINTER(_P)],
    "s1ifg_ctx_destroy": [_P],
    "s1ifg_ctx_set_stream": [_P, _P],
    "s1ifg_ctx_sync": [_P],
    "s1ifg_ctx_set_tile_stats": [_P, _P],
    "s1ifg_ctx_workspace": [_P, ctypes.POINTER(_I64)],
    "s1ifg_geom_create": [_P, ctypes.POINTER(Orbit), ctypes.POINTER(Burst), ctypes.POINTER(_P)],
    "s1ifg_geom_destroy": [_P],
    "s1ifg_geom_query": [_P, ctypes.POINTER(_D), _I],
    "s1ifg_geom_bounds": [_P, _D, _D, ctypes.POINTER(_D)],
    "s1ifg_geom_vectors": [_P, _P, _P, _I],
    "s1ifg_orbit_eval": [_P, _P, _P, _I64, _P, _P, _P],
    "s1ifg_deramp": [_P, _P, _P, _P],
    "s1ifg_deramp_ci16": [_P, _P, _P, _P],
    "s1ifg_geo2rdr": [_P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P, _P, _P],
    "s1ifg_geocode": [_P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P, _I, _P],
    "s1ifg_valid_rows": [_P, _P, _I, _I, _P],
    "s1ifg_crossmul_looks": [_P, _P, _I64, _P, _I64, _I, _I, _I, _I, _P, _P, _P],
    "s1ifg_looks": [_P, _P, _I64, _I, _I, _I, _I, _P],
    "s1ifg_phase": [_P, _P, _I64, _P],
    "s1ifg_amplitude": [_P, _P, _I64, _I, _I, _I, _I, _P],
    "s1ifg_coherence": [_P, _P, _P, _P, _I, _I, _I, _I, _U32, _P],
    "s1ifg_goldstein": [_P, _P, _I, _I, _I, _F, _I, _P],
    "s1ifg_pair": [_P, _P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P, _P, _I,
                   _I, _I, _P, _P, _P],
    "s1ifg_pair_raw": [_P, _P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P, _P, _I,
                       _I, _I, _P, _P, _P],
    "s1ifg_pair_noalloc": [_P, _P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P,
                           _P, _I, _I, _I, _P, _P, _P],
    "s1ifg_pair_backward": [_P, _P, _P, ctypes.POINTER(Grid), ctypes.POINTER(Dem), _P, _P, _P,
                            _P, _I, _I, _I, _P, _P, _P, _P, _P],
    "s1ifg_coherence_backward": [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P],
    "s1ifg_batch_create": [_P, ctypes.POINTER(PairTask), _I, ctypes.POINTER(BatchOpts),
                           ctypes.POINTER(_P)],
    "s1ifg_batch_destroy": [_P],
    "s1ifg_batch_run": [_P, _I],
    "s1ifg_batch_bytes": [_P, ctypes.POINTER(_I64)],
    "s1ifg_batch_get_opts": [_P, ctypes.POINTER(BatchOpts)],
    "s1ifg_batch_set_launch": [_P, ctypes.POINTER(BatchOpts)],
    "s1ifg_stitch_subswath": [_P, ctypes.POINTER(StitchBurst), _I, _I, _P, _I, _I, _I, _I, _P],
    "s1ifg_merge_subswath": [_P, _P, _I, _I, _I, _P, _I64],
    "s1ifg_ps_scan_offsets": [_I, _I, _P, _I],
    "s1ifg_ps_normalize": [_P, _P, _P, _I64],
    "s1ifg_ps_update": [_P, _P, _P, _F, _I, _I64],
    "s1ifg_ps_median_similarity": [_P, _P, _I, _I, _I, _P, _I, _I, _I, _P],
    "s1ifg_ps_maximum_similarity": [_P, _P, _I, _I, _I, _P, _I, _I, _I, _P],
    "s1ifg_ps_similarity": [_P, _P, _I, _I, _I, _P, _I, _I, _I, _F, _P, _P],
}
_VOID = {"s1ifg_ctx_destroy", "s1ifg_geom_destroy", "s1ifg_batch_destroy"}
_STR = {"s1ifg_last_error"}

_lib = None


def lib_path():
    return _build.lib_path()


def load():
    """Load libs1ifg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        if name in _VOID:
            fn.restype = None
        elif name in _STR:
            fn.restype = ctypes.c_char_p
        else:
            fn.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc):
    if rc != S1IFG_OK:
        msg = load().s1ifg_last_error().decode(errors="replace")
        raise S1IfgError(rc, msg)
    return rc


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("libs1ifg takes device tensors")
    if not t.is_contiguous():
        raise ValueError("libs1ifg takes contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def _dbl_array(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One library context bound to a device and (by default) to torch's
    current stream on that device."""

    def __init__(self, device=0, stream=None):
        import torch
        self.lib = load()
        if not torch.cuda.is_available():
            raise RuntimeError("no ROCm GPU visible: libs1ifg needs a GPU")
        self.device = int(device)
        h = ctypes.c_void_p()
        check(self.lib.s1ifg_ctx_create(self.device, ctypes.byref(h)))
        self.handle = h
        self.bind_stream(stream)

    def bind_stream(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        check(self.lib.s1ifg_ctx_set_stream(self.handle, ctypes.c_void_p(stream.cuda_stream)))

    def sync(self):
        check(self.lib.s1ifg_ctx_sync(self.handle))

    def set_tile_stats(self, counts):
        """Count the fused sinc pair kernel's tile classes into `counts` (a
        zeroed int64 device tensor of TILE_STATS elements; None = off)."""
        if counts is not None and (counts.numel() < TILE_STATS or counts.element_size() != 8):
            raise ValueError(f"tile stats need {TILE_STATS} 64-bit counters")
        self._tile_stats = counts
        check(self.lib.s1ifg_ctx_set_tile_stats(self.handle, ptr(counts)))

    def workspace_bytes(self):
        v = ctypes.c_int64()
        check(self.lib.s1ifg_ctx_workspace(self.handle, ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.s1ifg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_orbit(orbit):
    """orbit: (t[n], pos[n,3], vel[n,3]) -> (Orbit, keepalive)."""
    t, pos, vel = orbit
    ta, tp = _dbl_array(t)
    pa, pp = _dbl_array(pos)
    va, vp = _dbl_array(vel)
    o = Orbit(n=len(ta), pad_=0, t=tp, pos=pp, vel=vp)
    return o, (ta, pa, va)


def make_burst(b):
    fm, fmp = _dbl_array(b["fm"])
    dc, dcp = _dbl_array(b["dc"])
    s = Burst(lines=b["lines"], samples=b["samples"],
              first_valid_line=b["first_valid_line"], last_valid_line=b["last_valid_line"],
              first_valid_sample=b["first_valid_sample"],
              last_valid_sample=b["last_valid_sample"],
              t_start=b["t_start"], prf=b["prf"],
              azimuth_time_interval=b["azimuth_time_interval"],
              slant_range_time=b["slant_range_time"],
              range_sampling_rate=b["range_sampling_rate"], wavelength=b["wavelength"],
              radar_frequency=b["radar_frequency"],
              azimuth_steering_rate=b["azimuth_steering_rate"],
              n_fm=fm.shape[0], n_dc=dc.shape[0], fm=fmp, dc=dcp)
    return s, (fm, dc)


def make_grid(g):
    return Grid(nrow=g["nrow"], ncol=g["ncol"], lat0=g["lat0"], lon0=g["lon0"],
                dlat=g["dlat"], dlon=g["dlon"])


def make_dem(dem_tensor, d):
    import torch
    if dem_tensor.dtype == torch.float32:
        dt = DEM_F32
    elif dem_tensor.dtype == torch.int16:
        dt = DEM_I16
    else:
        raise ValueError("DEM must be float32 or int16")
    return Dem(data=ptr(dem_tensor).value, dtype=dt, nrow=dem_tensor.shape[0],
               ncol=dem_tensor.shape[1], pad_=0, lat0=d["lat0"], lon0=d["lon0"],
               dlat=d["dlat"], dlon=d["dlon"])


class Geometry:
    """Prepared burst geometry (device-resident orbit polynomials and deramp
    tables) for one burst of one acquisition."""

    def __init__(self, ctx, orbit, burst):
        self.ctx = ctx
        self.burst = dict(burst)
        o, keep1 = make_orbit(orbit)
        b, keep2 = make_burst(burst)
        h = ctypes.c_void_p()
        check(ctx.lib.s1ifg_geom_create(ctx.handle, ctypes.byref(o), ctypes.byref(b),
                                        ctypes.byref(h)))
        self.handle = h

    def query(self):
        out = (ctypes.c_double * 17)()
        check(self.ctx.lib.s1ifg_geom_query(self.handle, out, 17))
        v = list(out)
        return dict(ks=v[0], fm=v[1:5], dc=v[5:9], etac0=v[9], tmid=v[10], xmid=v[11:14],
                    vmid=v[14:17])

    def vectors(self):
        """(kt, etaref) per range sample, as the kernels read them."""
        n = self.burst["samples"]
        kt = np.empty(n, np.float64)
        er = np.empty(n, np.float64)
        check(self.ctx.lib.s1ifg_geom_vectors(self.handle, kt.ctypes.data_as(ctypes.c_void_p),
                                              er.ctypes.data_as(ctypes.c_void_p), n))
        return kt, er

    def bounds(self, hmin, hmax):
        out = (ctypes.c_double * 4)()
        check(self.ctx.lib.s1ifg_geom_bounds(self.handle, hmin, hmax, out))
        return tuple(out)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.s1ifg_geom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def offpoly_arg(p):
    if p is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(p, np.float64).reshape(6))
    return a, a.ctypes.data_as(ctypes.c_void_p)
