"""ctypes binding of the C/OpenMP restatement (oracle/c/s1cpu.c ->
oracle/libs1cpu.so), test infrastructure: used by tests/ and by bench.py's
CPU baseline only.  Arrays are numpy (complex64 / float32 / float64)."""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libs1cpu.so")

_P = ctypes.c_void_p
_D = ctypes.c_double


class Orbit(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("t", _P), ("pos", _P), ("vel", _P)]


class Burst(ctypes.Structure):
    _fields_ = [("lines", ctypes.c_int), ("samples", ctypes.c_int), ("flv", ctypes.c_int),
                ("llv", ctypes.c_int), ("fvs", ctypes.c_int), ("lvs", ctypes.c_int),
                ("t_start", _D), ("prf", _D), ("ati", _D), ("slant_range_time", _D),
                ("range_sampling_rate", _D), ("wavelength", _D), ("radar_frequency", _D),
                ("steering_rate", _D), ("n_fm", ctypes.c_int), ("n_dc", ctypes.c_int),
                ("fm", _P), ("dc", _P)]


class Grid(ctypes.Structure):
    _fields_ = [("nrow", ctypes.c_int), ("ncol", ctypes.c_int), ("lat0", _D), ("lon0", _D),
                ("dlat", _D), ("dlon", _D)]


class Dem(ctypes.Structure):
    _fields_ = [("data", _P), ("nrow", ctypes.c_int), ("ncol", ctypes.c_int), ("lat0", _D),
                ("lon0", _D), ("dlat", _D), ("dlon", _D)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: `make -C oracle all`")
        _lib = ctypes.CDLL(LIB)
        _lib.s1c_threads.restype = ctypes.c_int
        _lib.s1c_deramp_setup.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def threads():
    return load().s1c_threads()


class _Keep:
    """Holds the arrays a structure points to."""

    def __init__(self):
        self.items = []

    def arr(self, a, dtype):
        a = np.ascontiguousarray(np.asarray(a, dtype))
        self.items.append(a)
        return a


def _orbit(orbit, keep):
    t, pos, vel = (keep.arr(x, np.float64) for x in orbit)
    return Orbit(len(t), _p(t), _p(pos), _p(vel))


def _burst(b, keep):
    fm = keep.arr(b["fm"], np.float64)
    dc = keep.arr(b["dc"], np.float64)
    return Burst(b["lines"], b["samples"], b["first_valid_line"], b["last_valid_line"],
                 b["first_valid_sample"], b["last_valid_sample"], b["t_start"], b["prf"],
                 b["azimuth_time_interval"], b["slant_range_time"], b["range_sampling_rate"],
                 b["wavelength"], b["radar_frequency"], b["azimuth_steering_rate"],
                 fm.shape[0], dc.shape[0], _p(fm), _p(dc))


def deramp_setup(orbit, burst):
    """(kt, etaref, scal): scal = [ks, fm(4), dc(4), etac0]."""
    keep = _Keep()
    n = burst["samples"]
    kt = np.empty(n)
    er = np.empty(n)
    sc = np.empty(10)
    o, b = _orbit(orbit, keep), _burst(burst, keep)
    if load().s1c_deramp_setup(ctypes.byref(o), ctypes.byref(b), _p(kt), _p(er), _p(sc)):
        raise ValueError("bad orbit / burst metadata")
    return kt, er, sc


def deramp(burst, kt, etaref, slc):
    keep = _Keep()
    b = _burst(burst, keep)
    z = keep.arr(slc, np.complex64)
    out = np.empty_like(z)
    load().s1c_deramp(ctypes.byref(b), _p(kt), _p(etaref), _p(z), _p(out))
    return out


def reproject(orbit, burst, scal, grid, dem, deramped, interp="bilinear", offpoly=None):
    keep = _Keep()
    o, b = _orbit(orbit, keep), _burst(burst, keep)
    g = Grid(grid["nrow"], grid["ncol"], grid["lat0"], grid["lon0"], grid["dlat"], grid["dlon"])
    d = keep.arr(dem["data"], np.float32)
    dm = Dem(_p(d), d.shape[0], d.shape[1], dem["lat0"], dem["lon0"], dem["dlat"], dem["dlon"])
    z = keep.arr(deramped, np.complex64)
    op = None if offpoly is None else _p(keep.arr(offpoly, np.float64))
    sc = keep.arr(scal, np.float64)
    out = np.empty((grid["nrow"], grid["ncol"]), np.complex64)
    load().s1c_reproject(ctypes.byref(o), ctypes.byref(b), _p(sc), ctypes.byref(g),
                         ctypes.byref(dm), op, _p(z), 0 if interp == "bilinear" else 1, _p(out))
    return out


def crossmul_looks(a, b, rl, cl):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    nrow, ncol = a.shape
    nr, nc = nrow // rl, ncol // cl
    ifg = np.empty((nr, nc), np.complex64)
    p1 = np.empty((nr, nc), np.float32)
    p2 = np.empty((nr, nc), np.float32)
    load().s1c_crossmul_looks(_p(a), _p(b), nrow, ncol, rl, cl, _p(ifg), _p(p1), _p(p2))
    return ifg, p1, p2


def goldstein(img, alpha=0.5):
    img = np.ascontiguousarray(img, np.complex64)
    out = np.empty_like(img)
    load().s1c_goldstein(_p(img), img.shape[0], img.shape[1], ctypes.c_float(alpha), _p(out))
    return out


def coherence(ifg, p1, p2, win=5):
    ifg = np.ascontiguousarray(ifg, np.complex64)
    p1 = np.ascontiguousarray(p1, np.float32)
    p2 = np.ascontiguousarray(p2, np.float32)
    coh = np.empty(p1.shape, np.float32)
    load().s1c_coherence(_p(ifg), _p(p1), _p(p2), p1.shape[0], p1.shape[1], win, _p(coh))
    return coh


def pair_chain(sc, ref, sec, interp="bilinear", rl=1, cl=4, alpha=0.5, win=5):
    """The whole chain of one pair on the CPU: deramp x 2, reproject x 2,
    crossmul + looks, Goldstein, coherence.  Returns the products."""
    kr, er, sr = deramp_setup(sc["orbit_ref"], sc["burst_ref"])
    ks, es, ss = deramp_setup(sc["orbit_sec"], sc["burst_sec"])
    dr = deramp(sc["burst_ref"], kr, er, ref)
    ds = deramp(sc["burst_sec"], ks, es, sec)
    gr = reproject(sc["orbit_ref"], sc["burst_ref"], sr, sc["grid"], sc["dem"], dr, interp)
    gs = reproject(sc["orbit_sec"], sc["burst_sec"], ss, sc["grid"], sc["dem"], ds, interp,
                   sc.get("offpoly"))
    ifg, p1, p2 = crossmul_looks(gr, gs, rl, cl)
    filt = goldstein(ifg, alpha)
    coh = coherence(ifg, p1, p2, win)
    return dict(gr=gr, gs=gs, ifg=ifg, p1=p1, p2=p2, filt=filt, coh=coh)


# ------------------------------------------------------- PS phase similarity
def ps_scan(rdmin, rdmax):
    lib = load()
    n = lib.s1c_ps_scan(int(rdmin), int(rdmax), None, 0)
    out = np.empty((max(n, 0), 2), np.int32)
    if n > 0:
        lib.s1c_ps_scan(int(rdmin), int(rdmax), _p(out), n)
    return out


def ps_normalize(stack):
    z = np.ascontiguousarray(stack, np.complex64)
    out = np.empty_like(z)
    load().s1c_ps_normalize(_p(z), _p(out), ctypes.c_long(z.size))
    return out


def _ps_args(zn, ps, offsets):
    zn = np.ascontiguousarray(zn, np.complex64)
    ps = np.ascontiguousarray(ps, np.int32)
    off = np.ascontiguousarray(offsets, np.int32)
    return zn, ps, off


def ps_median(zn, ps, nneigh, offsets):
    zn, ps, off = _ps_args(zn, ps, offsets)
    nifg, nrow, ncol = zn.shape
    out = np.empty((nrow, ncol), np.float32)
    load().s1c_ps_median(_p(zn), nifg, nrow, ncol, _p(ps), int(nneigh), _p(off), len(off),
                         _p(out))
    return out


def ps_max(zn, ps, nneigh, offsets):
    zn, ps, off = _ps_args(zn, ps, offsets)
    nifg, nrow, ncol = zn.shape
    out = np.empty((nrow, ncol), np.float32)
    load().s1c_ps_max(_p(zn), nifg, nrow, ncol, _p(ps), int(nneigh), _p(off), len(off),
                      _p(out))
    return out


def ps_update(ps, sim, th, overwrite=True):
    ps = np.array(ps, np.int32, copy=True, order="C")
    sim = np.ascontiguousarray(sim, np.float32)
    load().s1c_ps_update(_p(ps), _p(sim), ctypes.c_float(th), int(bool(overwrite)),
                         ctypes.c_long(ps.size))
    return ps


def ps_similarity(stack, ps, nneigh, rdmin, rdmax, med_sim_th):
    """(med_sim, max_sim, updated candidates), as oracle.psps.similarity."""
    off = ps_scan(rdmin, rdmax)
    zn = ps_normalize(stack)
    med = ps_median(zn, ps, nneigh, off)
    ps2 = ps_update(ps, med, med_sim_th, True)
    return med, ps_max(zn, ps2, nneigh, off), ps2
