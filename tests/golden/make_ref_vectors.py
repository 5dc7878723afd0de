"""Write golden vectors from the reference's own host geometry code.

oracle/_ref/libs1ref.so is the reference's csrc/src/orbit.cu and
csrc/src/bounds.cu compiled as plain C++ from the reference tree
(`make -C oracle ref`).  This script feeds it seeded inputs and stores
inputs and outputs in tests/golden/ref_geometry.npz (data only).

Run: python tests/golden/make_ref_vectors.py
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libs1ref.so")

GM = 3.986004418e14


def circular_orbit(t, r=6378137.0 + 693e3, inc_deg=98.18, raan=0.3):
    """Keplerian circular orbit (inertial frame treated as ECEF)."""
    n = np.sqrt(GM / r ** 3)
    u = n * t
    inc = np.radians(inc_deg)
    cu, su = np.cos(u), np.sin(u)
    co, so = np.cos(raan), np.sin(raan)
    ci, si = np.cos(inc), np.sin(inc)
    pos = r * np.stack([co * cu - so * su * ci, so * cu + co * su * ci, su * si], axis=-1)
    vel = r * n * np.stack([-co * su - so * cu * ci, -so * su + co * cu * ci, cu * si], axis=-1)
    return pos, vel


def main():
    lib = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)

    def P(a):
        return a.ctypes.data_as(dp)

    rng = np.random.default_rng(20260101)
    out = {}
    # A1/A2
    llh = np.stack([rng.uniform(-89, 89, 400), rng.uniform(-179.9, 179.9, 400),
                    rng.uniform(-500, 9000, 400)], axis=-1)
    llh[:3] = [[0, 0, 0], [90, 0, 0], [0, 90, 1234.5]]
    xyz = np.zeros_like(llh)
    lib.ref_llh2xyz(P(llh), P(xyz), ctypes.c_long(len(llh)))
    back = np.zeros_like(llh)
    lib.ref_xyz2llh(P(xyz), P(back), ctypes.c_long(len(llh)))
    out.update(llh_in=llh, llh2xyz_out=xyz, xyz2llh_out=back)
    # A3/A4: curved orbit with 10 s nodes
    t = 1000.0 + 10.0 * np.arange(16)
    pos, vel = circular_orbit(t)
    tq = np.concatenate([rng.uniform(t[0] - 5, t[-1] + 5, 300), t, [t[0] - 3.0, t[-1] + 2.0]])
    px = np.zeros((len(tq), 3))
    pv = np.zeros((len(tq), 3))
    lib.ref_intp_orbit(ctypes.c_long(len(t)), P(t), P(pos), P(vel), P(tq), ctypes.c_long(len(tq)),
                       P(px), P(pv))
    out.update(orb_t=t, orb_pos=pos, orb_vel=vel, intp_t=tq, intp_pos=px, intp_vel=pv)
    # A5: targets on the ground to the right of the track around t = 1075 s
    tmid = 1075.0
    xmid = np.zeros(3)
    vmid = np.zeros(3)
    lib.ref_intp_orbit(ctypes.c_long(len(t)), P(t), P(pos), P(vel), P(np.array([tmid])),
                       ctypes.c_long(1), P(xmid), P(vmid))
    up = xmid / np.linalg.norm(xmid)
    right = np.cross(vmid / np.linalg.norm(vmid), up)
    tt = tmid + rng.uniform(-20, 20, 200)
    sat, _ = circular_orbit(tt)
    look = rng.uniform(0.45, 0.75, 200)
    X = sat * (6378137.0 / np.linalg.norm(sat, axis=-1))[:, None] + (
        np.tan(look) * 693e3)[:, None] * right + rng.uniform(-300, 3000, 200)[:, None] * up
    tl = np.zeros(len(X))
    dr = np.zeros_like(X)
    lib.ref_orbitrangetime(ctypes.c_long(len(t)), P(t), P(pos), P(vel), P(X), ctypes.c_long(len(X)),
                           ctypes.c_double(tmid), P(xmid), P(vmid), P(tl), P(dr))
    out.update(ort_X=X, ort_tmid=np.array([tmid]), ort_xmid=xmid, ort_vmid=vmid, ort_t=tl,
               ort_dr=dr)
    # A6
    ll = np.zeros(4)
    lib.ref_bounds(ctypes.c_double(1060.0), ctypes.c_double(1063.0), ctypes.c_double(800e3),
                   ctypes.c_double(858e3), ctypes.c_double(-100.0), ctypes.c_double(2600.0),
                   ctypes.c_long(len(t)), P(t), P(pos), P(vel), P(ll))
    out.update(bounds_args=np.array([1060.0, 1063.0, 800e3, 858e3, -100.0, 2600.0]),
               bounds_out=ll)
    np.savez_compressed(os.path.join(HERE, "ref_geometry.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_geometry.npz"))


if __name__ == "__main__":
    main()
