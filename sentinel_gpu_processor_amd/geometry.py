"""Drop-in for ``s1proc.geometry``: WGS-84 geodesy used for set-up (burst
footprints, look vectors).  Same names, signatures and conventions as the
reference module (geometry.py:1-356); host NumPy, like the reference's
numba CPU code.  The per-pixel geodesy of the hot path runs in the HIP
kernels (llh2xyz inside the geo2rdr kernel).
"""

import numpy as np

RA = 6378137.0  # semi-major axis of Earth (geometry.py:4)
RE2 = 0.0066943799901499996  # eccentricity squared (geometry.py:5)
SOL = 299792458.0  # speed of light (geometry.py:6)


def rotation3(d_theta, axis):
    """Rotation matrix about axis 1, 2 or 3 by d_theta degrees (geometry.py:9-31)."""
    t = np.radians(d_theta)
    c, s = np.cos(t), np.sin(t)
    if axis == 1:
        return np.array([[1, 0, 0], [0, c, s], [0, -s, c]])
    if axis == 2:
        return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])
    if axis == 3:
        return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])
    raise ValueError("axis must be 1, 2, or 3")


def geo_hdg(ll1, ll2, r_a=RA, r_e2=RE2):
    """Heading (radians, clockwise from north) of the geodesic from ll1 to
    ll2 (geometry.py:34-147; Sodano inverse solution on the ellipsoid,
    spherical formula when r_e2 == 0).  Returns None for coincident points
    on the ellipsoid, as the reference does."""
    rad = np.pi / 180.0
    if r_e2 == 0.0:
        sli, cli = np.sin(ll1[0] * rad), np.cos(ll1[0] * rad)
        tlf = np.tan(ll2[0] * rad)
        t1 = ll2[1] * rad - ll1[1] * rad
        if abs(t1) > np.pi:
            t1 = (2.0 * np.pi - abs(t1)) * np.copysign(1.0, -t1)
        return np.arctan2(np.sin(t1), cli * tlf - sli * np.cos(t1))
    e = np.sqrt(r_e2)
    sq = np.sqrt(1.0 - r_e2)
    f = 1.0 - sq
    ep = e * f / (r_e2 - f)
    nn = f / r_e2
    d = 16.0 * r_e2 * nn * nn + ep * ep
    k1 = d / ep ** 2
    k2 = (16.0 * r_e2 * nn * nn) / d
    k3 = (16.0 * r_e2 * nn * nn) / ep ** 2
    k4 = (16.0 * nn - ep * ep) / d
    k5 = 16.0 / (r_e2 * d)
    tli, tlf = np.tan(ll1[0] * rad), np.tan(ll2[0] * rad)
    l = abs(ll2[1] * rad - ll1[1] * rad)
    lsign = ll2[1] * rad - ll1[1] * rad
    sinlon, coslon = np.sin(l), np.cos(l)
    tbi, tbf = sq * tli, sq * tlf
    cbi, cbf = 1.0 / np.sqrt(1.0 + tbi ** 2), 1.0 / np.sqrt(1.0 + tbf ** 2)
    sbi, sbf = tbi * cbi, tbf * cbf
    ac, bc = sbi * sbf, cbi * cbf
    cosphi = ac + bc * coslon
    sinphi = np.copysign(1.0, sinlon) * np.sqrt(1.0 - min(cosphi ** 2, 1.0))
    phi = abs(np.arctan2(sinphi, cosphi))
    if not (r_a * abs(phi) > 1e-6):
        return None
    ca = (bc * sinlon) / sinphi
    cb = ca * ca
    cc = (cosphi * (1.0 - cb)) / k1
    cd = (-2.0 * ac) / k1
    ce = -ac * k2
    cf = k3 * cc
    cg = phi ** 2 / sinphi
    x = ((phi * (k4 + cb) + sinphi * (cc + cd) + cg * (cf + ce)) * ca) / k5
    lam = l + x
    t2 = tbf * cbi - np.cos(lam) * sbi
    return np.arctan2(np.sin(lam) * np.copysign(1.0, lsign), t2)


def llh2xyz(llh, r_a=RA, r_e2=RE2):
    """(lat deg, lon deg, h m) -> ECEF (geometry.py:150-162)."""
    lat = np.radians(llh[0])
    lon = np.radians(llh[1])
    h = llh[2]
    re = r_a / np.sqrt(1.0 - r_e2 * np.sin(lat) ** 2)
    return np.array([(re + h) * np.cos(lat) * np.cos(lon), (re + h) * np.cos(lat) * np.sin(lon),
                     (re * (1 - r_e2) + h) * np.sin(lat)])


def _xyz2llh_arrays(x, y, z, r_a, r_e2):
    q2 = 1.0 / (1.0 - r_e2)
    q = np.sqrt(q2)
    q3 = q2 - 1.0
    b = r_a * np.sqrt(1 - r_e2)
    lon = np.arctan2(y, x)
    lon = np.where((y < 0) & (lon > 0), lon - np.pi, lon)
    lon = np.where((y > 0) & (lon < 0), lon + np.pi, lon)
    p = np.sqrt(x ** 2 + y ** 2)
    th = np.arctan((z / p) * q)
    lat = np.arctan((z + q3 * b * np.sin(th) ** 3) / (p - r_e2 * r_a * np.cos(th) ** 3))
    re = r_a / np.sqrt(1.0 - r_e2 * np.sin(lat) ** 2)
    h = p / np.cos(lat) - re
    return np.degrees(lat), np.degrees(lon), h


def xyz2llh(xyz, r_a=RA, r_e2=RE2):
    """ECEF -> (lat, lon, h), one-step Bowring (geometry.py:165-192)."""
    lat, lon, h = _xyz2llh_arrays(xyz[0], xyz[1], xyz[2], r_a, r_e2)
    return np.array([lat, lon, h])


def llh2xyz_vec(llh, r_a=RA, r_e2=RE2):
    """Row-wise llh2xyz for an (N, 3) array (geometry.py:195-209)."""
    llh = np.asarray(llh, np.float64)
    lat = np.radians(llh[:, 0])
    lon = np.radians(llh[:, 1])
    re = r_a / np.sqrt(1.0 - r_e2 * np.sin(lat) ** 2)
    x = (re + llh[:, 2]) * np.cos(lat) * np.cos(lon)
    y = (re + llh[:, 2]) * np.cos(lat) * np.sin(lon)
    z = (re * (1 - r_e2) + llh[:, 2]) * np.sin(lat)
    return np.column_stack((x, y, z))


def xyz2llh_vec(xyz, r_a=RA, r_e2=RE2):
    """Row-wise xyz2llh for an (N, 3) array (geometry.py:212-241)."""
    xyz = np.asarray(xyz, np.float64)
    lat, lon, h = _xyz2llh_arrays(xyz[:, 0], xyz[:, 1], xyz[:, 2], r_a, r_e2)
    return np.column_stack((lat, lon, h))


def _rdir(r_lat, hdg, r_a=RA, r_e2=RE2):
    re = r_a / np.sqrt(1.0 - r_e2 * np.sin(r_lat) ** 2)
    rn = r_a * (1 - r_e2) / (1.0 - r_e2 * np.sin(r_lat) ** 2) ** 1.5
    return (re * rn) / (re * np.cos(hdg) ** 2 + rn * np.sin(hdg) ** 2)


def sch2xyz(llg, r_a=RA, r_e2=RE2):
    """SCH <-> XYZ transformation at peg (lat, lon, hdg) (geometry.py:259-297)."""
    lat, lon = np.radians(llg[0]), np.radians(llg[1])
    hdg = llg[2]
    clt, slt, clo, slo = np.cos(lat), np.sin(lat), np.cos(lon), np.sin(lon)
    chg, shg = np.cos(hdg), np.sin(hdg)
    mat = np.array([
        [clt * clo, -shg * slo - slt * clo * chg, slo * chg - slt * clo * shg],
        [clt * slo, clo * shg - slt * slo * chg, -clo * chg - slt * slo * shg],
        [slt, clt * chg, clt * shg],
    ])
    radcur = _rdir(lat, hdg, r_a, r_e2)
    xyz = llh2xyz([llg[0], llg[1], 0.0], r_a, r_e2)
    up = [clt * clo, clt * slo, slt]
    ov = [xyz[k] - radcur * up[k] for k in range(3)]
    return {"mat": mat, "matinv": mat.T, "radcur": radcur, "ov": ov}


def tcnbasis(pos, vel, look_dir="RIGHT", r_a=RA, r_e2=RE2):
    """Track / cross-track / nadir unit vectors (geometry.py:300-317)."""
    llh = xyz2llh(pos, r_a, r_e2)
    lat, lon = np.radians(llh[0]), np.radians(llh[1])
    n = -np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    t = np.cross(n, vel) if look_dir.lower() == "right" else -np.cross(n, vel)
    c = t / np.linalg.norm(t)
    t = np.cross(c, n)
    return t / np.linalg.norm(t), c, n


def tcnbasis_vec(pos, vel, look_dir="RIGHT", r_a=RA, r_e2=RE2):
    """Vectorised tcnbasis.  Like the reference (geometry.py:334-336) the
    normalisation divides by the norm of the whole (N, 3) array."""
    llh = xyz2llh_vec(pos, r_a, r_e2)
    lat, lon = np.radians(llh[:, 0]), np.radians(llh[:, 1])
    n = -np.column_stack((np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)))
    t = np.cross(n, vel) if look_dir.lower() == "right" else -np.cross(n, vel)
    c = t / np.linalg.norm(t)
    t = np.cross(c, n)
    return t / np.linalg.norm(t), c, n


def enu2xyz(xyz, lat, lon):
    """ENU vector -> ECEF vector at (lat, lon) (geometry.py:340-343)."""
    return np.dot(np.matmul(rotation3(-(90 + lon), 3), rotation3(-(90 - lat), 1)), xyz)


def xyz2enu(enu, lat, lon):
    """ECEF vector -> ENU at (lat, lon) (geometry.py:346-356)."""
    return np.dot(np.matmul(rotation3(90 - lat, 1), rotation3(90 + lon, 3)), enu)
