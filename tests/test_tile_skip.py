"""The pair kernel skips a 16 x 16-cell tile (16 rl x 16 cl target pixels)
of a burst when 8 points -- the tile's corners at the lowest and highest
DEM sample its pixels blend -- all lie beyond the same edge of the burst's
validity window by 16 samples or lines.  That is sound when no pixel of a
tile lies further out along an edge's direction than the furthest of the
8 points by 16 or more.  These tests measure that excess with the oracle
geometry over every tile of the test scenes, and check that the rule skips
no tile holding a valid pixel."""

import numpy as np
import pytest

from oracle import geom as G
from tests.scenes import scene

MARGIN = 16.0
TILE_R, TILE_C = 16, 64  # 16 x 16 look cells of 1 x 4 pixels


def _dem_index(f, n):
    f = np.clip(f, 0.0, n - 1.0)
    return np.minimum(np.floor(f).astype(np.int64), n - 2)


def _tile_heights(grid, dem, r0, r1, c0, c1):
    """Lowest / highest DEM sample blended by pixels of rows r0..r1, columns
    c0..c1 (with one sample of slack, as the kernel)."""
    d = np.asarray(dem["data"])
    ny, nx = d.shape
    fy = (grid["lat0"] + np.array([r0, r1]) * grid["dlat"] - dem["lat0"]) / dem["dlat"]
    fx = (grid["lon0"] + np.array([c0, c1]) * grid["dlon"] - dem["lon0"]) / dem["dlon"]
    iy, ix = _dem_index(fy, ny), _dem_index(fx, nx)
    y0, y1 = max(iy.min() - 1, 0), min(iy.max() + 2, ny - 1)
    x0, x1 = max(ix.min() - 1, 0), min(ix.max() + 2, nx - 1)
    box = d[y0:y1 + 1, x0:x1 + 1].astype(np.float64)
    return box.min(), box.max()


def _positions(orbit, burst, grid, rows, cols, heights, offpoly):
    lat = grid["lat0"] + np.asarray(rows, np.float64) * grid["dlat"]
    lon = grid["lon0"] + np.asarray(cols, np.float64) * grid["dlon"]
    torb, xx, vv = orbit
    L = burst["lines"]
    dtaz = 1.0 / burst["prf"]
    tmid = burst["t_start"] + (L - 1) * dtaz * 0.5
    xmid, vmid = G.intp_orbit(torb, xx, vv, tmid)
    X = G.llh2xyz(lat, lon, heights)
    tline, dr, _ = G.orbitrangetime(torb, xx, vv, X, tmid, xmid, vmid)
    rng = np.sqrt((dr ** 2).sum(-1))
    rgoff = (rng - burst["slant_range_time"] * G.SOL / 2.0) / (G.SOL / 2.0 / burst["range_sampling_rate"])
    azoff = (tline - burst["t_start"]) / dtaz
    drg, daz = G.offset_poly_shift(offpoly, rgoff, azoff, burst)
    return rgoff + drg, azoff + daz


@pytest.mark.parametrize("name", ["small_misreg", "c1"])
def test_tile_corner_bound(name):
    sc = scene(name)
    grid, dem = sc["grid"], sc["dem"]
    h = G.target_heights(grid, dem)
    nrow, ncol = grid["nrow"], grid["ncol"]
    skipped = tiles = 0
    worst = -np.inf
    for orbit, burst, op in ((sc["orbit_ref"], sc["burst_ref"], None),
                             (sc["orbit_sec"], sc["burst_sec"], sc["offpoly"])):
        rr, cc = np.meshgrid(np.arange(nrow), np.arange(ncol), indexing="ij")
        rg, az = _positions(orbit, burst, grid, rr, cc, h, op)
        valid = ((rg >= burst["first_valid_sample"] + 3) & (rg < burst["last_valid_sample"] - 4)
                 & (az >= burst["first_valid_line"] + 5) & (az < burst["last_valid_line"] - 5))
        nc = ncol - ncol % 4  # columns of whole 1 x 4 look cells
        for r0 in range(0, nrow, TILE_R):
            r1 = min(r0 + TILE_R, nrow) - 1
            for c0 in range(0, nc, TILE_C):
                c1 = min(c0 + TILE_C, nc) - 1
                lo, hi = _tile_heights(grid, dem, r0, r1, c0, c1)
                cr = np.array([r0, r0, r1, r1] * 2)
                ccol = np.array([c0, c1, c0, c1] * 2)
                ch = np.array([lo] * 4 + [hi] * 4)
                krg, kaz = _positions(orbit, burst, grid, cr, ccol, ch, op)
                prg = rg[r0:r1 + 1, c0:c1 + 1]
                paz = az[r0:r1 + 1, c0:c1 + 1]
                # excess of the pixels over the 8 points along each edge direction
                for pv, kv in ((prg, krg), (-prg, -krg), (paz, kaz), (-paz, -kaz)):
                    worst = max(worst, float(pv.max() - kv.max()))
                lr = burst["first_valid_sample"] + 3
                ur = burst["last_valid_sample"] - 4
                la = burst["first_valid_line"] + 5
                ua = burst["last_valid_line"] - 5
                out = ((krg < lr - MARGIN).all() or (krg >= ur + MARGIN).all()
                       or (kaz < la - MARGIN).all() or (kaz >= ua + MARGIN).all())
                tiles += 1
                if out:
                    skipped += 1
                    assert not valid[r0:r1 + 1, c0:c1 + 1].any()
    # the bound holds with a wide margin, and the rule does skip tiles
    assert worst < MARGIN / 8, worst
    print(name, "worst excess", worst, "skipped", skipped, "of", tiles)
    assert skipped > 0


def _tile_corners(grid, dem, nrow, ncol):
    """Per tile (row-major over the tile grid): the 8 corner (row, column,
    height) points of the skip test."""
    rs = list(range(0, nrow, TILE_R))
    cs = list(range(0, ncol, TILE_C))
    pts = []
    for r0 in rs:
        r1 = min(r0 + TILE_R, nrow) - 1
        for c0 in cs:
            c1 = min(c0 + TILE_C, ncol) - 1
            lo, hi = _tile_heights(grid, dem, r0, r1, c0, c1)
            pts.append((r0, r1, c0, c1, lo, hi))
    P = np.array(pts)
    cr = np.stack([P[:, 0], P[:, 0], P[:, 1], P[:, 1]] * 2, axis=1)
    cc = np.stack([P[:, 2], P[:, 3], P[:, 2], P[:, 3]] * 2, axis=1)
    ch = np.concatenate([np.repeat(P[:, 4:5], 4, 1), np.repeat(P[:, 5:6], 4, 1)], axis=1)
    return P.astype(np.int64), cr, cc, ch


@pytest.mark.parametrize("hmax", [2500.0, 3000.0])
def test_tile_corner_bound_high_relief(hmax):
    """The same measurement on the benchmark's burst geometry (1500 x 25000,
    misregistered secondary) with fractal relief of 2500 and 3000 m, over the
    west 1500 x 2048 pixels of the grid, where the footprint's edge crosses
    the grid (tiles outside, inside and straddling it)."""
    from sentinel_gpu_processor_amd import synthetic as S
    sc = S.make_scene(lines=1500, samples=25000, dem_shape=(2000, 3000), dem_kind="fractal",
                      hmax=hmax, seed=0, misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1),
                      fidelity="none")
    grid = dict(sc["grid"], ncol=2048)
    dem = sc["dem"]
    h = G.target_heights(grid, dem)
    assert np.ptp(dem["data"]) > 0.95 * hmax
    nrow, ncol = grid["nrow"], grid["ncol"]
    P, cr, cc, ch = _tile_corners(grid, dem, nrow, ncol)
    worst = -np.inf
    skipped = 0
    for orbit, burst, op in ((sc["orbit_ref"], sc["burst_ref"], None),
                             (sc["orbit_sec"], sc["burst_sec"], sc["offpoly"])):
        rr, ccol = np.meshgrid(np.arange(nrow), np.arange(ncol), indexing="ij")
        rg, az = _positions(orbit, burst, grid, rr, ccol, h, op)
        valid = ((rg >= burst["first_valid_sample"] + 3) & (rg < burst["last_valid_sample"] - 4)
                 & (az >= burst["first_valid_line"] + 5) & (az < burst["last_valid_line"] - 5))
        krg, kaz = _positions(orbit, burst, grid, cr, cc, ch, op)
        lr = burst["first_valid_sample"] + 3
        ur = burst["last_valid_sample"] - 4
        la = burst["first_valid_line"] + 5
        ua = burst["last_valid_line"] - 5
        for t, (r0, r1, c0, c1, _, _) in enumerate(P):
            prg = rg[r0:r1 + 1, c0:c1 + 1]
            paz = az[r0:r1 + 1, c0:c1 + 1]
            for pv, kv in ((prg, krg[t]), (-prg, -krg[t]), (paz, kaz[t]), (-paz, -kaz[t])):
                worst = max(worst, float(pv.max() - kv.max()))
            out = ((krg[t] < lr - MARGIN).all() or (krg[t] >= ur + MARGIN).all()
                   or (kaz[t] < la - MARGIN).all() or (kaz[t] >= ua + MARGIN).all())
            if out:
                skipped += 1
                assert not valid[r0:r1 + 1, c0:c1 + 1].any()
    print(hmax, "worst excess", worst, "skipped", skipped, "of", 2 * len(P))
    assert worst < MARGIN / 8, worst
    assert skipped > 0.2 * 2 * len(P)
