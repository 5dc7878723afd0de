"""PS phase-similarity oracle (test infrastructure): a restatement of the
reference's ``phase_similarity`` executable (csrc/src/phase_similarity.cu)
for whole images, in float32 like the reference, with plain loops (tiny
stacks only; oracle/c/s1cpu.c holds the same algorithm in C for larger
ones).

Functions and the reference code each follows (file:line):
  scan_offsets         scan_array, phase_similarity.cu:295-366 (the ordered
                       neighbour offsets: rings 1 .. rdmax-1 of a midpoint
                       circle with its gaps filled, kept for rings > rdmin)
  normalize            normalize, phase_similarity.cu:116-134
  median_similarity    median_similarity<N>, phase_similarity.cu:174-233
  update_ps            update_ps, phase_similarity.cu:145-157
  maximum_similarity   maximum_similarity, phase_similarity.cu:235-283
  similarity           similarity(), phase_similarity.cu:375-548 (one batch)

Semantics fixed here where the reference leaves them open:
  - median_similarity never writes pixels that are not candidates, so the
    reference leaves them as its device allocation held them; here they are
    0, which is also what update_ps then sees.
  - similarity() batches rows by 2e9 bytes and, with more than one batch,
    indexes its row window with the whole image's row count
    (phase_similarity.cu:202-203, 264-265 against 461); the one-batch result
    (the whole image resident) is the one restated.
  - nneigh > 128 overflows the reference's per-thread array (its dispatch
    tests nifg, not nneigh, at phase_similarity.cu:478); 1 <= nneigh <= 128
    here (nneigh 0 behaves as 1, as in the reference's loop).

Parity pins (tests/test_oracle_psps.py): scan_offsets against rings 1-4
traced by hand from the reference's steps and against invariants (each
offset once, 8-fold symmetry, ring bounds, hole-free union); the similarity
functions against an independent vectorised brute force (all pair
similarities, then first-N selection by cumulative counts, np.sort
medians) on tiny stacks, and special cases with closed forms (identical
phases -> 1, opposite -> -1, isolated candidates -> 0).
"""

import numpy as np


def scan_offsets(rdmin, rdmax):
    """(dr, dc) offsets in the order the reference explores them
    (phase_similarity.cu:295-366), int32 array of shape (n, 2)."""
    rdmax, rdmin = int(rdmax), int(rdmin)
    out = []
    if rdmax <= 0:
        return np.zeros((0, 2), np.int32)
    seen = np.zeros((rdmax, rdmax), bool)  # first quadrant, [x][y]
    seen[0, 0] = True
    for r in range(1, rdmax):
        keep = r > rdmin
        x, y, p = r, 0, 1 - r
        if keep:
            out += [(r, 0), (-r, 0), (0, r), (0, -r)]
        seen[r, 0] = seen[0, r] = True
        pending = 0  # columns stepped back to fill a gap, still to revisit
        while x > y:
            if pending == 0:
                y += 1
                if p <= 0:
                    p += 2 * y + 1
                else:
                    x -= 1
                    p += 2 * y - 2 * x + 1
            else:
                pending -= 1
            if x < y:
                break
            # step inwards while the point inside is still unvisited
            while x - 1 >= 0 and y < rdmax and not seen[x - 1, y]:
                x -= 1
                pending += 1
            seen[x, y] = seen[y, x] = True
            if keep:
                out += [(x, y), (-x, -y), (x, -y), (-x, y)]
                if x != y:
                    out += [(y, x), (-y, -x), (y, -x), (-y, x)]
            if pending > 0:
                x += 1
    return np.asarray(out, np.int32).reshape(-1, 2)


def normalize(stack):
    """Unit phasors, 0 where |z| <= 1e-8 (float32 arithmetic)."""
    z = np.asarray(stack, np.complex64)
    re = z.real.astype(np.float32)
    im = z.imag.astype(np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        mag = np.sqrt(re * re + im * im, dtype=np.float32)
        ok = mag > np.float32(1e-8)
        inv = np.where(ok, np.float32(1) / np.where(ok, mag, np.float32(1)), np.float32(0))
    out = np.zeros(z.shape, np.complex64)
    out.real = np.where(ok, re * inv, np.float32(0))
    out.imag = np.where(ok, im * inv, np.float32(0))
    return out


def _sim(stack, r0, c0, r, c):
    nifg = stack.shape[0]
    s = np.float32(0)
    for j in range(nifg):
        a, b = stack[j, r0, c0], stack[j, r, c]
        s = np.float32(s + np.float32(np.float32(a.real * b.real) + np.float32(a.imag * b.imag)))
    return np.float32(s / np.float32(nifg))


def _neighbours(ps, r0, c0, offsets, nneigh):
    nrow, ncol = ps.shape
    got = []
    for dr, dc in offsets:
        r, c = r0 + dr, c0 + dc
        if 0 <= r < nrow and 0 <= c < ncol and ps[r, c]:
            got.append((r, c))
            if len(got) >= nneigh:
                break
    return got


def median_similarity(stack, ps, nneigh, offsets):
    """Median (element counter/2 of the sorted list) of the similarities
    to the first nneigh candidate neighbours, for candidate pixels."""
    _, nrow, ncol = stack.shape
    out = np.zeros((nrow, ncol), np.float32)
    for r0 in range(nrow):
        for c0 in range(ncol):
            if not ps[r0, c0]:
                continue
            sims = sorted(_sim(stack, r0, c0, r, c)
                          for r, c in _neighbours(ps, r0, c0, offsets, nneigh))
            out[r0, c0] = sims[len(sims) // 2] if sims else np.float32(0)
    return out


def update_ps(ps, sim, threshold, overwrite=True):
    ps = np.array(ps, np.int32, copy=True)
    hit = sim >= np.float32(threshold)
    ps[hit] = 1
    if overwrite:
        ps[~hit] = 0
    return ps


def maximum_similarity(stack, ps, nneigh, offsets):
    """1 for candidates; otherwise the largest similarity to the first
    nneigh candidate neighbours (0 without any)."""
    _, nrow, ncol = stack.shape
    out = np.zeros((nrow, ncol), np.float32)
    for r0 in range(nrow):
        for c0 in range(ncol):
            if ps[r0, c0]:
                out[r0, c0] = 1
                continue
            nb = _neighbours(ps, r0, c0, offsets, nneigh)
            best = np.float32(-1)  # the reference's start value
            for r, c in nb:
                s = _sim(stack, r0, c0, r, c)
                if best < s:
                    best = s
            out[r0, c0] = best if nb else 0
    return out


def similarity(stack, ps, nneigh, rdmin, rdmax, med_sim_th):
    """The executable's sequence: normalise, median similarity of the
    candidates, candidates <- median >= threshold, maximum similarity.
    Returns (med_sim, max_sim, updated candidates)."""
    if not 0 <= nneigh <= 128:
        raise ValueError("nneigh must be in [0, 128]")
    offs = scan_offsets(rdmin, rdmax)
    n = max(int(nneigh), 1)
    z = normalize(stack)
    ps = np.asarray(ps, np.int32)
    med = median_similarity(z, ps, n, offs)
    ps2 = update_ps(ps, med, med_sim_th, True)
    mx = maximum_similarity(z, ps2, n, offs)
    return med, mx, ps2
