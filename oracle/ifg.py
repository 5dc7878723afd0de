"""FP64 NumPy restatement of crossmul, looks, coherence and stitching (test
infrastructure).

Parity pins (tests/test_oracle_ifg.py):
  conj_mul           brute force on tiny arrays; phase = phi_sec - phi_ref
  cpx_looks          constant stays constant; ramp mean is the midpoint;
                     brute-force loop on tiny arrays
  overlap            disjoint / nested / look-unaligned cases
  powlooks/multilook 1x1 quirk (sqrt|s|); constant amplitude
  coherence          identical SLCs give 1; statistical KAT for gamma0 pairs;
                     Cauchy-Schwarz bound
  coherence_window   window=1 equals the look-cell estimator; identical
                     SLCs give 1 (no reference counterpart)
  stitch_subswath    injected phase offsets across a scene boundary
"""

import numpy as np


def conj_mul(a, b):
    """sario.cu:199-211: c = conj(a) * b (a = reference, b = secondary)."""
    return np.conj(np.asarray(a, np.complex128)) * np.asarray(b, np.complex128)


def cpx_looks(img, rowlook, collook):
    """sario.cu:241-286 as launched at crossmul_daemon.cu:940-956: column
    mean over `collook`, then row mean over `rowlook` (FP64)."""
    img = np.asarray(img)
    nrow, ncol = img.shape
    nr = nrow // rowlook
    nc = ncol // collook
    x = img[: nr * rowlook, : nc * collook].astype(np.complex128)
    x = x.reshape(nr * rowlook, nc, collook).sum(axis=2) / collook
    x = x.reshape(nr, rowlook, nc).sum(axis=1) / rowlook
    return x


def overlap(h1, h2, rowlook, collook):
    """crossmul_daemon.cu:606-666: intersection of two .gslc windows snapped
    inward to look multiples, and the 6-int output header.

    h1/h2: (nrow0, ncol0, left, top, right, bottom).  Returns None when the
    overlap is empty, else dict(left, top, right, bottom, nrow, ncol, header).
    """
    left = max(h1[2], h2[2])
    left = (left + collook - 1) // collook * collook
    right = min(h1[4], h2[4])
    right = right // collook * collook
    top = max(h1[3], h2[3])
    top = (top + rowlook - 1) // rowlook * rowlook
    bottom = min(h1[5], h2[5])
    bottom = bottom // rowlook * rowlook
    nrow = bottom - top
    ncol = right - left
    if nrow <= 0 or ncol <= 0:
        return None
    header = [h1[0] // rowlook, h1[1] // collook, left // collook, top // rowlook,
              right // collook, bottom // rowlook]
    return dict(left=left, top=top, right=right, bottom=bottom, nrow=nrow, ncol=ncol,
                header=header)


def crossmul_pair(ref, href, sec, hsec, rowlook, collook):
    """The daemon's per-task computation (crossmul_daemon.cu:606-666, 873-878,
    921-975): crop both images to the overlap, conj-multiply, look.  Returns
    (ifg_ml complex128, header) or None."""
    ov = overlap(href, hsec, rowlook, collook)
    if ov is None:
        return None
    r = ref[ov["top"] - href[3]: ov["bottom"] - href[3], ov["left"] - href[2]: ov["right"] - href[2]]
    s = sec[ov["top"] - hsec[3]: ov["bottom"] - hsec[3], ov["left"] - hsec[2]: ov["right"] - hsec[2]]
    return cpx_looks(conj_mul(r, s), rowlook, collook), ov["header"]


def cpxlooks_np(imgbg, rowlook, collook):
    """sario.py:397-413: rows summed first, then columns, / collook / rowlook."""
    naz, nrg = imgbg.shape
    newnaz = naz // rowlook
    newnrg = nrg // collook
    x = np.asarray(imgbg, np.float64)
    if rowlook > 1:
        x = x[: newnaz * rowlook].reshape(newnaz, rowlook, nrg).sum(axis=1)
    if collook > 1:
        x = x[:, : newnrg * collook].reshape(x.shape[0], newnrg, collook).sum(axis=2)
    return x / collook / rowlook


def powlooks(imgbg, rowlook, collook):
    """sario.py:416-420 (1x1 quirk: returns |s|, so amp = sqrt|s|)."""
    if rowlook == 1 and collook == 1:
        return np.abs(np.asarray(imgbg, np.complex128))
    return np.abs(cpxlooks_np(np.abs(np.asarray(imgbg, np.complex128)) ** 2, rowlook, collook))


def amplitude(slc, rowlook, collook):
    """coherence.py:46-48 for one look-aligned image: sqrt(powlooks)."""
    return np.sqrt(powlooks(slc, rowlook, collook))


def coherence_lookcell(ifg_ml, amp1, amp2, fill="clamp"):
    """coherence.py:142-146: |ifg|/(amp1+1e-12)/(amp2+1e-12); pixels > 1 are
    filled (the reference uses unseeded random numbers; here 'clamp' -> 1.0
    or 'mask' -> NaN so that parity tests can exclude them)."""
    c = np.abs(ifg_ml) / (amp1 + 1e-12) / (amp2 + 1e-12)
    bad = c > 1
    if fill == "clamp":
        c[bad] = 1.0
    elif fill == "mask":
        c[bad] = np.nan
    return c


def boxsum(x, win):
    """Sum over the win x win window centred on each pixel, clipped at the
    borders: direct sums of shifted copies, so all-zero windows give exactly
    0 (a running-sum form leaves rounding residue there)."""
    x = np.asarray(x)
    h = win // 2
    p = np.pad(x, h)
    n0, n1 = x.shape
    out = np.zeros_like(p[:n0, :n1])
    for dy in range(win):
        for dx in range(win):
            out = out + p[dy:dy + n0, dx:dx + n1]
    return out


def coherence_window(ifg_ml, pow1, pow2, win):
    """Windowed coherence on the multilooked grid (no reference counterpart):
    |sum_W ifg| / sqrt(sum_W pow1 * sum_W pow2) over a win x win window
    clipped at the borders; 0 where the denominator is 0."""
    si = boxsum(np.asarray(ifg_ml, np.complex128), win)
    p1 = boxsum(np.asarray(pow1, np.float64), win)
    p2 = boxsum(np.asarray(pow2, np.float64), win)
    den = np.sqrt(p1 * p2)
    out = np.zeros(den.shape)
    ok = den > 0
    out[ok] = np.abs(si[ok]) / den[ok]
    return out


def cc_rmg(amp1, amp2, coh):
    """coherence.py:145-147 + sario.py:387-394: rmg rows [amp1*amp2 | coh]."""
    return np.concatenate([(amp1 * amp2).astype(np.float32), coh.astype(np.float32)], axis=1)


def match_bursts(ref_bursts, sec_bursts):
    """interfere.py:16-45 on (top, bottom) pairs: index of the best secondary
    burst for every reference burst (argmin |dtop| + |dbottom|, first wins)."""
    st = np.array([b[0] for b in sec_bursts])
    sb = np.array([b[1] for b in sec_bursts])
    out = []
    for t, b in ref_bursts:
        out.append(int(np.argmin(np.abs(st - t) + np.abs(sb - b))))
    return out


def stitch_subswath(bursts, ids, out_float=False):
    """interfere.py:81-150 on in-memory bursts.

    bursts: list of (header6, data) already sorted by top+bottom, data
    complex64 (out_float False) or float32 phase (out_float True); ids: list
    of (unique_idx1, unique_idx2) per burst.  Returns the stitched array
    covering the subswath bounds (first/last burst), the bounds, and the
    phase offset applied to each burst (the running mean_phase_diff).
    """
    dtype = np.float32 if out_float else np.complex64
    hs = [b[0] for b in bursts]
    left = min(hs[0][2], hs[-1][2])
    top = min(hs[0][3], hs[-1][3])
    right = max(hs[0][4], hs[-1][4])
    bottom = max(hs[0][5], hs[-1][5])
    out = np.zeros((bottom - top, right - left), dtype)
    mean_phase_diff = 0.0
    prev = None
    offsets = []
    for i, ((h, data), uid) in enumerate(zip(bursts, ids)):
        if i == 0:
            # CroppedImage.load_data -> resample into the subswath bounds
            _paste(out, h, data, left, top, right, bottom)
            prev = uid
            offsets.append(0.0)
            continue
        old = out[h[3] - top: h[5] - top, h[2] - left: h[4] - left]
        new = data
        if out_float:
            replace = (old == 0) & (new != 0)
        else:
            replace = (old.real == 0) & (new.real != 0)
        ovm = None
        if uid[0] != prev[0] or uid[1] != prev[1]:
            ovm = (old != 0) & (new != 0) if out_float else (old.real != 0) & (new != 0)
        prev = uid
        if ovm is not None and np.any(ovm):
            if out_float:
                pdiff = np.exp(1j * (-old[ovm] + new[ovm]))
            else:
                pdiff = np.conj(old[ovm]) * new[ovm]
            mean_phase_diff = np.angle(np.mean(pdiff))
            pdiff = pdiff * np.exp(-1j * mean_phase_diff)
            mean_phase_diff += np.median(np.angle(pdiff))
        if mean_phase_diff != 0:
            if out_float:
                new = new - mean_phase_diff
                new[new > np.pi] -= 2 * np.pi
                new[new < -np.pi] += 2 * np.pi
            else:
                new = new * np.exp(-1j * mean_phase_diff)
        old[replace] = new[replace]
        offsets.append(float(mean_phase_diff))
    return out, (left, top, right, bottom), offsets


def merge_subswaths(subs, nrow0, ncol0, out_float=False):
    """interfere.py:153-180 in memory: subswath mosaics (bounds4, data) into
    one (nrow0, ncol0) image; later subswaths overwrite where their value
    (real part) is non-zero."""
    dtype = np.float32 if out_float else np.complex64
    mm = np.zeros((nrow0, ncol0), dtype)
    for (left, top, right, bottom), data in subs:
        old = mm[top:bottom, left:right]
        rep = (data != 0) if out_float else (data.real != 0)
        old[rep] = data[rep]
    return mm


def _paste(dst, h, data, left, top, right, bottom):
    ol, orr = max(left, h[2]), min(right, h[4])
    ot, ob = max(top, h[3]), min(bottom, h[5])
    dst[ot - top: ob - top, ol - left: orr - left] = data[ot - h[3]: ob - h[3], ol - h[2]: orr - h[2]]
