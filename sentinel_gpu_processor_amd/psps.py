"""PS candidate selection by phase similarity: the reference's
``s1proc.psps._ps_selection`` (psps.py:11-78) and the ``phase_similarity``
executable it runs (phase_similarity.cu:375-569), in process on the GPU.

The file contract is the executable's: a list of raw interferograms
(nrow x ncol complex64 each, no header), an int32 candidate file (1 = PS
candidate), and two float32 outputs, the median similarity of the
candidates and the maximum similarity of the pixels that are not PS after
the median test.  The whole stack is read once and stays resident (no row
batches).  ``max_sim_th`` is accepted and unused, as in the reference
(the executable reads eight arguments, psps.py:63-76 passes nine).

Entry points:
  _ps_selection(ifg_list, nrow, ncol, ps_file, med_sim_outfile,
                max_sim_outfile, nneigh, rdmin, rdmax, med_sim_th, max_sim_th)
  phase_similarity(infile, psfile, med_sim_outfile, max_sim_outfile,
                   nneigh, rdmin, rdmax, med_sim_th)   (the executable)
  python -m sentinel_gpu_processor_amd.psps infile psfile med max N rdmin rdmax th
"""

import os
import sys

import numpy as np

from . import kernels as K
from .runtime import context, device


def read_ifg_list(infile):
    """(nifg, nrow, ncol, names) from the executable's list file: a line
    "nifg nrow ncol" then the names (phase_similarity.cu:68-88)."""
    with open(infile) as f:
        tok = f.read().split()
    nifg, nrow, ncol = int(tok[0]), int(tok[1]), int(tok[2])
    names = tok[3:3 + nifg]
    if len(names) != nifg:
        raise ValueError(f"{infile}: {nifg} interferograms announced, {len(names)} listed")
    return nifg, nrow, ncol, names


def load_stack(names, nrow, ncol, dev):
    """The interferograms as one device stack [nifg][nrow][ncol]."""
    import torch
    st = torch.empty((len(names), nrow, ncol), dtype=torch.complex64, device=dev)
    n = nrow * ncol
    for k, fn in enumerate(names):
        a = np.fromfile(fn, dtype=np.complex64, count=n)
        if a.size != n:
            raise IOError(f"{fn}: expected {n} complex samples, read {a.size}")
        st[k].copy_(torch.from_numpy(a.reshape(nrow, ncol)))
    return st


def similarity_arrays(stack, ps, nneigh, rdmin, rdmax, med_sim_th):
    """Device tensors in, (med_sim, max_sim) device tensors out; ps (int32)
    updated in place to the selected scatterers."""
    return K.ps_similarity(context(), stack, ps, nneigh, rdmin, rdmax, med_sim_th)


def phase_similarity(infile, psfile, med_sim_outfile, max_sim_outfile, nneigh, rdmin, rdmax,
                     med_sim_th):
    """The ``phase_similarity`` executable (phase_similarity.cu:375-569)."""
    import torch
    nifg, nrow, ncol, names = read_ifg_list(infile)
    dev = device()
    stack = load_stack(names, nrow, ncol, dev)
    ps = np.fromfile(psfile, dtype=np.int32, count=nrow * ncol)
    if ps.size != nrow * ncol:
        raise IOError(f"{psfile}: expected {nrow * ncol} int32 values, read {ps.size}")
    ps_t = torch.from_numpy(ps.reshape(nrow, ncol)).to(dev)
    med, mx = similarity_arrays(stack, ps_t, int(nneigh), int(rdmin), int(rdmax),
                                float(med_sim_th))
    K.to_numpy(med).astype(np.float32).tofile(med_sim_outfile)
    K.to_numpy(mx).astype(np.float32).tofile(max_sim_outfile)


def _ps_selection(ifg_list, nrow, ncol, ps_file, med_sim_outfile, max_sim_outfile, nneigh,
                  rdmin, rdmax, med_sim_th, max_sim_th):
    """Run PS selection for an interferogram list (psps.py:11-78): writes
    the same list file under ``temp/`` as the reference, then the
    executable's computation in process."""
    temp_dir = "temp"
    os.makedirs(temp_dir, exist_ok=True)
    b1 = os.path.basename(ifg_list[0])
    b2 = os.path.basename(ifg_list[1])
    infile = os.path.join(temp_dir, f"infile_{b1[0:17]}_{b2[0:17]}.txt")
    with open(infile, "w") as f:
        f.write(f"{len(ifg_list)} {nrow} {ncol}\n")
        f.write("\n".join(ifg_list))
    phase_similarity(infile, ps_file, med_sim_outfile, max_sim_outfile, nneigh, rdmin, rdmax,
                     med_sim_th)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 8:
        print("Usage: phase_similarity infile psfile med_sim_outfile max_sim_outfile N rdmin "
              "rdmax med_sim_th")
        return 0
    phase_similarity(argv[0], argv[1], argv[2], argv[3], int(argv[4]), int(argv[5]),
                     int(argv[6]), float(argv[7]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
