"""Writes tests/golden/ann_deck_240x1792.npz from the annulus reference
executable (oracle/_ref/ann_ref = velvor2dnewsch3.4.f compiled unmodified;
only in the container that holds the reference sources).

The deck is the reference's own input1e81e5pr1-1.s (dt 4e-8, nr 240,
np 1792, Ra 1e8, Re 1e5, Gr 0, Pr 0.1, pert 0, amp 1, nval 3) with these
changes: 50 steps instead of 5000, the program's own initial condition
instead of the restart files (which the reference does not ship), output
file names A..W and T, and no further iterations.  The fixture holds the
reference's outputs only:
  heat:   the heat-transfer row the program prints at step 50;
  stride, n_<F>, s_<F>: a sample of every stride-th value of each field
          file F (T, C, U, V; one value per line), with the file's length;
  norm_<F>: the maximum absolute value of the whole file;
  means:  zmtmtempvar.dat (the running means).

    python tests/golden/make_golden_anndeck.py     (about 4 minutes)
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "ann_ref")
NTS = 50
STRIDE = 97
DECK = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "4.d-8", "240", "1792", str(NTS),
                  "0.5d0,1.0d0,0.1d0", "1.d8", "1.d5", "0.d0", "0.1d0", "0.d0", "1.d0", "3", "no", "n"]) + "\n"


def heat_rows(out):
    rows = []
    for line in out.splitlines():
        t = line.split()
        if len(t) == 3 and t[-1].isdigit():
            rows.append([float(t[0]), float(t[1])])
    return np.array(rows)


def main():
    with tempfile.TemporaryDirectory() as d:
        res = subprocess.run([REF], input=DECK, text=True, capture_output=True, cwd=d, check=True,
                             env=dict(os.environ, OMP_NUM_THREADS="1"))
        out = {"heat": heat_rows(res.stdout), "stride": STRIDE, "nts": NTS}
        for f in ("T", "C", "U", "V"):
            v = np.loadtxt(os.path.join(d, f))
            out[f"n_{f}"] = v.size
            out[f"s_{f}"] = v[::STRIDE].copy()
            out[f"norm_{f}"] = float(np.max(np.abs(v)))
        out["means"] = np.loadtxt(os.path.join(d, "zmtmtempvar.dat"))
    np.savez_compressed(os.path.join(HERE, "ann_deck_240x1792.npz"), **out)
    print({k: (np.shape(v) if hasattr(v, "shape") else v) for k, v in out.items()})


if __name__ == "__main__":
    main()
