"""Converts the reference's COBAHH golden data (dev/tests/correctness/*.txt.gz:
two 0/1 connection matrices and the recorded (i, t) raster) into a compact
fixture: nonzero coordinates of Ce/Ci and the raster as (step, index).

Run once in the build container (reads /root/reference); the output
tests/golden/cobahh.npz is committed.  The data are the reference test's
inputs and expected outputs, stored in another encoding.
"""
import os

import numpy as np

SRC = "/root/reference/dev/tests/correctness"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cobahh.npz")


def main():
    Ce = np.loadtxt(os.path.join(SRC, "Ce_matrix.txt.gz"))
    Ci = np.loadtxt(os.path.join(SRC, "Ci_matrix.txt.gz"))
    sp = np.loadtxt(os.path.join(SRC, "COBAHH_spikes.txt.gz"))
    assert set(np.unique(Ce)) <= {0.0, 1.0} and set(np.unique(Ci)) <= {0.0, 1.0}
    ce_r, ce_c = np.nonzero(Ce)
    ci_r, ci_c = np.nonzero(Ci)
    idx = sp[:, 0].astype(np.int64)
    t = sp[:, 1]
    step = np.rint(t / 1e-4).astype(np.int64)
    assert np.all(np.abs(step * 1e-4 - t) < 1e-9)
    np.savez_compressed(OUT, ce_shape=np.array(Ce.shape), ce_r=ce_r.astype(np.int16),
                        ce_c=ce_c.astype(np.int16), ci_shape=np.array(Ci.shape),
                        ci_r=ci_r.astype(np.int16), ci_c=ci_c.astype(np.int16),
                        spike_idx=idx.astype(np.int16), spike_step=step.astype(np.int16),
                        spike_t=t)
    print(OUT, len(ce_r), len(ci_r), len(idx))


if __name__ == "__main__":
    main()
