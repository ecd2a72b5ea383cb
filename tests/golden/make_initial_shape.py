"""Writes tests/golden/initial_shape.npz from the reference's only input
file, demo/initialShape.mat (MATLAB 5.0: ndeIniMat 2 x 3131, lyrNdeNum,
lyrCntNum), read with scipy.io.loadmat (no code is executed from the file).

    python tests/golden/make_initial_shape.py /path/to/reference
"""
import os
import sys

import numpy as np
import scipy.io

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
d = scipy.io.loadmat(os.path.join(ref, "demo", "initialShape.mat"))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "initial_shape.npz")
np.savez_compressed(out, ndeIniMat=d["ndeIniMat"].astype(np.float64),
                    lyrNdeNum=int(d["lyrNdeNum"][0, 0]), lyrCntNum=int(d["lyrCntNum"][0, 0]))
print(out)
