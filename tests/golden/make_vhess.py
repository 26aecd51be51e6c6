"""Writes tests/golden/v_hess_example.npz: the reference's only data fixture,
examples/modeling/V_Hess.mat (a 400 x 1000 velocity model), as the example
driver uses it (examples/modeling/main_Acoustic_Data.m:50-53): every second
row and column from the second on, scaled by 1500/5000, giving a 200 x 500
model of 1500.0-4440.0 m/s with 12 distinct values.

    python tests/golden/make_vhess.py

Run in the build container (the reference is not on the GPU hosts).  The
.mat file is read with scipy.io.loadmat, which reads MATLAB 5 arrays as
data."""
import os

import numpy as np
import scipy.io

SRC = "/root/reference/examples/modeling/V_Hess.mat"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    d = scipy.io.loadmat(SRC)
    V = np.asarray(d["V"], dtype=np.float64)
    Vd = (V[1::2, 1::2] / 5000.0) * 1500.0
    np.savez_compressed(os.path.join(HERE, "v_hess_example.npz"), V=Vd, source_shape=np.array(V.shape))
    print(Vd.shape, Vd.min(), Vd.max(), len(np.unique(Vd)))
