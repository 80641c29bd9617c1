"""Writes cornell_fp64_4096spp.npz: the FP64 converged render of the Cornell
box benchmark scene (256x256, UV sphere 128x256, max path length 8) made by
the oracle's independent FP64 estimator (oracle/pt_fp64.cpp) at 4096 samples
per pixel. bench.py states the RMSE of the GPU's 64-iteration image against
it. Stored as float16 (relative rounding 2^-11, far below the RMSE measured)
with the mean squared standard error of the converged pixels, the noise floor
that RMSE figure carries from the reference itself.

Usage: python tests/golden/make_cornell_fp64.py  (about 6 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mufflon_amd import scenes as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

SPP = 4096
SEED = 20
MAX_PATH_LENGTH = 8


def main():
    c = S.cornell_box(128, 256, 256, 256)
    t0 = time.time()
    mean, var = O.fp64_render(c["verts"], c["normals"], c["tris"], c["mat_idx"], c["materials"], c["camera"],
                              MAX_PATH_LENGTH, SPP, seed=SEED)
    dt = time.time() - t0
    out = os.path.join(HERE, "cornell_fp64_4096spp.npz")
    np.savez_compressed(out, mean=mean.astype(np.float16), spp=SPP, seed=SEED, max_path_length=MAX_PATH_LENGTH,
                        mean_sq_std_error=float((var / SPP).mean()))
    print("wrote %s (%.0f s, %d bytes)" % (out, dt, os.path.getsize(out)))


if __name__ == "__main__":
    main()
