"""Measurement (not collected by pytest; about 9 minutes on 16 host threads):
the oracle's own max |div B| dx over the interior after 1, 50, 100, 150 and
200 steps of the 256^3 shearing box (seed 1002), absolute and relative to
max |B|.  Constrained transport keeps div B at zero in exact arithmetic; in
FP32 each face update rounds, and the reference's arithmetic accumulates
these roundings as a random walk.  The device's value after the same 200
steps (tests/test_gpu_full_size.py::test_c2_full_size_200_steps_div_b) is
compared with these figures.

    python tests/measure_divb_oracle_256.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, os.path.join(ROOT, "oracle"), HERE):
    sys.path.insert(0, p)

import pyoracle as po  # noqa: E402
from mhd_checks import div_b  # noqa: E402
from nmhd_amd import problems as P  # noqa: E402


def main():
    cfg = P.config_c2((256, 256, 256))
    o = po.Oracle(cfg, P.shearing_box_mri(cfg, seed=1002))
    m, ny = cfg["m"][1], cfg["n"][1]
    t0 = time.time()
    done = 0
    for target in (1, 50, 100, 150, 200):
        o.run(1e30, target - done)
        done = target
        S = o.get_state()
        dv = np.abs(div_b(S, cfg["d"])[:, m:ny - m, :]).max() * cfg["d"][0]
        print("oracle step %d: div B dx = %.3g, / max|B| = %.3g  (%.0f s)"
              % (done, dv, dv / np.abs(S[5:8]).max(), time.time() - t0), flush=True)
    o.close()


if __name__ == "__main__":
    main()
