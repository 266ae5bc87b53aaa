"""Write tests/golden/hlld_ref.npz: HLLD face states and the fluxes that the
reference's own hlld_flux (library/hlld.inl, compiled for the host into
oracle/_ref/libref_hlld.so by oracle/build.py) returns for them.

The states cover every fan region of the solver (supersonic left/right,
single-star and double-star on both sides), Bx = 0 (signBx = 0), equal
left/right states and strong jumps.  Run from the repository root after
`python oracle/build.py`:  python tests/golden/make_hlld_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hlld_ref.npz")


def face_states(seed=20240, n=1200):
    rng = np.random.default_rng(seed)
    x = np.zeros((15, n), np.float32)
    x[0] = rng.uniform(-2, 2, n)
    ranges = [(0.1, 10), (-3, 3), (-3, 3), (-3, 3), (0.01, 10), (-2, 2), (-2, 2)]
    for i, (lo, hi) in enumerate(ranges):
        x[1 + 2 * i] = rng.uniform(lo, hi, n)
        x[2 + 2 * i] = rng.uniform(lo, hi, n)
    k = n // 6
    x[0, :k] = 0.0                       # Bx = 0
    x[3, k:2 * k] += 15.0                # supersonic to the right
    x[4, k:2 * k] += 15.0
    x[3, 2 * k:3 * k] -= 15.0            # supersonic to the left
    x[4, 2 * k:3 * k] -= 15.0
    sl = slice(3 * k, 4 * k)             # equal left/right states
    for i in range(7):
        x[2 + 2 * i, sl] = x[1 + 2 * i, sl]
    sl = slice(4 * k, 5 * k)             # strong jumps
    x[1, sl] *= 100.0
    x[9, sl] *= 1000.0
    return x


def main():
    if pyoracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref_hlld.so missing: run python oracle/build.py")
    x = face_states()
    out = {}
    for name, gamma in (("g53", 5.0 / 3.0), ("g75", 7.0 / 5.0)):
        out["flux_" + name] = pyoracle.hlld_batch(x, gamma=gamma, use_ref=True)
    np.savez_compressed(OUT, states=x, gammas=np.array([5.0 / 3.0, 7.0 / 5.0], np.float32),
                        **out)
    print("wrote", OUT, x.shape)


if __name__ == "__main__":
    main()
