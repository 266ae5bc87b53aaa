"""On-device run diagnostics (nmhd_get_diagnostics) against the same
quantities computed with numpy from the state copied back: <B_d^2> over every
face with FP32 squares summed in FP64 (as the Kepler-disk monitor sums them,
application/kepler_disk_demo/mhd_solver.curb:107-130), interior totals and
max |div B|."""
import numpy as np
import pytest

from mhd_checks import div_b
from nmhd_amd import capi, problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["ot", "shear"])
def test_diagnostics_match_host(gpu_ready, case):
    print("\n[case] diagnostics " + case, flush=True)
    if case == "ot":
        cfg = P.config_c1(16)
        W = P.orszag_tang_noisy(cfg, seed=3)
    else:
        cfg = P.config_c2((16, 16, 8))
        W = P.shearing_box_mri(cfg, seed=4)
    g = capi.Context(cfg)
    g.set_state(W)
    g.run(1e30, 3)
    S = g.get_state()
    d = g.diagnostics()
    for k in range(3):
        want = (S[5 + k] * S[5 + k]).astype(np.float64).mean()
        assert d["b2_mean"][k] == pytest.approx(want, rel=1e-12)
    m = cfg["m"]
    ny, nz = cfg["n"][1], cfg["n"][2]
    sl = (slice(m[2], nz - m[2]), slice(m[1], ny - m[1]), slice(None))
    X = S.astype(np.float64)
    rho, v, p, b = X[0], X[1:4], X[4], X[5:8]
    bb = 0.25 * (b[0] ** 2 + np.roll(b[0], -1, 2) ** 2 + b[1] ** 2 + np.roll(b[1], -1, 1) ** 2 +
                 b[2] ** 2 + np.roll(b[2], -1, 0) ** 2)
    e = p / (cfg["gamma"] - 1.0) + 0.5 * rho * (v ** 2).sum(0) + bb
    want = [rho, rho * v[0], rho * v[1], rho * v[2], e, b[0], b[1], b[2]]
    for k in range(8):
        w = want[k][sl].sum()
        scale = np.abs(want[k][sl]).sum() + 1e-300
        assert abs(d["total"][k] - w) <= 1e-12 * scale, (k, d["total"][k], w)
    div = np.abs(div_b(S, cfg["d"]))[sl].max() * min(cfg["d"])
    assert d["divb_max"] == pytest.approx(div, rel=1e-6, abs=1e-12)
