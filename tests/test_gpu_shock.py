"""The reference's MHD shock-tube demo (application/mhd_shock_demo: 512 x 1
x 1 cells, run to t = 0.1) on the GPU against the oracle: the strict path
bit for bit; the default (tiled) path, for each shock normal and both the
second- and first-order schemes, within the oracle's own sensitivity to a
one-ulp change of the initial state (a shock amplifies round-off)."""
import numpy as np
import pytest

import pyoracle as po
from mhd_checks import bitwise_equal, group_rel_err, perturb_ulps
from nmhd_amd import capi, problems as P

pytestmark = pytest.mark.gpu


def announce(msg):
    print("\n[case] " + msg, flush=True)


def test_strict_shock_tube_bitwise(gpu_ready):
    announce("strict path, shock tube 512x1x1 to t = 0.1, bitwise")
    cfg, W = P.shock_tube(0)
    o = po.Oracle(cfg, W)
    g = capi.Context(cfg, path=capi.PATH_STRICT)
    g.set_state(W)
    ko, done_o = o.run(0.1, 100000)
    kg, done_g = g.run(0.1, 100000)
    assert done_o and done_g and ko == kg
    assert g.get_time()[0] == o.time
    assert bitwise_equal(g.get_state(), o.get_state())


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("scheme", [(0, 0), (1, 1)])
def test_tiled_shock_tube_within_oracle_envelope(gpu_ready, axis, scheme):
    announce("tiled path, shock tube along axis %d, recon/integrator %s" % (axis, scheme))
    cfg, W = P.shock_tube(axis, 512, *scheme)
    o = po.Oracle(cfg, W)
    o2 = po.Oracle(cfg, perturb_ulps(W))
    g = capi.Context(cfg, path=capi.PATH_TILED)
    g.set_state(W)
    ko, _ = o.run(0.1, 100000)
    o2.run(0.1, 100000)
    kg, done = g.run(0.1, 100000)
    assert done and abs(kg - ko) <= 1
    a, b, c = g.get_state(), o.get_state(), o2.get_state()
    env, err = group_rel_err(c, b), group_rel_err(a, b)
    l1 = np.abs(a.astype(np.float64) - b).mean() / np.abs(b).mean()
    print("group err %s, oracle 1-ulp envelope %s, L1 rel %.3g"
          % (np.array2string(err, precision=2), np.array2string(env, precision=2), l1))
    assert np.all(err <= 3.0 * env + 1e-5)
    assert l1 < 1e-5
    assert np.all(a[5 + axis] == np.float32(0.75))
