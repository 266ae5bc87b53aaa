"""nmhd_amd -- MI355X-native ideal-MHD step (HIP kernels for gfx950 behind a
C ABI, include/nmhd.h), with the reference's solver interface on top.

    from nmhd_amd import capi, problems
    cfg = problems.config_c2((256, 256, 256))
    ctx = capi.Context(cfg)
    ctx.set_state(problems.shearing_box_mri(cfg))
    ctx.run(goal=1e30, max_steps=200)
"""
from . import problems  # noqa: F401

__all__ = ["problems", "capi", "adjoint", "postprocess"]
