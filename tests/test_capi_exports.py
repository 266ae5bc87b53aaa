"""The C-ABI library loads on a CPU-only host and exports every entry point
that include/nmhd.h declares (no compute calls: no GPU here)."""
import os
import re

from nmhd_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(nmhd_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = capi.load()
    decl = declared_symbols()
    assert "nmhd_create" in decl and len(decl) >= 20
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) == decl


def test_create_rejects_bad_descriptor_without_gpu():
    """Descriptor validation happens before any device call."""
    from nmhd_amd import problems as P
    cfg = P.make_config((8, 8, 8))
    d = capi.desc_from_config(cfg)
    d.margin[0] = 2
    import ctypes
    h = ctypes.c_void_p()
    rc = capi.load().nmhd_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == 1 and not h.value
    assert b"margin[0]" in capi.load().nmhd_last_error(None)


def test_batch_create_validates_every_box_without_gpu():
    """Each box's descriptor is checked as a context's is, and the batch's
    own limits (one integrator, 32767 cells per axis with ghosts), before
    any device call."""
    from nmhd_amd import problems as P
    good = P.make_config((16, 16, 16))
    import pytest
    with pytest.raises(capi.NmhdError, match="integrator"):
        capi.Batch([good, P.make_config((8, 8, 8), integrator=1)])
    with pytest.raises(capi.NmhdError, match="32767"):
        capi.Batch([good, P.make_config((4, 32764, 4), margin=(0, 2, 0))])
    bad = dict(P.config_c2((16, 16, 16)), fixes=8)
    with pytest.raises(capi.NmhdError, match="box 1: unknown bit in fixes"):
        capi.Batch([good, bad])


def _bad_desc_cases():
    from nmhd_amd import problems as P
    base = P.make_config((8, 8, 8))
    cases = []

    def case(name, edit, msg):
        cases.append((name, edit, msg))

    case("zero size", lambda d: d.global_size.__setitem__(1, 0), b"global_size")
    case("negative margin", lambda d: d.margin.__setitem__(2, -1), b"margin")
    case("zero extent", lambda d: d.global_extent.__setitem__(0, 0.0), b"global_extent")
    case("gamma 1", lambda d: setattr(d, "gamma", 1.0), b"gamma")
    case("cfl 0", lambda d: setattr(d, "cfl_factor", 0.0), b"cfl_factor")
    case("source 7", lambda d: setattr(d, "source", 7), b"source")
    case("recon 5", lambda d: setattr(d, "recon", 5), b"recon")
    case("integrator 3", lambda d: setattr(d, "integrator", 3), b"integrator")
    case("path 9", lambda d: setattr(d, "path", 9), b"path")
    case("shear without margin", lambda d: setattr(d, "shear_y", 1), b"margin[1]")
    case("margin beyond interior", lambda d: d.margin.__setitem__(1, 9), b"margin[1]")
    case("2^31 cells", lambda d: [d.global_size.__setitem__(i, 1291) for i in range(3)],
         b"2^31")
    case("2^30-cell plane", lambda d: [d.global_size.__setitem__(i, v)
                                       for i, v in enumerate((32768, 32768, 1))], b"2^30")
    case("fixes", lambda d: setattr(d, "fixes", 8), b"fixes")
    return base, cases


def test_create_rejects_every_invalid_descriptor_without_gpu():
    """Each field of the descriptor is range-checked on the host before any
    device call, with a message naming the field."""
    import ctypes
    base, cases = _bad_desc_cases()
    L = capi.load()
    for name, edit, msg in cases:
        d = capi.desc_from_config(base)
        edit(d)
        h = ctypes.c_void_p()
        rc = L.nmhd_create(ctypes.byref(d), ctypes.byref(h))
        assert rc == 1 and not h.value, name
        assert msg in L.nmhd_last_error(None), (name, L.nmhd_last_error(None))


def test_noalloc_step_rejects_before_any_device_call():
    """nmhd_adj_forward_noalloc checks its descriptor, the box size against
    its fixed scratch (2^20 cells) and aliasing before it touches a device;
    the pointers below are never dereferenced."""
    import ctypes
    from nmhd_amd import adjoint as A, problems as P
    L = A._lib()
    w_in = (ctypes.c_void_p * 8)(*[0x1000 * (i + 1) for i in range(8)])
    w_out = (ctypes.c_void_p * 8)(*[0x100000 * (i + 1) for i in range(8)])
    cases = [
        (P.make_config((128, 128, 65)), w_out, b"exceeds the no-allocation scratch"),
        (P.config_c2((16, 16, 16)), w_out, b"periodic boxes"),
        (P.make_config((16, 16, 16)), w_in, b"alias"),
    ]
    for cfg, out, msg in cases:
        d = capi.desc_from_config(cfg)
        assert L.nmhd_adj_forward_noalloc(ctypes.byref(d), w_in, 1e-3, out, None) == 1
        assert msg in L.nmhd_adj_last_error(None)
