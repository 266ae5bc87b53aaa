"""Post-processing of state dumps: the reference's gnuplot cross-section
tool (tools/parallel/to_gnuplot_3d.cpp:122-291) for the files that
`MHDSolver::save_state` / `dump` write (`<dir>/<tag>/<rank>.bin` and
`<rank>_setup.txt`), reference dumps included.

    python -m nmhd_amd.postprocess <dir>/<tag> <output_prefix>

For every rank's dump: cell centres of the whole array (ghosts included) are
shifted along the azimuth (X) by the shearing displacement integrated over
the radial (Y) ranks below, v_X is boosted by the integrated relative
velocity, and each value is averaged into an output grid of at most 120
cells per axis over the global box; points that fall outside the box along
the radial axis (the radial ghosts) are dropped, X and Z wrap periodically.
Writes the three mid-plane cross sections `<prefix>_yz.txt`, `_xz.txt`,
`_xy.txt` as gnuplot `splot` blocks: `x y z rho vx vy vz p Bx By Bz`.
"""
import os
import sys

import numpy as np

F32 = np.float32
MAX_RESOLUTION = 120
ITEMS = 8


def read_block(tokens, pos):
    """One setup block: a count, then that many `key value` pairs."""
    n = int(tokens[pos])
    pos += 1
    out = {}
    for _ in range(n):
        out[tokens[pos]] = tokens[pos + 1]
        pos += 2
    return out, pos


def read_setup(path):
    """(global, local) key -> value-string dictionaries of `<rank>_setup.txt`
    (GlobalSetup, a spacer line, LocalSetup)."""
    tokens = open(path).read().split()
    glob, pos = read_block(tokens, 0)
    pos += 1  # spacer
    loc, _ = read_block(tokens, pos)
    return glob, loc


def read_bin(path):
    """(sx, sy, sz, fields) of a `.bin` dump: 3 int32 sizes, then 8 FP32
    arrays of sx*sy*sz values (x fastest)."""
    with open(path, "rb") as f:
        sx, sy, sz = (int(v) for v in np.fromfile(f, dtype=np.int32, count=3))
        data = np.fromfile(f, dtype=F32, count=ITEMS * sx * sy * sz)
    if data.size != ITEMS * sx * sy * sz:
        raise ValueError("%s: truncated dump" % path)
    return sx, sy, sz, data.reshape(ITEMS, sz, sy, sx)


class Grid:
    """Cell grid over a box (the tool's Accessor)."""

    def __init__(self, size, corner, extent):
        self.size = [int(v) for v in size]
        self.corner = [F32(v) for v in corner]
        self.extent = [F32(v) for v in extent]

    def centres(self, axis):
        n = self.size[axis]
        i = np.arange(n, dtype=F32)
        return self.corner[axis] + (i + F32(0.5)) / F32(n) * self.extent[axis]

    def index(self, r, axis):
        """floor((r - corner) / extent * size) in FP32."""
        v = (r.astype(F32) - self.corner[axis]) / self.extent[axis] * F32(self.size[axis])
        return np.floor(v).astype(np.int64)


def accumulate(snapshot_dir, max_resolution=MAX_RESOLUTION):
    """Average every rank's dump onto the output grid; returns (grid, arrays)."""
    glob, loc0 = read_setup(os.path.join(snapshot_dir, "0000_setup.txt"))

    def get(h1, h2, key, cast):
        if key in h1:
            return cast(h1[key])
        if key in h2:
            return cast(h2[key])
        raise KeyError("setup key not found: " + key)

    if get(loc0, glob, "real_size", int) != 4:
        raise ValueError("only FP32 dumps are supported")
    out = Grid([min(max_resolution, get({}, glob, "global_size_" + a, int)) for a in "xyz"],
               [get({}, glob, "global_corner_" + a, float) for a in "xyz"],
               [get({}, glob, "global_extent_" + a, float) for a in "xyz"])
    n_out = out.size[0] * out.size[1] * out.size[2]
    num = np.zeros((ITEMS, n_out))
    den = np.zeros(n_out)
    npy = get({}, glob, "parallel_size_y", int)
    nranks = get({}, glob, "parallel_size", int)
    disp = [0.0] * npy
    velo = [0.0] * npy
    for rank in range(nranks):
        _, loc = read_setup(os.path.join(snapshot_dir, "%04d_setup.txt" % rank))
        src = Grid([get(loc, glob, "computational_size_" + a, int) for a in "xyz"],
                   [get(loc, glob, "computational_corner_" + a, float) for a in "xyz"],
                   [get(loc, glob, "computational_extent_" + a, float) for a in "xyz"])
        rank_y = get(loc, glob, "mpi_rank", int) % npy
        if rank_y < npy - 1:
            disp[rank_y + 1] = disp[rank_y] + float(F32(get(loc, glob, "relative_displacement",
                                                            float)))
            velo[rank_y + 1] = velo[rank_y] + float(F32(get(loc, glob, "relative_velocity",
                                                            float)))
        sx, sy, sz, fields = read_bin(os.path.join(snapshot_dir, "%04d.bin" % rank))
        if [sx, sy, sz] != src.size:
            raise ValueError("rank %d: dump sizes differ from the setup" % rank)
        rz, ry, rx = np.meshgrid(src.centres(2), src.centres(1), src.centres(0), indexing="ij")
        rx = rx + F32(disp[rank_y])
        ix = out.index(rx, 0) % out.size[0]
        iy0 = out.index(ry, 1)
        iy = iy0 % out.size[1]
        iz = out.index(rz, 2) % out.size[2]
        keep = (iy == iy0).ravel()  # no radial carry: inside the box
        addr = ((iz * out.size[1] + iy) * out.size[0] + ix).ravel()[keep]
        den += np.bincount(addr, minlength=n_out)
        for item in range(ITEMS):
            vals = fields[item].ravel()[keep].astype(np.float64)
            if item == 1:
                vals = vals + velo[rank_y]
            num[item] += np.bincount(addr, weights=vals, minlength=n_out)
    with np.errstate(invalid="ignore", divide="ignore"):
        arrays = (num / den).reshape(ITEMS, out.size[2], out.size[1], out.size[0])
    return out, arrays


def write_cross_sections(grid, arrays, prefix):
    """The three mid-plane cross sections in the tool's gnuplot layout."""
    nx, ny, nz = grid.size
    cx, cy, cz = grid.centres(0), grid.centres(1), grid.centres(2)
    names = []
    for axis, tag in ((0, "yz"), (1, "xz"), (2, "xy")):
        xr = range(nx // 2, nx // 2 + 1) if axis == 0 else range(nx)
        yr = range(ny // 2, ny // 2 + 1) if axis == 1 else range(ny)
        zr = range(nz // 2, nz // 2 + 1) if axis == 2 else range(nz)
        out = []  # the tool's stream: a newline after a point only when the
        # loop it closes spans more than one index
        for z in zr:
            for y in yr:
                for x in xr:
                    vals = " ".join("%g" % arrays[i, z, y, x] for i in range(ITEMS))
                    out.append("%g %g %g %s" % (cx[x], cy[y], cz[z], vals))
                    if len(xr) > 1:
                        out.append("\n")
                if len(yr) > 1:
                    out.append("\n")
            if len(zr) > 1:
                out.append("\n")
        name = "%s_%s.txt" % (prefix, tag)
        with open(name, "w") as f:
            f.write("".join(out))
        names.append(name)
    return names


def main(argv):
    if len(argv) < 3:
        print("usage: python -m nmhd_amd.postprocess folder/of/snapshot output/prefix",
              file=sys.stderr)
        return 2
    grid, arrays = accumulate(argv[1])
    for name in write_cross_sections(grid, arrays, argv[2]):
        print(name)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
