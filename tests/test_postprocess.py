"""The dump post-processor (nmhd_amd/postprocess.py, the reference's
tools/parallel/to_gnuplot_3d.cpp) on synthetic two-rank dumps with a known
shearing displacement: radial ghosts are dropped, the outer rank is shifted
along the azimuth by its integrated displacement and its v_x boosted, and the
cross sections have the tool's layout."""
import os

import numpy as np

from nmhd_amd import postprocess as PP

F32 = np.float32


def write_rank(d, rank, glob, loc, fields):
    def block(h):
        return "%d\n" % len(h) + "".join("%s %s\n" % (k, v) for k, v in h.items())
    with open(os.path.join(d, "%04d_setup.txt" % rank), "w") as f:
        f.write(block(glob) + "################\n" + block(loc))
    sz, sy, sx = fields.shape[1:]
    with open(os.path.join(d, "%04d.bin" % rank), "wb") as f:
        np.array([sx, sy, sz], np.int32).tofile(f)
        fields.astype(F32).tofile(f)


def test_two_radial_ranks_with_displacement(tmp_path):
    d = str(tmp_path)
    nx, ny_loc, nz, m = 8, 4, 4, 1
    dx = 1.0 / nx
    glob = {"generation": 3, "real_size": 4, "real_precision": 6, "current_time": 0.5,
            "parallel_size": 2, "global_size_x": nx, "parallel_size_x": 1,
            "global_extent_x": 1.0, "global_corner_x": 0.0, "global_size_y": 2 * ny_loc,
            "parallel_size_y": 2, "global_extent_y": 1.0, "global_corner_y": -0.5,
            "global_size_z": nz, "parallel_size_z": 1, "global_extent_z": 1.0,
            "global_corner_z": 0.0}
    dy = 1.0 / (2 * ny_loc)
    shift_cells, vrel = 2, 0.75
    truth = np.zeros((8, nz, 2 * ny_loc, nx))
    rng = np.random.default_rng(0)
    truth[:] = rng.normal(size=truth.shape)
    for rank in range(2):
        sy = ny_loc + 2 * m
        loc = {"mpi_rank": rank, "relative_displacement": shift_cells * dx,
               "relative_velocity": vrel, "computational_size": nx * sy * nz,
               "computational_size_x": nx, "computational_corner_x": 0.0,
               "computational_extent_x": 1.0, "computational_size_y": sy,
               "computational_corner_y": -0.5 + (ny_loc * rank - m) * dy,
               "computational_extent_y": sy * dy, "computational_size_z": nz,
               "computational_corner_z": 0.0, "computational_extent_z": 1.0}
        # the rank's rows incl. ghosts: ghosts inside the global box hold the
        # neighbour's values (they are averaged in, as the tool does), the
        # ones outside it hold garbage (they must be dropped)
        f = np.full((8, nz, sy, nx), 1e6)
        lo = rank * ny_loc - m
        for j in range(sy):
            g = lo + j
            if 0 <= g < 2 * ny_loc:
                f[:, :, j, :] = truth[:, :, g, :]
        if rank == 1:  # stored in the rank's own frame
            f = np.roll(f, -shift_cells, axis=3)
            f[1] -= vrel
        write_rank(d, rank, glob, loc, f)
    grid, arrays = PP.accumulate(d)
    assert grid.size == [nx, 2 * ny_loc, nz]
    assert np.allclose(arrays, truth.astype(F32), rtol=0, atol=1e-6)
    names = PP.write_cross_sections(grid, arrays, os.path.join(d, "out"))
    xy = open(names[2]).read().split("\n")
    # xy section: one line per point, a blank line after each row of x
    assert sum(1 for l in xy if l.strip()) == nx * 2 * ny_loc
    assert xy[nx] == ""
    first = [float(v) for v in xy[0].split()]
    assert len(first) == 11 and abs(first[0] - 0.5 * dx) < 1e-6
