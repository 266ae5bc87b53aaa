"""The C++ facade (include/nmhd/): addressing laws on the CPU, and the
MHDSolver drop-in (proceed / save_state / load_state / read / write) on the
GPU through libnmhd.so.  The programs are built by __graft_entry__.build()."""
import os

import numpy as np
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def run(exe, *args, timeout=120):
    path = os.path.join(BIN, exe) if os.sep not in exe else exe
    if not os.path.exists(path):
        pytest.fail("%s not built (run __graft_entry__.build())" % path)
    return subprocess.run([path, *args], capture_output=True, text=True, timeout=timeout)


def test_crystal_addressing_laws():
    r = run("crystal_test")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_setup_text_known_answer():
    """save_state's setup text, byte for byte against a file written by hand
    from the reference's encoder rules, and load_state's reader on a
    reference-format file with %-escaped strings (tests/golden/setup_*.txt)."""
    r = run("setup_format_test", os.path.join(ROOT, "tests", "golden"))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_facade_header_compiles_standalone(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "nmhd/mhd_solver.h"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_facade_restart_and_io(gpu_ready, tmp_path):
    print("\n[case] C++ facade restart / io", flush=True)
    r = run("facade_test", str(tmp_path))
    print(r.stdout, r.stderr)
    assert r.returncode == 0


@pytest.mark.gpu
def test_kepler_disk_app_continue(gpu_ready, tmp_path):
    """The shearing-box driver: a run stopped by its step limit saves the
    'c' checkpoint and a --continue run resumes from it."""
    print("\n[case] kepler_disk app", flush=True)
    app = os.path.join(ROOT, "nmhd_amd", "bin", "kepler_disk")
    r1 = run(app, "--new", "-o", str(tmp_path), "-n", "16", "--steps", "3")
    assert r1.returncode == 0, r1.stdout + r1.stderr
    assert "ToBeContinued" in r1.stdout and "generation = 3" in r1.stdout
    r2 = run(app, "--continue", "-o", str(tmp_path), "-n", "16", "--steps", "2")
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert "generation = 5" in r2.stdout
    assert os.path.getsize(os.path.join(str(tmp_path), "c", "0000.bin")) == 12 + 8 * 4 * 16 * 24 * 16
    # the monitor's log lines (reduced on the GPU), one per step of each run
    log = open(os.path.join(str(tmp_path), "log", "0000.txt")).read().splitlines()
    mon = [l for l in log if "Bx_squared_average:" in l]
    assert len(mon) >= 5
    by2 = [float(l.split("By_squared_average:")[1]) for l in mon]
    assert all(v >= 0 for v in by2) and max(by2) < 1e-3
    # the dump post-processor reads the checkpoint back: one radial rank, so
    # the output grid holds the interior (ghost rows fall outside the box)
    from nmhd_amd import postprocess as PP
    grid, arrays = PP.accumulate(os.path.join(str(tmp_path), "c"))
    sx, sy, sz, fields = PP.read_bin(os.path.join(str(tmp_path), "c", "0000.bin"))
    assert grid.size == [16, 16, 16] and (sx, sy, sz) == (16, 24, 16)
    assert np.allclose(arrays, fields[:, :, 4:20, :], rtol=0, atol=0)


@pytest.mark.gpu
def test_kepler_disk_app_with_fixes(gpu_ready, tmp_path):
    """SimulationConfig::fixes reaches the device: the driver with all three
    corrections (--fixes 7) runs, and its state after 3 steps differs from
    the reference-behaviour run of the same setup."""
    print("\n[case] kepler_disk app --fixes", flush=True)
    app = os.path.join(ROOT, "nmhd_amd", "bin", "kepler_disk")
    out = {}
    for fx in ("0", "7"):
        d = tmp_path / ("f" + fx)
        r = run(app, "--new", "-o", str(d), "-n", "16", "--steps", "3", "--fixes", fx)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "generation = 3" in r.stdout
        out[fx] = open(os.path.join(str(d), "c", "0000.bin"), "rb").read()
    assert len(out["0"]) == len(out["7"]) and out["0"] != out["7"]
