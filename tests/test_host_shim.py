"""The C++ host classes (graphics_amd/host/SphGpu.h) that keep the reference's
SphBase/DFSPH call pattern: compile a Demo-Main-style program against them
and run it (error path without a device, three steps with one)."""
import os
import subprocess

import pytest

from graphics_amd import sphgpu as sg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = str(tmp_path / "host_shim_demo")
    lib = os.path.join(ROOT, "graphics_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", ROOT,
           os.path.join(ROOT, "tests", "host_shim_demo.cpp"), "-o", exe,
           "-L", lib, "-lsphgpu", f"-Wl,-rpath,{lib}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_shim_compiles_and_reports_missing_device(tmp_path):
    sg.load()   # the library must be built
    exe = _build(tmp_path)
    if sg.device_available():
        pytest.skip("device present: covered by the gpu test")
    r = subprocess.run([exe, "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "error -3" in r.stdout
    assert "fluids 1 solids 1" in r.stdout


@pytest.mark.gpu
def test_shim_steps_on_device(tmp_path, gpu):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "3", str(tmp_path / "frame.pos")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tok = r.stdout.split()
    vals = dict(zip(tok[0::2], tok[1::2]))
    assert vals["frame"] == "3"
    assert vals["pos_floats"] == str(3 * 216) and vals["pos_mask_ok"] == "1" and vals["first_x_equal"] == "1"
    assert vals["fluid"] == "216"
    # a 6^3 block: a third of its particles sit on the free surface
    assert 800.0 < float(vals["mean_rho"]) < 1010.0
    assert int(vals["iters"].split("/")[0]) >= 2


@pytest.mark.gpu
def test_shim_rigid_host_step_on_device(tmp_path, gpu):
    """updateSolids() runs once per step inside sphStep, reads the fluid force
    (rigidForceJoin) and moves the box (setSolidTransform); the boundary
    mirror follows the last transform."""
    exe = _build(tmp_path)
    r = subprocess.run([exe, "40", "rigid"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    tok = r.stdout.split()
    vals = dict(zip(tok[0::2], tok[1::2]))
    assert vals["frame"] == "40" and vals["calls"] == "40"
    fy, y, ymin = float(vals["force_y"]), float(vals["box_y"]), float(vals["sample_ymin"])
    assert fy > 0.0                              # the water pushes the submerged box up
    assert abs(ymin - (y - 0.04)) < 1e-5         # samples sit where the last pose put them
