"""The C++ facade with the reference's Plasma interface (include/plasma.hpp)
and the main.cu-style driver built on it (0buf_amd/examples/plasma_main.cpp).

CPU: the driver links against libpic_amd.so and, without a device, ends with
the library's PIC_ERR_NO_DEVICE status and a message instead of a crash.
GPU: Initialize() from a parameter file + Compute() writes energy.dat lines
in the reference's format ("%10d %25.15e \\n", gpu_plasma.h:681) with the
energies of the same seeded run driven through the Python binding (to 1e-12:
the plain FP64 form's atomics leave the last bits run-dependent).
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "0buf_amd", "_build", "plasma_main")
GRID = ((16, 4, 4), (1.6, 0.4, 0.4), 0.05)


def small_param_file(picmod, path):
    pf = picmod.read_param_file(os.path.join(ROOT, "tests", "golden", "params_ref.dat"))
    pf.lp = 4
    picmod.write_param_file(path, pf)
    return pf


def driver_args(pfile, first, last):
    (nx, ny, nz), (lx, ly, lz), tau = GRID
    return [EXE, pfile, str(first), str(last), str(nx), str(ny), str(nz), repr(lx), repr(ly),
            repr(lz), "1.0", "4", "1.0", repr(tau)]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a HIP device is present")
def test_driver_without_device_reports_status(picmod, tmp_path):
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    pfile = str(tmp_path / "000_params.dat")
    small_param_file(picmod, pfile)
    r = subprocess.run(driver_args(pfile, 1, 3), cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 3
    assert "no HIP device" in r.stderr


@pytest.mark.gpu
def test_driver_energy_file_matches_binding(picmod, tmp_path):
    pfile = str(tmp_path / "000_params.dat")
    pf = small_param_file(picmod, pfile)
    r = subprocess.run(driver_args(pfile, 1, 8), cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = open(tmp_path / "energy.dat").read().splitlines()
    # the same run through the Python binding
    n, L, tau = GRID
    params = picmod.make_params(n, L, tau)
    _, specs = picmod.species_from_param_file(pf)
    rng = picmod.PicRng()
    picmod.load_library().pic_rng_seed(picmod.C.byref(rng), 1, 255, 25555)
    g = picmod.Pic(params)
    for sp in specs:
        x, y, z, px, py, pz, m = picmod.generate_species(params, sp, rng, 0, 0, 1)
        g.load_particles(sp.sort, x, y, z, px, py, pz, m, sp.q_m)
    e = [g.step(nt) for nt in range(1, 9)]
    g.close()
    # the reference's line format, then the values: FP64 atomics add J in a
    # run-dependent order, so two runs may differ in the last digits
    assert len(lines) == 8
    for nt, (line, v) in enumerate(zip(lines, e), start=1):
        assert len(line) == len("%10d %25.15e " % (nt, v)) and line.endswith(" ")
        step, val = line.split()
        assert int(step) == nt
        assert float(val) == pytest.approx(v, rel=1e-12)
    assert np.all(np.array(e) >= 0)
