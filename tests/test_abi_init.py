"""C-ABI surface and the host-side seeded initializer (no GPU needed).

The library is loaded on the CPU: every function include/pic_amd.h declares
must be exported; only host-only entry points are called here.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pic_amd.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef\s+struct\s+\w*\s*\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
    text = re.sub(r"enum\s+\w*\s*\{.*?\}\s*;", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", text)
    return sorted(set(names))


def test_header_declares_expected_entry_points():
    names = header_functions()
    for must in ("MemoryCopy", "MemoryAllocate", "getErrorString", "pic_create", "pic_step",
                 "pic_push_deposit", "pic_migrate", "pic_get_particles", "pic_generate_species"):
        assert must in names
    assert len(names) >= 39


def test_library_exports_every_declared_symbol(picmod):
    lib = picmod.load_library()
    names = header_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", picmod.LIB_PATH],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported
    assert set(picmod.EXPORTS) == set(names)


def test_archapi_constants_match_reference(picmod):
    # archAPI.h:8-11
    assert (picmod.HOST_TO_DEVICE, picmod.HOST_TO_HOST, picmod.DEVICE_TO_HOST,
            picmod.DEVICE_TO_DEVICE) == (-131313, -131314, -131315, -131316)
    text = open(HEADER).read()
    for name, val in (("HOST_TO_DEVICE", -131313), ("DEVICE_TO_DEVICE", -131316)):
        assert re.search(rf"#define {name}\s+{val}\b", text)


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of a header struct, from the C compiler."""
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src.write_text(f"#include <stdio.h>\n#include <stddef.h>\n#include \"{HEADER}\"\n"
                   f"int main(void){{printf(\"%zu\\n\", sizeof({struct}));{body}return 0;}}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                                           check=True).stdout.split()]
    return vals[0], vals[1:]


@pytest.mark.parametrize("struct,cls", [("pic_params", "PicParams"),
                                        ("pic_step_status", "PicStepStatus"),
                                        ("pic_adjoint_io", "PicAdjointIO"),
                                        ("pic_species_spec", "PicSpeciesSpec"),
                                        ("pic_param_file", "PicParamFile"),
                                        ("pic_dump_sort", "PicDumpSort")])
def test_struct_layout_matches_header(picmod, tmp_path, struct, cls):
    ct = getattr(picmod, cls)
    fields = [f for f, _ in ct._fields_]
    size, offs = _c_layout(struct, fields, tmp_path)
    assert C.sizeof(ct) == size
    assert [getattr(ct, f).offset for f in fields] == offs


def test_params_defaults(picmod):
    p = picmod.default_params()
    assert (p.nx, p.ny, p.nz, p.departure_cap) == (32, 8, 8, 50)
    assert p.tau == 0.05 and p.precision == picmod.FP64 and p.differentiable == 0


def test_host_error_strings(picmod):
    lib = picmod.load_library()
    assert lib.pic_error_string(4).decode().startswith("too many")
    assert lib.pic_error_string(0).decode() == "success"


def test_generate_single_stream_draw_order(picmod):
    """stream_mode 0 starts from the reference seeds (1,255,25555) and draws
    x,y,z uniforms first: the first particle's x is (0 + u1) * hx."""
    g = np.load(os.path.join(GOLDEN, "rng_ref.npz"))
    cfg = picmod.beam_plasma_config(0)
    params = picmod.make_params(cfg["n"], cfg["l"], cfg["tau"])
    x, y, z, px, py, pz, m = picmod.generate_species(params, cfg["species"][0], None, 0)
    hx = params.lx / params.nx
    hy = params.ly / params.ny
    assert x[0] == (0.0 + g["uniform"][0]) * hx
    assert y[0] == (0.0 + g["uniform"][1]) * hy
    assert m == pytest.approx(1.01 / 16, rel=1e-15)
    assert np.all(px == 0.0)     # cold ions


def test_generate_per_cell_streams_thread_independent(picmod):
    cfg = picmod.beam_plasma_config(0)
    params = picmod.make_params(cfg["n"], cfg["l"], cfg["tau"])
    a = picmod.generate_species(params, cfg["species"][1], None, 1, seed=99, nthreads=1)
    b = picmod.generate_species(params, cfg["species"][1], None, 1, seed=99, nthreads=5)
    for u, v in zip(a[:6], b[:6]):
        assert np.array_equal(u, v)
    # every particle inside its generating cell (cell-major order)
    n = cfg["n"]
    cells = np.repeat(np.arange(n[0] * n[1] * n[2]), 16)
    ci = np.floor(a[0] / (params.lx / n[0])).astype(int)
    assert np.array_equal(ci, cells // (n[1] * n[2]))


def test_generate_beam_momentum_and_weights(picmod):
    cfg = picmod.beam_plasma_config(0)
    params = picmod.make_params(cfg["n"], cfg["l"], cfg["tau"])
    x, y, z, px, py, pz, m = picmod.generate_species(params, cfg["species"][2], None, 1, seed=1)
    # u = 0.3 +- 0.005 -> p = u / sqrt(1 - u^2) (maxwell.cu:521-531)
    assert px.mean() == pytest.approx(0.3 / np.sqrt(1 - 0.09), rel=2e-3)
    assert py.std() == pytest.approx(0.005 / np.sqrt(1 - 0.09), rel=5e-2)
    assert m == pytest.approx(-0.01 / 16, rel=1e-15)


def test_generate_gaussian_profile_counts(picmod):
    """Profile 1: ppc is the mean count per cell; the transverse Gaussian
    (sigma 8 cells) puts A = ppc ny nz / sum(f) on the axis and
    floor(A f + 1/2) in each cell; the weight carries the axis density with
    the axis count, and the config-4 ions neutralise the mean beam charge."""
    cfg = picmod.beam_plasma_config(4)
    params = picmod.make_params((4, 64, 64), (0.4, 6.4, 6.4), 0.05)
    spec = cfg["species"][2]
    lib = picmod.load_library()
    n = lib.pic_species_count(C.byref(params), C.byref(spec))
    yc = np.arange(64) + 0.5 - 32
    f = np.exp(-(yc[:, None] ** 2 + yc[None, :] ** 2) / (2 * 8.0 ** 2))
    A = spec.ppc * f.size / f.sum()
    counts = np.floor(A * f + 0.5)
    assert n == 4 * int(counts.sum())
    assert abs(counts.mean() / spec.ppc - 1) < 1e-3
    assert counts.max() > 9 * spec.ppc and counts.min() == 0
    x, y, z, px, py, pz, m = picmod.generate_species(params, spec, None, 1, seed=1)
    assert len(x) == n
    assert m == pytest.approx(-0.05 / A, rel=1e-14)
    ions = cfg["species"][0]
    assert ions.density - 1.0 == pytest.approx(0.05 * f.mean(), rel=1e-12)


def test_create_without_gpu_reports_no_device(picmod):
    """pic_create only counts devices before allocating; on a host without a
    GPU it must fail with PIC_ERR_NO_DEVICE, not crash."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = picmod.load_library()
    p = picmod.default_params()
    ctx = C.c_void_p()
    assert lib.pic_create(C.byref(p), C.byref(ctx)) == 3


def test_create_refuses_fp32_step_over_half_a_cell(picmod):
    """The FP32 push moves a particle at most one cell and puts its end in the
    start's dual cell or the next: |v| tau / h < 1/2 on every axis, which
    |v| < 1 guarantees only for tau <= h/2.  A longer step is refused before
    any device call (so the answer is the same with and without a GPU); at
    tau = h/2 creation gets past the check."""
    import torch
    lib = picmod.load_library()
    ok_code = 0 if torch.cuda.is_available() else 3
    # default grid: 32 x 8 x 8 cells of h = 0.1, tau = 0.05 = h/2; shrink one
    # axis's h (or lengthen tau) by 0.2% to cross the limit on that axis
    for axis, want in (("tau", 1), ("x", 1), ("y", 1), ("z", 1), (None, ok_code)):
        p = picmod.default_params()
        p.precision = picmod.FP32
        assert (p.lx / p.nx, p.ly / p.ny, p.lz / p.nz, p.tau) == (0.1, 0.1, 0.1, 0.05)
        if axis == "tau":
            p.tau = 0.0501
        elif axis == "x":
            p.lx *= 0.998
        elif axis == "y":
            p.ly *= 0.998
        elif axis == "z":
            p.lz *= 0.998
        ctx = C.c_void_p()
        rc = lib.pic_create(C.byref(p), C.byref(ctx))
        if rc == 0:
            lib.pic_destroy(ctx)
        assert rc == want, (axis, rc)
    p = picmod.default_params()
    p.precision = picmod.FP64
    p.tau = 0.0501          # the FP64 form checks ranges on the device instead
    ctx = C.c_void_p()
    rc = lib.pic_create(C.byref(p), C.byref(ctx))
    if rc == 0:
        lib.pic_destroy(ctx)
    assert rc == ok_code


def test_param_file_roundtrip_matches_reference_writer(picmod, tmp_path):
    """params_ref.dat was written by the reference's writeParamsFile
    (maxwell.cu:330-377); the reader must recover its values and the writer
    must reproduce it byte for byte."""
    ref = os.path.join(GOLDEN, "params_ref.dat")
    pf = picmod.read_param_file(ref)
    assert (pf.tex0, pf.tey0, pf.tez0, pf.rimp, pf.Tb, pf.rbd, pf.ni, pf.B0) == \
        (1e-3, 2e-3, 3e-3, 0.2, 0.14, 2e-3, 1.0, 0.5)
    assert (pf.lx, pf.ly, pf.lz, pf.py, pf.pz) == (1.1424, 0.05, 0.06, 0.04, 0.05)
    assert (pf.lp, pf.nx, pf.ny, pf.nz, pf.tau) == (1000, 100, 4, 5, 1e-3)
    assert (pf.beam_plasma, pf.start_from_file, pf.phase, pf.ts, pf.ms) == (1, 0, 0, 5, 2)
    out = tmp_path / "000_params.dat"
    picmod.write_param_file(str(out), pf)
    assert out.read_bytes() == open(ref, "rb").read()


def test_species_from_param_file(picmod):
    pf = picmod.read_param_file(os.path.join(GOLDEN, "params_ref.dat"))
    grid, (ion, el, beam) = picmod.species_from_param_file(pf)
    assert (grid.nx, grid.ny, grid.nz, grid.tau) == (100, 4, 5, 1e-3)
    # getMassCharge (maxwell.cu:606-619): weights ni/lp/2, (ni+rbd)/lp, rbd/lp
    assert ion.ppc == 1000 and el.ppc == 2000 and beam.ppc == 1000
    assert ion.charge * ion.density / ion.ppc == pytest.approx((1.0 + 2e-3) / 1000)
    assert el.charge * el.density / el.ppc == pytest.approx(-1.0 / 1000 / 2)
    assert beam.charge * beam.density / beam.ppc == pytest.approx(-2e-3 / 1000)
    assert beam.drift[0] == 0.2 and beam.vth[0] == pytest.approx(0.14 * 0.2)
    assert el.drift[0] == pytest.approx(-2e-3 * 0.2) and el.vth[2] == 3e-3
