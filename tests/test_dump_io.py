"""State dumps in the reference's Fortran-record format (mumu files).

The record layout is restated here independently of the library, from the
reference's reader: readFortranBinaryArray (load_data.cu:35-64) takes an
int32 byte count, the payload and the int32 count again;
readBinaryParticleArraysOneSort (load_data.cu:160-201) reads per sort an
int32, three doubles (count, q_m, m), an int32, then six such arrays; the
file holds the 12 ghosted field arrays first (readPreliminary3Darrays,
load_data.cu:66-130) and the sorts ION, PLASMA_ELECTRON, BEAM_ELECTRON
(load_data.cu:307-317).  A file written by this restatement must read back
exactly, and a file written by the library must parse with it.
"""
import struct

import numpy as np
import pytest

FIELDS = 12


def write_ref_layout(path, fields, sorts):
    with open(path, "wb") as f:
        def rec(b):
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(b)))
        for a in fields:
            rec(np.asarray(a, dtype="<f8").tobytes())
        for d in sorts:
            rec(struct.pack("<3d", float(len(d["x"])), d["q_m"], d["m"]))
            for k in ("x", "y", "z", "px", "py", "pz"):
                rec(np.asarray(d[k], dtype="<f8").tobytes())


def parse_ref_layout(path, G):
    data = open(path, "rb").read()
    pos = 0

    def rec():
        nonlocal pos
        n, = struct.unpack_from("<i", data, pos)
        body = data[pos + 4:pos + 4 + n]
        m, = struct.unpack_from("<i", data, pos + 4 + n)
        assert m == n
        pos += 8 + n
        return body
    fields = [np.frombuffer(rec(), dtype="<f8") for _ in range(FIELDS)]
    sorts = []
    for _ in range(3):
        cnt, qm, m = struct.unpack("<3d", rec())
        d = {"q_m": qm, "m": m}
        for k in ("x", "y", "z", "px", "py", "pz"):
            d[k] = np.frombuffer(rec(), dtype="<f8")
        assert len(d["x"]) == int(cnt)
        sorts.append(d)
    assert pos == len(data)
    assert all(len(a) == G for a in fields)
    return fields, sorts


def random_state(rng, G, counts):
    fields = [rng.normal(size=G) for _ in range(FIELDS)]
    sorts = []
    for n, qm, m in zip(counts, (1 / 1836.0, -1.0, -1.0), (0.1, -0.05, -0.001)):
        d = {k: rng.uniform(0, 1, n) for k in ("x", "y", "z", "px", "py", "pz")}
        d["q_m"], d["m"] = qm, m
        sorts.append(d)
    return fields, sorts


def test_file_name(picmod):
    # gpu_plasma.h:1101 sprintf(fn, "mumu%03d%08d.dat", num, nt)
    assert picmod.dump_file_name(5, 600) == "mumu00500000600.dat"
    assert picmod.dump_file_name(0, 1) == "mumu00000000001.dat"


@pytest.mark.parametrize("counts", [(7, 13, 5), (0, 9, 0)])
def test_reader_reads_reference_layout(picmod, tmp_path, counts):
    rng = np.random.default_rng(1)
    G = 5 * 6 * 7
    fields, sorts = random_state(rng, G, counts)
    path = str(tmp_path / "mumu00000000001.dat")
    write_ref_layout(path, fields, sorts)
    f2, s2 = picmod.dump_read(path, G)
    for a, b in zip(fields, f2):
        assert np.array_equal(a, b)
    for d, e in zip(sorts, s2):
        assert d["q_m"] == e["q_m"] and d["m"] == e["m"]
        for k in ("x", "y", "z", "px", "py", "pz"):
            assert np.array_equal(d[k], e[k])


def test_writer_matches_reference_layout(picmod, tmp_path):
    rng = np.random.default_rng(2)
    G = 4 * 4 * 5
    fields, sorts = random_state(rng, G, (3, 8, 2))
    path = str(tmp_path / "mumu00100000002.dat")
    picmod.dump_write(path, fields, sorts)
    f2, s2 = parse_ref_layout(path, G)
    for a, b in zip(fields, f2):
        assert np.array_equal(a, b)
    for d, e in zip(sorts, s2):
        assert (d["q_m"], d["m"]) == (e["q_m"], e["m"])
        for k in ("x", "y", "z", "px", "py", "pz"):
            assert np.array_equal(d[k], e[k])


def test_reader_rejects_damaged_files(picmod, tmp_path):
    rng = np.random.default_rng(3)
    G = 60
    fields, sorts = random_state(rng, G, (2, 2, 2))
    path = tmp_path / "d.dat"
    write_ref_layout(str(path), fields, sorts)
    raw = path.read_bytes()
    (tmp_path / "t.dat").write_bytes(raw[:-9])                     # truncated
    with pytest.raises(picmod.PicError):
        picmod.dump_read(str(tmp_path / "t.dat"), G)
    bad = bytearray(raw)
    bad[4 + G * 8] ^= 1                                           # trailing count of record 0
    (tmp_path / "b.dat").write_bytes(bytes(bad))
    with pytest.raises(picmod.PicError):
        picmod.dump_read(str(tmp_path / "b.dat"), G)
    with pytest.raises(picmod.PicError):                          # wrong grid size
        picmod.dump_read(str(path), G + 1)


@pytest.mark.gpu
def test_gpu_dump_round_trip_and_oracle_step(picmod, orc, tmp_path):
    """A dump of the GPU state after 10 config-0 steps, loaded into a fresh
    context and into the oracle: the next steps agree with the original
    context (to the last bits that FP64 atomics leave run-dependent, counts
    exactly), and the oracle's migration from the dump matches."""
    from helpers import load_workload
    params, parts = picmod.generate_workload(0)
    a = picmod.Pic(params)
    load_workload(a, parts)
    for nt in range(1, 11):
        a.step(nt)
    path = str(tmp_path / picmod.dump_file_name(0, 10))
    a.save_dump(path)
    b = picmod.Pic(params)
    assert b.load_dump(path) == 0
    assert not b.control_point(path)[:12].any()
    n = (params.nx, params.ny, params.nz)
    o = orc.Oracle(n, (params.lx, params.ly, params.lz), params.tau)
    G = b.G
    fields, sorts = picmod.dump_read(path, G)
    for i in range(12):
        o.load_field(i, fields[i])
    for s, d in enumerate(sorts):
        if len(d["x"]):
            o.load_particles(s, d["x"], d["y"], d["z"], d["px"], d["py"], d["pz"], d["m"],
                             d["q_m"])
    for nt in range(11, 16):
        ea, eb = a.step(nt), b.step(nt)
        eo, _ = o.step(nt)
        assert ea == pytest.approx(eb, rel=1e-13)   # FP64 atomic summation order
        assert np.array_equal(a.cell_counts(), b.cell_counts())
        assert np.array_equal(b.cell_counts(), o.cell_counts())
        assert eb == pytest.approx(eo, rel=1e-12)
    report = str(tmp_path / "control_points.dat")
    cp = b.control_point(path, nt=15, num=0, report=report)
    assert cp[:6].max() > 0
    for k in range(12):                    # CheckArraySilent as the oracle restates it
        assert cp[k] == pytest.approx(orc.rms(b.get_field(k), fields[k]), rel=1e-14, abs=0)
    line = open(report).read()
    assert line.startswith("nt    15 num   0 Ex ") and line.rstrip().split()[-2] == "particles"
    a.close()
    b.close()
