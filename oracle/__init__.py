"""ctypes binding of oracle/libpic_oracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is a CPU FP64 restatement of the reference's Plasma::Step
(gpu_plasma.h:605-620), see pic_oracle.cpp.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it, and only
as the checker; the product (0buf_amd/) never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpic_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "libpic_oracle.so"], check=True)
    lib = C.CDLL(LIB_PATH)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    sig = {
        "orc_create": (vp, [C.c_int] * 3 + [C.c_double] * 4 + [C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_threads": (None, [vp, C.c_int]),
        "orc_load_particles": (C.c_long, [vp, C.c_int, dp, dp, dp, dp, dp, dp, C.c_long,
                                          C.c_double, C.c_double]),
        "orc_set_field": (None, [vp, C.c_int, dp]),
        "orc_get_field": (None, [vp, C.c_int, dp]),
        "orc_h_half": (None, [vp]), "orc_push": (None, [vp]), "orc_deposit": (None, [vp]),
        "orc_migrate": (C.c_int, [vp]), "orc_fold_j": (None, [vp]), "orc_h2": (None, [vp]),
        "orc_e": (None, [vp]), "orc_e_periodic": (None, [vp]),
        "orc_energy": (C.c_double, [vp]),
        "orc_step": (C.c_int, [vp, C.c_int, dp]),
        "orc_total_particles": (C.c_long, [vp]),
        "orc_cell_counts": (None, [vp, ip]),
        "orc_get_particles": (C.c_long, [vp, dp, dp, dp, dp, dp, dp, dp, dp, ip, ip, C.c_long]),
        "orc_moments": (None, [vp, dp]),
        "orc_compare": (C.c_double, [dp, dp, C.c_int, C.c_double]),
        "orc_rms": (C.c_double, [dp, dp, C.c_int]),
        "orc_boris": (None, [dp, dp, C.c_double, C.c_double, dp, dp, dp]),
        "orc_relay": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """CPU restatement of Plasma with the same data shapes as 0buf_amd.Pic."""

    def __init__(self, n, l, tau, departure_cap=50, threads=1):
        self.lib = load()
        self.n = tuple(n)
        self.G = (n[0] + 2) * (n[1] + 2) * (n[2] + 2)
        self.ncell = n[0] * n[1] * n[2]
        self.h = self.lib.orc_create(n[0], n[1], n[2], l[0], l[1], l[2], tau, departure_cap)
        self.lib.orc_set_threads(self.h, threads)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_particles(self, sort, x, y, z, px, py, pz, m, q_m):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, px, py, pz)]
        return self.lib.orc_load_particles(self.h, sort, *[_d(a) for a in arrs], len(arrs[0]),
                                           float(m), float(q_m))

    def load_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        assert a.size == self.G
        self.lib.orc_set_field(self.h, which, _d(a))

    def get_field(self, which):
        a = np.empty(self.G, dtype=np.float64)
        self.lib.orc_get_field(self.h, which, _d(a))
        return a

    def step(self, nt=1):
        e = C.c_double(0.0)
        st = self.lib.orc_step(self.h, nt, C.byref(e))
        return e.value, st

    def stage(self, name):
        return getattr(self.lib, "orc_" + name)(self.h)

    def energy(self):
        return self.lib.orc_energy(self.h)

    def particle_count(self):
        return int(self.lib.orc_total_particles(self.h))

    def relay(self, cell, x, x1):
        """cell.h:1306-1403: crossing code m and relay point of a particle of
        cell (i, l, k) moving x -> x1 (x1 itself when m = 0)."""
        a = np.array(x, dtype=np.float64)
        b = np.array(x1, dtype=np.float64)
        r = np.empty(3)
        m = self.lib.orc_relay(self.h, int(cell[0]), int(cell[1]), int(cell[2]), _d(a), _d(b),
                               _d(r))
        return int(m), r

    def cell_counts(self):
        a = np.empty(self.ncell, dtype=np.int32)
        self.lib.orc_cell_counts(self.h, _i(a))
        return a

    def particles(self):
        n = self.particle_count()
        cap = max(n, 1)
        out = {k: np.empty(cap) for k in ("x", "y", "z", "px", "py", "pz", "m", "q_m")}
        sort = np.empty(cap, dtype=np.int32)
        cell = np.empty(cap, dtype=np.int32)
        got = self.lib.orc_get_particles(
            self.h, *[_d(out[k]) for k in ("x", "y", "z", "px", "py", "pz", "m", "q_m")],
            _i(sort), _i(cell), cap)
        assert got == n
        res = {k: v[:n] for k, v in out.items()}
        res["sort"] = sort[:n]
        res["cell"] = cell[:n]
        return res

    def moments(self):
        a = np.empty(27)
        self.lib.orc_moments(self.h, _d(a))
        return a.reshape(3, 9)


def compare(a, b, tol):
    """compare.cxx:9-36: fraction of entries with |a-b| < tol."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return load().orc_compare(_d(a), _d(b), a.size, tol)


def rms(a, b):
    """service_functions.cu:108-126 CheckArraySilent."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return load().orc_rms(_d(a), _d(b), a.size)


def boris(E, H, tau, q_m, x, p):
    lib = load()
    E = np.ascontiguousarray(E, dtype=np.float64)
    H = np.ascontiguousarray(H, dtype=np.float64)
    x = np.array(x, dtype=np.float64)
    p = np.array(p, dtype=np.float64)
    x1 = np.empty(3)
    lib.orc_boris(_d(E), _d(H), tau, q_m, _d(x), _d(p), _d(x1))
    return p, x1
