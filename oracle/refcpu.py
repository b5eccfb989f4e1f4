"""ORACLE -- test infrastructure, not product code.

ctypes binding of the CPU restatement of MaHyCo's VNR Lagrange + remap loop
(oracle/refcpu.cc, built into oracle/build/librefcpu.so by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

from mahyco_amd.abi import MhyConfig, MhyTimeState, MHY_OK, CELL_FIELDS, MAT_FIELDS, NODE_FIELDS, mat_field

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "librefcpu.so")
RUNNER = os.path.join(HERE, "build", "refcpu_run")
_lib = None


def build(quiet=True):
    """Compiles the oracle with its Makefile (g++, -ffp-contract=off)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.ref_read_donnees.argtypes = [ctypes.c_char_p, ctypes.POINTER(MhyConfig)]
        L.ref_config_default.argtypes = [ctypes.POINTER(MhyConfig)]
        L.ref_create.argtypes = [ctypes.POINTER(MhyConfig), ctypes.POINTER(ctypes.c_int)]
        L.ref_create.restype = ctypes.c_void_p
        L.ref_destroy.argtypes = [ctypes.c_void_p]
        L.ref_step.argtypes = [ctypes.c_void_p]
        L.ref_simulate.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.ref_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_longlong]
        L.ref_get_time_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(MhyTimeState)]
        L.ref_set_threads.argtypes = [ctypes.c_int]
        d = ctypes.c_double
        L.ref_probe_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int] + [d] * 8
        L.ref_probe_gradient.restype = d
        L.ref_probe_int2y.argtypes = [d] * 5
        L.ref_probe_int2y.restype = d
        L.ref_probe_flux_pp.argtypes = [ctypes.c_void_p, ctypes.c_int] + [d] * 9 + [ctypes.c_int]
        L.ref_probe_flux_pp.restype = d
        dp, ip = ctypes.POINTER(d), ctypes.POINTER(ctypes.c_int)
        L.ref_probe_cell_pp.argtypes = [ctypes.c_void_p] + [dp] * 5 + [ip] * 2 + [d] * 7 + [dp] * 4
        L.ref_probe_cell_pp.restype = ctypes.c_int
        L.ref_probe_flux_o3.argtypes = [ctypes.c_void_p, ctypes.POINTER(d), ctypes.POINTER(d), d]
        L.ref_probe_flux_o3.restype = d
        _lib = L
    return _lib


def read_donnees(path):
    cfg = MhyConfig()
    st = lib().ref_read_donnees(path.encode(), ctypes.byref(cfg))
    if st != MHY_OK:
        raise RuntimeError(f"cannot read {path}: {st}")
    return cfg


def probe_int2y(X, x0, y0, x1, y1):
    """INT2Y (remap/UtilesRemap.cc:177-188)."""
    return lib().ref_probe_int2y(X, x0, y0, x1, y1)


def set_threads(n):
    return lib().ref_set_threads(n)


class Oracle:
    """One problem on the CPU oracle."""

    def __init__(self, cfg):
        st = ctypes.c_int(0)
        self.cfg = cfg
        self.h = lib().ref_create(ctypes.byref(cfg), ctypes.byref(st))
        if not self.h:
            raise RuntimeError(f"oracle init failed: {st.value}")
        self.nx, self.ny = cfg.nx, cfg.ny

    def close(self):
        if self.h:
            lib().ref_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        for _ in range(n):
            st = lib().ref_step(self.h)
            if st != MHY_OK:
                return st
        return MHY_OK

    def simulate(self, outdir=None, max_steps=0):
        return lib().ref_simulate(self.h, (outdir or "none").encode(), max_steps)

    def time_state(self):
        ts = MhyTimeState()
        lib().ref_get_time_state(self.h, ctypes.byref(ts))
        return ts

    def field_id(self, name, k=0):
        if name in CELL_FIELDS:
            return CELL_FIELDS[name], self.nx * self.ny
        if name in MAT_FIELDS:
            return mat_field(MAT_FIELDS[name], k), self.nx * self.ny
        return NODE_FIELDS[name], (self.nx + 1) * (self.ny + 1)

    def probe_gradient(self, lim, gp, gm, phi, phip, phim, h0, hp, hm):
        """Limited gradient of one component (remap/UtilesRemap-Impl.h:99-126)."""
        return lib().ref_probe_gradient(self.h, lim, gp, gm, phi, phip, phim, h0, hp, hm)

    def probe_flux_pp(self, lim, grad, phi, phip, phim, h0, hp, hm, fnv, dt, face):
        """PenteBorne swept flux of one component; face 0 = back, 1 = front."""
        return lib().ref_probe_flux_pp(self.h, lim, grad, phi, phip, phim, h0, hp, hm, fnv, dt, face)

    def probe_cell_pp(self, gp, gm, phi, phip, phim, mixte, pure, h0, hp, hm, fm, fp, fdual, dt):
        """The cell stage of computeGradPhi on one cell (13-component vectors;
        mixte / pure: flags of the cell, its plus and its minus neighbour).
        Returns (limited gradient, Ar, Av, PB dual flux)."""
        def arr(v):
            a = np.zeros(13, dtype=np.float64)
            a[:len(v)] = v
            return a
        ins = [arr(v) for v in (gp, gm, phi, phip, phim)]
        outs = [np.zeros(13, dtype=np.float64) for _ in range(4)]
        P = ctypes.POINTER(ctypes.c_double)
        mix3 = (ctypes.c_int * 3)(*[int(v) for v in mixte])
        pure3 = (ctypes.c_int * 3)(*[int(v) for v in pure])
        lib().ref_probe_cell_pp(self.h, *[a.ctypes.data_as(P) for a in ins], mix3, pure3, float(h0), float(hp),
                                float(hm), float(fm), float(fp), float(fdual), float(dt),
                                *[a.ctypes.data_as(P) for a in outs])
        return tuple(outs)

    def probe_flux_o3(self, y, h, vdt):
        """Third-order face value of one component (remap/UtilesRemap.cc:197-320);
        y, h: the six cell means and widths around the face, minus side first."""
        ya = (ctypes.c_double * 6)(*[float(v) for v in y])
        ha = (ctypes.c_double * 6)(*[float(v) for v in h])
        return lib().ref_probe_flux_o3(self.h, ya, ha, float(vdt))

    def get(self, name, k=0):
        fid, n = self.field_id(name, k)
        out = np.empty(n, dtype=np.float64)
        st = lib().ref_get_field(self.h, fid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        if st != MHY_OK:
            raise RuntimeError(f"get {name}: {st}")
        return out
