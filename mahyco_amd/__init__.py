"""MI355X-native VNR Lagrange + directional remap hot path of MaHyCo.

Python access to libmahyco_hip.so (the C ABI of include/mahyco.h) through
ctypes: configuration, one Context per problem on the GPU, field access.
The library is the product; it fails loudly when it is missing.
"""
import ctypes
import os

import numpy as np

from .abi import (MhyConfig, MhyTimeState, MHY_OK, CELL_FIELDS, MAT_FIELDS, NODE_FIELDS, mat_field)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmahyco_hip.so")

# every entry point declared in include/mahyco.h
EXPORTS = [
    "mhy_read_donnees", "mhy_read_json", "mhy_config_default", "mhy_create", "mhy_destroy", "mhy_last_error", "mhy_step",
    "mhy_run", "mhy_simulate", "mhy_output_open", "mhy_output_write", "mhy_compute_dt", "mhy_lagrange",
    "mhy_prepare_remap", "mhy_remap_sweep", "mhy_remap_stage",
    "mhy_remap_final", "mhy_end_iteration", "mhy_synchronize", "mhy_get_time_state", "mhy_get_mesh_size",
    "mhy_get_field", "mhy_set_field", "mhy_conservation", "mhy_conservation_noalloc", "mhy_kernel_count",
    "mhy_kernel_name", "mhy_tuned_kernels",
    "mhy_kernel_bytes", "mhy_enable_kernel_timing", "mhy_set_kernel_timing_mask", "mhy_kernel_time",
    "mhy_batch_create", "mhy_batch_destroy", "mhy_batch_last_error", "mhy_batch_run", "mhy_batch_context",
    "mhy_batch_cells",
]

_lib = None


class MahycoError(RuntimeError):
    pass


def library():
    """Loads libmahyco_hip.so (built by mahyco_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MahycoError(f"{LIB_PATH} is missing: run `python -m mahyco_amd.build` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.mhy_read_donnees.argtypes = [ctypes.c_char_p, ctypes.POINTER(MhyConfig)]
        L.mhy_read_json.argtypes = [ctypes.c_char_p, ctypes.POINTER(MhyConfig), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
        L.mhy_config_default.argtypes = [ctypes.POINTER(MhyConfig)]
        L.mhy_config_default.restype = None
        L.mhy_create.argtypes = [ctypes.POINTER(MhyConfig), ctypes.POINTER(P)]
        L.mhy_destroy.argtypes = [P]
        L.mhy_last_error.argtypes = [P]
        L.mhy_last_error.restype = ctypes.c_char_p
        L.mhy_step.argtypes = [P, ctypes.c_int]
        L.mhy_run.argtypes = [P, ctypes.c_int]
        L.mhy_simulate.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
        for name in ("mhy_compute_dt", "mhy_lagrange", "mhy_prepare_remap", "mhy_remap_final",
                     "mhy_end_iteration", "mhy_synchronize"):
            getattr(L, name).argtypes = [P]
        L.mhy_remap_sweep.argtypes = [P, ctypes.c_int]
        L.mhy_remap_stage.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.mhy_output_open.argtypes = [P, ctypes.c_char_p]
        L.mhy_output_write.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.mhy_get_time_state.argtypes = [P, ctypes.POINTER(MhyTimeState)]
        L.mhy_get_mesh_size.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.mhy_get_field.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_longlong]
        L.mhy_set_field.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_longlong]
        L.mhy_conservation.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.mhy_conservation_noalloc.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.mhy_kernel_count.argtypes = [P]
        L.mhy_tuned_kernels.argtypes = [P]
        L.mhy_tuned_kernels.restype = ctypes.c_char_p
        L.mhy_kernel_name.argtypes = [P, ctypes.c_int]
        L.mhy_kernel_name.restype = ctypes.c_char_p
        L.mhy_kernel_bytes.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.mhy_enable_kernel_timing.argtypes = [P, ctypes.c_int]
        L.mhy_set_kernel_timing_mask.argtypes = [P, ctypes.c_uint]
        L.mhy_kernel_time.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        L.mhy_batch_create.argtypes = [ctypes.POINTER(MhyConfig), ctypes.c_int, ctypes.POINTER(P)]
        L.mhy_batch_destroy.argtypes = [P]
        L.mhy_batch_last_error.argtypes = [P]
        L.mhy_batch_last_error.restype = ctypes.c_char_p
        L.mhy_batch_run.argtypes = [P, ctypes.c_int]
        L.mhy_batch_context.argtypes = [P, ctypes.c_int, ctypes.POINTER(P)]
        L.mhy_batch_cells.argtypes = [P]
        L.mhy_batch_cells.restype = ctypes.c_longlong
        _lib = L
    return _lib


def read_donnees(path):
    """LectureDonnees: parses a reference donnees.txt into an MhyConfig."""
    cfg = MhyConfig()
    st = library().mhy_read_donnees(path.encode(), ctypes.byref(cfg))
    if st != MHY_OK:
        raise MahycoError(f"cannot read {path}: status {st}")
    return cfg


def read_json(path):
    """The JSON input (mhy_read_json): a list of MhyConfig, one per problem."""
    n = ctypes.c_int(0)
    L = library()
    st = L.mhy_read_json(path.encode(), None, 0, ctypes.byref(n))
    if st != MHY_OK:
        raise MahycoError(f"cannot read {path}: status {st}")
    arr = (MhyConfig * n.value)()
    st = L.mhy_read_json(path.encode(), arr, n.value, ctypes.byref(n))
    if st != MHY_OK:
        raise MahycoError(f"cannot read {path}: status {st}")
    return list(arr)


class Context:
    """One problem on the GPU (mhy_create .. mhy_destroy)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.nx, self.ny = cfg.nx, cfg.ny
        h = ctypes.c_void_p()
        st = library().mhy_create(ctypes.byref(cfg), ctypes.byref(h))
        self.h = h
        if st != MHY_OK:
            msg = library().mhy_last_error(h).decode() if h else ""
            self.close()
            raise MahycoError(f"mhy_create failed ({st}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            library().mhy_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st, what):
        if st != MHY_OK:
            raise MahycoError(f"{what} failed ({st}): {library().mhy_last_error(self.h).decode()}")
        return st

    def step(self, n=1, sync=True):
        L = library()
        for _ in range(n):
            self._check(L.mhy_step(self.h, 1 if sync else 0), "mhy_step")

    def run(self, n):
        self._check(library().mhy_run(self.h, n), "mhy_run")

    def simulate(self, outdir=None, verbose=False):
        return self._check(library().mhy_simulate(self.h, (outdir or "none").encode(), int(verbose)), "mhy_simulate")

    def synchronize(self):
        self._check(library().mhy_synchronize(self.h), "mhy_synchronize")

    def time_state(self):
        ts = MhyTimeState()
        self._check(library().mhy_get_time_state(self.h, ctypes.byref(ts)), "mhy_get_time_state")
        return ts

    def field_id(self, name, k=0):
        if name in CELL_FIELDS:
            return CELL_FIELDS[name], self.nx * self.ny
        if name in MAT_FIELDS:
            return mat_field(MAT_FIELDS[name], k), self.nx * self.ny
        return NODE_FIELDS[name], (self.nx + 1) * (self.ny + 1)

    def get(self, name, k=0):
        fid, n = self.field_id(name, k)
        out = np.empty(n, dtype=np.float64)
        self._check(library().mhy_get_field(self.h, fid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n),
                    f"get {name}")
        return out

    def set(self, name, values, k=0):
        fid, n = self.field_id(name, k)
        arr = np.ascontiguousarray(values, dtype=np.float64)
        assert arr.size == n
        self._check(library().mhy_set_field(self.h, fid, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n),
                    f"set {name}")

    def conservation(self):
        out = (ctypes.c_double * (2 + 3))()
        self._check(library().mhy_conservation(self.h, out), "mhy_conservation")
        return list(out)[: 2 + self.cfg.nbmat]

    def tuned_kernels(self):
        """The y-direction kernel variants the launch-shape tuner kept."""
        return library().mhy_tuned_kernels(self.h).decode()

    def enable_timing(self, on=True):
        self._check(library().mhy_enable_kernel_timing(self.h, int(on)), "timing")

    def set_timing_mask(self, mask):
        """Record events only around the kernels whose bit (1 << index) is set."""
        self._check(library().mhy_set_kernel_timing_mask(self.h, int(mask)), "timing mask")

    def kernel_stats(self):
        L = library()
        out = {}
        for k in range(L.mhy_kernel_count(self.h)):
            ms = ctypes.c_double()
            n = ctypes.c_int()
            b = ctypes.c_double()
            L.mhy_kernel_time(self.h, k, ctypes.byref(ms), ctypes.byref(n))
            L.mhy_kernel_bytes(self.h, k, ctypes.byref(b))
            out[L.mhy_kernel_name(self.h, k).decode()] = {"id": k, "avg_ms": ms.value, "launches": n.value,
                                                         "bytes": b.value}
        return out


class _BatchMember(Context):
    """Problem i of a Batch: field and time-state access; owned by the batch."""

    def __init__(self, batch, h, cfg):
        self._batch = batch
        self.cfg = cfg
        self.nx, self.ny = cfg.nx, cfg.ny
        self.h = h

    def close(self):
        self.h = None  # the batch owns the context


class Batch:
    """An ensemble of independent problems stepped together
    (mhy_batch_create .. mhy_batch_destroy)."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (MhyConfig * len(self.cfgs))(*self.cfgs)
        h = ctypes.c_void_p()
        st = library().mhy_batch_create(arr, len(self.cfgs), ctypes.byref(h))
        self.h = h
        if st != MHY_OK:
            msg = library().mhy_batch_last_error(h).decode() if h else ""
            self.close()
            raise MahycoError(f"mhy_batch_create failed ({st}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            library().mhy_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def run(self, n):
        st = library().mhy_batch_run(self.h, n)
        if st != MHY_OK:
            raise MahycoError(f"mhy_batch_run failed ({st}): {library().mhy_batch_last_error(self.h).decode()}")

    @property
    def cells(self):
        return library().mhy_batch_cells(self.h)

    def problem(self, i):
        h = ctypes.c_void_p()
        st = library().mhy_batch_context(self.h, i, ctypes.byref(h))
        if st != MHY_OK:
            raise MahycoError(f"mhy_batch_context({i}) failed ({st})")
        return _BatchMember(self, h, self.cfgs[i])
