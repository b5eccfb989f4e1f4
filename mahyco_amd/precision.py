"""Per-field error of the mixed FP32 / FP64 form against an FP64 run.

The FP64 form of this library reproduces the reference's FP64 arithmetic
operation by operation, so an FP64 context stepped alongside a mixed one
gives the reference fields on the same input.  Errors are relative to the
FP64 field: L1 = sum|a - b| / sum|b| and Linf = max|a - b| / max|b|.
"""
import numpy as np


def report_fields(nbmat):
    """(label, field name, material) of every field of the error report:
    density, specific internal energy, pressure, node velocity components
    and the volume fraction of each material."""
    out = [("rho", "density", 0), ("e", "energy", 0), ("p", "pressure", 0),
           ("ux", "velocity_x", 0), ("uy", "velocity_y", 0)]
    out += [(f"f_{k + 1}", "fracvol", k) for k in range(nbmat)]
    return out


def relative_errors(a, b, scale=None):
    """Relative L1 and Linf error of a against the reference b; scale (same
    shape as b) replaces |b| as the normalisation (velocity components are
    normalised by the speed, which a nearly-zero component alone would not
    measure)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b) if scale is None else np.asarray(scale, dtype=np.float64)
    d = np.abs(a - b)
    s1 = float(np.sum(s))
    sinf = float(np.max(s)) if s.size else 0.0
    l1 = float(np.sum(d)) / s1 if s1 > 0 else float(np.sum(d))
    linf = float(np.max(d)) / sinf if sinf > 0 else (float(np.max(d)) if d.size else 0.0)
    return l1, linf


def field_errors(mixed, ref, nbmat):
    """{label: (L1, Linf)} of the mixed context's fields against ref's."""
    out = {}
    speed = np.hypot(ref.get("velocity_x"), ref.get("velocity_y"))
    for label, name, k in report_fields(nbmat):
        scale = speed if name.startswith("velocity") else None
        out[label] = relative_errors(mixed.get(name, k), ref.get(name, k), scale)
    return out
