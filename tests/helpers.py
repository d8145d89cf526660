"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_max_err(a, b):
    """max |a - b| / max |b|: the field-relative max-norm error used by every
    tolerance in these tests (0 when both fields are identically zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max()
    diff = np.abs(a - b).max()
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def field_errors(gpu_fields, ref_fields, names=("rho", "u", "v", "w", "T")):
    """Relative max-norm errors of (rho, u, v, w, T).  rho and T are scaled
    by their own maximum; each velocity component is scaled by the velocity
    scale max(|u|, |v|, |w|) of the reference, so that a component that is
    nearly zero by symmetry (the spanwise u of the 3D cavity) is judged
    against the flow's velocity and not against its own tiny maximum."""
    ref = dict(zip(names, ref_fields))
    vel_scale = max(float(np.abs(ref[k]).max()) for k in ("u", "v", "w"))
    out = {}
    for name, g, r in zip(names, gpu_fields, ref_fields):
        if name in ("u", "v", "w") and vel_scale > 0.0:
            out[name] = float(np.abs(np.asarray(g) - np.asarray(r)).max() / vel_scale)
        else:
            out[name] = rel_max_err(g, r)
    return out
