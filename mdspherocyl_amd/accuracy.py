"""Error metrics of one force evaluation against a reference evaluation.

Per-particle vectors (forces, torques) are compared row by row:

* ``rel_l2``      ||a - b|| / ||b|| over all components;
* ``max_rel_rms`` max_i |a_i - b_i| / rms_j |b_j|, the worst particle against
  the typical magnitude;
* ``max_rel_own`` max_i |a_i - b_i| / |b_i|, the worst particle against its
  own magnitude (particles with b_i = 0 are skipped);
* ``max_abs``     max_i |a_i - b_i|.

Scalars (energies) are compared as |a - b| / |b|.  A comparison against a
reference that is zero or not finite is degenerate: ``check_reference``
raises, so that a dead evaluation can never report perfect agreement.
"""
import math

import numpy as np


class DegenerateReference(ValueError):
    """The reference is all zero or not finite."""


def check_reference(name, b):
    b = np.asarray(b, dtype=np.float64)
    if not np.all(np.isfinite(b)):
        raise DegenerateReference("%s: reference holds values that are not finite" % name)
    if b.size and not np.any(b != 0):
        raise DegenerateReference("%s: reference is all zero" % name)


def vector_errors(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(b, dtype=np.float64).reshape(-1, 3)
    d = np.linalg.norm(a - b, axis=1)
    nb = np.linalg.norm(b, axis=1)
    rms = math.sqrt(float(np.mean(nb * nb)))
    own = d[nb > 0] / nb[nb > 0]
    out = {
        "rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
        "max_rel_rms": float(d.max() / rms),
        "max_rel_own": float(own.max()) if own.size else 0.0,
        "max_abs": float(d.max()),
        "ref_rms": rms,
    }
    if not all(math.isfinite(v) for v in out.values()):
        raise DegenerateReference("error of a vector set is not finite")
    return out


def scalar_error(a, b):
    if not (math.isfinite(a) and math.isfinite(b)) or b == 0:
        raise DegenerateReference("energy comparison against %r" % b)
    return abs(a - b) / abs(b)


def force_errors(F, T, U, Fr, Tr, Ur):
    """Errors of (F, T, U) against the reference (Fr, Tr, Ur)."""
    check_reference("F", Fr)
    check_reference("T", Tr)
    return {"F": vector_errors(F, Fr), "T": vector_errors(T, Tr), "U": scalar_error(U, Ur)}
