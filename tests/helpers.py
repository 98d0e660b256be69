"""Shared helpers of the parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


def interior(w, ng=3):
    return w[..., ng:-ng, ng:-ng, ng:-ng]


def fill_ghosts_wrap(w, ng=3):
    """Give every ghost point (edges and corners too) the periodic image of
    the interior, so no ghost holds a zero density."""
    out = w.copy()
    inner = w[:, ng:-ng, ng:-ng, ng:-ng]
    out[:] = np.pad(inner, ((0, 0), (ng, ng), (ng, ng), (ng, ng)), mode="wrap")
    return out
