"""Shared helpers of the test suite."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def vel_rel_l2(st_a, st_b, ux="ux_b", uy="uy_b"):
    a = np.concatenate([np.ravel(st_a[ux]), np.ravel(st_a[uy])])
    b = np.concatenate([np.ravel(st_b[ux]), np.ravel(st_b[uy])])
    return rel_l2(a, b)
