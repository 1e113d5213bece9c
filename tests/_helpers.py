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


def reverse_line_columns(a_ptr1, a_idx1, a_val, line_len):
    """The system with the cells of every grid line numbered backwards.

    Cell n (unknowns 2n and 2n+1) lies on line n // line_len at position
    n % line_len; it moves to position line_len - 1 - n % line_len of the
    same line.  Returns the 1-based CSR of P A P^T with sorted columns and the
    permutation perm (new index = perm[old index]), so that P b is
    `bp[perm] = b` and a solution maps back as `x = xp[perm]`."""
    ptr = np.asarray(a_ptr1, dtype=np.int64) - 1
    n = len(ptr) - 1
    idx = np.arange(n)
    cell, comp = idx // 2, idx % 2
    perm = 2 * ((cell // line_len) * line_len + line_len - 1 - cell % line_len) + comp
    rows = np.repeat(idx, np.diff(ptr))
    cols = np.asarray(a_idx1, dtype=np.int64) - 1
    nr, nc = perm[rows], perm[cols]
    order = np.lexsort((nc, nr))
    counts = np.bincount(nr, minlength=n)
    ptr_new = np.concatenate([[0], np.cumsum(counts)])
    return ptr_new + 1, nc[order] + 1, np.asarray(a_val)[order], perm
