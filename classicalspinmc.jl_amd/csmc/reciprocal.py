"""Momentum-space helpers of reciprocal.jl (host side, numpy).

reciprocal           reciprocal.jl:3-16
get_allowed_wavevectors  reciprocal.jl:18-29
get_k_plane          reciprocal.jl:31-43 (marked untested in the reference)
get_k_path           reciprocal.jl:45-101 (direction form and high-symmetry-path form)

Momenta are returned as the reference's D x n_k matrices.  Julia's
Iterators.product varies its first factor fastest; the same order is kept.
"""
import itertools

import numpy as np


def reciprocal(a1, a2, a3=None):
    """reciprocal(a1, a2) / reciprocal(a1, a2, a3) (reciprocal.jl:3-16)."""
    a1 = np.asarray(a1, dtype=np.float64)
    a2 = np.asarray(a2, dtype=np.float64)
    if a3 is None:
        mag = 2 * np.pi / (a1[0] * a2[1] - a1[1] * a2[0])
        return mag * np.array([a2[1], -a2[0]]), mag * np.array([-a1[1], a1[0]])
    a3 = np.asarray(a3, dtype=np.float64)
    mag = 2 * np.pi / np.dot(a1, np.cross(a2, a3))
    return mag * np.cross(a2, a3), mag * np.cross(a3, a1), mag * np.cross(a1, a2)


def get_allowed_wavevectors(uc, shape, min=0, max=1):
    """get_allowed_wavevectors(uc, shape; min, max) (reciprocal.jl:18-29):
    b-vector combinations with steps (min L .. max L) / L along each
    dimension, both ends included."""
    bs = reciprocal(*uc.lattice_vectors)
    rs = [np.arange(min * L, max * L + 1) / L for L in shape]
    # Iterators.product: the first dimension varies fastest
    steps = [tuple(reversed(p)) for p in itertools.product(*reversed(rs))]
    return np.column_stack(bs) @ np.array(steps, dtype=np.float64).T


def _in_bounds(temp, lo, hi):
    keep = np.ones(temp.shape[1], dtype=bool)
    for i in range(temp.shape[0]):
        v = np.round(temp[i, :], 6)
        keep &= (np.round(lo[i], 6) <= v) & (v <= np.round(hi[i], 6))
    return keep


def get_k_plane(uc, shape, min=-2, max=2):
    """get_k_plane (reciprocal.jl:31-43): the allowed momenta whose every
    component lies in [min 2 pi, max 2 pi] (to 6 digits)."""
    ks = get_allowed_wavevectors(uc, shape, min=min, max=max)
    D = ks.shape[0]
    return ks[:, _in_bounds(ks, [min * 2 * np.pi] * D, [max * 2 * np.pi] * D)]


def _parallel(ks, origin, line):
    ks_ = ks - np.asarray(origin, dtype=np.float64)[:, None]
    norms = np.linalg.norm(ks_, axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.abs(line @ (ks_ / norms))
    return np.round(c, 6) == 1.0


def get_k_path(uc, direction_or_hsp, *args, **kw):
    """get_k_path (reciprocal.jl:45-101).

    get_k_path(uc, direction, shape, min=-2, max=2): the allowed momenta
    parallel to `direction` (measured from the tip of the normalised
    direction, as the reference does) within [min 2 pi, max 2 pi].

    get_k_path(uc, hsp, path, shape): along the high-symmetry points
    hsp[path[0]] -> hsp[path[1]] -> ...; returns (point_count, kpath)."""
    if isinstance(direction_or_hsp, dict):
        return _k_path_hsp(uc, direction_or_hsp, *args, **kw)
    direction = np.asarray(direction_or_hsp, dtype=np.float64)
    shape = args[0] if args else kw["shape"]
    lo = args[1] if len(args) > 1 else kw.get("min", -2)
    hi = args[2] if len(args) > 2 else kw.get("max", 2)
    ks = get_allowed_wavevectors(uc, shape, min=lo, max=hi)
    direction = direction / np.linalg.norm(direction)
    temp = ks[:, _parallel(ks, direction, direction)]
    D = ks.shape[0]
    return temp[:, _in_bounds(temp, [lo * 2 * np.pi] * D, [hi * 2 * np.pi] * D)]


def _k_path_hsp(uc, hsp, path, shape):
    ks = get_allowed_wavevectors(uc, shape)
    kpath = np.asarray(hsp[path[0]], dtype=np.float64).reshape(-1, 1)
    point_count = np.zeros(len(path), dtype=np.int64)
    for ind in range(len(path) - 1):
        p1 = np.asarray(hsp[path[ind]], dtype=np.float64)
        p2 = np.asarray(hsp[path[ind + 1]], dtype=np.float64)
        line = (p2 - p1) / np.linalg.norm(p2 - p1)
        temp = ks[:, _parallel(ks, p1, line)]
        new_path = temp[:, _in_bounds(temp, np.minimum(p1, p2), np.maximum(p1, p2))]
        if new_path.shape[1] and not np.array_equal(np.round(new_path[:, -1], 5), np.round(p2, 5)):
            new_path = new_path[:, ::-1]
        kpath = np.hstack([kpath, new_path])
        point_count[ind + 1] = new_path.shape[1] + point_count[ind]
    return point_count, kpath
