"""NumPy restatement of the reference's Solution post-processing.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``Solution.y`` (pararealml/solution.py:114-145) evaluates the snapshot stack
at arbitrary points through ``scipy.interpolate.interpn(...,
bounds_error=False, fill_value=None)``, a third-party routine (SciPy 1.15.3
in this environment, ``scipy/interpolate/_rgi.py``).  ``interp_regular``
restates its published algorithm for the "linear" and "nearest" methods:

* per axis the interval ``i`` with ``x[i] <= v < x[i+1]`` (``0`` below the
  first coordinate, ``n-2`` at and above the last: outside the grid the
  boundary interval extrapolates), an axis of one coordinate gives index 0
  with distance 0;
* the normalised distance ``y = (v - x[i]) / (x[i+1] - x[i])``;
* "linear": ``sum over the 2^D corners of v[corner] * prod(1 - y or y)``,
  the corners in lexicographic order (axis 0 slowest, lower first), each
  weight multiplied up from 1 over the axes in order and the sum started
  from 0 -- the operation order of ``RegularGridInterpolator.
  _evaluate_linear``;
* "nearest": the lower corner where ``y <= 0.5``, else the upper;
* a query point with a NaN coordinate gives NaN.

``matching_indices`` / ``diff`` restate ``Solution.diff``
(solution.py:182-258): the time points of the solution with the fewest,
matched in every other solution by ``round((t - t_0) / d_t)`` within
``atol``, and the differences other - self at the matched snapshots.

Pinning: ``interp_regular`` against ``scipy.interpolate.interpn`` itself
(the reference's call, bit for bit in FP64, inside and outside the grid,
1-3 axes, single-coordinate axes, NaN points), against exactness on
multilinear fields and, for "nearest", against the vertex values at jittered
vertices; ``diff`` against hand-built time grids whose matches are known
(every other point of a grid of half the step, a grid shifted by more and by
less than ``atol``, no match at all) -- tests/test_oracle_solution.py.
"""

from typing import List, Sequence, Tuple

import numpy as np


def find_interval(x: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Interval index per value (see the module docstring); NaN gives 0."""
    n = len(x)
    if n == 1:
        return np.zeros(v.shape, dtype=np.int64)
    i = np.searchsorted(x, v, side="right") - 1
    i = np.clip(i, 0, n - 2)
    return np.where(np.isnan(v), 0, i).astype(np.int64)


def interp_regular(axes: Sequence[np.ndarray], values: np.ndarray, xi: np.ndarray,
                   method: str = "linear") -> np.ndarray:
    """Snapshot stack ``values`` (T, *grid, C) at the points ``xi`` (..., D):
    returns (T, *xi.shape[:-1], C), the layout Solution.y returns
    (solution.py:132-145)."""
    d = len(axes)
    xi = np.asarray(xi, dtype=np.float64)
    pts = xi.reshape(-1, d)
    lo, hi, y = [], [], []
    for a, x in enumerate(axes):
        v = pts[:, a]
        i = find_interval(x, v)
        if len(x) == 1:
            lo.append(i)
            hi.append(i)
            y.append(np.zeros(len(v)))
        else:
            lo.append(i)
            hi.append(i + 1)
            y.append((v - x[i]) / (x[i + 1] - x[i]))
    vals = np.moveaxis(values, 0, -2)  # (*grid, T, C)
    if method == "nearest":
        idx = tuple(np.where(y[a] <= 0.5, lo[a], hi[a]) for a in range(d))
        res = vals[idx]
    elif method == "linear":
        res = np.zeros((len(pts),) + vals.shape[d:])
        for k in range(1 << d):
            corner = []
            weight = np.ones(len(pts))
            for a in range(d):
                upper = (k >> (d - 1 - a)) & 1
                corner.append(hi[a] if upper else lo[a])
                weight = weight * (y[a] if upper else 1.0 - y[a])
            res = res + vals[tuple(corner)] * weight[:, None, None]
    else:
        raise ValueError(f"method {method} is not restated")
    res[np.isnan(pts).any(axis=1)] = np.nan
    res = np.moveaxis(res, -2, 0)  # (T, P, C)
    return res.reshape((values.shape[0],) + xi.shape[:-1] + (values.shape[-1],))


def matching_indices(all_t: Sequence[np.ndarray], all_dt: Sequence[float],
                     atol: float) -> Tuple[List[float], List[List[int]]]:
    """Matched time points and, per solution, the matched snapshot indices
    (solution.py:203-240)."""
    shortest = 0
    for j, t in enumerate(all_t):
        if len(t) < len(all_t[shortest]):
            shortest = j
    times: List[float] = []
    idx: List[List[int]] = [[] for _ in all_t]
    for i, t in enumerate(all_t[shortest]):
        found = []
        for j, t_j in enumerate(all_t):
            if j == shortest:
                found.append(i)
                continue
            k = int(round((t - t_j[0]) / all_dt[j]))
            if 0 <= k < len(t_j) and abs(t - t_j[k]) <= atol:
                found.append(k)
            else:
                found = None
                break
        if found is None:
            continue
        times.append(t)
        for j, k in enumerate(found):
            idx[j].append(k)
    return times, idx


def diff(self_y: np.ndarray, others_y: Sequence[np.ndarray], all_t, all_dt,
         atol: float = 1e-8):
    """(matched times, [other - self at the matched snapshots])
    (solution.py:241-258); the arrays are (n_matched, *y_shape), or shape
    (0,) without a match, as the reference's ``np.array([])``."""
    times, idx = matching_indices(all_t, all_dt, atol)
    out = []
    for j, other in enumerate(others_y):
        if times:
            out.append(other[np.array(idx[j + 1])] - self_y[np.array(idx[0])])
        else:
            out.append(np.array([]))
    return np.array(times), out
