"""Spatial meshes (host-side problem setup).

Semantics follow pararealml/mesh.py:28-107 (construction and validation),
mesh.py:305-376 (vertex/cell shapes, axis coordinates, ``ij`` coordinate
grids) and mesh.py:379-628 (volume, boundary sizes and coordinate system
conversions).  The mesh is only ever evaluated on the host; the device
kernels receive the vertex counts and step sizes it derives.
"""

from enum import Enum
from typing import Sequence, Tuple, TypeVar

import numpy as np

SpatialDomainInterval = Tuple[float, float]


class CoordinateSystem(Enum):
    """Supported coordinate systems (pararealml/mesh.py:9-17)."""

    CARTESIAN = 0
    POLAR = 1
    CYLINDRICAL = 2
    SPHERICAL = 3


class Mesh:
    """A uniform hyper-rectangular grid (pararealml/mesh.py:20-107).

    Vertex counts are ``round((hi - lo) / d_x + 1)`` per axis and vertex
    coordinates are ``linspace(lo, hi, n)``; cell centres are shifted by half
    a step and there is one fewer of them per axis.
    """

    def __init__(
        self,
        x_intervals: Sequence[SpatialDomainInterval],
        d_x: Sequence[float],
        coordinate_system_type: CoordinateSystem = CoordinateSystem.CARTESIAN,
    ):
        if len(x_intervals) == 0:
            raise ValueError(
                "number of spatial domain intervals must be greater than 0"
            )
        if len(x_intervals) != len(d_x):
            raise ValueError(
                f"number of spatial domain intervals ({len(x_intervals)}) "
                f"must match number of spatial step sizes ({len(d_x)})"
            )
        if any(hi <= lo for lo, hi in x_intervals):
            raise ValueError(
                "upper bound of every spatial domain interval must be greater "
                "than its lower bound"
            )
        if any(step <= 0.0 for step in d_x):
            raise ValueError("all spatial step sizes must be greater than 0")

        dims = len(x_intervals)
        if coordinate_system_type != CoordinateSystem.CARTESIAN:
            if x_intervals[0][0] < 0:
                raise ValueError(
                    f"lower bound of r interval ({x_intervals[0][0]}) "
                    "must be non-negative"
                )
            if x_intervals[1][0] < 0.0 or x_intervals[1][1] > 2.0 * np.pi:
                raise ValueError(
                    f"lower bound of theta ({x_intervals[1][0]}) must be "
                    f"non-negative and upper bound ({x_intervals[1][1]}) must "
                    "be no more than two Pi"
                )
            if coordinate_system_type == CoordinateSystem.POLAR:
                if dims != 2:
                    raise ValueError(
                        f"number of dimensions ({dims}) of polar mesh must "
                        "be 2"
                    )
            else:
                if dims != 3:
                    raise ValueError(
                        f"number of dimensions ({dims}) of cylindrical and "
                        "spherical meshes must be 3"
                    )
                if coordinate_system_type == CoordinateSystem.SPHERICAL and (
                    x_intervals[2][0] < 0.0 or x_intervals[2][1] > np.pi
                ):
                    raise ValueError(
                        f"lower bound of phi ({x_intervals[2][0]}) must be "
                        f"non-negative and upper bound ({x_intervals[2][1]}) "
                        "must be no more than Pi"
                    )

        self._x_intervals = tuple(tuple(iv) for iv in x_intervals)
        self._d_x = tuple(d_x)
        self._coordinate_system_type = coordinate_system_type
        self._dimensions = dims

        self._volume = self._compute_volume()
        self._boundary_sizes = tuple(self._compute_boundary_sizes())
        self._vertices_shape = self._shape_for(True)
        self._cells_shape = self._shape_for(False)
        self._vertex_axis_coordinates = self._axis_coords_for(True)
        self._cell_center_axis_coordinates = self._axis_coords_for(False)
        self._vertex_coordinate_grids = self._grids_for(True)
        self._cell_center_coordinate_grids = self._grids_for(False)

    # ------------------------------------------------------------------ props
    @property
    def x_intervals(self) -> Sequence[SpatialDomainInterval]:
        return self._x_intervals

    @property
    def d_x(self) -> Sequence[float]:
        return self._d_x

    @property
    def coordinate_system_type(self) -> CoordinateSystem:
        return self._coordinate_system_type

    @property
    def dimensions(self) -> int:
        return self._dimensions

    @property
    def volume(self) -> float:
        return self._volume

    @property
    def boundary_sizes(self) -> Sequence[Tuple[float, float]]:
        return self._boundary_sizes

    @property
    def vertices_shape(self) -> Tuple[int, ...]:
        return self._vertices_shape

    @property
    def cells_shape(self) -> Tuple[int, ...]:
        return self._cells_shape

    @property
    def vertex_axis_coordinates(self) -> Tuple[np.ndarray, ...]:
        return self._vertex_axis_coordinates

    @property
    def cell_center_axis_coordinates(self) -> Tuple[np.ndarray, ...]:
        return self._cell_center_axis_coordinates

    @property
    def vertex_coordinate_grids(self) -> Tuple[np.ndarray, ...]:
        return self._vertex_coordinate_grids

    @property
    def cell_center_coordinate_grids(self) -> Tuple[np.ndarray, ...]:
        return self._cell_center_coordinate_grids

    # ---------------------------------------------------------------- methods
    def shape(self, vertex_oriented: bool) -> Tuple[int, ...]:
        return self._vertices_shape if vertex_oriented else self._cells_shape

    def axis_coordinates(self, vertex_oriented: bool) -> Tuple[np.ndarray, ...]:
        if vertex_oriented:
            return self._vertex_axis_coordinates
        return self._cell_center_axis_coordinates

    def coordinate_grids(self, vertex_oriented: bool) -> Tuple[np.ndarray, ...]:
        if vertex_oriented:
            return self._vertex_coordinate_grids
        return self._cell_center_coordinate_grids

    def cartesian_coordinate_grids(
        self, vertex_oriented: bool
    ) -> Tuple[np.ndarray, ...]:
        return tuple(
            to_cartesian_coordinates(
                self.coordinate_grids(vertex_oriented),
                self._coordinate_system_type,
            )
        )

    def all_index_coordinates(
        self, vertex_oriented: bool, flatten: bool = False
    ) -> np.ndarray:
        """Stacked coordinates of every vertex / cell centre, last axis = x."""
        coords = np.stack(self.coordinate_grids(vertex_oriented), axis=-1)
        if flatten:
            coords = coords.reshape((-1, self._dimensions))
        return coords

    def unit_vector_grids(
        self, vertex_oriented: bool
    ) -> Tuple[np.ndarray, ...]:
        grids = self.coordinate_grids(vertex_oriented)
        return tuple(
            np.stack(vec, axis=-1)
            for vec in unit_vectors_at(grids, self._coordinate_system_type)
        )

    # -------------------------------------------------------------- internals
    def _shape_for(self, vertex_oriented: bool) -> Tuple[int, ...]:
        extra = 1 if vertex_oriented else 0
        return tuple(
            int(round((hi - lo) / step + extra))
            for (lo, hi), step in zip(self._x_intervals, self._d_x)
        )

    def _axis_coords_for(self, vertex_oriented: bool) -> Tuple[np.ndarray, ...]:
        shape = self._vertices_shape if vertex_oriented else self._cells_shape
        out = []
        for axis, (lo, hi) in enumerate(self._x_intervals):
            if not vertex_oriented:
                half = self._d_x[axis] / 2.0
                lo, hi = lo + half, hi - half
            coords = np.linspace(lo, hi, shape[axis])
            coords.setflags(write=False)
            out.append(coords)
        return tuple(out)

    def _grids_for(self, vertex_oriented: bool) -> Tuple[np.ndarray, ...]:
        grids = np.meshgrid(*self.axis_coordinates(vertex_oriented), indexing="ij")
        for g in grids:
            g.setflags(write=False)
        return tuple(grids)

    def _compute_volume(self) -> float:
        cs = self._coordinate_system_type
        if cs == CoordinateSystem.CARTESIAN:
            return float(np.prod([hi - lo for lo, hi in self._x_intervals]))
        r_lo, r_hi = self._x_intervals[0]
        th_lo, th_hi = self._x_intervals[1]
        if cs == CoordinateSystem.SPHERICAL:
            ph_lo, ph_hi = self._x_intervals[2]
            return (
                (r_hi**3 - r_lo**3)
                / 3.0
                * (th_hi - th_lo)
                * (np.cos(ph_lo) - np.cos(ph_hi))
            )
        base = (r_hi**2 - r_lo**2) * (th_hi - th_lo) / 2.0
        if self._dimensions == 2:
            return base
        z_lo, z_hi = self._x_intervals[2]
        return base * (z_hi - z_lo)

    def _compute_boundary_sizes(self) -> Sequence[Tuple[float, float]]:
        cs = self._coordinate_system_type
        if cs == CoordinateSystem.CARTESIAN:
            lengths = [hi - lo for lo, hi in self._x_intervals]
            vol = float(np.prod(lengths))
            return [(vol / ln, vol / ln) for ln in lengths]
        r_lo, r_hi = self._x_intervals[0]
        th_span = self._x_intervals[1][1] - self._x_intervals[1][0]
        if cs == CoordinateSystem.SPHERICAL:
            ph_lo, ph_hi = self._x_intervals[2]
            cos_diff = np.cos(ph_lo) - np.cos(ph_hi)
            annulus = (r_hi**2 - r_lo**2) / 2.0
            return [
                (r_lo**2 * th_span * cos_diff, r_hi**2 * th_span * cos_diff),
                (annulus * (ph_hi - ph_lo),) * 2,
                (
                    annulus * th_span * np.sin(ph_lo),
                    annulus * th_span * np.sin(ph_hi),
                ),
            ]
        r_sizes = (r_lo * th_span, r_hi * th_span)
        th_sizes = (r_hi - r_lo, r_hi - r_lo)
        if self._dimensions == 2:
            return [r_sizes, th_sizes]
        z_span = self._x_intervals[2][1] - self._x_intervals[2][0]
        return [
            (r_sizes[0] * z_span, r_sizes[1] * z_span),
            (th_sizes[0] * z_span, th_sizes[1] * z_span),
            ((r_hi**2 - r_lo**2) * th_span / 2.0,) * 2,
        ]


Coordinate = TypeVar("Coordinate", float, np.ndarray)
Coordinates = Sequence[Coordinate]


def unit_vectors_at(
    x: Coordinates, coordinate_system_type: CoordinateSystem
) -> Sequence[Coordinates]:
    """Cartesian components of the coordinate system's unit vectors at x
    (pararealml/mesh.py:487-545)."""
    cs = coordinate_system_type
    if cs == CoordinateSystem.CARTESIAN:
        vecs = []
        for i in range(len(x)):
            comps = [np.zeros_like(x[i])] * len(x)
            comps[i] = np.ones_like(x[i])
            vecs.append(comps)
        return vecs
    if cs == CoordinateSystem.POLAR:
        s, c = np.sin(x[1]), np.cos(x[1])
        return [[c, s], [-s, c]]
    if cs == CoordinateSystem.CYLINDRICAL:
        th = x[1]
        z, o = np.zeros_like(th), np.ones_like(th)
        s, c = np.sin(th), np.cos(th)
        return [[c, s, z], [-s, c, z], [z, z, o]]
    if cs == CoordinateSystem.SPHERICAL:
        th, ph = x[1], x[2]
        z = np.zeros_like(th)
        st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
        return [[sp * ct, sp * st, cp], [-st, ct, z], [cp * ct, cp * st, -sp]]
    raise ValueError(f"unsupported coordinate system type ({cs.name})")


def to_cartesian_coordinates(
    x: Coordinates, from_coordinate_system_type: CoordinateSystem
) -> Coordinates:
    """pararealml/mesh.py:548-582."""
    cs = from_coordinate_system_type
    if cs == CoordinateSystem.CARTESIAN:
        return x
    if cs == CoordinateSystem.POLAR:
        return [x[0] * np.cos(x[1]), x[0] * np.sin(x[1])]
    if cs == CoordinateSystem.CYLINDRICAL:
        return [x[0] * np.cos(x[1]), x[0] * np.sin(x[1]), x[2]]
    if cs == CoordinateSystem.SPHERICAL:
        return [
            x[0] * np.sin(x[2]) * np.cos(x[1]),
            x[0] * np.sin(x[2]) * np.sin(x[1]),
            x[0] * np.cos(x[2]),
        ]
    raise ValueError(f"unsupported coordinate system type ({cs.name})")


def from_cartesian_coordinates(
    x: Coordinates, to_coordinate_system_type: CoordinateSystem
) -> Coordinates:
    """pararealml/mesh.py:585-628."""
    cs = to_coordinate_system_type
    if cs == CoordinateSystem.CARTESIAN:
        return x
    rho = np.sqrt(x[0] ** 2 + x[1] ** 2)
    if cs == CoordinateSystem.POLAR:
        return [rho, np.arctan2(x[1], x[0])]
    if cs == CoordinateSystem.CYLINDRICAL:
        return [rho, np.arctan2(x[1], x[0]), x[2]]
    if cs == CoordinateSystem.SPHERICAL:
        return [
            np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2),
            np.arctan2(x[1], x[0]),
            np.arctan2(rho, x[2]),
        ]
    raise ValueError(f"unsupported coordinate system type ({cs.name})")
