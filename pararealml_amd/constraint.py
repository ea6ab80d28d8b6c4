"""Masked value constraints (host-side problem setup).

Semantics follow pararealml/constraint.py:7-131: a constraint is a read-only
pair of a boolean mask and the values the masked elements take; applying it
writes the values in place.  The device kernels never see masks: the boundary
constraints of a problem are flattened into per-edge value arrays where NaN
marks an unconstrained element (``pararealml_amd._device.device_problem``).
"""

from typing import Optional, Sequence, Union

import numpy as np


class Constraint:
    """Mask + values (pararealml/constraint.py:7-41)."""

    def __init__(self, values: np.ndarray, mask: np.ndarray):
        values = np.asarray(values)
        mask = np.asarray(mask)
        if values.size != mask.sum():
            raise ValueError(
                f"number of values ({values.size}) must match number "
                f"of True elements in mask ({mask.sum()})"
            )
        self._values = np.array(values, copy=True)
        self._mask = np.array(mask, copy=True)
        self._values.setflags(write=False)
        self._mask.setflags(write=False)

    @property
    def values(self) -> np.ndarray:
        return self._values

    @property
    def mask(self) -> np.ndarray:
        return self._mask

    def apply(self, array: np.ndarray) -> np.ndarray:
        """Writes the values into the masked elements of ``array`` in place
        (pararealml/constraint.py:43-58)."""
        if array.shape[-self._mask.ndim :] != self._mask.shape:
            raise ValueError(
                f"input shape {array.shape} incompatible with mask shape "
                f"{self._mask.shape}"
            )
        array[..., self._mask] = self._values
        return array

    def multiply_and_add(
        self,
        addend: np.ndarray,
        multiplier: Union[float, np.ndarray],
        result: np.ndarray,
    ) -> np.ndarray:
        """``result[mask] = addend[mask] + multiplier * values`` in place
        (pararealml/constraint.py:60-101)."""
        if addend.shape != result.shape:
            raise ValueError(
                f"addend shape {addend.shape} must match result shape "
                f"{result.shape}"
            )
        if result.shape[-self._mask.ndim :] != self._mask.shape:
            raise ValueError(
                f"result shape {result.shape} incompatible with mask shape "
                f"{self._mask.shape}"
            )
        if (
            not isinstance(multiplier, float)
            and np.shape(multiplier) != self._values.shape
        ):
            raise ValueError(
                f"multiplier shape {np.shape(multiplier)} must match values "
                f"shape {self._values.shape}"
            )
        result[..., self._mask] = addend[..., self._mask] + multiplier * self._values
        return result


def apply_constraints_along_last_axis(
    constraints: Optional[Union[Sequence[Optional[Constraint]], np.ndarray]],
    array: np.ndarray,
) -> np.ndarray:
    """Applies one constraint per element of the last axis in place
    (pararealml/constraint.py:104-131)."""
    if constraints is None:
        return array
    if array.ndim <= 1:
        raise ValueError(f"input dimensions ({array.ndim}) must be at least 2")
    if len(constraints) != array.shape[-1]:
        raise ValueError(
            f"number of constraints ({len(constraints)}) must match the "
            f"size of the input array's last axis ({array.shape[-1]})"
        )
    for idx, constraint in enumerate(constraints):
        if constraint is not None:
            constraint.apply(array[..., idx : idx + 1])
    return array
