"""Conversions between the compact layout and the reference's padded layout.

Compact: interior voxels only, numpy shape ``(nz, ny, nx)`` (x fastest).
Face velocities: ``vx (nz, ny, nx+1)``, ``vy (nz, ny+1, nx)``,
``vz (nz+1, ny, nx)``; ``vx[k, j, i]`` is the face between cells i-1 and i.

Padded: the reference's arrays of ``(nz+2, ny+2, nx+2)`` with one ghost cell
per side (src/NmrRelaxationSolve.cpp:3284-3287); the mask holds the chars
'0'/'1'; ``velocity.x`` at padded (i, j, k) is the face between i-1 and i,
read for i = 1..nx+1 (cpp:1466-1482).
"""
from __future__ import annotations

import numpy as np


def pad_field(f: np.ndarray) -> np.ndarray:
    """Compact cell field -> padded FP64 array with the zero-derivative ghost
    copy of BoundaryConditionsForVectorField (cpp:1664-1754)."""
    return np.pad(np.asarray(f, dtype=np.float64), 1, mode="edge")


def pad_mask(mask: np.ndarray) -> np.ndarray:
    """Compact 0/1 mask -> padded '0'/'1' chars with the ghost copy of
    BoundaryConditionsForActiveCells (cpp:880-928)."""
    m = np.pad(np.asarray(mask, dtype=np.uint8), 1, mode="edge")
    return np.where(m != 0, ord("1"), ord("0")).astype(np.uint8)


def pad_velocity(vx, vy, vz, shape):
    """Compact face velocities -> the reference's padded arrays (faces at
    padded index 1..n+1 along their own axis, 0 elsewhere)."""
    nz, ny, nx = shape
    px = np.zeros((nz + 2, ny + 2, nx + 2))
    py = np.zeros_like(px)
    pz = np.zeros_like(px)
    if vx is not None:
        px[1 : nz + 1, 1 : ny + 1, 1 : nx + 2] = vx
    if vy is not None:
        py[1 : nz + 1, 1 : ny + 2, 1 : nx + 1] = vy
    if vz is not None:
        pz[1 : nz + 2, 1 : ny + 1, 1 : nx + 1] = vz
    return px, py, pz


def unpad(f: np.ndarray) -> np.ndarray:
    """Padded array -> its interior."""
    return np.ascontiguousarray(f[1:-1, 1:-1, 1:-1])
