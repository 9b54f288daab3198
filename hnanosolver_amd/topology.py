"""Host-side leaf-order helpers (numpy) for leaf-dense coordinate payloads.

The leaf order is the one NanoVDB's voxelsToGrid gives an IndexGrid built from the
payload's coordinates (nanovdb/tools/cuda/PointsToGrid.cuh:596-600 root key, :640-651
voxel key, NanoVDB.h:3568 internal-node child offsets). OpenVDB's LeafManager order, which
the reference's IndexGridBuilder writes coordinates in (src/Utils/GridBuilder.hpp:164-174),
is the same order, so a payload in this order is what the reference's SOPs hand over.
"""
from __future__ import annotations

import numpy as np

LEAF_LOCAL = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(512, 3).astype(np.int32)
"""Local voxel coordinates of leaf offsets 0..511: offset = x<<6 | y<<3 | z (NanoVDB.h:4598)."""


def root_key(origins: np.ndarray) -> np.ndarray:
    o = origins.astype(np.int64) + (1 << 31)
    k = (o >> 12).astype(np.uint64)
    return (k[:, 0] << np.uint64(42)) | (k[:, 1] << np.uint64(21)) | k[:, 2]


def tile_local_key(origins: np.ndarray) -> np.ndarray:
    o = origins.astype(np.int64)
    up = (((o[:, 0] & 4095) >> 7) << 10) | (((o[:, 1] & 4095) >> 7) << 5) | ((o[:, 2] & 4095) >> 7)
    lo = (((o[:, 0] & 127) >> 3) << 8) | (((o[:, 1] & 127) >> 3) << 4) | ((o[:, 2] & 127) >> 3)
    return (up << 12) | lo


def leaf_order(origins: np.ndarray) -> np.ndarray:
    """Permutation that sorts leaf origins (L, 3) into voxelsToGrid order."""
    origins = np.asarray(origins)
    if len(origins) == 0:
        return np.zeros(0, np.int64)
    return np.lexsort((tile_local_key(origins), root_key(origins)))


def coords_from_leaves(origins: np.ndarray) -> np.ndarray:
    """Leaf-dense coordinates (512 L, 3) int32: each leaf's voxels in offset order."""
    origins = np.asarray(origins, dtype=np.int32)
    return (origins[:, None, :] + LEAF_LOCAL[None, :, :]).reshape(-1, 3).astype(np.int32)


def leaf_origins(coords: np.ndarray) -> np.ndarray:
    return np.asarray(coords, dtype=np.int32).reshape(-1, 512, 3)[:, 0, :]
