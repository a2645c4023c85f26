"""Snapshot files in the layout of the reference's sample drivers, and their
drift-aware loader.

The reference's drivers (sample/*/main.c, e.g. sample/3d_hllc/main.c:51-82)
write one text file per rank and output step, ``<prefix><time_step>_<rank>.dat``,
one line per cell in loop order (z outermost, x innermost), each value as
``%.17e`` (exact for doubles), with the cell's position from ``to_pos_<axis>``
in the first columns.  Because the generated code never un-shifts, the file
holds the drifting frame: line order is index order, and the positions wrap
around once per axis where ``(i + offset) mod N`` returns to 0.

``load_data`` follows sample/*/formura_data_load.py:11-43: it reads every
rank's file, reshapes each to ``(nz, ny, nx, columns)`` of a node, assembles
the ranks along the MPI shape (rank = rx + ry*Px + rz*Px*Py), and rolls each
axis so that the position columns increase from the origin -- the first index
where the position decreases (the wrap) becomes index 0.
"""
from __future__ import annotations

import os
from typing import Iterable, List, Optional, Sequence

import numpy as np


def dat_path(prefix: str, time_step: int, rank: int = 0) -> str:
    """``<prefix><time_step>_<rank>.dat`` (sample/*/main.c file_open)."""
    return "%s%d_%d.dat" % (prefix, time_step, rank)


def write_dat(path: str, columns: Sequence[np.ndarray]) -> None:
    """One line per cell, the columns' values as ``%.17e`` separated by spaces.
    Columns are arrays of one node's grid in C order (x slowest); lines run
    with the last grid axis outermost, as the drivers' loops do (z, y, x)."""
    cols = [np.asarray(c, dtype=np.float64) for c in columns]
    flat = [c.transpose(tuple(reversed(range(c.ndim)))).reshape(-1) for c in cols]
    with open(path, "w") as f:
        for row in zip(*flat):
            f.write(" ".join("%.17e" % v for v in row) + "\n")


def origin_index(x: np.ndarray) -> int:
    """First index after the wrap of a position column (0 without a wrap)."""
    edge = np.where(np.diff(x) < 0.0)[0]
    return int(edge[0] + 1) if len(edge) else 0


def load_data(prefix: str, step: int, grid_per_node: Iterable[int],
              mpi_shape: Optional[Iterable[int]] = None) -> np.ndarray:
    """The snapshot of ``step`` as an array ``(Z, Y, X, columns)`` over the whole
    grid, positions increasing along every axis (columns 0, 1, 2 are x, y, z)."""
    g = list(grid_per_node)
    p = list(mpi_shape) if mpi_shape is not None else [1] * len(g)
    while len(g) < 3:
        g.append(1)
        p.append(1)
    blocks: List = []
    for rz in range(p[2]):
        plane = []
        for ry in range(p[1]):
            row = []
            for rx in range(p[0]):
                rank = rx + ry * p[0] + rz * p[0] * p[1]
                d = np.loadtxt(os.path.join(prefix, "%d_%d.dat" % (step, rank)), ndmin=2)
                row.append(d.reshape(g[2], g[1], g[0], d.shape[1]))
            plane.append(np.concatenate(row, axis=2))
        blocks.append(np.concatenate(plane, axis=1))
    data = np.concatenate(blocks, axis=0)
    oi = origin_index(data[0, 0, :, 0])
    oj = origin_index(data[0, :, 0, 1])
    ok = origin_index(data[:, 0, 0, 2])
    data = np.roll(data, -oi, axis=2)
    data = np.roll(data, -oj, axis=1)
    return np.roll(data, -ok, axis=0)


def read_config_shape(yaml_path: str):
    """(mpi_shape, grid_per_node) of a numerical config (safe YAML loader)."""
    import yaml
    with open(yaml_path) as f:
        cfg = yaml.safe_load(f)
    return list(cfg.get("mpi_shape", [1] * len(cfg["grid_per_node"]))), list(cfg["grid_per_node"])
