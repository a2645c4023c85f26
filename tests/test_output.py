"""Snapshot files and the drift-aware loader (formura_amd.output), in the
layout of the reference's drivers (sample/3d_hllc/main.c:51-82) and loader
(sample/3d_hllc/formura_data_load.py:11-43).

A known field on a 3-D grid split over an MPI shape is written the way each
rank's driver writes it -- its node's cells in the drifting frame (index i at
position (i + offset) mod N), lines z-outermost -- and the loader must return
the field in position order over the whole grid, bit for bit.
"""
import numpy as np
import pytest

from formura_amd import output


@pytest.mark.parametrize("grid,shape,offs", [((6, 4, 2), (1, 1, 1), (5, 0, 1)),
                                             ((4, 3, 2), (2, 3, 1), (3, 7, 1)),
                                             ((8,), (1,), (3,))])
def test_drifting_frame_round_trip(tmp_path, grid, shape, offs):
    g = list(grid) + [1] * (3 - len(grid))
    p = list(shape) + [1] * (3 - len(shape))
    o = list(offs) + [0] * (3 - len(offs))
    tot = [a * b for a, b in zip(g, p)]
    dx = [0.5, 0.25, 0.125]
    rng = np.random.default_rng(1)
    field = rng.standard_normal(tot)  # [X, Y, Z] in position order
    for rz in range(p[2]):
        for ry in range(p[1]):
            for rx in range(p[0]):
                rank = rx + ry * p[0] + rz * p[0] * p[1]
                lo = [rx * g[0], ry * g[1], rz * g[2]]
                # the node's index i holds global position (lo + i + offset) mod total
                pos = [(lo[a] + np.arange(g[a]) + o[a]) % tot[a] for a in range(3)]
                P = np.meshgrid(*pos, indexing="ij")
                cols = [dx[a] * P[a] for a in range(3)] + [field[P[0], P[1], P[2]]]
                output.write_dat(output.dat_path(str(tmp_path) + "/", 7, rank), cols)
    data = output.load_data(str(tmp_path), 7, grid, shape)
    assert data.shape == (tot[2], tot[1], tot[0], 4)
    np.testing.assert_array_equal(data[..., 3], field.transpose(2, 1, 0))
    np.testing.assert_array_equal(data[0, 0, :, 0], dx[0] * np.arange(tot[0]))


def test_dat_values_round_trip_exactly(tmp_path):
    x = np.array([1.0 / 3.0, -2.5e-300, 7e300, 0.1, np.nextafter(1.0, 2.0)])
    path = str(tmp_path / "0_0.dat")
    output.write_dat(path, [x, x * 2])
    back = np.loadtxt(path, ndmin=2)
    np.testing.assert_array_equal(back[:, 0], x)
    np.testing.assert_array_equal(back[:, 1], x * 2)
