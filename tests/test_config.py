"""Numerical config decoding and checking.

The valid/invalid cases are the reference's own test vectors
(test/NumericalConfigSpec.hs:13-86); the conversion cases follow
convertConfig (src/Formura/NumericalConfig.hs:111-143).
"""
from fractions import Fraction

import pytest

from formura_amd.config import (ConfigException, NumericalConfig, TemporalBlocking,
                                convert_config, decode_config)


def test_valid_1d_temporal_blocking():
    cfg = decode_config("length_per_node: [1.0]\ngrid_per_node: [10]\ngrid_per_block: [10]\n"
                        "mpi_shape: [1]\ntemporal_blocking_interval: 5\n")
    assert cfg == NumericalConfig([Fraction(1)], [10], [10], 5, None, [1], None)


def test_valid_3d_temporal_blocking():
    cfg = decode_config("length_per_node: [1.0,2.0,3.0]\nmpi_shape: [2,2,2]\ntemporal_blocking_interval: 5\n"
                        "grid_per_node: [10,10,10]\ngrid_per_block: [10,10,10]\nfilter_interval: 100\nwith_omp: 1\n")
    assert cfg == NumericalConfig([Fraction(1), Fraction(2), Fraction(3)], [10, 10, 10], [10, 10, 10],
                                  5, 100, [2, 2, 2], 1)


def test_valid_3d_without_temporal_blocking():
    cfg = decode_config("length_per_node: [1.0,2.0,3.0]\nmpi_shape: [2,2,2]\ngrid_per_node: [10,10,10]\n")
    assert cfg == NumericalConfig([Fraction(1), Fraction(2), Fraction(3)], [10, 10, 10], None, None, None,
                                  [2, 2, 2], None)


@pytest.mark.parametrize("text", [
    "length_per_node: [1.0]\ngrid_per_block: [10]\n",                      # missing MUST field
    "length_per_node: [1.0]\ngrid_per_node: [10]\ngrid_per_block: []\nmpi_shape: []\ntemporal_blocking_interval: 5\n",
    "length_per_node: [1.0]\ngrid_per_node: [10]\ngrid_per_block: [10]\nmpi_shape: [1]\ntemporal_blocking_interval: -5\n",
])
def test_invalid(text):
    with pytest.raises(ConfigException):
        decode_config(text)


def test_exact_decimal_lengths():
    cfg = decode_config("length_per_node: [20.48, 0.1]\ngrid_per_node: [2048, 10]\n")
    assert cfg.length_per_node == [Fraction(2048, 100), Fraction(1, 10)]


def test_convert_temporal_blocking_geometry():
    nc = decode_config("length_per_node: [1.0,1.0,1.0]\ngrid_per_node: [64,64,64]\n"
                       "grid_per_block: [8,8,8]\ntemporal_blocking_interval: 4\n")
    ic = convert_config(1, None, None, nc)
    assert ic.blocking == TemporalBlocking([8, 8, 8], [9, 9, 9], 4)
    assert ic.steps_per_forward == 4
    # diffusion3.yaml: (100 + 2*1*5) / 10 = 11 blocks
    nc = decode_config("length_per_node: [1.0,1.0,1.0]\ngrid_per_node: [100,100,100]\n"
                       "grid_per_block: [10,10,10]\ntemporal_blocking_interval: 5\n")
    assert convert_config(1, None, None, nc).blocking.block_per_node == [11, 11, 11]


def test_convert_rejects_inconsistent_blocks():
    nc = decode_config("length_per_node: [1.0]\ngrid_per_node: [64]\ngrid_per_block: [7]\n"
                       "temporal_blocking_interval: 4\n")
    with pytest.raises(ConfigException, match="Inconsistent config"):
        convert_config(1, None, None, nc)


def test_convert_rejects_filter_interval_not_multiple():
    nc = decode_config("length_per_node: [1.0]\ngrid_per_node: [64]\ngrid_per_block: [8]\n"
                       "temporal_blocking_interval: 4\nfilter_interval: 6\n")
    with pytest.raises(ConfigException):
        convert_config(1, None, None, nc)


def test_space_interval_is_double_division():
    nc = decode_config("length_per_node: [20.48]\ngrid_per_node: [2048]\n")
    ic = convert_config(1, None, None, nc)
    assert ic.space_interval == [20.48 / 2048]
    assert ic.blocking is None and ic.steps_per_forward == 1
