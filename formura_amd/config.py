"""Numerical configuration (the ``.yaml`` beside a ``.fmr`` program).

Follows ``src/Formura/NumericalConfig.hs``:

* ``decode_config`` mirrors ``decodeConfig`` (:52-64): the YAML keys are the
  snake-case field names (:37-42), ``length_per_node`` and ``grid_per_node``
  are required, the optional lists must be non-empty, and the interval / omp
  fields are range-checked with the same messages.
* ``convert_config`` mirrors ``convertConfig`` (:111-143): it derives the
  blocking type and the temporal-blocking geometry
  ``block_per_node = (N + 2*s*nt) / grid_per_block`` and rejects the same
  inconsistent configurations.

Keys that the reference does not know (it ignores unknown keys, as aeson's
derived decoder does) are kept in ``NumericalConfig.extra``; the HIP backend
reads its optional tuning keys from there (``hip_tile``, ``hip_steps_per_launch``).
"""
from __future__ import annotations

import dataclasses
from decimal import Decimal
from fractions import Fraction
from typing import Any, Dict, List, Optional

import yaml


class ConfigException(Exception):
    """Raised for an invalid numerical config (``ConfigException``, NumericalConfig.hs:44-50)."""


class _ExactLoader(yaml.SafeLoader):
    """SafeLoader that keeps YAML floats as exact decimal rationals.

    The reference decodes ``length_per_node`` as ``Scientific`` (an exact
    decimal), and the front-end injects ``dx = L/N`` as an exact rational
    (Parser.hs:506), so the decimal text must not pass through a binary float.
    """


def _construct_exact_float(loader, node):
    text = loader.construct_scalar(node).replace("_", "")
    low = text.lower()
    if low in (".inf", "+.inf", "-.inf", ".nan"):
        raise ConfigException("non-finite number in numerical config: %s" % text)
    return Fraction(Decimal(text))


_ExactLoader.add_constructor("tag:yaml.org,2002:float", _construct_exact_float)


@dataclasses.dataclass
class NumericalConfig:
    """User-facing config (``NumericalConfig``, NumericalConfig.hs:24-32)."""

    length_per_node: List[Fraction]
    grid_per_node: List[int]
    grid_per_block: Optional[List[int]] = None
    temporal_blocking_interval: Optional[int] = None
    filter_interval: Optional[int] = None
    mpi_shape: Optional[List[int]] = None
    with_omp: Optional[int] = None
    extra: Dict[str, Any] = dataclasses.field(default_factory=dict)

    def key(self):
        return (self.length_per_node, self.grid_per_node, self.grid_per_block,
                self.temporal_blocking_interval, self.filter_interval,
                self.mpi_shape, self.with_omp)

    def __eq__(self, other):  # extra keys do not take part, as in the reference
        return isinstance(other, NumericalConfig) and self.key() == other.key()


_FIELDS = ("length_per_node", "grid_per_node", "grid_per_block",
           "temporal_blocking_interval", "filter_interval", "mpi_shape",
           "with_omp")


def _as_int(name, v):
    if isinstance(v, bool) or not isinstance(v, int):
        if isinstance(v, Fraction) and v.denominator == 1:
            return int(v)
        raise ConfigException("Error in $: %s: expected Int, got %r" % (name, v))
    return v


def _as_rational(name, v):
    if isinstance(v, bool):
        raise ConfigException("Error in $: %s: expected Number, got %r" % (name, v))
    if isinstance(v, int):
        return Fraction(v)
    if isinstance(v, Fraction):
        return v
    raise ConfigException("Error in $: %s: expected Number, got %r" % (name, v))


def _as_list(name, v, conv):
    if not isinstance(v, list):
        raise ConfigException("Error in $: %s: expected Array, got %r" % (name, v))
    return [conv(name, x) for x in v]


def decode_config(text: str) -> NumericalConfig:
    """Decode and check a YAML config (``decodeConfig``, NumericalConfig.hs:52-64)."""
    try:
        doc = yaml.load(text, Loader=_ExactLoader)
    except yaml.YAMLError as e:
        raise ConfigException(str(e))
    if not isinstance(doc, dict):
        raise ConfigException("Error in $: expected Object")
    for req in ("length_per_node", "grid_per_node"):
        if req not in doc or doc[req] is None:
            raise ConfigException("Error in $: key \"%s\" not found" % req)
    opt = lambda k: doc.get(k)
    cfg = NumericalConfig(
        length_per_node=_as_list("length_per_node", doc["length_per_node"], _as_rational),
        grid_per_node=_as_list("grid_per_node", doc["grid_per_node"], _as_int),
        grid_per_block=None if opt("grid_per_block") is None else _as_list("grid_per_block", opt("grid_per_block"), _as_int),
        temporal_blocking_interval=None if opt("temporal_blocking_interval") is None else _as_int("temporal_blocking_interval", opt("temporal_blocking_interval")),
        filter_interval=None if opt("filter_interval") is None else _as_int("filter_interval", opt("filter_interval")),
        mpi_shape=None if opt("mpi_shape") is None else _as_list("mpi_shape", opt("mpi_shape"), _as_int),
        with_omp=None if opt("with_omp") is None else _as_int("with_omp", opt("with_omp")),
        extra={k: v for k, v in doc.items() if k not in _FIELDS},
    )
    # MUST fields and range checks, in the order of NumericalConfig.hs:57-64
    if len(cfg.length_per_node) == 0:
        raise ConfigException("length_per_node should be a nonempty list")
    if len(cfg.grid_per_node) == 0:
        raise ConfigException("grid_per_node should be a nonempty list")
    if cfg.grid_per_block is not None and len(cfg.grid_per_block) == 0:
        raise ConfigException("grid_per_block should be a nonempty list")
    if cfg.mpi_shape is not None and len(cfg.mpi_shape) == 0:
        raise ConfigException("mpi_shape should be a nonempty list")
    if cfg.temporal_blocking_interval is not None and cfg.temporal_blocking_interval < 1:
        raise ConfigException("temporal_blocking_interval should be a positive integer")
    if cfg.filter_interval is not None and cfg.filter_interval < 1:
        raise ConfigException("filter_interval should be a positive integer")
    if cfg.with_omp is not None and cfg.with_omp not in (0, 1, 2):
        raise ConfigException("with_omp should be 0, 1 or 2")
    return cfg


def read_config(path: str) -> NumericalConfig:
    with open(path) as f:
        return decode_config(f.read())


@dataclasses.dataclass
class TemporalBlocking:
    """``TemporalBlocking`` (NumericalConfig.hs:74-80)."""

    grid_per_block: List[int]
    block_per_node: List[int]
    interval: int


@dataclasses.dataclass
class InternalConfig:
    """Config for code generation (``InternalConfig``, NumericalConfig.hs:82-93)."""

    length_per_node: List[Fraction]
    grid_per_node: List[int]
    space_interval: List[float]
    blocking: Optional[TemporalBlocking]  # None == NoBlocking
    sleeve: int
    sleeve0: Optional[int]
    filter_sleeve: Optional[int]
    filter_interval: Optional[int]
    mpi_shape: Optional[List[int]]
    with_omp: int
    extra: Dict[str, Any]

    @property
    def steps_per_forward(self) -> int:
        return self.blocking.interval if self.blocking else 1


def _zip_trunc(a, b):
    # Vec arithmetic zips and truncates to the shorter list (Vec.hs:189-194)
    return list(zip(a, b))


def convert_config(s: int, s0: Optional[int], sf: Optional[int], nc: NumericalConfig) -> InternalConfig:
    """``convertConfig`` (NumericalConfig.hs:111-143)."""
    nt = nc.temporal_blocking_interval if nc.temporal_blocking_interval is not None else 1
    total = [n + 2 * s * nt for n in nc.grid_per_node]
    blocking = None
    ms = None
    if nc.grid_per_block is not None:
        ms = [t % g if g != 0 else None for t, g in _zip_trunc(total, nc.grid_per_block)]
        if nc.temporal_blocking_interval is not None:
            if any(g == 0 for g in nc.grid_per_block):
                raise ConfigException("grid_per_block should not contain zero")
            bpn = [t // g for t, g in _zip_trunc(total, nc.grid_per_block)]
            blocking = TemporalBlocking(list(nc.grid_per_block), bpn, nc.temporal_blocking_interval)
    space = [float(l) / float(n) if n != 0 else float("inf")
             for l, n in _zip_trunc(nc.length_per_node, nc.grid_per_node)]
    ic = InternalConfig(
        length_per_node=list(nc.length_per_node),
        grid_per_node=list(nc.grid_per_node),
        space_interval=space,
        blocking=blocking,
        sleeve=s,
        sleeve0=s0,
        filter_sleeve=sf,
        filter_interval=nc.filter_interval,
        mpi_shape=None if nc.mpi_shape is None else list(nc.mpi_shape),
        with_omp=nc.with_omp if nc.with_omp is not None else 0,
        extra=dict(nc.extra),
    )
    if any(l < 0 for l in ic.length_per_node):
        raise ConfigException("the element of length_per_node should be a positive number")
    if any(n < 1 for n in ic.grid_per_node):
        raise ConfigException("the element of grid_per_node should be a positive integer")
    if ic.mpi_shape is not None and any(m < 1 for m in ic.mpi_shape):
        raise ConfigException("the element of mpi_shape should be a positive integer")
    if ic.filter_interval is not None and ic.filter_interval % nt != 0:
        raise ConfigException("the filter interval is a multiple of temporal blocking interval")
    if ms is not None and any(m is None or m != 0 for m in ms):
        raise ConfigException("Inconsistent config")
    return ic
