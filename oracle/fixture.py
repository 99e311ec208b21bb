"""Seeded fixture words for the array-path fingerprint cases (test infrastructure only).

This module is part of the ORACLE: it exists so that tests can rebuild, word for word, the
seeded states the reference's behavioural fingerprint is minted from, and check the CPU
restatement and the HIP engine against the recorded SHA-256 digests. Nothing in the
product imports it.

What is restated here, from ``meep_gpu/array_path_oracle.py`` (reference):

* ``seed``            <- ``_seed``          :190-192  (first 8 bytes of sha256, little endian)
* ``splitmix64``      <- ``_splitmix64``    :195-204  (SplitMix64 in uint64 arithmetic)
* ``unit_float32``    <- ``_unit_float32``  :206-210  (23-bit mantissa times 2**-23)
* ``state_words``     <- ``_state_words``   :213-231  (signed zeros and subnormals planted)
* ``volume``          <- ``_volume``        :234-243  (material volume with planted zeros)
* ``stored_currents`` <- ``_StoredCurrents`` :246-280 (per-step complex current words)

Every function is integer or float32 arithmetic with NumPy's own elementwise rounding, so
the words are the same on every IEEE host.
"""

from __future__ import annotations

import hashlib
import math
from typing import List, Sequence, Tuple

import numpy

FIXTURE_SET = "array_path_oracle/real-v6"
STEPS = 6
RESOLUTION = 8.0
COURANT = 0.5
PML_LN_REFLECTION = 34.5
LORENTZ_ANGULAR = 3.0
DRUDE_ANGULAR = 2.0

_MASK64 = (1 << 64) - 1


def seed(*parts: str) -> int:
    """First 8 bytes (little endian) of sha256 over ``FIXTURE_SET/part/part/...``."""
    text = "/".join((FIXTURE_SET,) + tuple(parts)).encode("utf-8")
    return int.from_bytes(hashlib.sha256(text).digest()[:8], "little")


def splitmix64(value: int, count: int) -> numpy.ndarray:
    """``count`` SplitMix64 outputs after ``value`` (Steele, Lea & Flood 2014), as uint64."""
    index = numpy.arange(1, count + 1, dtype=numpy.uint64)
    with numpy.errstate(over="ignore"):
        z = numpy.uint64(value & _MASK64) + index * numpy.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> numpy.uint64(30))) * numpy.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> numpy.uint64(27))) * numpy.uint64(0x94D049BB133111EB)
        z = z ^ (z >> numpy.uint64(31))
    return z


def unit_float32(value: int, shape: Tuple[int, ...]) -> numpy.ndarray:
    count = int(numpy.prod(shape))
    mantissa = (splitmix64(value, count) >> numpy.uint64(41)).astype(numpy.float32)
    return (mantissa * numpy.float32(2.0 ** -23)).reshape(shape)


def state_words(value: int, shape: Tuple[int, ...]) -> numpy.ndarray:
    """Normal words of either sign in [2**-6, 2**2) with ~2% +0, ~2% -0, ~1% subnormals."""
    count = int(numpy.prod(shape))
    words = splitmix64(value, count)
    mantissa = words & numpy.uint64(0x7FFFFF)
    exponent = numpy.uint64(121) + ((words >> numpy.uint64(23)) & numpy.uint64(7))
    sign = (words >> numpy.uint64(31)) & numpy.uint64(1)
    bits = (sign << numpy.uint64(31)) | (exponent << numpy.uint64(23)) | mantissa
    selector = (words >> numpy.uint64(40)) & numpy.uint64(0xFF)
    bits = numpy.where(selector < 5, numpy.uint64(0), bits)
    bits = numpy.where((selector >= 5) & (selector < 10), numpy.uint64(0x80000000), bits)
    subnormal = (sign << numpy.uint64(31)) | (mantissa | numpy.uint64(1))
    bits = numpy.where((selector >= 10) & (selector < 13), subnormal, bits)
    return bits.astype(numpy.uint32).view(numpy.float32).reshape(shape)


def volume(value: int, shape: Tuple[int, ...], low: float, high: float,
           zero_fraction: float = 0.0) -> numpy.ndarray:
    unit = unit_float32(value, shape)
    values = (numpy.float32(low) + numpy.float32(high - low) * unit).astype(numpy.float32)
    if zero_fraction > 0.0:
        selector = unit_float32(value ^ 0x5A5A5A5A, shape)
        values = numpy.where(selector < numpy.float32(zero_fraction),
                             numpy.float32(0.0), values).astype(numpy.float32)
    return values


def stored_currents(value: int, count: int = STEPS + 4) -> List[complex]:
    """Per-step current words: multiples of 2**-19 in [-1, 1), magnitude >= 2**-4."""
    words = splitmix64(value, 2 * count)
    parts = ((words >> numpy.uint64(44)).astype(numpy.int64) - (1 << 19)).astype(float)
    parts = parts * 2.0 ** -19
    parts = numpy.where(numpy.abs(parts) < 2.0 ** -4,
                        numpy.copysign(0.5, parts + 2.0 ** -30), parts)
    return [complex(float(parts[2 * i]), float(parts[2 * i + 1])) for i in range(count)]


def stored_current_at(table: Sequence[complex], time_value: float, dt: float) -> complex:
    """The table word for a slot time ``n*dt`` or ``(n+1/2)*dt``: ``floor(t/dt + 1/4)``."""
    step = int(math.floor(float(time_value) / dt + 0.25))
    if not 0 <= step < len(table):
        raise IndexError(f"stored current asked for step {step} outside {len(table)} words")
    return table[step]
