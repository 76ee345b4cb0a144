"""End-of-series marker, as wildboar's ``utils/variable_len.py:17-115``.

``EOS`` is the float32 NaN with bit pattern ``0x7F800009`` (quieted to
``0x7FC00009`` when stored as float32, ``0x7FF8000120000000`` as float64).
A value is an end-of-series marker when it is NaN and the low four bits of
its float32 bit pattern are 9.  The length of a padded series is the index of
its first marker.
"""

import numpy as np

EOS = eos = EoS = np.array(0x7F800009, dtype=np.uint32).view(np.float32).item()

_MASK = np.uint64(0xF)


def _f32_bits(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)


def is_end_of_series(x):
    """Element-wise EOS test (``variable_len.py:29-46``)."""
    x = np.asarray(x)
    return np.logical_and(np.isnan(x), np.bitwise_and(_f32_bits(x), _MASK) == 9)


def is_variable_length(x):
    return bool(is_end_of_series(x).any())


def get_variable_length(x):
    """Index of the first EOS along the last axis, or the full length."""
    e = is_end_of_series(x)
    if e.ndim == 1:
        hit = np.flatnonzero(e)
        return int(hit[0]) if hit.size else e.shape[0]
    if e.ndim > 3:
        raise ValueError(f"Found array with {e.ndim} dimensions but expected <= 3")
    first = np.argmax(e, axis=-1)
    return np.where(e.any(axis=-1), first, e.shape[-1])
