"""Host restatement of the device counter-hash fill (csrc/stream.hip).

``uniform(seed, index, dtype)`` reproduces ``sb_stream_fill_random``
element by element: the splitmix64 finaliser of ``seed + (index + 1) *
0x9E3779B97F4A7C15`` scaled to [0, 1) with 53 (float64) or 24 (float32)
mantissa bits.  Integer arithmetic is exact, so values match bit for bit.
"""

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def mix64(seed, index):
    index = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (index + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed, index, dtype=np.float64):
    z = mix64(seed, index)
    if np.dtype(dtype) == np.float64:
        return (z >> np.uint64(11)).astype(np.float64) * 2.0**-53
    return ((z >> np.uint64(40)).astype(np.float32) * np.float32(2.0**-24)).astype(np.float32)
