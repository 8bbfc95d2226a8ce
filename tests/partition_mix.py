"""murmur3's 32-bit finaliser, the community mix of the hub and giant kernels,
and its inverse (numpy, uint64 arrays holding 32-bit values).

The kernels cut a row's communities into partitions by the mix's top bits
and a partition into pieces by its lower bits; a piece that overflows an LDS
table is halved until it fits, which ends because the mix is a bijection (a
piece of one mix value holds one community).  The inverse pins that.
"""
import numpy as np

M32 = 0xFFFFFFFF
C1, C2 = 0x85EBCA6B, 0xC2B2AE35


def fmix32(x):
    x = np.asarray(x, np.uint64) & M32
    x = x ^ (x >> 16)
    x = (x * C1) & M32
    x = x ^ (x >> 13)
    x = (x * C2) & M32
    x = x ^ (x >> 16)
    return x


def _unxorshift(y, s):
    x = y.copy()
    k = s
    while k < 32:           # x ^ (x >> s) undone by folding in y >> s, y >> 2s, ...
        x = x ^ (y >> k)
        k += s
    return x


def fmix32_inverse(h):
    y = np.asarray(h, np.uint64) & M32
    y = _unxorshift(y, 16)
    y = (y * pow(C2, -1, 1 << 32)) & M32
    y = _unxorshift(y, 13)
    y = (y * pow(C1, -1, 1 << 32)) & M32
    y = _unxorshift(y, 16)
    return y
