"""Host-side MRG32k3a helpers in exact integer arithmetic.

L'Ecuyer's combined multiple-recursive generator as vendored by CaSim
(rngstream.cpp): two order-3 recurrences modulo m1 = 2^32 - 209 and
m2 = 2^32 - 22853; streams start 2^127 steps apart, substreams 2^76 apart.
Jump matrices are computed here by repeated squaring rather than tabulated.
RandU01 of the vendored copy uses increased precision (rngstream.cpp:306),
so one draw consumes two generator steps (U01d, rngstream.cpp:271-283).
"""
import numpy as np

M1 = 4294967087
M2 = 4294944443
A1 = ((0, 1, 0), (0, 0, 1), (M1 - 810728, 1403580, 0))
A2 = ((0, 1, 0), (0, 0, 1), (M2 - 1370589, 0, 527612))
NORM = 1.0 / (4294967087.0 + 1.0)
FACT = 5.9604644775390625e-8
DEFAULT_SEED = (12345,) * 6  # RngStream::nextSeed default (rngstream.cpp:294-297)


def matmul(a, b, m):
    return tuple(tuple(sum(a[i][k] * b[k][j] for k in range(3)) % m for j in range(3)) for i in range(3))


def matvec(a, s, m):
    return tuple(sum(a[i][k] * s[k] for k in range(3)) % m for i in range(3))


def matpow(a, n, m):
    r = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    while n:
        if n & 1:
            r = matmul(a, r, m)
        a = matmul(a, a, m)
        n >>= 1
    return r


_J127 = (matpow(A1, 1 << 127, M1), matpow(A2, 1 << 127, M2))
_J76 = (matpow(A1, 1 << 76, M1), matpow(A2, 1 << 76, M2))


def check_seed(seed):
    """CheckSeed (rngstream.cpp:200-234)."""
    s = [int(v) for v in seed]
    if len(s) != 6 or any(v < 0 for v in s):
        return False
    if any(v >= M1 for v in s[:3]) or any(v >= M2 for v in s[3:]):
        return False
    return any(s[:3]) and any(s[3:])


def stream_states(seed=DEFAULT_SEED, n=1, first=0):
    """Start states of streams first..first+n-1 after SetPackageSeed(seed)."""
    if not check_seed(seed):
        raise ValueError("invalid MRG32k3a seed %r" % (seed,))
    s1, s2 = tuple(int(v) for v in seed[:3]), tuple(int(v) for v in seed[3:])
    if first:
        s1 = matvec(matpow(_J127[0], first, M1), s1, M1)
        s2 = matvec(matpow(_J127[1], first, M2), s2, M2)
    out = np.zeros((n, 6), dtype=np.uint64)
    for k in range(n):
        out[k, :3] = s1
        out[k, 3:] = s2
        s1 = matvec(_J127[0], s1, M1)
        s2 = matvec(_J127[1], s2, M2)
    return out


def next_substream(state):
    """ResetNextSubstream applied to a stream start (rngstream.cpp:347-353)."""
    s1 = matvec(_J76[0], tuple(int(v) for v in state[:3]), M1)
    s2 = matvec(_J76[1], tuple(int(v) for v in state[3:]), M2)
    return np.array(s1 + s2, dtype=np.uint64)


class Stream:
    """One MRG32k3a stream drawing RandU01 (= U01d) values."""

    def __init__(self, state):
        self.s = [int(v) for v in state]

    def _u01(self):
        s = self.s
        p1 = (1403580 * s[1] - 810728 * s[0]) % M1
        s[0], s[1], s[2] = s[1], s[2], p1
        p2 = (527612 * s[5] - 1370589 * s[3]) % M2
        s[3], s[4], s[5] = s[4], s[5], p2
        d = p1 - p2 if p1 > p2 else p1 - p2 + M1
        return float(d) * NORM

    def rand_u01(self):
        u = self._u01()
        u += self._u01() * FACT
        return u if u < 1.0 else u - 1.0

    def state(self):
        return np.array(self.s, dtype=np.uint64)
