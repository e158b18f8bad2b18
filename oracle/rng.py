"""Restatement of ``Matrix::Random`` (include/matrix.h:50-57) — oracle only.

The reference fills a fresh ``Matrix(r, c)`` in storage order (column-major,
``data_[j*rows_+i]``, matrix.h:36) from a thread-local ``std::mt19937`` seeded
by ``std::random_device`` through ``std::uniform_real_distribution<double>
(0.0, 1.0)``.  With libstdc++ that distribution is
``std::generate_canonical<double, 53>``: two engine calls g1, g2 and

    u = (double(g1) + double(g2) * 2**32) / 2**64,   u := nextafter(1, 0) if u >= 1

(the sum is rounded to double before the exact power-of-two division).

``std::random_device`` makes the reference nondeterministic; parity runs fix
the seed here and hand the recorded stream to the engine under test.
"""
import numpy as np

_N, _M = 624, 397
_MATRIX_A = np.uint32(0x9908B0DF)
_UPPER = np.uint32(0x80000000)
_LOWER = np.uint32(0x7FFFFFFF)


class MT19937:
    """std::mt19937 (32-bit Mersenne Twister, C++ [rand.predef])."""

    def __init__(self, seed=5489):
        mt = np.zeros(_N, dtype=np.uint64)
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, _N):
            prev = int(mt[i - 1])
            mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF
        self._mt = mt.astype(np.uint32)
        self._idx = _N
        self._buf = None

    def _twist(self):
        mt = self._mt
        # Element i uses mt[i+1] and mt[(i+397) % 624]; chunks of 227 rows
        # only ever read rows already final for this generation.
        for lo, hi in ((0, 227), (227, 454), (454, 623)):
            i = np.arange(lo, hi)
            y = (mt[i] & _UPPER) | (mt[i + 1] & _LOWER)
            mag = np.where((y & 1) == 1, _MATRIX_A, np.uint32(0))
            mt[i] = mt[(i + _M) % _N] ^ (y >> np.uint32(1)) ^ mag
        y = (mt[623] & _UPPER) | (mt[0] & _LOWER)
        mt[623] = mt[_M - 1] ^ (y >> np.uint32(1)) ^ (_MATRIX_A if (y & 1) else np.uint32(0))
        # tempering for the whole generation at once
        t = mt.copy()
        t ^= t >> np.uint32(11)
        t ^= (t << np.uint32(7)) & np.uint32(0x9D2C5680)
        t ^= (t << np.uint32(15)) & np.uint32(0xEFC60000)
        t ^= t >> np.uint32(18)
        self._buf = t
        self._idx = 0

    def integers(self, count):
        """Next ``count`` raw 32-bit outputs as a uint64 array."""
        out = np.empty(count, dtype=np.uint64)
        pos = 0
        while pos < count:
            if self._idx >= _N:
                self._twist()
            take = min(count - pos, _N - self._idx)
            out[pos:pos + take] = self._buf[self._idx:self._idx + take]
            self._idx += take
            pos += take
        return out


def canonical(gen, count):
    """``count`` draws of uniform_real_distribution<double>(0,1) (libstdc++)."""
    raw = gen.integers(2 * count).reshape(count, 2)
    s = raw[:, 0].astype(np.float64) + raw[:, 1].astype(np.float64) * 4294967296.0
    u = s / 18446744073709551616.0
    u[u >= 1.0] = np.nextafter(1.0, 0.0)
    return u


class RandomStream:
    """The reference's ``Matrix::Random`` on one thread's engine.

    ``random(r, c)`` returns an r x c Fortran-ordered array filled in storage
    order, exactly as matrix.h:55 walks ``data_``.  ``log`` keeps every block
    drawn (in draw order) so the same Omega can be handed to the GPU engine.
    """

    def __init__(self, seed, record=False):
        self.gen = MT19937(seed)
        self.log = [] if record else None

    def random(self, r, c):
        vals = canonical(self.gen, r * c)
        m = vals.reshape((c, r)).T.copy(order="F")
        if self.log is not None:
            self.log.append(m)
        return m

    def flat_log(self):
        """All recorded blocks, each column-major, concatenated."""
        if not self.log:
            return np.zeros(0)
        return np.concatenate([b.reshape(-1, order="F") for b in self.log])


class ReplayStream:
    """Hands back a recorded Omega sequence in the same order."""

    def __init__(self, blocks):
        self.blocks = list(blocks)
        self.pos = 0

    def random(self, r, c):
        b = self.blocks[self.pos]
        self.pos += 1
        assert b.shape == (r, c), (b.shape, (r, c))
        return b
