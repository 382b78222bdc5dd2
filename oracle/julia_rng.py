"""Oracle (test infrastructure only): Julia's ``rand(MersenneTwister(seed), n, n)``.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this module, and only as a checker.

Restates the noise source of ``core_shell.jl:32``
(``noise = 0.01 .* (2 .* rand(MersenneTwister(seed), n, n) .- 1)``).  The RNG is
not in the reference; it is Julia 1.12.6's stdlib ``Random`` (``Manifest.toml:3``),
whose published algorithm is restated here:

* seeding (Julia >= 1.11): the integer seed is hashed with SHA-256 over its
  little-endian 32-bit limbs (least significant first; a non-negative seed
  that fits one limb hashes exactly 4 bytes); the 32-byte digest, read as 8
  little-endian ``UInt32`` words, is the key of dSFMT's ``init_by_array``;
* generator: dSFMT 2.2 with Mersenne exponent 19937 (Saito & Matsumoto),
  parameters POS1=117, SL1=19, SR=12, MSK1/MSK2, FIX1/FIX2, PCV1/PCV2 below;
* array fill (``rand!(::MersenneTwister, ::Array{Float64})``): the first
  ``n2 = (len-2)÷2·2`` values come from one ``dsfmt_fill_array_close_open``
  call on the raw state (when ``n2 >= 382``; below that, from the cache), the
  remaining 1 or 2 values are scalar draws from a freshly generated 1002-value
  cache, which is filled by ``dsfmt_fill_array_close1_open2`` on the state left
  by the bulk fill; every value is in [1,2) minus 1.0.

Pinned by: F(t=0) of ``check-results.txt:2`` (64², R=16, w=4) and ``:7``
(256², R=70, w=4), which depend on the noise field at ~1e-5 relative, so a
match to 13+ digits pins this restatement bit-for-bit (tests/test_oracle.py).
This pure-Python version is meant for N <= 1024.
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np

M64 = (1 << 64) - 1
M32 = (1 << 32) - 1

DSFMT_MEXP = 19937
DSFMT_N = (DSFMT_MEXP - 128) // 104 + 1          # 191 128-bit words
DSFMT_N64 = DSFMT_N * 2                          # 382 doubles per block
POS1 = 117
SL1 = 19
SR = 12
MSK1 = 0x000FFAFFFFFFFB3F
MSK2 = 0x000FFDFFFC90FFFD
FIX1 = 0x90014964B32F4329
FIX2 = 0x3B8D12AC548A7C7A
PCV1 = 0x3D84E1AC0DC82880
PCV2 = 0x0000000000000001
LOW_MASK = 0x000FFFFFFFFFFFFF
HIGH_CONST = 0x3FF0000000000000
MT_CACHE_F = 1002                                # Julia's MersenneTwister float cache


def hash_seed(seed: int) -> bytes:
    """SHA-256 of the seed's 32-bit limbs (Julia >= 1.11 ``Random.hash_seed``)."""
    ctx = hashlib.sha256()
    neg = seed < 0
    if neg:
        seed = ~seed
    while True:
        ctx.update(struct.pack("<I", seed & M32))
        seed >>= 32
        if seed == 0:
            break
    if neg:
        ctx.update(b"\x01")
    return ctx.digest()


def _ini1(x):
    return ((x ^ (x >> 27)) * 1664525) & M32


def _ini2(x):
    return ((x ^ (x >> 27)) * 1566083941) & M32


class DSFMT:
    """dSFMT-19937 state: ``st`` holds DSFMT_N+1 words of two uint64 each."""

    def __init__(self, key_u32):
        size = (DSFMT_N + 1) * 4
        lag = 11
        mid = (size - lag) // 2
        s = [0x8B8B8B8B] * size
        key_length = len(key_u32)
        count = max(key_length + 1, size)
        r = _ini1(s[0] ^ s[mid % size] ^ s[(size - 1) % size])
        s[mid % size] = (s[mid % size] + r) & M32
        r = (r + key_length) & M32
        s[(mid + lag) % size] = (s[(mid + lag) % size] + r) & M32
        s[0] = r
        count -= 1
        i = 1
        j = 0
        while j < count and j < key_length:
            r = _ini1(s[i] ^ s[(i + mid) % size] ^ s[(i + size - 1) % size])
            s[(i + mid) % size] = (s[(i + mid) % size] + r) & M32
            r = (r + key_u32[j] + i) & M32
            s[(i + mid + lag) % size] = (s[(i + mid + lag) % size] + r) & M32
            s[i] = r
            i = (i + 1) % size
            j += 1
        while j < count:
            r = _ini1(s[i] ^ s[(i + mid) % size] ^ s[(i + size - 1) % size])
            s[(i + mid) % size] = (s[(i + mid) % size] + r) & M32
            r = (r + i) & M32
            s[(i + mid + lag) % size] = (s[(i + mid + lag) % size] + r) & M32
            s[i] = r
            i = (i + 1) % size
            j += 1
        for _ in range(size):
            r = _ini2((s[i] + s[(i + mid) % size] + s[(i + size - 1) % size]) & M32)
            s[(i + mid) % size] ^= r
            r = (r - i) & M32
            s[(i + mid + lag) % size] ^= r
            s[i] = r
            i = (i + 1) % size
        # 32-bit words -> 64-bit words (little endian), pairs -> 128-bit words
        u64 = [s[2 * k] | (s[2 * k + 1] << 32) for k in range(size // 2)]
        for k in range(DSFMT_N * 2):               # initial_mask
            u64[k] = (u64[k] & LOW_MASK) | HIGH_CONST
        self.st = [[u64[2 * k], u64[2 * k + 1]] for k in range(DSFMT_N + 1)]
        self._period_certification()

    def _period_certification(self):
        lung = self.st[DSFMT_N]
        t0 = lung[0] ^ FIX1
        t1 = lung[1] ^ FIX2
        inner = (t0 & PCV1) ^ (t1 & PCV2)
        i = 32
        while i > 0:
            inner ^= inner >> i
            i >>= 1
        if inner & 1:
            return
        lung[1] ^= 1

    @staticmethod
    def _rec(a, b, lung):
        """One dSFMT recursion; updates ``lung`` in place, returns the new word."""
        t0, t1 = a
        L0, L1 = lung
        n0 = ((t0 << SL1) & M64) ^ (L1 >> 32) ^ ((L1 << 32) & M64) ^ b[0]
        n1 = ((t1 << SL1) & M64) ^ (L0 >> 32) ^ ((L0 << 32) & M64) ^ b[1]
        lung[0] = n0
        lung[1] = n1
        return [(n0 >> SR) ^ (n0 & MSK1) ^ t0, (n1 >> SR) ^ (n1 & MSK2) ^ t1]

    def fill_c1o2(self, size_words: int):
        """``gen_rand_array_c1o2``: returns ``size_words`` 128-bit words (2 doubles
        each, as uint64 bit patterns in [1,2)) and advances the state."""
        N = DSFMT_N
        assert size_words >= N
        st = self.st
        lung = list(st[N])
        arr = [None] * size_words
        rec = self._rec
        arr[0] = rec(st[0], st[POS1], lung)
        i = 1
        while i < N - POS1:
            arr[i] = rec(st[i], st[i + POS1], lung)
            i += 1
        while i < N:
            arr[i] = rec(st[i], arr[i + POS1 - N], lung)
            i += 1
        while i < size_words - N:
            arr[i] = rec(arr[i - N], arr[i + POS1 - N], lung)
            i += 1
        j = 0
        while j < 2 * N - size_words:
            st[j] = list(arr[j + size_words - N])
            j += 1
        while i < size_words:
            arr[i] = rec(arr[i - N], arr[i + POS1 - N], lung)
            st[j] = list(arr[i])
            i += 1
            j += 1
        st[N] = lung
        return arr


def _words_to_f64(words) -> np.ndarray:
    flat = np.empty(2 * len(words), dtype=np.uint64)
    flat[0::2] = [w[0] for w in words]
    flat[1::2] = [w[1] for w in words]
    return flat.view(np.float64)


class MersenneTwister:
    """Julia's ``MersenneTwister(seed)`` for ``rand(rng, Float64, dims...)``."""

    def __init__(self, seed: int):
        digest = hash_seed(seed)
        key = list(struct.unpack("<8I", digest))
        self.state = DSFMT(key)
        self.cache = np.empty(0)
        self.idx = 0

    def _refill_cache(self):
        self.cache = _words_to_f64(self.state.fill_c1o2(MT_CACHE_F // 2))
        self.idx = 0

    def _scalar(self) -> float:
        if self.idx >= len(self.cache):
            self._refill_cache()
        v = self.cache[self.idx]
        self.idx += 1
        return v - 1.0

    def rand(self, *dims) -> np.ndarray:
        """Column-major fill order, returned as a Fortran-ordered array."""
        n = int(np.prod(dims))
        n2 = (n - 2) // 2 * 2
        out = np.empty(n, dtype=np.float64)
        if n2 < DSFMT_N64:
            # small arrays: served element by element from the cache
            for k in range(n):
                out[k] = self._scalar()
        else:
            bulk = _words_to_f64(self.state.fill_c1o2(n2 // 2))
            out[:n2] = bulk - 1.0
            for k in range(n2, n):
                out[k] = self._scalar()
        return out.reshape(dims, order="F")
