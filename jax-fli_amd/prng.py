"""JAX's counter-based generator on the host, for the few integers the reference draws with it:
OnionTiler's cube symmetries, jax.random.randint(jax.random.PRNGKey(42), shape=(27,), minval=0,
maxval=47) (src/jax_fli/pm/interp.py:334-335).

Threefry-2x32, 20 rounds (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC11), with JAX's conventions for jax_threefry_partitionable=True, the default since
JAX 0.5 (the reference requires jax >= 0.6):
  key(seed)         the word pair (seed >> 32, seed & 0xFFFFFFFF)
  split(key, n)[i]  threefry(key, (hi(i), lo(i))), i the flat index as a 64-bit count
  bits32(key, n)[i] y0 ^ y1 of threefry(key, (hi(i), lo(i))); bits64 is y0 << 32 | y1
  randint           k1, k2 = split(key, 2); high, low = bits(k1), bits(k2) of the dtype's width w;
                    offset = ((high % span) * m + low % span) % span with m = (2^(w/2) % span)^2 % span,
                    i.e. the 2w-bit number high * 2^w + low reduced modulo span = maxval - minval.
The reference's simulation path runs with jax_enable_x64 off (src/jax_fli/scripts/fli_simulate.py:530),
so randint draws int32 (w = 32); x64=True gives the int64 draw (w = 64).
"""

from __future__ import annotations

import numpy as np

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = np.uint32(0x1BD11BDA)


def _rotl(x: np.ndarray, r: int) -> np.ndarray:
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def threefry2x32(key, x0, x1):
    """Threefry-2x32 (20 rounds) of the counter words (x0, x1) under key = (k0, k1); uint32 arrays."""
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    ks = (k0, k1, k0 ^ k1 ^ _PARITY)
    x0 = np.asarray(x0, dtype=np.uint32).copy()
    x1 = np.asarray(x1, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        x0 += ks[0]
        x1 += ks[1]
        for g in range(5):  # five groups of four rounds, a key injection after each
            for r in _ROT[g % 2]:
                x0 += x1
                x1 = _rotl(x1, r) ^ x0
            x0 += ks[(g + 1) % 3]
            x1 += ks[(g + 2) % 3] + np.uint32(g + 1)
    return x0, x1


def prng_key(seed: int):
    """jax.random.PRNGKey(seed) as its two uint32 words."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return (np.uint32(s >> 32), np.uint32(s & 0xFFFFFFFF))


def _counts(n: int):
    i = np.arange(n, dtype=np.uint64)
    return (i >> np.uint64(32)).astype(np.uint32), (i & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def split(key, n: int = 2):
    """jax.random.split(key, n) (partitionable): n keys as an (n, 2) uint32 array."""
    y0, y1 = threefry2x32(key, *_counts(n))
    return np.stack([y0, y1], axis=1)


def random_bits(key, n: int, width: int = 32) -> np.ndarray:
    """n random words of `width` (32 or 64) bits (partitionable)."""
    y0, y1 = threefry2x32(key, *_counts(n))
    if width == 32:
        return y0 ^ y1
    if width == 64:
        return (y0.astype(np.uint64) << np.uint64(32)) | y1.astype(np.uint64)
    raise ValueError("width must be 32 or 64")


def randint(key, n: int, minval: int, maxval: int, x64: bool = False) -> np.ndarray:
    """jax.random.randint(key, (n,), minval, maxval) for int32 (x64=False) or int64 (x64=True), with
    minval < maxval inside the dtype's range."""
    if not minval < maxval:
        raise ValueError("randint needs minval < maxval")
    width = 64 if x64 else 32
    udt = np.uint64 if x64 else np.uint32
    k1, k2 = split(key, 2)
    high, low = random_bits(k1, n, width), random_bits(k2, n, width)
    span = udt(maxval - minval)
    m = udt(pow(2, width // 2, int(span)))
    m = udt((int(m) * int(m)) % int(span))
    with np.errstate(over="ignore"):
        off = ((high % span) * m + low % span) % span
    return (np.int64(minval) + off.astype(np.int64)).astype(np.int64 if x64 else np.int32)
