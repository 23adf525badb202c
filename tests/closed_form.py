"""Closed-form group-by answers for the generators TGA_GEN_MOD_U32 /
TGA_GEN_MOD_U64 (keys) and TGA_GEN_SEQ_I32 (values).

Row i has key mix(i mod D) and value lo + i, where mix is a bijection (the
fmix32 finalizer for 32-bit keys, the SplitMix64 finalizer for 64-bit ones,
of (i mod D) ^ salt).  Group k (k < min(n, D)) therefore holds

    c_k   = floor((n - 1 - k) / D) + 1           rows
    SUM_k = c_k (lo + k) + D c_k (c_k - 1) / 2
    MIN_k = lo + k
    MAX_k = lo + k + D (c_k - 1)

(values as int32 while lo + n - 1 stays within it), so every group of a
full-size run is checked without aggregating the rows.
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z):
    """SplitMix64 finalizer (Steele, Lea & Flood 2014) on a uint64 array."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def fmix32(h):
    """MurmurHash3 fmix32 on a uint32 array."""
    h = np.asarray(h, dtype=np.uint32)
    with np.errstate(over="ignore"):
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x85EBCA6B)
        h = h ^ (h >> np.uint32(13))
        h = h * np.uint32(0xC2B2AE35)
    return h ^ (h >> np.uint32(16))


def keys_of(k, bits, salt):
    """Key column value of group index k."""
    k = np.asarray(k, dtype=np.uint64)
    if bits == 32:
        return fmix32((k & np.uint64(0xFFFFFFFF)).astype(np.uint32) ^ np.uint32(salt & 0xFFFFFFFF))
    return mix64(k ^ np.uint64(salt))


def expected(n, D, bits=64, salt=0, lo=0):
    """Sorted (keys, [COUNT, SUM, MIN, MAX]) of the closed-form workload."""
    G = min(n, D)
    k = np.arange(G, dtype=np.int64)
    c = (n - 1 - k) // D + 1
    base = lo + k
    cnt = c.astype(np.int64)
    sm = c * base + D * c * (c - 1) // 2
    mn = base.astype(np.int32)
    mx = (base + D * (c - 1)).astype(np.int32)
    keys = keys_of(k, bits, salt)
    order = np.argsort(keys, kind="stable")
    return keys[order], [cnt[order], sm[order], mn[order], mx[order]]


def columns(n, D, bits=64, salt=0, lo=0):
    """The host columns themselves (small n only)."""
    i = np.arange(n, dtype=np.uint64)
    keys = keys_of(i % np.uint64(D), bits, salt)
    vals = (np.uint64(lo & 0xFFFFFFFFFFFFFFFF) + i).astype(np.uint32).view(np.int32)
    return keys, vals
