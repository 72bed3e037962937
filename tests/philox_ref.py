"""numpy Philox4x32-10 (Salmon et al., SC 2011) for checking the device
generator: the round function, key schedule and the 53-bit uniform the
device uses, written from the paper."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr [n, 4], key [n, 2] (uint32 values) -> [n, 4] uint32."""
    c = [np.asarray(ctr[:, j], np.uint64) for j in range(4)]
    k0 = np.asarray(key[:, 0], np.uint64)
    k1 = np.asarray(key[:, 1], np.uint64)
    for r in range(10):
        if r:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=1).astype(np.uint32)


def counters(first, n, w2, w3):
    i = np.arange(n, dtype=np.uint64) + np.uint64(first)
    return np.stack([i & MASK, i >> np.uint64(32), np.full(n, w2, np.uint64),
                     np.full(n, w3, np.uint64)], axis=1)


def keys(seed, n):
    s = np.uint64(seed)
    return np.tile(np.array([s & MASK, s >> np.uint64(32)], np.uint64), (n, 1))


def u53(a, b):
    return ((a >> np.uint32(5)).astype(np.float64) * 67108864.0
            + (b >> np.uint32(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def uniform_positions(n, lo, hi, seed, first=0):
    """Particles first .. first+n-1 of stream seed, as the device draws them."""
    a = philox4x32_10(counters(first, n, 0, 0), keys(seed, n))
    b = philox4x32_10(counters(first, n, 1, 0), keys(seed, n))
    lo = np.asarray(lo, np.float64)
    span = np.asarray(hi, np.float64) - lo
    return np.stack([lo[0] + u53(a[:, 0], a[:, 1]) * span[0],
                     lo[1] + u53(a[:, 2], a[:, 3]) * span[1],
                     lo[2] + u53(b[:, 0], b[:, 1]) * span[2]], axis=1)
