"""CPU restatement of the device Gaussian stream (test infrastructure).

Philox-4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplications
(M0 = 0xD2511F53 on word 0, M1 = 0xCD9E8D57 on word 2), outputs
(hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0), key bumped by the Weyl constants
(0x9E3779B9, 0xBB67AE85) between rounds.  Pinned by the published
known-answer vectors (tests/test_rng.py).

The stream the library documents (bd_set_rng): value e of a draw starting
at stream position `offset` is the Box-Muller output of pair
(offset + e) // 2 ... as the device builds it: counter (p_lo, p_hi,
0x243F6A88, 0x85A308D3) with p = (offset + 2 g) >> 1 for pair g, key
(seed_lo, seed_hi); u1 = ((w0 w1 >> 11) + 1) 2^-53, u2 = (w2 w3 >> 11) 2^-53,
z = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2).
"""
import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """Vectorised over the leading axis: ctr (..., 4), key (..., 2) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + np.uint64(W0)) & np.uint64(MASK)
        k1 = (k1 + np.uint64(W1)) & np.uint64(MASK)
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def gaussian_stream(seed, offset, count):
    """count values of the device stream from position offset (offset even,
    as the library keeps it: every draw advances it by an even number)."""
    assert offset % 2 == 0
    npair = (count + 1) // 2
    p = (np.uint64(offset) + 2 * np.arange(npair, dtype=np.uint64)) >> np.uint64(1)
    ctr = np.stack([p & np.uint64(MASK), p >> np.uint64(32), np.full(npair, 0x243F6A88, np.uint64),
                    np.full(npair, 0x85A308D3, np.uint64)], axis=-1)
    key = np.tile(np.array([seed & MASK, seed >> 32], dtype=np.uint64), (npair, 1))
    w = philox4x32_10(ctr, key).astype(np.uint64)
    a = (w[:, 0] << np.uint64(32)) | w[:, 1]
    b = (w[:, 2] << np.uint64(32)) | w[:, 3]
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * npair)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:count]
