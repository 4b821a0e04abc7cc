"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC'11), vectorised in NumPy.

Test infrastructure (oracle).  The reference draws from NumPy's legacy global
MT19937 (``np.random.*`` in ``Python/physics_engine.py:1018-1099``,
``:1400-1416``, ``:1719-1734``); the MI355X backend replaces that with a
counter-based stream so every draw is a pure function of
(seed, purpose, particle id, step, block) and therefore independent of thread
order.  This module is the CPU restatement used to check the HIP generator.

Pinned by: the published Random123 known-answer vectors and by rocRAND's
``philox4x32_10_engine`` (tests/golden/philox_kat.json, regenerated by
tests/golden/make_philox_kat.py).
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)


def philox4x32_10(ctr, key):
    """Return Philox4x32-10 blocks.

    ctr : uint32 array of shape (..., 4)
    key : uint32 array broadcastable to (..., 2)
    returns uint32 array of shape (..., 4)
    """
    ctr = np.asarray(ctr, dtype=np.uint32)
    key = np.asarray(key, dtype=np.uint32)
    c0 = ctr[..., 0].astype(np.uint64)
    c1 = ctr[..., 1].astype(np.uint64)
    c2 = ctr[..., 2].astype(np.uint64)
    c3 = ctr[..., 3].astype(np.uint64)
    k0 = np.broadcast_to(key[..., 0].astype(np.uint64), c0.shape)
    k1 = np.broadcast_to(key[..., 1].astype(np.uint64), c0.shape)
    for r in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> _S32, p0 & _MASK
        hi1, lo1 = p1 >> _S32, p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        if r < 9:
            k0 = (k0 + W0) & _MASK
            k1 = (k1 + W1) & _MASK
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def u01_pair(block):
    """Two f64 uniforms in [0, 1) from one block: 53 high bits of the 64-bit
    words (w1:w0) and (w3:w2).  Exact in f64, identical on the GPU."""
    b = block.astype(np.uint64)
    a = (b[..., 1] << _S32) | b[..., 0]
    c = (b[..., 3] << _S32) | b[..., 2]
    scale = 2.0 ** -53
    return ((a >> np.uint64(11)).astype(np.float64) * scale,
            (c >> np.uint64(11)).astype(np.float64) * scale)
