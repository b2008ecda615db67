"""Philox4x32-10 and the noise samplers, restated on the CPU (test infrastructure only).

Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11.  Round function (per round r, key bumped by the Weyl constants
between rounds):
    (hi0, lo0) = M0 * c0 ; (hi1, lo1) = M1 * c2
    c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0)
with M0 = 0xD2511F53, M1 = 0xCD9E8D57, W0 = 0x9E3779B9, W1 = 0xBB67AE85.

Counter layout used by the noise stage of the SH WFS (one stream per element):
    ctr = (element index, block index j, call counter, domain), key = (seed lo, seed hi)
domain 0: Poisson photon noise of the spots, 1: Gaussian readout noise of the
spots, 2: Gaussian readout noise of the detector background.
This replaces the reference's per-thread XORWOW curandState
(libcarma/src.cu/carma_rng.cu:35-50,98-141,287-303); only the *semantics*
(bincube = Poisson(bincube); bincube += sigma*N(0,1); binimg = sigma*N(0,1)) are
the reference's, the stream is not bit-comparable to cuRAND.

Samplers (float64 arithmetic, identical operation order on the GPU):
* uniform(x) = (x + 0.5) * 2^-32 in (0, 1)
* normal: Box-Muller, z = sqrt(-2 ln u0) * cos(2 pi u1), block 0 words 0, 1
* poisson(lam): lam <= 0 -> 0; lam < 10 -> sequential-search inversion on
  u = word 0 of block 0 (at most 200 terms); lam >= 10 -> PTRS, the transformed
  rejection method with squeeze of W. Hoermann, "The transformed rejection
  method for generating Poisson random variables", Insurance: Mathematics and
  Economics 12 (1993) 39-45; attempt t uses words (2*(t%2), 2*(t%2)+1) of block
  t//2.
"""
import math

import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32(ctr, key, rounds=10):
    """Vectorised Philox4x32: ctr (..., 4) uint32, key (..., 2) uint32 -> (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k0 = np.asarray(key[..., 0], dtype=np.uint64)
    k1 = np.asarray(key[..., 1], dtype=np.uint64)
    for r in range(rounds):
        if r > 0:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = c[0] * np.uint64(M0)
        p1 = c[2] * np.uint64(M1)
        hi0, lo0 = p0 >> np.uint64(32), p0 & np.uint64(MASK)
        hi1, lo1 = p1 >> np.uint64(32), p1 & np.uint64(MASK)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def block(elem, j, call, domain, seed):
    elem = np.asarray(elem, dtype=np.uint64)
    ctr = np.stack(np.broadcast_arrays(elem & MASK, np.uint64(j), np.uint64(call & MASK),
                                       np.uint64(domain)), axis=-1)
    key = np.broadcast_to(np.array([seed & MASK, (seed >> 32) & MASK], dtype=np.uint64),
                          ctr.shape[:-1] + (2,))
    return philox4x32(ctr, key)


def uniform(words):
    return (np.asarray(words, dtype=np.float64) + 0.5) * (1.0 / 4294967296.0)


def normal(elem, call, domain, seed):
    w = block(elem, 0, call, domain, seed)
    u0 = uniform(w[..., 0])
    u1 = uniform(w[..., 1])
    return np.sqrt(-2.0 * np.log(u0)) * np.cos(2.0 * math.pi * u1)


def _poisson_one(lam, elem, call, domain, seed):
    if not lam > 0.0:
        return 0.0
    if lam < 10.0:
        u = uniform(block(elem, 0, call, domain, seed)[0])
        p = math.exp(-lam)
        F = p
        k = 0
        while u > F and k < 200:
            k += 1
            p = p * lam / k
            F = F + p
        return float(k)
    slam = math.sqrt(lam)
    loglam = math.log(lam)
    b = 0.931 + 2.53 * slam
    a = -0.059 + 0.02483 * b
    invalpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    t = 0
    w = None
    while t < 4096:            # same bound as the device loop; never reached in practice
        if t % 2 == 0:
            w = block(elem, t // 2, call, domain, seed)
        U = uniform(w[2 * (t % 2)]) - 0.5
        V = uniform(w[2 * (t % 2) + 1])
        t += 1
        us = 0.5 - abs(U)
        k = math.floor((2.0 * a / us + b) * U + lam + 0.43)
        if us >= 0.07 and V <= vr:
            return float(k)
        if k < 0 or (us < 0.013 and V > us):
            continue
        if (math.log(V) + math.log(invalpha) - math.log(a / (us * us) + b)
                <= -lam + k * loglam - math.lgamma(k + 1.0)):
            return float(k)
    return float(math.floor(lam))


def poisson(lam, elem0, call, domain, seed):
    """Element-wise Poisson draws for a flat array ``lam`` with element indices elem0+i."""
    lam = np.asarray(lam, dtype=np.float64).ravel()
    return np.array([_poisson_one(float(l), elem0 + i, call, domain, seed)
                     for i, l in enumerate(lam)])
