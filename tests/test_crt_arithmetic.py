"""CPU checks of the exactness arguments behind the CRT (Ozaki scheme II)
GEMM kernels, on the arithmetic the GPU code performs (emulated here in
IEEE float32 / float64 with numpy) against exact integer arithmetic:

  - the symmetric residue of a weighted byte sum s < 2^20 modulo each odd
    modulus, computed with the 2^23 bit-pattern accumulator, the fused
    quotient and the 1.5 * 2^23 remainder (crt_residue_kernel), for every s;
  - the byte weights (2^(8 i) mod m) and offsets that turn a biased 32- or
    64-bit integer into such a sum congruent to it;
  - the explicit-CRT reconstruction (crt_combine_kernel): 32-bit limbs of
    w_t and P, exact FP64 partial sums, q = rint(S / P), and C rounded once
    (two limbs) or summed in double-double (more limbs), returning C
    correctly rounded for |C| <= 0.4965 P, including the range ends.

The GPU tests check the kernels end to end; these pin the numerical claims.
"""
import math
import random

import numpy as np
import pytest

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211,
          199, 197, 193, 191, 181, 179, 173]
ODD = [m for m in MODULI if m % 2]
F32, F64 = np.float32, np.float64


def _fma32(a, b, c):
    """float32 fma emulated exactly: the float64 product of two float32
    values is exact and, for the magnitudes used here (< 2^25, steps of at
    least 2^-40), so is the sum before the one rounding to float32."""
    return (a.astype(F64) * b.astype(F64) + c.astype(F64)).astype(F32)


def test_moduli_pairwise_coprime():
    for i, a in enumerate(MODULI):
        for b in MODULI[i + 1:]:
            assert math.gcd(a, b) == 1, (a, b)


@pytest.mark.parametrize("m", ODD)
def test_residue_of_byte_sum_exhaustive(m):
    s = np.arange(1 << 20, dtype=np.int64)
    acc = (s + 0x4B000000).astype(np.uint32)          # bits of 2^23 + s
    sf = acc.view(F32)
    inv = F32(1.0) / F32(m)
    c23 = F32(8388608.0) * inv                        # exact: power of two
    qf = np.rint(_fma32(sf, np.full_like(sf, inv), np.full_like(sf, -c23)))
    sf2 = (acc | np.uint32(0x400000)).view(F32)       # 1.5 * 2^23 + s
    rf = _fma32(-qf, np.full_like(qf, F32(m)), sf2)
    byte = rf.view(np.uint32) & 0xFF
    r = (s + (m - 1) // 2) % m - (m - 1) // 2         # symmetric residue
    assert np.array_equal(byte.astype(np.int64), r & 0xFF)
    # the quotient is the direct rint(fl(s * (1/m)))
    q_direct = np.rint((s.astype(F32) * inv).astype(F32))
    assert np.array_equal(qf, q_direct)


def _byte_weights(m, shift):
    return [(1 << (shift + 8 * i)) % m for i in range(4)]


@pytest.mark.parametrize("m", ODD)
def test_byte_weights_give_congruent_small_sum(m):
    rng = random.Random(m)
    for wide, bias in ((False, 30), (True, 62)):
        k = (m - (1 << bias) % m) % m
        for _ in range(2000):
            lim = 1 << bias
            v = rng.randint(-lim, lim)
            if wide:
                lo = (v + (1 << 62)) & 0xFFFFFFFF        # v mod 2^32, biased
                hi = ((v + (1 << 62)) >> 32)              # floor(v/2^32) + 2^30
                words = [(lo, 0), (hi, 32)]
            else:
                words = [(v + (1 << 30), 0)]
            s = k
            for w, sh in words:
                wb = _byte_weights(m, sh)
                s += sum(((w >> (8 * i)) & 0xFF) * wb[i] for i in range(4))
            assert s < (1 << 20)
            assert (s - v) % m == 0


def _recon_constants(n):
    P = 1
    for m in MODULI[:n]:
        P *= m
    npart = (P.bit_length() + 31) // 32
    w = []
    for m in MODULI[:n]:
        Mt = P // m
        w.append(Mt * pow(Mt % m, -1, m))
    limbs = lambda x: [float((x >> (32 * j)) & 0xFFFFFFFF) for j in range(npart)]
    f = [float(F64(2.0) ** (32 * j) / F64(float(P)))
         for j in range(npart)]
    return P, npart, [limbs(x) for x in w], limbs(P), f


def _two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _crt_value(u, n, consts):
    """The reconstruction as the kernel computes it, in float64."""
    P, npart, w, p, f = consts
    S = [0.0] * npart
    for t in range(n):
        for j in range(npart):
            S[j] = S[j] + float(u[t]) * w[t][j]      # exact: < 2^46
    qf = 0.0
    for j in range(npart):
        qf = math.fsum([S[j] * f[j], qf])             # within 2^-38 of S/P
    q = float(np.rint(qf))
    D = [S[j] - q * p[j] for j in range(npart)]       # exact
    if npart == 2:
        return math.fsum([D[1] * 4294967296.0, D[0]])  # one rounding (fma)
    hi, lo = D[npart - 1], 0.0
    for j in range(npart - 2, -1, -1):
        hi *= 4294967296.0
        lo *= 4294967296.0
        sm, e = _two_sum(hi, D[j])
        l = lo + e
        hi = sm + l
        lo = l - (hi - sm)
    return hi + lo


@pytest.mark.parametrize("n", [7, 8, 12, 16, 20])
def test_explicit_crt_reconstruction_correctly_rounded(n):
    consts = _recon_constants(n)
    P = consts[0]
    lim = int(P * 0.4965)
    rng = random.Random(n)
    cases = [0, 1, -1, lim, -lim, lim - 1, -(lim - 1)]
    cases += [rng.randint(-lim, lim) for _ in range(1500)]
    cases += [rng.randint(-(1 << 40), 1 << 40) for _ in range(200)]
    for C in cases:
        u = [C % m for m in MODULI[:n]]
        got = _crt_value(u, n, consts)
        assert got == float(C), (n, C, got)
