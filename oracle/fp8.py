"""CPU restatement of the FP8 low-fidelity complex GEMM (test infrastructure only).

The GPU form (qflex_amd/csrc/qfx_fp8.hip) computes C = A B for complex64
A (m x k) and B (k x n) as the real product A_real [m x 2k] * Bhat [2k x 2n],
with A_real = A's interleaved (re, im) storage and Bhat holding each b as
the block [[br, bi], [-bi, br]].  Both real operands are cut into 128 x 128
tiles along (rows, K) for A_real and (K, columns) for Bhat; each tile is
multiplied by 2^-s with s = e - 8 for the tile's max |x| = f 2^e, f in
[0.5, 1) (s = 0 for an all-zero tile), rounded to the nearest OCP E4M3 value
(ties to even), and the products of the rounded values are summed with the
scales 2^(sA + sB) applied.  This module does the same quantization and
sums in float64; the GPU sums in FP32, which is the only difference.

There is no reference counterpart: the reference computes in FP32
(src/tensor.cpp:960-980, cblas_cgemm).
"""
import numpy as np

TILE = 128
E4M3_MAX = 448.0


def e4m3_values():
    """All 256 OCP E4M3 codes decoded (NaN for 0x7f / 0xff)."""
    out = np.empty(256)
    for c in range(256):
        s, e, m = c >> 7, (c >> 3) & 15, c & 7
        if e == 15 and m == 7:
            v = np.nan
        elif e == 0:
            v = m * 2.0 ** -9
        else:
            v = (1 + m / 8) * 2.0 ** (e - 7)
        out[c] = -v if s else v
    return out


def e4m3_round(x):
    """Round float64 values of magnitude <= 448 to the nearest E4M3 value,
    ties to even (the even code has mantissa bit 0): exponent of the
    binade, quantum 2^(e-3) (2^-9 below 2^-6), np.round (half to even) on
    the scaled value."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    if np.any(a > E4M3_MAX):
        raise ValueError("value beyond E4M3 range")
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6)
    q = 2.0 ** (e - 3)
    return np.sign(x) * np.round(a / q) * q


def tile_shift(amax):
    """s with amax 2^-s < 256: amax = f 2^e, f in [0.5, 1) -> s = e - 8."""
    if not amax > 0:
        return 0
    _, e = np.frexp(amax)
    return int(min(max(e - 8, -127), 127))


def real_a(a):
    """A_real: m x 2k, (re, im) interleaved."""
    a = np.asarray(a, dtype=np.complex128)
    out = np.empty((a.shape[0], 2 * a.shape[1]))
    out[:, 0::2], out[:, 1::2] = a.real, a.imag
    return out


def real_bhat(b):
    """Bhat: 2k x 2n, each b as [[br, bi], [-bi, br]]."""
    b = np.asarray(b, dtype=np.complex128)
    k, n = b.shape
    out = np.empty((2 * k, 2 * n))
    out[0::2, 0::2], out[0::2, 1::2] = b.real, b.imag
    out[1::2, 0::2], out[1::2, 1::2] = -b.imag, b.real
    return out


def quantize(x):
    """Per-128 x 128 tile scaled and E4M3-rounded values, dequantized
    (float64, same shape as x), and the tile shifts."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    shifts = np.zeros(((x.shape[0] + TILE - 1) // TILE, (x.shape[1] + TILE - 1) // TILE), int)
    for ti in range(shifts.shape[0]):
        for tj in range(shifts.shape[1]):
            blk = x[ti * TILE:(ti + 1) * TILE, tj * TILE:(tj + 1) * TILE]
            s = tile_shift(float(np.max(np.abs(blk))) if blk.size else 0.0)
            shifts[ti, tj] = s
            out[ti * TILE:(ti + 1) * TILE, tj * TILE:(tj + 1) * TILE] = \
                e4m3_round(blk * 2.0 ** -s) * 2.0 ** s
    return out, shifts


def cgemm_fp8(a, b):
    """The FP8 form's result in float64 (complex128, m x n)."""
    qa, _ = quantize(real_a(a))
    qb, _ = quantize(real_bhat(b))
    c = qa @ qb
    return c[:, 0::2] + 1j * c[:, 1::2]


def fidelity(ref, got):
    """|<ref|got>|^2 / (|ref|^2 |got|^2)."""
    ref = np.asarray(ref, dtype=np.complex128).ravel()
    got = np.asarray(got, dtype=np.complex128).ravel()
    return float(abs(np.vdot(ref, got)) ** 2 / (np.vdot(ref, ref).real * np.vdot(got, got).real))
