"""The 16-bit storage codecs as the device kernels run them: every code is
written by the seeding kernel (the step kernels' encoder) and read back by
m3lb_storage (their decoder), and values between codes must come back as the
nearest code.  The E2M13 decoder inserts the sign with inline assembly on the
device only, so the exhaustive host test of narrow_float.hpp does not cover
that path.

The esoteric-pull seeding of a periodic box is a bijection between the
logical populations and the slots, and a slot of direction q holds direction
q or its opposite, which share a scale: the checks compare multisets of bit
patterns per opposite pair, which does not depend on the slot map (the slot
map itself is checked bit for bit in test_gpu_parity.py)."""
import numpy as np
import pytest

from m3lb_amd import abi
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu

SHAPE = (16, 16, 16)  # 4096 nodes x 27 slots: room for all 65536 codes


def e2m13_magnitudes():
    """|v| of the magnitude codes m = 0 .. 32767 from the format's definition:
    2 exponent bits e, 13 mantissa bits f; f 2^-8 for e = 0, else
    2^(4 + e) (1 + f / 2^13)."""
    m = np.arange(32768)
    e, f = m >> 13, m & 8191
    return np.where(e == 0, f / 256.0, np.ldexp(1.0 + f / 8192.0, 4 + e))


def e2m13_decode(codes):
    mag = e2m13_magnitudes()[codes >> 1]
    return np.where(codes & 1, -mag, mag)


def decode(dtype, codes):
    """Value of each 16-bit code at scale 1; NaN for codes that hold none."""
    c = codes.astype(np.uint16)
    if dtype == abi.E2M13:
        return e2m13_decode(codes)
    if dtype == abi.F16:
        return c.view(np.float16).astype(np.float64)
    if dtype == abi.BF16:
        # kept to normal FP32 values and zero, well inside the FP32 range at
        # every scale used here (an FP32 subnormal or overflow is not a code
        # round trip)
        with np.errstate(invalid="ignore"):  # NaN codes
            v = (c.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return np.where((v == 0) | ((np.abs(v) >= 2.0 ** -100) & (np.abs(v) < 2.0 ** 100)),
                        v, np.nan)
    v = c.view(np.int16).astype(np.float64)  # I16: saturates at +-32767
    return np.where(v < -32767, np.nan, v)


def pair_of(Q):
    """Opposite-pair index of each direction (q and its opposite share it)."""
    return np.array([0] + [(q + 1) // 2 for q in range(1, Q)])


def seed_and_read(dtype, g, scales):
    """Seed the shifted populations g (n, q) at the given per-direction
    scales; return the stored values m3lb_storage reports (n, q)."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_BGK, abi.SHIFTED, dtype, SHAPE)
    with FluidSolver(cfg) as s:
        assert (s.np, s.Q) == (np.prod(SHAPE), 27)  # periodic: no halo nodes
        s.set_scales(scales)
        s.seed_populations(g)
        return s.storage()


def assert_same_per_pair(got, want):
    """Equal multisets of float64 bit patterns within each opposite pair."""
    p = pair_of(want.shape[1])
    for k in range(p.max() + 1):
        cols = p == k
        a = np.sort(got[:, cols].reshape(-1).view(np.int64))
        b = np.sort(want[:, cols].reshape(-1).view(np.int64))
        assert np.array_equal(a, b), f"pair {k}"


def pair_scales(seed):
    """A power-of-two scale per opposite pair, 2^-3 .. 2^3."""
    k = np.random.default_rng(seed).integers(-3, 4, size=14)
    return np.ldexp(1.0, k[pair_of(27)])


@pytest.mark.parametrize("dtype", [abi.E2M13, abi.F16, abi.BF16, abi.I16])
def test_every_code_round_trips(dtype):
    """Every code that holds a finite value (BF16: every normal one below
    2^100 and zero), at scale 1 and at mixed scales: seeded as its value, it
    is stored as that code and read back exactly, the sign of zero included,
    except for FP16: its encoder compiles to one fused multiply-and-convert
    (v_fma_mixlo_f16) with a +0 addend, which stores -0 as +0, so FP16 zeros
    are compared by value."""
    rng = np.random.default_rng(1)
    n = int(np.prod(SHAPE)) * 27
    codes = np.arange(65536)
    codes = codes[np.isfinite(decode(dtype, codes))]
    codes = np.concatenate([codes, rng.choice(codes, size=n - codes.size)])
    rng.shuffle(codes)
    v = decode(dtype, codes).reshape(-1, 27)
    for scales in (np.ones(27), pair_scales(2)):
        g = v / scales
        got = seed_and_read(dtype, g, scales)
        if dtype == abi.F16:
            got, g = got + 0.0, g + 0.0  # -0 + 0 = +0: zeros by value
        assert_same_per_pair(got, g)


def test_e2m13_ties_go_to_the_even_code():
    """Each midpoint between two neighbouring magnitudes, both signs, is
    stored as the neighbour with the even magnitude code."""
    mag = e2m13_magnitudes()
    m = np.arange(32767)
    mid = (mag[m] + mag[m + 1]) / 2  # exact in FP32: at most 15 significant bits
    assert np.array_equal(mid.astype(np.float32).astype(np.float64), mid)
    want = mag[np.where(m % 2 == 0, m, m + 1)]
    n = int(np.prod(SHAPE)) * 27
    rng = np.random.default_rng(3)
    extra = rng.integers(0, 32767, n - 2 * m.size)
    sign = np.concatenate([np.ones(m.size), -np.ones(m.size),
                           np.where(rng.random(extra.size) < 0.5, 1.0, -1.0)])
    idx = np.concatenate([m, m, extra])
    perm = rng.permutation(n)
    sign, idx = sign[perm], idx[perm]
    scales = pair_scales(4)
    g = (sign * mid[idx]).reshape(-1, 27) / scales
    expect = (sign * want[idx]).reshape(-1, 27) / scales
    assert_same_per_pair(seed_and_read(abi.E2M13, g, scales), expect)


def test_e2m13_rounds_to_the_nearest_code_and_saturates():
    """Random FP32 values over the whole range and beyond it: each is stored
    as the nearest magnitude (values within 2^-10 of a step of a tie below 32
    are left out, where the encoder may take either neighbour), and
    magnitudes past the largest code saturate to it, keeping the sign."""
    mag = e2m13_magnitudes()
    n = int(np.prod(SHAPE)) * 27
    rng = np.random.default_rng(5)
    a = np.exp(rng.uniform(np.log(1e-4), np.log(400.0), size=n))
    a = a.astype(np.float32).astype(np.float64)
    j = np.clip(np.searchsorted(mag, a), 1, 32767)
    lo, hi = mag[j - 1], mag[j]
    near_tie = (np.abs(a - (lo + hi) / 2) <= (hi - lo) / 1024) & (a < 32)
    exact_tie = a == (lo + hi) / 2
    pick = np.where(a - lo < hi - a, j - 1, j)
    pick = np.where(exact_tie, np.where((j - 1) % 2 == 0, j - 1, j), pick)
    want = np.where(a >= mag[-1], mag[-1], mag[pick])
    keep = ~near_tie
    a, want = a[keep], want[keep]
    fill = rng.integers(0, 32768, size=n - a.size)  # exact codes fill the rest
    a = np.concatenate([a, mag[fill]])
    want = np.concatenate([want, mag[fill]])
    sign = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    scales = pair_scales(6)
    g = (sign * a).reshape(-1, 27) / scales
    expect = (sign * want).reshape(-1, 27) / scales
    assert_same_per_pair(seed_and_read(abi.E2M13, g, scales), expect)
