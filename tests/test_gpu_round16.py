"""The 16-bit store rounding of the step kernels (lbm3d_round16) against its
closed forms: round to nearest even, and stochastic rounding as stated in
include/lbm3d.h (BF16: the 16 random bits added below the last place, then
truncated; FP16: the top bits of the same 16 added below the last place, then
truncated), restated here on the bit patterns with numpy.  Enumerating all
65,536 values of the random bits gives each stochastic rounding's exact
expectation: the input itself for BF16 and for FP16 normal results; FP16
results below 2^-14 (its subnormal range, where the hardware's carry comes up
slightly short) average low by 1e-5 to 1e-4 of a step, held to 2^-12."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F16, BF16 = 1, 2


def round16(gpu, dtype, stochastic, vals, rnd=None):
    import ctypes
    import torch
    from taichi_lbm3d_amd import _lib
    x = torch.from_numpy(np.ascontiguousarray(vals, np.float32)).to(gpu)
    r = None
    if rnd is not None:
        r = torch.from_numpy(np.ascontiguousarray(rnd, np.uint32).view(np.int32)).to(gpu)
    out = torch.empty(x.numel(), dtype=torch.int16, device=gpu)
    _lib.lib().call("lbm3d_round16", None, dtype, int(stochastic), ctypes.c_void_p(x.data_ptr()),
                    ctypes.c_void_p(r.data_ptr() if r is not None else 0), ctypes.c_void_p(out.data_ptr()),
                    x.numel())
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint16)


def sample_values():
    rng = np.random.default_rng(16)
    mags = 10.0 ** rng.uniform(-9, 4.8, 4000)
    v = (mags * rng.choice([-1.0, 1.0], mags.size)).astype(np.float32)
    special = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.0, -65519.0, 7e4, 2.0 ** -24, 2.0 ** -14,
                        1.5 * 2.0 ** -14, 3.0 * 2.0 ** -25, 2.0 ** -26, 1e-38, 3.0e38], np.float32)
    return np.concatenate([v, special])


def f16_sr_rule(v, r16):
    """FP16 stochastic rounding for |v| >= 2^-24, on the FP32 bit pattern."""
    bits = v.view(np.uint32).astype(np.uint64)
    ex = (bits >> 23) & 0xFF
    drop = np.where(ex >= 113, 13, 126 - ex).astype(np.uint64)
    word = r16.astype(np.uint64) << 16
    dither = word >> (np.uint64(32) - drop)
    mask = (np.uint64(1) << drop) - np.uint64(1)
    d = ((bits + dither) & ~mask & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    with np.errstate(over="ignore"):
        return d.view(np.float32).astype(np.float16).view(np.uint16)


def test_round_to_nearest(gpu):
    v = sample_values()
    fin = np.isfinite(v)
    h = round16(gpu, F16, 0, v)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(h, v.astype(np.float16).view(np.uint16))
    b = round16(gpu, BF16, 0, v)
    bits = v.view(np.uint32).astype(np.uint64)
    rne = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)
    np.testing.assert_array_equal(b[fin], rne[fin])


def test_stochastic_rule(gpu):
    v = sample_values()
    rng = np.random.default_rng(17)
    r = rng.integers(0, 1 << 32, v.size, dtype=np.uint64).astype(np.uint32)
    r16 = r & 0xFFFF
    # BF16: the 16 bits below the last place plus the random bits, truncated
    b = round16(gpu, BF16, 1, v, r)
    exp_b = ((v.view(np.uint32).astype(np.uint64) + r16) >> 16).astype(np.uint16)
    np.testing.assert_array_equal(b, exp_b)
    # FP16 from 2^-24 up (normal and subnormal results): the rule on the bit pattern
    h = round16(gpu, F16, 1, v, r)
    big = np.abs(v) >= 2.0 ** -24
    np.testing.assert_array_equal(h[big], f16_sr_rule(v[big], r16[big]))
    # below 2^-24: 0 or the smallest subnormal, with the sign of v
    small = ~big
    hs = h[small]
    sign = (v[small].view(np.uint32) >> 16).astype(np.uint16) & 0x8000
    assert np.all((hs == sign) | (hs == (sign | 1)))
    # the two values around v, never farther
    hv = h.view(np.float16).astype(np.float64)
    fin = np.abs(v) < 65504.0
    lo = np.nextafter(v[fin].astype(np.float16), np.float16(-np.inf)).astype(np.float64)
    hi = np.nextafter(v[fin].astype(np.float16), np.float16(np.inf)).astype(np.float64)
    assert np.all((hv[fin] >= lo) & (hv[fin] <= hi))


@pytest.mark.parametrize("dtype", [F16, BF16])
def test_stochastic_rounding_is_unbiased(gpu, dtype):
    """Every value of the 16 random bits once: the mean of the results is the
    exact expectation.  It equals the input for BF16 and FP16 normal results
    and is within 2^-12 of the step between the two values around the input
    (below it) for FP16 results below 2^-14."""
    vals = np.array([1.0 + 1.0 / 3.0 * 2.0 ** -10, -2.718281828, 3.1e-5, 4.7e-6, 1.2345e-7, 3.0e-8, 1.1e-8,
                     2.5e-9, 1234.567, 0.333333], np.float32)
    rnd = np.arange(1 << 16, dtype=np.uint32)
    errs = []
    for v in vals:
        out = round16(gpu, dtype, 1, np.full(rnd.size, v, np.float32), rnd)
        if dtype == F16:
            got = out.view(np.float16).astype(np.float64)
            a = np.float16(v)
            step = float(np.abs(np.nextafter(a, np.float16(np.inf)) - np.nextafter(a, np.float16(-np.inf)))) / 2
            step = max(step, 2.0 ** -24)
        else:
            got = (out.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            step = 2.0 ** (np.floor(np.log2(abs(float(v)))) - 7)
        errs.append((float(v), (got.mean() - float(v)) / step))
        assert np.unique(got).size <= 2
    for v, e in errs:
        print(f"dtype {dtype} v {v:.6e} bias {e:+.3e} steps")
        sub = dtype == F16 and abs(v) < 2.0 ** -14
        assert (-(2.0 ** -12) <= e <= 0.0) if sub else e == 0.0, (dtype, v, e)
