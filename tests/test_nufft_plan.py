"""The NUFFT plan of the mixed-precision channel sums, checked on the CPU.

The plan (grid size, kernel width, frequency offset, precorrection factors and
weight polynomials) comes from the library's host code; here it is used by an
independent numpy evaluation of the type-2 NUFFT and compared with what the
mathematics fixes:
- the channel sum sum_c v_c exp(i c b) computed directly (a closed form of the
  quantity the NUFFT approximates), relative to sum_c |v_c|;
- the weight polynomials against the "exponential of semicircle" kernel
  exp(beta (sqrt(1 - z^2) - 1)) evaluated directly;
- the precorrection factors against the kernel's Fourier transform from an
  independent Gauss-Legendre rule with twice the nodes.
A wrong sign of the frequency offset, a transposed tap order or a dropped
precorrection fails the first check by orders of magnitude.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre

CHANNELS = [1, 7, 48, 63, 64, 65, 77, 96, 127, 128, 129, 200, 256, 257, 300, 384, 447, 448, 512]


def es(z, beta):
    out = np.zeros_like(z, dtype=float)
    m = np.abs(z) < 1
    out[m] = np.exp(beta * (np.sqrt((1 - z[m]) * (1 + z[m])) - 1))
    return out


def weights(plan, s):
    """W tap weights at fractional offsets s (taps j < W/2 from the plan's
    polynomials, tap W-1-j as the mirror image P_j(-s))."""
    W = plan["W"]
    out = np.zeros((W,) + np.shape(s))
    for j in range(W // 2):
        c = plan["coef"][j]
        out[j] = np.polynomial.polynomial.polyval(s, c)
        out[W - 1 - j] = np.polynomial.polynomial.polyval(-s, c)
    return out


def nufft_numpy(plan, v, b):
    """sum_c v_c exp(i c b) at each b through the plan: precorrect, inverse FFT on
    N points, W-point interpolation, phase exp(i h b)."""
    N, W, h = plan["N"], plan["W"], plan["h"]
    C = len(v)
    ft = np.zeros(N, complex)
    ft[(np.arange(C) - h) % N] = v * plan["scale"]
    g = np.fft.ifft(ft) * N
    x = b / (2 * np.pi)
    u = (x - np.floor(x)) * N
    m0 = np.ceil(u - W / 2)
    s = (u - m0) - (W - 1) / 2
    w = weights(plan, s)
    acc = np.zeros(len(b), complex)
    for j in range(W):
        acc += g[(m0 + j).astype(int) % N] * w[j]
    return acc * np.exp(1j * h * b)


# (fp32 plan, kernel widths, bound on the kernel's own error relative to sum |v|)
PLANS = [(False, (10, 12), 2e-10), (True, (8, 10), 1e-8)]
PLAN_IDS = ["fp64-transform", "fp32-transform"]


@pytest.mark.parametrize("fp32,widths,bound", PLANS, ids=PLAN_IDS)
@pytest.mark.parametrize("C", CHANNELS)
def test_plan_shape(leap, C, fp32, widths, bound):
    p = leap.nufft_plan(C, fp32)
    N, W, h = p["N"], p["W"], p["h"]
    assert N >= max(256, 2 * C) and N & (N - 1) == 0 and N <= 1024
    assert h % 64 == 0 and abs(h - C / 2) <= 32
    kmax = max(h, C - 1 - h)
    assert 4 * kmax <= N  # oversampling of at least 2 over the shifted band
    assert W in widths
    assert p["scale"].shape == (C,) and np.all(p["scale"] > 0)


@pytest.mark.parametrize("fp32,widths,bound", PLANS, ids=PLAN_IDS)
@pytest.mark.parametrize("C", CHANNELS)
def test_plan_matches_direct_channel_sum(leap, C, fp32, widths, bound):
    p = leap.nufft_plan(C, fp32)
    rng = np.random.default_rng(C)
    v = rng.normal(size=C) + 1j * rng.normal(size=C)
    b = np.concatenate([rng.uniform(-400, 400, 1500), [0.0, -1e-17, 2 * np.pi, np.pi, -np.pi, 1e-300]])
    direct = np.exp(1j * np.outer(b, np.arange(C))) @ v
    err = np.abs(nufft_numpy(p, v, b) - direct).max() / np.abs(v).sum()
    assert err < bound, err


@pytest.mark.parametrize("C", [48, 200, 384, 512])
def test_fp32_transform_error_envelope(leap, C):
    """The FP32 plan evaluated with the transform, the grid and the W-tap sum
    rounded to FP32 (numpy's single-precision FFT, not the kernel's) stays
    below 1e-7 of sum |v|, half the per-row error the FP32 route's GPU tests
    allow (2e-7)."""
    import scipy.fft as sfft

    p = leap.nufft_plan(C, True)
    N, W, h = p["N"], p["W"], p["h"]
    rng = np.random.default_rng(C + 1)
    worst = 0.0
    for _ in range(10):
        v = rng.normal(size=C) + 1j * rng.normal(size=C)
        b = rng.uniform(-400, 400, 500)
        ft = np.zeros(N, np.complex64)
        ft[(np.arange(C) - h) % N] = (v * p["scale"]).astype(np.complex64)
        g = (sfft.ifft(ft) * np.float32(N)).astype(np.complex64)
        x = b / (2 * np.pi)
        u = (x - np.floor(x)) * N
        m0 = np.ceil(u - W / 2)
        s = ((u - m0) - (W - 1) / 2).astype(np.float32)
        w = weights(p, s).astype(np.float32)
        acc = np.zeros(len(b), np.complex64)
        for j in range(W):
            acc += g[(m0 + j).astype(int) % N] * w[j]
        got = acc.astype(np.complex128) * np.exp(1j * h * b)
        direct = np.exp(1j * np.outer(b, np.arange(C))) @ v
        worst = max(worst, np.abs(got - direct).max() / np.abs(v).sum())
    assert worst < 1e-7, worst


@pytest.mark.parametrize("fp32", [False, True], ids=PLAN_IDS)
@pytest.mark.parametrize("C", [48, 200, 384, 512])
def test_weight_polynomials_match_kernel(leap, C, fp32):
    p = leap.nufft_plan(C, fp32)
    W, beta = p["W"], p["beta"]
    s = np.linspace(-0.5, 0.5, 2001)
    w = weights(p, s)
    for j in range(W):
        exact = es((s + (W - 1) / 2 - j) / (W / 2), beta)
        # Chebyshev interpolation error sits at the kernel's truncation jump exp(-beta)
        assert np.abs(w[j] - exact).max() < 20 * np.exp(-beta), j


@pytest.mark.parametrize("fp32", [False, True], ids=PLAN_IDS)
@pytest.mark.parametrize("C", [48, 129, 384, 512])
def test_precorrection_matches_quadrature(leap, C, fp32):
    p = leap.nufft_plan(C, fp32)
    N, W, h, beta = p["N"], p["W"], p["h"], p["beta"]
    z, wq = legendre.leggauss(200)
    half = W * np.pi / N
    k = np.arange(C) - h
    psihat = half * (wq[None, :] * es(z[None, :], beta) * np.cos(np.outer(k, z) * half)).sum(1)
    want = (2 * np.pi / N) / psihat
    assert np.abs(p["scale"] / want - 1).max() < 1e-12


def test_plan_rejects_unsupported(leap):
    for C in (0, 513, 4096):
        for fp32 in (False, True):
            with pytest.raises(leap.LeapError, match="NUFFT"):
                leap.nufft_plan(C, fp32)
        with pytest.raises(leap.LeapError, match="NUFFT"):
            leap.nufft_plan(C, w16=True)


# ---- the width-16 plan (LEAP_CHANNEL_SUM_NUFFT_W16) ---------------------------

PI_LD = np.longdouble("3.14159265358979323846264338327950288")


def exact_sum_turns(v, x):
    """sum_c v_c exp(2 pi i c x) in extended precision for evaluation points x
    given in turns (c x reduced modulo 1 before the trigonometric functions)."""
    ld = np.longdouble
    k = np.arange(len(v)).astype(ld)
    vr, vi = v.real.astype(ld), v.imag.astype(ld)
    out = np.empty(len(x), complex)
    for i, xx in enumerate(x):
        ph = 2 * PI_LD * ((ld(xx) * k) % 1)
        c, s = np.cos(ph), np.sin(ph)
        out[i] = complex(float((vr * c - vi * s).sum()), float((vr * s + vi * c).sum()))
    return out


def nufft_numpy_turns(plan, v, x):
    """nufft_numpy with the evaluation points in turns, as the kernels form
    them (grid coordinate N frac(x), phase exp(2 pi i h x) with h x reduced
    modulo 1 exactly)."""
    N, W, h = plan["N"], plan["W"], plan["h"]
    C = len(v)
    ft = np.zeros(N, complex)
    ft[(np.arange(C) - h) % N] = v * plan["scale"]
    g = np.fft.ifft(ft) * N
    u = (x - np.floor(x)) * N
    m0 = np.ceil(u - W / 2)
    s = (u - m0) - (W - 1) / 2
    w = weights(plan, s)
    acc = np.zeros(len(x), complex)
    for j in range(W):
        acc += g[(m0 + j).astype(int) % N] * w[j]
    hx = (np.longdouble(h) * x.astype(np.longdouble)) % 1
    return acc * np.exp(2j * np.pi * hx.astype(float))


def horner_turns(v, x):
    """The FP64 Horner evaluation of the same sums (the batched FP64 kernels'
    method): z = exp(2 pi i x), P(z) by Horner's rule."""
    out = np.empty(len(x), complex)
    for i, xx in enumerate(x):
        z = np.exp(2j * np.pi * (xx % 1.0))
        acc = 0j
        for c in range(len(v) - 1, -1, -1):
            acc = acc * z + v[c]
        out[i] = acc
    return out


@pytest.mark.parametrize("C", CHANNELS)
def test_w16_plan_shape(leap, C):
    p = leap.nufft_plan(C, w16=True)
    q = leap.nufft_plan(C)
    assert p["W"] == 16 and p["N"] == q["N"] and p["h"] == q["h"]
    assert np.exp(-p["beta"]) < 2e-16
    assert p["coef"].shape == (8, 16) and p["scale"].shape == (C,) and np.all(p["scale"] > 0)


@pytest.mark.parametrize("C", [7, 96, 200, 300, 384, 447, 512])
def test_w16_plan_is_as_accurate_as_fp64_horner(leap, C):
    """With the evaluation point given in turns (as the kernels form it), the
    width-16 plan's channel sum is within 2e-15 of sum |v| of the extended-
    precision sum: the kernel's truncation lies below FP64 rounding, and the
    FP64 Horner chains (n eps) are not closer."""
    p = leap.nufft_plan(C, w16=True)
    rng = np.random.default_rng(C + 7)
    v = rng.normal(size=C) + 1j * rng.normal(size=C)
    x = np.concatenate([rng.uniform(-10, 10, 150), [0.0, 0.5, -0.5, 1e-300, 0.25 + 1e-12]])
    exact = exact_sum_turns(v, x)
    scale = np.abs(v).sum()
    err = np.abs(nufft_numpy_turns(p, v, x) - exact).max() / scale
    assert err < 2e-15, err
    if C >= 96:
        herr = np.abs(horner_turns(v, x) - exact).max() / scale
        assert err < herr, (err, herr)


@pytest.mark.parametrize("C", [48, 200, 384, 512])
def test_w16_weight_polynomials_and_precorrection(leap, C):
    p = leap.nufft_plan(C, w16=True)
    W, N, h, beta = p["W"], p["N"], p["h"], p["beta"]
    s = np.linspace(-0.5, 0.5, 2001)
    w = weights(p, s)
    for j in range(W):
        exact = es((s + (W - 1) / 2 - j) / (W / 2), beta)
        # at FP64 rounding: exp(beta (sqrt(1 - z^2) - 1)) with beta near 40
        # carries about 40 eps of relative rounding itself, in the plan's fit
        # and in this reference alike
        assert np.abs(w[j] - exact).max() < 1e-14, j
    z, wq = legendre.leggauss(300)
    half = W * np.pi / N
    k = np.arange(C) - h
    psihat = half * (wq[None, :] * es(z[None, :], beta) * np.cos(np.outer(k, z) * half)).sum(1)
    assert np.abs(p["scale"] / ((2 * np.pi / N) / psihat) - 1).max() < 1e-13


def test_w16_planted_narrow_kernel_fails(leap):
    """The 2e-15 bound separates the width-16 plan from the width-10 one on
    the same points, so a route that took the narrower plan fails it."""
    C = 384
    rng = np.random.default_rng(11)
    v = rng.normal(size=C) + 1j * rng.normal(size=C)
    x = rng.uniform(-10, 10, 100)
    exact = exact_sum_turns(v, x)
    err10 = np.abs(nufft_numpy_turns(leap.nufft_plan(C), v, x) - exact).max() / np.abs(v).sum()
    assert err10 > 1e-13


def test_planted_mistakes_fail(leap):
    """The direct-sum check catches a plan used with the wrong frequency offset
    or without precorrection."""
    p = leap.nufft_plan(384)
    rng = np.random.default_rng(5)
    v = rng.normal(size=384) + 1j * rng.normal(size=384)
    b = rng.uniform(-50, 50, 200)
    direct = np.exp(1j * np.outer(b, np.arange(384))) @ v
    bad_offset = dict(p, h=p["h"] + 1)
    bad_scale = dict(p, scale=np.ones(384))
    for bad in (bad_offset, bad_scale):
        err = np.abs(nufft_numpy(bad, v, b) - direct).max() / np.abs(v).sum()
        assert err > 1e-4
