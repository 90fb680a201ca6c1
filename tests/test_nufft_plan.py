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
