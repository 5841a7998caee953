"""Host emulation of the lane maps of the split line transforms in
inq_amd/csrc/fft_reg.hpp, with each stage's exchange and twiddle restated
here on numpy arrays (one row per part of the line):

  fft_split<N>::dit  a line over a lane pair: part h holds x[2n + h]; after the
                     transform it holds X[k + h N/2];
  fft_quad<N>::dit   a line over four lanes (part q = lane / 16): part q holds
                     x[4n + q]; the transform runs the quarter-length FFT, then
                     combines parts q and q ^ 2 (the 32-lane exchange, twiddles
                     of length N/2) and parts q and q ^ 1 (the 16-lane exchange,
                     twiddles of length N, times SIGN i on parts 2 and 3); part
                     q then holds X[k + (N/4) quad_rev(q)];
  fft_quad<N>::dif   its mirror (the pair line pass's inverse): part q holds
                     x[m + (N/4) quad_rev(q)]; the 16-lane exchange combines
                     parts q and q ^ 1, the 32-lane exchange parts q and q ^ 2
                     (times SIGN i on odd parts), then the twiddle W_N^(m q) and
                     the quarter-length FFT; part q then holds X[4k + q].

All are checked against numpy's FFT in both directions for every line length
the plane passes are built for, so a wrong exchange partner, twiddle sign or
output order fails here, on the CPU, before any GPU run.
"""
import numpy as np
import pytest

SIZES = [12, 16, 18, 20, 24, 28, 30, 32, 36, 40, 48, 54, 60, 64, 72, 80, 96, 100, 108, 120, 128]


def dft(y, sign):
    m = len(y)
    n = np.arange(m)
    return np.exp(sign * 2j * np.pi * np.outer(n, n) / m) @ y


def quad_rev(q):
    return ((q & 1) << 1) | (q >> 1)


def split_dit(x, sign):
    n = len(x)
    h = n // 2
    k = np.arange(h)
    part = [dft(x[p::2], sign) for p in (0, 1)]  # each lane's half-length transform
    a, b = part[0], part[1]                        # pair_vals: lower and upper values
    t = b * np.exp(sign * 2j * np.pi * k / n)
    return [a + t, a - t]                          # lower: X[k], upper: X[k + N/2]


def quad_dit(x, sign):
    n = len(x)
    qn = n // 4
    k = np.arange(qn)
    y = [dft(x[q::4], sign) for q in range(4)]
    # parts q, q ^ 2: the 32-lane exchange, length-N/2 twiddles
    z = []
    for q in range(4):
        a, b = y[q & 1], y[(q & 1) + 2]
        t = b * np.exp(sign * 2j * np.pi * k / (n // 2))
        z.append(a + t if q < 2 else a - t)
    # parts q, q ^ 1: the 16-lane exchange, length-N twiddles, SIGN i on parts 2, 3
    out = []
    for q in range(4):
        odd, hi = q & 1, q >= 2
        e = z[q ^ 1] if odd else z[q]
        o = z[q] if odd else z[q ^ 1]
        t = o * np.exp(sign * 2j * np.pi * k / n)
        if hi:
            t = t * (1j * sign)
        out.append(e - t if odd else e + t)
    return out


def quad_dif(parts, sign):
    """parts[q] = x[m + Q quad_rev(q)], m < Q; returns parts[q] = X[4k + q]"""
    qn = len(parts[0])
    n = 4 * qn
    m = np.arange(qn)
    # parts q, q ^ 1 (the 16-lane exchange): u = x[m + Q p0] + (-1)^t0 x[m + Q (p0 + 2)]
    u = []
    for q in range(4):
        own, w = parts[q], parts[q ^ 1]
        u.append(w - own if q & 1 else own + w)
    # parts q, q ^ 2 (the 32-lane exchange): lower value a, upper value b
    z = []
    for q in range(4):
        a, b = u[q & 1], u[(q & 1) + 2]
        if q & 1:
            b = b * (1j * sign)
        z.append(a - b if q >= 2 else a + b)
    return [dft(z[q] * np.exp(sign * 2j * np.pi * m * q / n), sign) for q in range(4)]


def reference(x, sign):
    return np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * len(x)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("sign", [-1, 1])
def test_lane_pair_split(n, sign):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    lo, hi = split_dit(x, sign)
    ref = reference(x, sign)
    assert np.allclose(lo, ref[: n // 2]) and np.allclose(hi, ref[n // 2:])


@pytest.mark.parametrize("n", [s for s in SIZES if s % 4 == 0])
@pytest.mark.parametrize("sign", [-1, 1])
def test_four_lane_split(n, sign):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 7).standard_normal(n)
    parts = quad_dit(x, sign)
    ref = reference(x, sign)
    qn = n // 4
    for q in range(4):
        assert np.allclose(parts[q], ref[np.arange(qn) + qn * quad_rev(q)]), q


@pytest.mark.parametrize("n", [s for s in SIZES if s % 4 == 0])
@pytest.mark.parametrize("sign", [-1, 1])
def test_four_lane_split_dif(n, sign):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 5).standard_normal(n)
    qn = n // 4
    parts = quad_dif([x[np.arange(qn) + qn * quad_rev(q)] for q in range(4)], sign)
    ref = reference(x, sign)
    for q in range(4):
        assert np.allclose(parts[q], ref[q::4]), q


def test_four_lane_round_trip():
    """dit then dif of the opposite sign, as the pair line pass applies them,
    returns N x in the input's lane order"""
    n = 128
    x = np.random.default_rng(11).standard_normal(n) + 1j * np.random.default_rng(12).standard_normal(n)
    back = quad_dif(quad_dit(x, -1), 1)
    for q in range(4):
        assert np.allclose(back[q], n * x[q::4]), q


def test_a_wrong_exchange_partner_fails():
    """the check has teeth: combining parts q and q ^ 3 in the last stage (a
    plausible slip) does not reproduce the transform"""
    n, sign = 72, -1
    x = np.random.default_rng(3).standard_normal(n) + 0j
    qn = n // 4
    k = np.arange(qn)
    y = [dft(x[q::4], sign) for q in range(4)]
    z = [y[q & 1] + (1 if q < 2 else -1) * y[(q & 1) + 2] * np.exp(sign * 2j * np.pi * k / (n // 2))
         for q in range(4)]
    bad = []
    for q in range(4):
        e, o = (z[q ^ 3], z[q]) if q & 1 else (z[q], z[q ^ 3])
        t = o * np.exp(sign * 2j * np.pi * k / n) * ((1j * sign) if q >= 2 else 1)
        bad.append(e - t if q & 1 else e + t)
    ref = reference(x, sign)
    assert not all(np.allclose(bad[q], ref[k + qn * quad_rev(q)]) for q in range(4))
