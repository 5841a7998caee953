// fft_reg.hpp -- mixed-radix FFT of one line held in registers (gfx950).
//
// A line of N complex values lives in a register array x[N] of one lane.
// The transform is a recursive decimation-in-time split N = R*M, fully
// unrolled at compile time: every index and every twiddle is a constant, so
// the array is promoted to VGPRs and the twiddles become instruction
// immediates.  Twiddles come from a constexpr table evaluated by the
// compiler (Taylor series after quadrant reduction, exact at multiples of
// pi/2), so there is no table in memory and no sincos at run time.
//
// SIGN = -1: forward, unnormalised, exp(-2 pi i j k / N)
// SIGN = +1: backward, unnormalised, exp(+2 pi i j k / N)
// These are the conventions of the reference's to_fourier / to_real before
// its separate 1/N pass (src/operations/transform.hpp:79,141,175-181).
#pragma once

#include <hip/hip_runtime.h>

namespace inqx {

// plain aggregate complex (trivially copyable, so it can be loaded and
// stored through address-space-qualified pointers)
template <typename T> struct alignas(2 * sizeof(T)) cplx_pod {
  T x, y;
};
template <typename T> using cplx = cplx_pod<T>;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {  // conj(a) * b
  return C{a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename C, typename S> __device__ __forceinline__ C cscale(C a, S s) { return C{a.x * s, a.y * s}; }

// ---------------------------------------------------------------- twiddles
struct tw_cs {
  double c, s;
};

// exp(+2 pi i k / n), evaluated at compile time
constexpr tw_cs tw_value(long k, long n) {
  k %= n;
  if (k < 0) k += n;
  long q = (4 * k) / n;
  long r = 4 * k - q * n;  // angle = pi/2 * (q + r/n)
  double phi = 1.5707963267948966 * (double)r / (double)n;
  double p2 = phi * phi;
  double c = 1.0, s = phi, tc = 1.0, ts = phi;
  for (int m = 1; m < 24; ++m) {
    tc *= -p2 / (double)((2 * m - 1) * (2 * m));
    ts *= -p2 / (double)((2 * m) * (2 * m + 1));
    c += tc;
    s += ts;
  }
  if (r == 0) {
    c = 1.0;
    s = 0.0;
  }
  switch (q) {
    case 0: return tw_cs{c, s};
    case 1: return tw_cs{-s, c};
    case 2: return tw_cs{-c, -s};
    default: return tw_cs{s, -c};
  }
}

template <int N> struct tw_table {
  double c[N], s[N];
  constexpr tw_table() : c(), s() {
    for (int k = 0; k < N; ++k) {
      tw_cs v = tw_value(k, N);
      c[k] = v.c;
      s[k] = v.s;
    }
  }
};
template <int N> inline constexpr tw_table<N> kTw{};

// multiply by exp(SIGN * 2 pi i e / N); e is a constant after unrolling and
// the exact 0 / +-1 cases fold to moves and negations.
template <int N, int SIGN, typename C>
__device__ __forceinline__ C twmul(C v, int e) {
  using T = decltype(v.x);
  const int k = ((e % N) + N) % N;
  const double dc = kTw<N>.c[k];
  const double ds = SIGN * kTw<N>.s[k];
  if (ds == 0.0) {
    if (dc == 1.0) return v;
    if (dc == -1.0) return C{-v.x, -v.y};
  }
  if (dc == 0.0) {
    if (ds == 1.0) return C{-v.y, v.x};
    if (ds == -1.0) return C{v.y, -v.x};
  }
  const T c = (T)dc, s = (T)ds;
  return C{v.x * c - v.y * s, v.x * s + v.y * c};
}

// ---------------------------------------------------------------- small DFTs
template <int R, int SIGN, typename C> struct dft_small {
  // generic R-point DFT, t[q] = sum_r W_R^{SIGN r q} t[r]
  static __device__ __forceinline__ void run(C (&t)[R]) {
    C o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      C acc = t[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cadd(acc, twmul<R, SIGN>(t[r], r * q));
      o[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) t[q] = o[q];
  }
};

template <int SIGN, typename C> struct dft_small<2, SIGN, C> {
  static __device__ __forceinline__ void run(C (&t)[2]) {
    C a = t[0], b = t[1];
    t[0] = cadd(a, b);
    t[1] = csub(a, b);
  }
};

template <int SIGN, typename C> struct dft_small<4, SIGN, C> {
  static __device__ __forceinline__ void run(C (&t)[4]) {
    C a0 = cadd(t[0], t[2]), a1 = csub(t[0], t[2]);
    C b0 = cadd(t[1], t[3]), b1 = csub(t[1], t[3]);
    // b1 * (SIGN i)
    C jb1 = (SIGN > 0) ? C{-b1.y, b1.x} : C{b1.y, -b1.x};
    t[0] = cadd(a0, b0);
    t[2] = csub(a0, b0);
    t[1] = cadd(a1, jb1);
    t[3] = csub(a1, jb1);
  }
};

template <int SIGN, typename C> struct dft_small<3, SIGN, C> {
  static __device__ __forceinline__ void run(C (&t)[3]) {
    using T = decltype(t[0].x);
    const T c = (T)-0.5, s = (T)(SIGN * 0.86602540378443864676);
    C a = t[0];
    C u = cadd(t[1], t[2]);
    C v = csub(t[1], t[2]);
    C m = C{a.x + c * u.x, a.y + c * u.y};
    C js = C{-s * v.y, s * v.x};
    t[0] = cadd(a, u);
    t[1] = cadd(m, js);
    t[2] = csub(m, js);
  }
};

template <int SIGN, typename C> struct dft_small<5, SIGN, C> {
  static __device__ __forceinline__ void run(C (&t)[5]) {
    using T = decltype(t[0].x);
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)(SIGN * 0.95105651629515357212), s2 = (T)(SIGN * 0.58778525229247312917);
    C a = t[0];
    C u1 = cadd(t[1], t[4]), v1 = csub(t[1], t[4]);
    C u2 = cadd(t[2], t[3]), v2 = csub(t[2], t[3]);
    C m1 = C{a.x + c1 * u1.x + c2 * u2.x, a.y + c1 * u1.y + c2 * u2.y};
    C m2 = C{a.x + c2 * u1.x + c1 * u2.x, a.y + c2 * u1.y + c1 * u2.y};
    // i*(s1 v1 + s2 v2) and i*(s2 v1 - s1 v2)
    C n1 = C{-(s1 * v1.y + s2 * v2.y), s1 * v1.x + s2 * v2.x};
    C n2 = C{-(s2 * v1.y - s1 * v2.y), s2 * v1.x - s1 * v2.x};
    t[0] = cadd(a, cadd(u1, u2));
    t[1] = cadd(m1, n1);
    t[4] = csub(m1, n1);
    t[2] = cadd(m2, n2);
    t[3] = csub(m2, n2);
  }
};

// ---------------------------------------------------------------- radix choice
constexpr int fft_radix(int n) {
  if (n % 4 == 0 && n != 8) return 4;
  if (n % 2 == 0) return 2;
  if (n % 3 == 0) return 3;
  if (n % 5 == 0) return 5;
  if (n % 7 == 0) return 7;
  return n;
}

constexpr bool fft_supported(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  while (n % 7 == 0) n /= 7;
  return n == 1;
}

// ---------------------------------------------------------------- recursive FFT
template <int N, int SIGN, typename C> struct fft_line {
  static_assert(fft_supported(N), "FFT length must be a product of 2, 3, 5, 7");
  static __device__ __forceinline__ void run(C (&x)[N]) {
    constexpr int R = fft_radix(N);
    constexpr int M = N / R;
    if constexpr (M == 1) {
      dft_small<R, SIGN, C>::run(x);
    } else {
      C sub[R][M];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int m = 0; m < M; ++m) sub[r][m] = x[r + R * m];
#pragma unroll
      for (int r = 0; r < R; ++r) fft_line<M, SIGN, C>::run(sub[r]);
#pragma unroll
      for (int k1 = 0; k1 < M; ++k1) {
        C t[R];
        t[0] = sub[0][k1];
#pragma unroll
        for (int r = 1; r < R; ++r) t[r] = twmul<N, SIGN>(sub[r][k1], r * k1);
        dft_small<R, SIGN, C>::run(t);
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) x[k1 + M * k2] = t[k2];
      }
    }
  }
};

// ---------------------------------------------------------------- lines split over lane pairs
// A line of even length N is shared by lane l < 32 ("lower") and lane l + 32
// ("upper") of a wave, N/2 values each, so a lane holds half the registers of
// fft_line.  One radix-2 stage crosses the pair; v_permlane32_swap hands each
// lane both values of its pair (no LDS traffic).

// the values held by the lower and the upper lane of the pair, on both lanes
__device__ __forceinline__ void pair_vals(float v, float &lo, float &hi) {
  const unsigned u = __float_as_uint(v);
  const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
  lo = __uint_as_float(r[0]);
  hi = __uint_as_float(r[1]);
}
__device__ __forceinline__ void pair_vals(double v, double &lo, double &hi) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned a = (unsigned)u, b = (unsigned)(u >> 32);
  const auto ra = __builtin_amdgcn_permlane32_swap(a, a, false, false);
  const auto rb = __builtin_amdgcn_permlane32_swap(b, b, false, false);
  lo = __longlong_as_double((long long)(((unsigned long long)rb[0] << 32) | ra[0]));
  hi = __longlong_as_double((long long)(((unsigned long long)rb[1] << 32) | ra[1]));
}
template <typename C> __device__ __forceinline__ void pair_vals(C v, C &lo, C &hi) {
  pair_vals(v.x, lo.x, hi.x);
  pair_vals(v.y, lo.y, hi.y);
}

template <int N, int SIGN, typename C> struct fft_split {
  static_assert(N % 2 == 0, "split lines need an even length");
  static constexpr int H = N / 2;
  using T = decltype(C{}.x);
  // decimation in time: lower lanes hold x[2n], upper lanes x[2n+1] (n < H);
  // afterwards lower lanes hold X[k], upper lanes X[k + H] (k < H)
  static __device__ __forceinline__ void dit(C (&v)[H], bool upper) {
    fft_line<H, SIGN, C>::run(v);
    const T sg = upper ? (T)-1 : (T)1;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      C a, b;
      pair_vals(v[k], a, b);
      const C t = twmul<N, SIGN>(b, k);
      v[k] = C{a.x + sg * t.x, a.y + sg * t.y};
    }
  }
  // decimation in frequency: lower lanes hold x[n], upper lanes x[n + H];
  // afterwards lower lanes hold X[2k], upper lanes X[2k + 1]
  static __device__ __forceinline__ void dif(C (&v)[H], bool upper) {
    const T sg = upper ? (T)-1 : (T)1;
#pragma unroll
    for (int n = 0; n < H; ++n) {
      C a, b;
      pair_vals(v[n], a, b);
      const C d{a.x + sg * b.x, a.y + sg * b.y};
      const C t = twmul<N, SIGN>(d, n);
      v[n] = upper ? t : d;
    }
    fft_line<H, SIGN, C>::run(v);
  }
};

// ---------------------------------------------------------------- lines split over four lanes
// A line of length N (a multiple of 4) is shared by the lanes l, l + 16, l + 32
// and l + 48 of a wave (part q = lane / 16), N/4 values each: a quarter of the
// registers of fft_line.  Two radix-2 stages cross the lanes: parts q and q ^ 2
// (32 lanes apart, v_permlane32_swap as in fft_split) and parts q and q ^ 1
// (16 lanes apart, a lane shuffle within each half-wave).

// the value of lane l ^ 16
__device__ __forceinline__ float xor16(float v) { return __shfl_xor(v, 16); }
__device__ __forceinline__ double xor16(double v) { return __shfl_xor(v, 16); }

// output position of part q: the values of part q are X[k + (N/4) quad_rev(q)]
__host__ __device__ constexpr int quad_rev(int q) { return ((q & 1) << 1) | (q >> 1); }

template <int N, int SIGN, typename C> struct fft_quad {
  static_assert(N % 4 == 0, "four-lane lines need a length divisible by 4");
  static constexpr int Q = N / 4;
  using T = decltype(C{}.x);
  // decimation in time: part q holds x[4n + q] (n < Q); afterwards it holds
  // X[k + Q quad_rev(q)] (k < Q)
  static __device__ __forceinline__ void dit(C (&v)[Q], int q) {
    fft_line<Q, SIGN, C>::run(v);  // Y_q[k], the transform of x[4n + q]
    // parts q, q ^ 2 hold the even and odd samples of x[2m + (q & 1)]: their
    // length-N/2 transform, E (q & 1 = 0) or O (q & 1 = 1), at k (q < 2) and k + Q (q >= 2)
    const T sa = q >= 2 ? (T)-1 : (T)1;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      C a, b;
      pair_vals(v[k], a, b);
      const C t = twmul<N / 2, SIGN>(b, k);
      v[k] = C{a.x + sa * t.x, a.y + sa * t.y};
    }
    // parts q, q ^ 1 hold E[m] and O[m], m = k + Q (q >> 1): X[m] = E + W^m O on
    // the even part, X[m + N/2] = E - W^m O on the odd part
    const bool odd = (q & 1) != 0, hi = q >= 2;
    const T sb = odd ? (T)-1 : (T)1;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const C w{xor16(v[k].x), xor16(v[k].y)};
      const C e = odd ? w : v[k], o = odd ? v[k] : w;
      C t = twmul<N, SIGN>(o, k);
      if (hi) t = SIGN > 0 ? C{-t.y, t.x} : C{t.y, -t.x};  // W^(k + N/4) = W^k (SIGN i)
      v[k] = C{e.x + sb * t.x, e.y + sb * t.y};
    }
  }
  // decimation in frequency, the mirror of dit: part q holds x[m + Q quad_rev(q)]
  // (m < Q); afterwards it holds X[4k + q] (k < Q).  With chunk p = p0 + 2 p1
  // in part q (p0 = q >> 1, p1 = q & 1) and frequency t = t0 + 2 t1,
  // X[4k + t] = sum_m W_Q^(mk) W_N^(mt) sum_p x[m + Q p] W_4^(pt), the inner sum
  // taken over p1 (parts q, q ^ 1) and then over p0 (parts q, q ^ 2)
  static __device__ __forceinline__ void dif(C (&v)[Q], int q) {
    const bool odd = (q & 1) != 0, hi = q >= 2;
    // u = x[m + Q p0] + (-1)^t0 x[m + Q (p0 + 2)], t0 = q & 1
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const C w{xor16(v[k].x), xor16(v[k].y)};
      const C a = odd ? w : v[k], b = odd ? v[k] : w;
      v[k] = odd ? C{a.x - b.x, a.y - b.y} : C{a.x + b.x, a.y + b.y};
    }
    // z = u(p0 = 0) + (-1)^t1 W_4^t0 u(p0 = 1), t1 = q >> 1
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      C a, b;
      pair_vals(v[k], a, b);
      if (odd) b = SIGN > 0 ? C{-b.y, b.x} : C{b.y, -b.x};  // W_4 = SIGN i
      v[k] = hi ? C{a.x - b.x, a.y - b.y} : C{a.x + b.x, a.y + b.y};
    }
    // W_N^(m q), then the length-Q transform
#pragma unroll
    for (int m = 0; m < Q; ++m) {
      const T c = q == 0 ? (T)1 : q == 1 ? (T)kTw<N>.c[m % N] : q == 2 ? (T)kTw<N>.c[(2 * m) % N] : (T)kTw<N>.c[(3 * m) % N];
      const T sn = q == 0 ? (T)0
                          : (T)SIGN * (q == 1   ? (T)kTw<N>.s[m % N]
                                       : q == 2 ? (T)kTw<N>.s[(2 * m) % N]
                                                : (T)kTw<N>.s[(3 * m) % N]);
      v[m] = C{v[m].x * c - v[m].y * sn, v[m].x * sn + v[m].y * c};
    }
    fft_line<Q, SIGN, C>::run(v);
  }
};

}  // namespace inqx
