// eigensolver.hip -- the per-state passes of the steepest-descent
// eigensolver and of the Teter-Payne-Allan preconditioner on band-major sphere
// states [nst][npw].
//
//   dots       d_k(s) = vol sum_G w_k(G) conj(a_k(s,G)) b_k(s,G), up to four
//              pairs per pass (operations::overlap_diagonal,
//              src/operations/overlap_diagonal.hpp:25-60); FP64 accumulation,
//              partial sums over fixed 4096-point chunks added in chunk order,
//              so the result does not depend on the launch.
//   residual   r = H phi - (e / Re n) phi  (steepest_descent.hpp:37-49)
//   tpa        r(G) *= K((1/2)|G|^2 / <T>)  (preconditioner.hpp:31-58)
//   update     lambda from the six state overlaps, phi += lambda r and
//              H phi += lambda H r  (steepest_descent.hpp:53-83)
#include "internal.hpp"

#include <algorithm>

#if defined(__HIP_DEVICE_COMPILE__)
#define GAS __attribute__((address_space(1)))
#else
#define GAS
#endif

namespace inqx {

namespace {

template <typename T> __device__ __forceinline__ const GAS T *sgp(const T *p) { return (const GAS T *)p; }
template <typename T> __device__ __forceinline__ GAS T *sgw(T *p) { return (GAS T *)p; }

constexpr int kChunk = 4096;
constexpr int kThreads = 256;

struct DotArgs {
  const void *a[4];
  const void *b[4];
  const double *w;  // weight per sphere point for the pairs with bit k of wmask set
  int wmask;
  int n;
};

// partial sums of one (state, chunk): part[(s * nch + c) * n + k] (re, im)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_sd_dots(DotArgs d, long K, int nch, double *part) {
  using C = cplx<T>;
  __shared__ double red[8][kThreads];
  const int tid = threadIdx.x;
  const int c = blockIdx.x % nch, s = blockIdx.x / nch;
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  const long p0 = (long)c * kChunk;
  const long p1 = p0 + kChunk < K ? p0 + kChunk : K;
  for (long p = p0 + tid; p < p1; p += kThreads) {
    const double wp = d.w ? sgp(d.w)[p] : 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < d.n) {
        const C x = sgp(reinterpret_cast<const C *>(d.a[k]))[(size_t)s * K + p];
        const C y = sgp(reinterpret_cast<const C *>(d.b[k]))[(size_t)s * K + p];
        const double wk = (d.wmask >> k) & 1 ? wp : 1.0;
        acc[2 * k] += wk * ((double)x.x * y.x + (double)x.y * y.y);
        acc[2 * k + 1] += wk * ((double)x.x * y.y - (double)x.y * y.x);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (tid < h)
#pragma unroll
      for (int k = 0; k < 8; ++k) red[k][tid] += red[k][tid + h];
    __syncthreads();
  }
  if (tid < 2 * d.n) sgw(part)[((size_t)s * nch + c) * 2 * d.n + tid] = red[tid][0];
}

// out[s * n + k] = vol * sum over chunks, in chunk order
__global__ void __launch_bounds__(kThreads) k_sd_dots_combine(const double *part, int nst, int nch, int n, double vol,
                                                              double *out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= nst * 2 * n) return;
  const int s = i / (2 * n), r = i - s * 2 * n;
  double acc = 0.0;
  for (int c = 0; c < nch; ++c) acc += sgp(part)[((size_t)s * nch + c) * 2 * n + r];
  sgw(out)[i] = vol * acc;
}

// r = hphi - (e / Re n) phi with e = en[2s], n = en[2s+1] (complex pairs)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_sd_residual(const cplx<T> *phi, const cplx<T> *hphi, cplx<T> *res,
                                                          const double *en, long K) {
  using C = cplx<T>;
  const int s = blockIdx.y;
  const long p = (long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= K) return;
  const GAS double *q = sgp(en) + 4 * (size_t)s;
  const double nr = q[2];
  const T er = (T)(q[0] / nr), ei = (T)(q[1] / nr);
  const C f = sgp(phi)[(size_t)s * K + p], h = sgp(hphi)[(size_t)s * K + p];
  sgw(res)[(size_t)s * K + p] = C{h.x - (er * f.x - ei * f.y), h.y - (er * f.y + ei * f.x)};
}

__device__ __forceinline__ double tpa_k(double x) {
  const double num = ((8.0 * x + 12.0) * x + 18.0) * x + 27.0;
  const double den = (((16.0 * x + 8.0) * x + 12.0) * x + 18.0) * x + 27.0;
  return num / den;
}

// r(G) *= K(g2h(G) / <T>_s), <T>_s = Re(kn[2s] / kn[2s+1]) (complex pairs)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_tpa_scale(cplx<T> *res, const double *g2h, const double *kn, long K) {
  using C = cplx<T>;
  const int s = blockIdx.y;
  const long p = (long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= K) return;
  const GAS double *q = sgp(kn) + 4 * (size_t)s;
  // Re(a / b) for complex a, b
  const double kin = (q[0] * q[2] + q[1] * q[3]) / (q[2] * q[2] + q[3] * q[3]);
  const T f = (T)tpa_k(sgp(g2h)[p] / kin);
  C v = sgp(res)[(size_t)s * K + p];
  v.x *= f;
  v.y *= f;
  sgw(res)[(size_t)s * K + p] = v;
}

__device__ __forceinline__ void cmul_d(double ar, double ai, double br, double bi, double &cr, double &ci) {
  cr = ar * br - ai * bi;
  ci = ar * bi + ai * br;
}

// lambda per state from m0..m3 = (r,r), (phi,r), (r,Hr), (phi,Hr) and m4, m5 = (Hphi,phi), (phi,phi)
__global__ void __launch_bounds__(kThreads) k_sd_lambda(const double *m03, const double *m45, int nst, double *lam) {
  const int s = blockIdx.x * kThreads + threadIdx.x;
  if (s >= nst) return;
  const GAS double *a = sgp(m03) + 8 * (size_t)s;
  const GAS double *b = sgp(m45) + 4 * (size_t)s;
  double m[6][2];
  for (int k = 0; k < 4; ++k) {
    m[k][0] = a[2 * k];
    m[k][1] = a[2 * k + 1];
  }
  for (int k = 0; k < 2; ++k) {
    m[4 + k][0] = b[2 * k];
    m[4 + k][1] = b[2 * k + 1];
  }
  double x, y, u, v;
  cmul_d(m[0][0], m[0][1], m[3][0], m[3][1], x, y);
  cmul_d(m[2][0], m[2][1], m[1][0], m[1][1], u, v);
  const double ca = x - u;
  cmul_d(m[5][0], m[5][1], m[2][0], m[2][1], x, y);
  cmul_d(m[4][0], m[4][1], m[0][0], m[0][1], u, v);
  const double cb = x - u;
  cmul_d(m[4][0], m[4][1], m[1][0], m[1][1], x, y);
  cmul_d(m[3][0], m[3][1], m[5][0], m[5][1], u, v);
  const double cc = x - u;
  const double den = cb + sqrt(cb * cb - 4.0 * ca * cc);
  sgw(lam)[s] = fabs(den) < 1e-15 ? 0.0 : 2.0 * cc / den;
}

// phi += lambda r; hphi += lambda hr when hphi is given
template <typename T>
__global__ void __launch_bounds__(kThreads) k_sd_update(cplx<T> *phi, cplx<T> *hphi, const cplx<T> *res,
                                                        const cplx<T> *hres, const double *lam, long K) {
  using C = cplx<T>;
  const int s = blockIdx.y;
  const long p = (long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= K) return;
  const T l = (T)sgp(lam)[s];
  const size_t o = (size_t)s * K + p;
  const C r = sgp(res)[o];
  C f = sgp(phi)[o];
  f.x += l * r.x;
  f.y += l * r.y;
  sgw(phi)[o] = f;
  if (hphi) {
    const C hr = sgp(hres)[o];
    C h = sgp(hphi)[o];
    h.x += l * hr.x;
    h.y += l * hr.y;
    sgw(hphi)[o] = h;
  }
}

inline dim3 point_grid(int nst, long K) { return dim3((unsigned)((K + kThreads - 1) / kThreads), (unsigned)nst); }

}  // namespace

size_t sd_dots_partial_bytes(int nst, long K) { return (size_t)nst * ((K + kChunk - 1) / kChunk) * 8 * sizeof(double); }

void launch_sd_dots(Ctx &cx, int prec, int nst, long K, int n, const void *const *a, const void *const *b,
                    const double *w, int wmask, double vol, double *part, double *out) {
  INQX_CHECK(n >= 1 && n <= 4, INQX_ERR_ARG, "one to four overlaps per pass");
  DotArgs d{};
  for (int k = 0; k < n; ++k) {
    d.a[k] = a[k];
    d.b[k] = b[k];
  }
  d.w = w;
  d.wmask = w ? wmask : 0;
  d.n = n;
  const int nch = (int)((K + kChunk - 1) / kChunk);
  ProfScope ps(cx, "sd_dots");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_sd_dots<double>, dim3((unsigned)(nst * nch)), dim3(kThreads), 0, cx.stream, d, K, nch, part);
  else
    hipLaunchKernelGGL(k_sd_dots<float>, dim3((unsigned)(nst * nch)), dim3(kThreads), 0, cx.stream, d, K, nch, part);
  INQX_HIP(hipGetLastError());
  const int tot = nst * 2 * n;
  hipLaunchKernelGGL(k_sd_dots_combine, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     cx.stream, part, nst, nch, n, vol, out);
  INQX_HIP(hipGetLastError());
}

void launch_sd_residual(Ctx &cx, int prec, int nst, long K, const void *phi, const void *hphi, void *res,
                        const double *en) {
  ProfScope ps(cx, "sd_residual");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_sd_residual<double>, point_grid(nst, K), dim3(kThreads), 0, cx.stream,
                       (const cplx<double> *)phi, (const cplx<double> *)hphi, (cplx<double> *)res, en, K);
  else
    hipLaunchKernelGGL(k_sd_residual<float>, point_grid(nst, K), dim3(kThreads), 0, cx.stream,
                       (const cplx<float> *)phi, (const cplx<float> *)hphi, (cplx<float> *)res, en, K);
  INQX_HIP(hipGetLastError());
}

void launch_tpa_scale(Ctx &cx, int prec, int nst, long K, void *res, const double *g2h, const double *kn) {
  ProfScope ps(cx, "tpa_scale");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_tpa_scale<double>, point_grid(nst, K), dim3(kThreads), 0, cx.stream, (cplx<double> *)res,
                       g2h, kn, K);
  else
    hipLaunchKernelGGL(k_tpa_scale<float>, point_grid(nst, K), dim3(kThreads), 0, cx.stream, (cplx<float> *)res, g2h,
                       kn, K);
  INQX_HIP(hipGetLastError());
}

void launch_sd_lambda(Ctx &cx, int nst, const double *m03, const double *m45, double *lam) {
  hipLaunchKernelGGL(k_sd_lambda, dim3((unsigned)((nst + kThreads - 1) / kThreads)), dim3(kThreads), 0, cx.stream,
                     m03, m45, nst, lam);
  INQX_HIP(hipGetLastError());
}

void launch_sd_update(Ctx &cx, int prec, int nst, long K, void *phi, void *hphi, const void *res, const void *hres,
                      const double *lam) {
  ProfScope ps(cx, "sd_update");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_sd_update<double>, point_grid(nst, K), dim3(kThreads), 0, cx.stream, (cplx<double> *)phi,
                       (cplx<double> *)hphi, (const cplx<double> *)res, (const cplx<double> *)hres, lam, K);
  else
    hipLaunchKernelGGL(k_sd_update<float>, point_grid(nst, K), dim3(kThreads), 0, cx.stream, (cplx<float> *)phi,
                       (cplx<float> *)hphi, (const cplx<float> *)res, (const cplx<float> *)hres, lam, K);
  INQX_HIP(hipGetLastError());
}

}  // namespace inqx
