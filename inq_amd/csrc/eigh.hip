// eigh.hip -- FP64 Hermitian eigensolver on the device, for the subspace
// step (matrix::diagonalize, src/matrix/diagonalize.hpp:35-189, which calls
// zheev / cuSOLVER zheevd: eigenvalues ascending, eigenvectors orthonormal).
//
// Three stages, no host round trip of the matrix:
//  1. k_tridiag: Householder reduction A = Q T Q^H in one 1024-lane
//     work-group (matrix in global memory, reflector, A v and the rank-2
//     update parallel over rows and columns); T has a real diagonal and a
//     complex sub-diagonal e_k, made real by the phases delta (D^H T D).
//  2. k_tridiag_eig: eigenvalues by Sturm-count bisection (one lane per
//     eigenvalue), eigenvectors by inverse iteration on T - lambda I with a
//     partial-pivoting tridiagonal LU (one lane per eigenvector, scratch laid
//     out [component][vector] so that the lanes' accesses coalesce); vectors
//     of a cluster (eigenvalue gaps below 1e-3 ||T||_1, the criterion of
//     LAPACK's dstein) are computed in rounds and re-orthogonalised (modified
//     Gram-Schmidt) against the cluster's earlier vectors at every iteration.
//  3. k_eig_back: eigenvectors of A = Q D z, eight per work-group, reflectors
//     applied in reverse order; written as rows U[k][j] (component j of the
//     k-th eigenvector).
#include "internal.hpp"

#include <cmath>

namespace inqx {
namespace {

using cd = cplx<double>;

constexpr int kTriThreads = 1024;
constexpr int kBackCols = 8;
constexpr int kInvIters = 3;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// sum over the work-group, the same value on every lane (fixed order)
__device__ double block_sum(double v, double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------------ stage 1
// A [n][n] Hermitian (full storage, destroyed); V [n][n]: reflector k in row k,
// components 0..n-k-2 acting on indices k+1..n-1 (zero: identity); d [n] real
// diagonal of T, e [n-1] its complex sub-diagonal T[k+1][k]
__global__ void __launch_bounds__(kTriThreads) k_tridiag(cd *A, int n, cd *V, double *d, cd *e) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  cd *v = reinterpret_cast<cd *>(sm_raw);
  cd *p = v + n;
  cd *w = p + n;
  double *red = reinterpret_cast<double *>(w + n);
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1, o = k + 1;
    double s = 0.0;
    for (int i = tid; i < m; i += nt) {
      const cd x = A[(size_t)(o + i) * n + k];
      v[i] = x;
      s += x.x * x.x + x.y * x.y;
    }
    const double xn2 = block_sum(s, red);  // also orders the writes of v
    const cd x0 = v[0];
    __syncthreads();  // every lane has read v[0] before it is overwritten below
    const double a0 = x0.x * x0.x + x0.y * x0.y;
    const double tail2 = xn2 - a0;
    if (!(tail2 > 0.0)) {  // the column is already reduced: identity reflector
      for (int i = tid; i < m; i += nt) V[(size_t)k * n + i] = cd{0.0, 0.0};
      if (tid == 0) {
        e[k] = x0;
        d[k] = A[(size_t)k * n + k].x;
      }
      __syncthreads();
      continue;
    }
    const double xnorm = sqrt(xn2), ax0 = sqrt(a0);
    const cd ph = ax0 > 0.0 ? cd{x0.x / ax0, x0.y / ax0} : cd{1.0, 0.0};
    const cd alpha{-ph.x * xnorm, -ph.y * xnorm};
    // v = (x - alpha e1) / |x - alpha e1|; v0 = phase (|x0| + |x|), no cancellation
    const cd v0{x0.x - alpha.x, x0.y - alpha.y};
    const double inv = 1.0 / sqrt(tail2 + v0.x * v0.x + v0.y * v0.y);
    for (int i = tid; i < m; i += nt) {
      const cd x = i == 0 ? v0 : v[i];
      const cd u{x.x * inv, x.y * inv};
      v[i] = u;
      V[(size_t)k * n + i] = u;
    }
    if (tid == 0) {
      e[k] = alpha;
      d[k] = A[(size_t)k * n + k].x;
    }
    __syncthreads();
    // p = A22 v, one wave per row
    for (int i = wave; i < m; i += nw) {
      const cd *row = A + (size_t)(o + i) * n + o;
      double pr = 0.0, pi = 0.0;
      for (int j = lane; j < m; j += 64) {
        const cd a = row[j], b = v[j];
        pr += a.x * b.x - a.y * b.y;
        pi += a.x * b.y + a.y * b.x;
      }
      pr = wave_sum(pr);
      pi = wave_sum(pi);
      if (lane == 0) p[i] = cd{pr, pi};
    }
    __syncthreads();
    double b = 0.0;  // beta = v^H p (real for a Hermitian A)
    for (int i = tid; i < m; i += nt) b += v[i].x * p[i].x + v[i].y * p[i].y;
    const double beta = block_sum(b, red);
    for (int i = tid; i < m; i += nt) w[i] = cd{2.0 * (p[i].x - beta * v[i].x), 2.0 * (p[i].y - beta * v[i].y)};
    __syncthreads();
    // A22 -= v w^H + w v^H
    for (int i = wave; i < m; i += nw) {
      cd *row = A + (size_t)(o + i) * n + o;
      const cd vi = v[i], wi = w[i];
      for (int j = lane; j < m; j += 64) {
        const cd vj = v[j], wj = w[j];
        cd a = row[j];
        a.x -= vi.x * wj.x + vi.y * wj.y + wi.x * vj.x + wi.y * vj.y;
        a.y -= vi.y * wj.x - vi.x * wj.y + wi.y * vj.x - wi.x * vj.y;
        row[j] = a;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (n >= 2) {
      d[n - 2] = A[(size_t)(n - 2) * n + (n - 2)].x;
      e[n - 2] = A[(size_t)(n - 1) * n + (n - 2)];
    }
    d[n - 1] = A[(size_t)(n - 1) * n + (n - 1)].x;
  }
}

// ------------------------------------------------------------------ stage 2
__device__ __forceinline__ double start_value(int i, int j) {
  unsigned long long z = 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1) + (unsigned long long)j * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // U(-1, 1)
}

// number of eigenvalues of the real tridiagonal (d, er) below x
__device__ int sturm_count(const double *d, const double *er, int n, double x, double pivmin) {
  double q = d[0] - x;
  int c = q < 0.0;
  for (int j = 1; j < n; ++j) {
    if (fabs(q) < pivmin) q = -pivmin;
    q = (d[j] - x) - er[j - 1] * er[j - 1] / q;
    c += q < 0.0;
  }
  return c;
}

// d, e from stage 1; lam [n] ascending eigenvalues; delta [n] phases (T = D T_r D^H);
// Zr [n][n] eigenvectors of T_r as rows; scr [6][n][n] doubles of scratch
__global__ void __launch_bounds__(kTriThreads) k_tridiag_eig(const double *dd, const cd *e, int n, double *lam, cd *delta,
                                                          double *Zr, double *scr) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  double *d = reinterpret_cast<double *>(sm_raw);
  double *er = d + n;
  double *lv = er + n;
  int *rank = reinterpret_cast<int *>(lv + n);
  int *cstart = rank + n;
  __shared__ double s_gl, s_gu, s_tnorm;
  __shared__ int s_maxrank;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < n; j += nt) {
    d[j] = dd[j];
    er[j] = j + 1 < n ? sqrt(e[j].x * e[j].x + e[j].y * e[j].y) : 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    cd dl{1.0, 0.0};
    delta[0] = dl;
    double gl = d[0], gu = d[0], tn = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j + 1 < n) {
        const double a = er[j];
        const cd nx = a > 0.0 ? cd{(dl.x * e[j].x - dl.y * e[j].y) / a, (dl.x * e[j].y + dl.y * e[j].x) / a} : dl;
        dl = nx;
        delta[j + 1] = dl;
      }
      const double r = (j > 0 ? er[j - 1] : 0.0) + (j + 1 < n ? er[j] : 0.0);
      gl = fmin(gl, d[j] - r);
      gu = fmax(gu, d[j] + r);
      tn = fmax(tn, fabs(d[j]) + r);
    }
    s_gl = gl;
    s_gu = gu;
    s_tnorm = tn;
  }
  __syncthreads();
  double emax2 = 0.0;
  for (int j = 0; j + 1 < n; ++j) emax2 = fmax(emax2, er[j] * er[j]);
  const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emax2);
  const double tnorm = s_tnorm;
  const double eps = 2.220446049250313e-16;
  // bisection: lv[i] = i-th smallest eigenvalue.  count(lo) <= i < count(hi)
  // keeps lambda_i in [lo, hi); the bracket shrinks to adjacent doubles (or
  // below pivmin), and lo is then the largest double not above lambda_i
  // (a 1 x 1 or diagonal T gives its diagonal exactly)
  for (int i = tid; i < n; i += nt) {
    double lo = s_gl - 2.0 * eps * tnorm - pivmin, hi = s_gu + 2.0 * eps * tnorm + pivmin;
    bool done = false;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi || hi - lo <= pivmin) {
        done = true;
        break;
      }
      if (sturm_count(d, er, n, mid, pivmin) > i) hi = mid;
      else lo = mid;
    }
    lv[i] = done ? lo : 0.5 * (lo + hi);
    lam[i] = lv[i];
  }
  __syncthreads();
  // clusters (gap below 1e-3 ||T||_1) and the rank of each eigenvalue in its cluster
  if (tid == 0) {
    const double ortol = 1e-3 * tnorm;
    int r = 0, mr = 0, st = 0;
    for (int i = 0; i < n; ++i) {
      if (i > 0 && lv[i] - lv[i - 1] > ortol) {
        r = 0;
        st = i;
      }
      rank[i] = r;
      cstart[i] = st;
      mr = max(mr, r);
      ++r;
    }
    s_maxrank = mr;
  }
  __syncthreads();
  const size_t nn = (size_t)n * n;
  double *U0 = scr, *U1 = scr + nn, *U2 = scr + 2 * nn, *L = scr + 3 * nn, *X = scr + 4 * nn, *P = scr + 5 * nn;
  const double pert = eps * fmax(tnorm, pivmin);
  for (int round = 0; round <= s_maxrank; ++round) {
    for (int i = tid; i < n; i += nt) {
      if (rank[i] != round) continue;
      const double x = lv[i];
      // LU with partial pivoting of T - x I (columns of this lane: index j * n + i)
      U0[i] = d[0] - x;
      if (n > 1) U1[i] = er[0];
      for (int j = 0; j + 1 < n; ++j) {
        const size_t aj = (size_t)j * n + i, aj1 = (size_t)(j + 1) * n + i;
        const double b = er[j], a1 = d[j + 1] - x, c1 = j + 2 < n ? er[j + 1] : 0.0;
        const double u0 = U0[aj];
        if (fabs(u0) >= fabs(b)) {
          const double l = u0 != 0.0 ? b / u0 : 0.0;
          L[aj] = l;
          P[aj] = 0.0;
          U2[aj] = 0.0;
          U0[aj1] = a1 - l * U1[aj];
          U1[aj1] = c1;
        } else {
          const double l = u0 / b;
          L[aj] = l;
          P[aj] = 1.0;
          const double u1 = U1[aj];
          U0[aj] = b;
          U1[aj] = a1;
          U2[aj] = c1;
          U0[aj1] = u1 - l * a1;
          U1[aj1] = -l * c1;
        }
      }
      for (int j = 0; j < n; ++j) {
        const size_t aj = (size_t)j * n + i;
        if (fabs(U0[aj]) < pert) U0[aj] = U0[aj] >= 0.0 ? pert : -pert;
      }
      for (int j = 0; j < n; ++j) X[(size_t)j * n + i] = start_value(i, j);
      for (int it = 0; it < kInvIters; ++it) {
        // forward: the same row operations on the right-hand side
        for (int j = 0; j + 1 < n; ++j) {
          const size_t aj = (size_t)j * n + i, aj1 = (size_t)(j + 1) * n + i;
          double xj = X[aj], xj1 = X[aj1];
          if (P[aj] != 0.0) {
            const double t = xj;
            xj = xj1;
            xj1 = t;
          }
          X[aj] = xj;
          X[aj1] = xj1 - L[aj] * xj;
        }
        // back substitution with the two super-diagonals
        double y2 = 0.0, y1 = 0.0, nrm = 0.0;
        for (int j = n - 1; j >= 0; --j) {
          const size_t aj = (size_t)j * n + i;
          const double y = (X[aj] - U1[aj] * y1 - U2[aj] * y2) / U0[aj];
          X[aj] = y;
          y2 = y1;
          y1 = y;
          nrm = fmax(nrm, fabs(y));
        }
        const double sc = nrm > 0.0 ? 1.0 / nrm : 1.0;
        for (int j = 0; j < n; ++j) X[(size_t)j * n + i] *= sc;
        // re-orthogonalise against the finished vectors of the cluster
        for (int k = cstart[i]; k < i; ++k) {
          double dot = 0.0;
          for (int j = 0; j < n; ++j) dot += Zr[(size_t)k * n + j] * X[(size_t)j * n + i];
          for (int j = 0; j < n; ++j) X[(size_t)j * n + i] -= dot * Zr[(size_t)k * n + j];
        }
      }
      double s2 = 0.0;
      for (int j = 0; j < n; ++j) s2 += X[(size_t)j * n + i] * X[(size_t)j * n + i];
      const double inv = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
      // sign convention: the largest component positive
      double big = 0.0;
      for (int j = 0; j < n; ++j) {
        const double xv = X[(size_t)j * n + i];
        if (fabs(xv) > fabs(big)) big = xv;
      }
      const double sg = big < 0.0 ? -inv : inv;
      for (int j = 0; j < n; ++j) Zr[(size_t)i * n + j] = X[(size_t)j * n + i] * sg;
    }
    __threadfence_block();
    __syncthreads();
  }
}

// ------------------------------------------------------------------ stage 3
// U[i][:] = Q D z_i for eigenvectors i of this work-group
__global__ void __launch_bounds__(256) k_eig_back(const cd *V, int n, const cd *delta, const double *Zr, cd *U) {
  extern __shared__ __align__(16) unsigned char sm_raw[];
  cd *y = reinterpret_cast<cd *>(sm_raw);                     // [kBackCols][n]
  double *red = reinterpret_cast<double *>(y + (size_t)kBackCols * n);  // [kBackCols][2][4]
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.x * kBackCols;
  const int nc = min(kBackCols, n - i0);
  for (int idx = tid; idx < nc * n; idx += nt) {
    const int c = idx / n, j = idx - c * n;
    const double z = Zr[(size_t)(i0 + c) * n + j];
    const cd dl = delta[j];
    y[(size_t)c * n + j] = cd{dl.x * z, dl.y * z};
  }
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1, o = k + 1;
    const cd *v = V + (size_t)k * n;
    // s_c = v^H y_c[o..]: wave c (of 4) handles vectors c, c + 4
    for (int c = wave; c < nc; c += 4) {
      double sr = 0.0, si = 0.0;
      for (int j = lane; j < m; j += 64) {
        const cd a = v[j], b = y[(size_t)c * n + o + j];
        sr += a.x * b.x + a.y * b.y;
        si += a.x * b.y - a.y * b.x;
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) {
        red[2 * c] = sr;
        red[2 * c + 1] = si;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < nc * m; idx += nt) {
      const int c = idx / m, j = idx - c * m;
      const cd a = v[j];
      const double sr = 2.0 * red[2 * c], si = 2.0 * red[2 * c + 1];
      cd &t = y[(size_t)c * n + o + j];
      t.x -= a.x * sr - a.y * si;
      t.y -= a.x * si + a.y * sr;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nc * n; idx += nt) {
    const int c = idx / n, j = idx - c * n;
    U[(size_t)(i0 + c) * n + j] = y[(size_t)c * n + j];
  }
}

__global__ void __launch_bounds__(256) k_conj(cd *a, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) a[i].y = -a[i].y;
}

}  // namespace

void launch_conj(Ctx &cx, void *a, long n) {
  hipLaunchKernelGGL(k_conj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.stream, reinterpret_cast<cd *>(a), n);
  INQX_HIP(hipGetLastError());
}

size_t eigh_workspace_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return sizeof(cd) * nn          // V
         + sizeof(double) * nn    // Zr
         + sizeof(double) * 6 * nn  // scratch
         + sizeof(double) * 2 * n + sizeof(cd) * 2 * n + 1024;
}

// A [n][n] Hermitian (device, destroyed) -> lam [n] ascending (device), U [n][n]
// rows = eigenvectors (device); ws of eigh_workspace_bytes(n)
void launch_eigh(Ctx &cx, int n, void *A, double *lam, void *U, void *ws) {
  INQX_CHECK(n >= 1 && n <= kMaxEigN, INQX_ERR_UNSUPPORTED, "the device eigensolver takes n <= 1024");
  const size_t nn = (size_t)n * n;
  unsigned char *p = static_cast<unsigned char *>(ws);
  cd *V = reinterpret_cast<cd *>(p);
  p += sizeof(cd) * nn;
  double *Zr = reinterpret_cast<double *>(p);
  p += sizeof(double) * nn;
  double *scr = reinterpret_cast<double *>(p);
  p += sizeof(double) * 6 * nn;
  double *d = reinterpret_cast<double *>(p);
  p += sizeof(double) * n;
  cd *e = reinterpret_cast<cd *>(p);
  p += sizeof(cd) * n;
  cd *delta = reinterpret_cast<cd *>(p);
  if (n == 1) {
    // 1 x 1: lam = Re A, U = 1
    hipLaunchKernelGGL(k_tridiag, dim3(1), dim3(64), 3 * sizeof(cd) + 16 * sizeof(double), cx.stream,
                       reinterpret_cast<cd *>(A), n, V, d, e);
  } else {
    ProfScope ps(cx, "eigh_tridiag");
    hipLaunchKernelGGL(k_tridiag, dim3(1), dim3(kTriThreads), 3 * n * sizeof(cd) + 16 * sizeof(double), cx.stream,
                       reinterpret_cast<cd *>(A), n, V, d, e);
  }
  INQX_HIP(hipGetLastError());
  {
    ProfScope ps(cx, "eigh_tridiag_eig");
    hipLaunchKernelGGL(k_tridiag_eig, dim3(1), dim3(kTriThreads), 3 * n * sizeof(double) + 2 * n * sizeof(int),
                       cx.stream, d, e, n, lam, delta, Zr, scr);
  }
  INQX_HIP(hipGetLastError());
  {
    ProfScope ps(cx, "eigh_back");
    const int nb = (n + kBackCols - 1) / kBackCols;
    const size_t lds = sizeof(cd) * kBackCols * n + sizeof(double) * 2 * kBackCols;
    static DeviceOnce attr;
    if (!attr.test_and_set(cx.device))
      INQX_HIP(hipFuncSetAttribute((const void *)k_eig_back, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(cd) * kBackCols * kMaxEigN + sizeof(double) * 2 * kBackCols)));
    hipLaunchKernelGGL(k_eig_back, dim3(nb), dim3(256), lds, cx.stream, V, n, delta, Zr, reinterpret_cast<cd *>(U));
  }
  INQX_HIP(hipGetLastError());
}

}  // namespace inqx
