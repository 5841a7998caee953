// linalg.hip -- contractions over band-major sphere states on MFMA, and the
// FP64 small-matrix steps of orthogonalisation.
//
//   overlap   S = vol * A^H B   (operations::overlap, src/operations/overlap.hpp:26-41)
//             A [n1][K], B [n2][K] complex; split-K over K with
//             v_mfma_{f32,f64}_16x16x4 on the interleaved (re, im) storage read
//             as a real K' = 2K contraction:
//               Re S = sum_k' A'[i][k'] B'[j][k'],
//               Im S = sum_k' A'[i][k'] B''[j][k'],  B''[j][k'] = (+/-) B'[j][k' ^ 1]
//             Partial tiles are reduced in FP64 in a fixed order.
//   rotate    Out = M C (+ beta Out): C [n][K] complex, M [n][n] complex
//             (operations::rotate, src/operations/rotate.hpp:26-156), MFMA with
//             the real operands A' = [Mr, -Mi], A'' = [Mi, Mr] interleaved along j.
//   cholesky  FP64 lower Cholesky S = L L^H, upper zeroed (matrix::cholesky,
//             src/matrix/cholesky.hpp:28-100), blocked in 32-wide panels.
//   trinv     FP64 inverse of a lower-triangular matrix (matrix::invert_triangular,
//             src/matrix/invert_triangular.hpp:20-50), blocked by 32x32 block rows.
//
// MFMA 16x16x4 lane layout (CDNA4): lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives column l%16 of D in its four accumulators; the
// accumulator-to-row map differs between the f32 and f64 forms (mfma_t::row),
// and tests/test_gpu_linalg.py checks both against numpy.
#include "internal.hpp"

#include <algorithm>
#include <cstring>

#if defined(__HIP_DEVICE_COMPILE__)
#define GAS __attribute__((address_space(1)))
#else
#define GAS
#endif

namespace inqx {

template <typename T> __device__ __forceinline__ const GAS T *lgp(const T *p) { return (const GAS T *)p; }
template <typename T> __device__ __forceinline__ GAS T *lgw(T *p) { return (GAS T *)p; }

typedef float floatx4 __attribute__((ext_vector_type(4)));
struct alignas(16) V4f {
  float x, y, z, w;
};
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef double doublex2 __attribute__((ext_vector_type(2)));

// row of the 16x16 output tile held in accumulator r of a lane (column = lane % 16)
template <typename T> struct mfma_t;
template <> struct mfma_t<float> {
  using acc = floatx4;
  using pair = floatx2;
  using v16 = floatx4;
  static __device__ __forceinline__ acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct mfma_t<double> {
  using acc = doublex4;
  using pair = doublex2;
  using v16 = doublex2;
  static __device__ __forceinline__ acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// ------------------------------------------------------------------------ overlap
constexpr int kOvTile = 64;   // output tile edge per work-group
constexpr int kOvKC = 32;     // real K per LDS stage
constexpr int kOvChunkMax = 2048;  // real K per work-group (split K), upper bound
constexpr int kOvChunkMin = 256;   // lower bound when few output tiles call for more chunks
constexpr int kOvGridTarget = 1024;  // work-groups that keep every CU busy with 4 of them

constexpr size_t kOvPartialMax = size_t(1) << 30;  // bytes of partial tiles at most

// real K per work-group: the full 2048 when the output has enough tiles, else
// shorter chunks (a multiple of the LDS stage) so that the grid fills the chip;
// longer ones when the partial tiles (nchunk x tiles x 64 x 64 complex) would
// pass kOvPartialMax (thousands of states over a long K)
static long overlap_chunk(long K2, int tiles, size_t es) {
  const long want = std::max<long>(1, (kOvGridTarget + tiles - 1) / tiles);
  long c = (K2 + want - 1) / want;
  c = std::min<long>(kOvChunkMax, std::max<long>(kOvChunkMin, c));
  const size_t tile_bytes = es * (size_t)tiles * kOvTile * kOvTile * 2;
  const long nchunk_max = std::max<long>(1, (long)(kOvPartialMax / tile_bytes));
  c = std::max<long>(c, (K2 + nchunk_max - 1) / nchunk_max);
  return (c + kOvKC - 1) / kOvKC * kOvKC;
}

// partial[chunk][tile_i][tile_j] holds a 64x64 complex tile in T.  Each stage
// loads (re, im) pairs (rows start at even real offsets, so pairs are aligned),
// and the next stage's pairs are loaded into registers while the matrix cores
// work on the current one.
template <typename T>
__global__ void __launch_bounds__(256) k_overlap_partial(const T *A, const T *B, int n1, int n2, long K2,
                                                         int tiles_i, int tiles_j, int kchunk, T *partial) {
  using M = mfma_t<T>;
  using P2 = typename M::pair;
  __shared__ T At[kOvTile][kOvKC + 1];
  __shared__ T Bt[kOvTile][kOvKC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tj = blockIdx.x % tiles_j, ti = (blockIdx.x / tiles_j) % tiles_i, chunk = blockIdx.x / (tiles_j * tiles_i);
  const long k_begin = (long)chunk * kchunk;
  const long k_end = min(k_begin + kchunk, K2);
  const int i0 = ti * kOvTile, j0 = tj * kOvTile;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;  // wave sub-tile 32x32
  typename M::acc re[2][2], im[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      re[a][b] = typename M::acc{0, 0, 0, 0};
      im[a][b] = typename M::acc{0, 0, 0, 0};
    }
  const GAS T *Ag = lgp(A), *Bg = lgp(B);
  // stage = 64 rows x 32 real k' of A and of B = 64 x 16 pairs each, 4 pairs per lane;
  // K2 and kchunk are even, so k < k_end implies k + 1 < k_end
  constexpr int kPairs = kOvTile * kOvKC / 2 / 256;
  P2 ra[kPairs], rb[kPairs];
  auto load_stage = [&](long kb) {
#pragma unroll
    for (int q = 0; q < kPairs; ++q) {
      const int e = tid + 256 * q, r = e >> 4;
      const long k = kb + 2 * (e & 15);
      const int ia = i0 + r, jb = j0 + r;
      ra[q] = (ia < n1 && k < k_end) ? *(const GAS P2 *)(Ag + (long)ia * K2 + k) : P2{0, 0};
      rb[q] = (jb < n2 && k < k_end) ? *(const GAS P2 *)(Bg + (long)jb * K2 + k) : P2{0, 0};
    }
  };
  load_stage(k_begin);
  for (long kb = k_begin; kb < k_end; kb += kOvKC) {
#pragma unroll
    for (int q = 0; q < kPairs; ++q) {
      const int e = tid + 256 * q, r = e >> 4, c = 2 * (e & 15);
      At[r][c] = ra[q].x;
      At[r][c + 1] = ra[q].y;
      Bt[r][c] = rb[q].x;
      Bt[r][c + 1] = rb[q].y;
    }
    __syncthreads();
    if (kb + kOvKC < k_end) load_stage(kb + kOvKC);
#pragma unroll
    for (int ks = 0; ks < kOvKC; ks += 4) {
      const int kk = ks + (lane >> 4);
      const T sgn = (kk & 1) ? (T)-1 : (T)1;
      T a[2], bp[2], bq[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = At[wr + 16 * t + (lane & 15)][kk];
        bp[t] = Bt[wc + 16 * t + (lane & 15)][kk];
        bq[t] = sgn * Bt[wc + 16 * t + (lane & 15)][kk ^ 1];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          re[x][y] = M::run(a[x], bp[y], re[x][y]);
          im[x][y] = M::run(a[x], bq[y], im[x][y]);
        }
    }
    __syncthreads();
  }
  // write the 64x64 complex partial tile: row i = wr + 16x + 4(l/16) + r, col j = wc + 16y + l%16
  GAS T *P = lgw(partial) + ((long)blockIdx.x) * kOvTile * kOvTile * 2;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = wr + 16 * x + M::row(lane, r), j = wc + 16 * y + (lane & 15);
        P[(i * kOvTile + j) * 2 + 0] = re[x][y][r];
        P[(i * kOvTile + j) * 2 + 1] = im[x][y][r];
      }
}

// S[i][j] = vol * sum_chunks partial in FP64, in a fixed order: work-group = one
// 64-wide row of one output tile; lane group g (of 4) sums chunks g, g+4, g+8, ...
// in ascending order, and the four group sums are added g = 0..3.
constexpr int kOvRedGroups = 4;
template <typename T>
__global__ void __launch_bounds__(256) k_overlap_reduce(const T *partial, int n1, int n2, int tiles_i, int tiles_j,
                                                        int nchunk, double vol, double *S) {
  using P2 = typename mfma_t<T>::pair;
  __shared__ double red[kOvRedGroups][kOvTile][2];
  const int jj = threadIdx.x & (kOvTile - 1), g = threadIdx.x >> 6;
  const int ii = blockIdx.x % kOvTile, tile = blockIdx.x / kOvTile;
  const int tj = tile % tiles_j, ti = tile / tiles_j;
  const long tile_stride = (long)tiles_i * tiles_j * kOvTile * kOvTile;  // P2 per chunk
  const GAS P2 *P = (const GAS P2 *)lgp(partial) + (long)tile * kOvTile * kOvTile + ii * kOvTile + jj;
  double sr = 0.0, si = 0.0;
#pragma unroll 8
  for (int c = g; c < nchunk; c += kOvRedGroups) {
    const P2 v = P[(long)c * tile_stride];
    sr += (double)v.x;
    si += (double)v.y;
  }
  red[g][jj][0] = sr;
  red[g][jj][1] = si;
  __syncthreads();
  const int i = ti * kOvTile + ii, j = tj * kOvTile + jj;
  if (g == 0 && i < n1 && j < n2) {
    double tr = red[0][jj][0], tim = red[0][jj][1];
#pragma unroll
    for (int q = 1; q < kOvRedGroups; ++q) {
      tr += red[q][jj][0];
      tim += red[q][jj][1];
    }
    GAS double *Sg = lgw(S);
    Sg[2 * ((long)i * n2 + j)] = vol * tr;
    Sg[2 * ((long)i * n2 + j) + 1] = vol * tim;
  }
}

// ------------------------------------------------------------------------ rotate
// operand matrices Ap = [Mr, -Mi], Aq = [Mi, Mr] interleaved along j: [nr_pad][ld] in T,
// rows padded to a multiple of 64 and ld = 2 nc rounded up to 32, the padding zero
// mode 0: M = conj(R) (Phi <- Phi R^H);  mode 1: M = alpha R^T (5-argument rotate,
// R is [nc][nr]); mode 2: M = R
template <typename T>
__global__ void k_rot_prep(const double *R, int nr, int nc, int mode, double ar, double ai, int nr_pad, int ld, T *Ap,
                           T *Aq) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int half = ld / 2;
  if (idx >= (long)nr_pad * half) return;
  const int i = (int)(idx / half), j = (int)(idx - (long)i * half);
  double mr = 0.0, mi = 0.0;
  if (i < nr && j < nc) {
    if (mode == 0) {
      mr = R[2 * ((long)i * nc + j)];
      mi = -R[2 * ((long)i * nc + j) + 1];
    } else if (mode == 1) {
      const double rr = R[2 * ((long)j * nr + i)], ri = R[2 * ((long)j * nr + i) + 1];
      mr = ar * rr - ai * ri;
      mi = ar * ri + ai * rr;
    } else {
      mr = R[2 * ((long)i * nc + j)];
      mi = R[2 * ((long)i * nc + j) + 1];
    }
  }
  Ap[(long)i * ld + 2 * j] = (T)mr;
  Ap[(long)i * ld + 2 * j + 1] = (T)-mi;
  Aq[(long)i * ld + 2 * j] = (T)mi;
  Aq[(long)i * ld + 2 * j + 1] = (T)mr;
}

constexpr int kRotJC = 16;  // complex j per LDS stage (32 real k')

// Out[i][p] = sum_j M[i][j] C[j][p] (+ beta Out[i][p]), i < nr, j < n; work-group = 64 rows x 64 columns p.
// Within a stage, lane group q = lane / 16 of an MFMA step s holds real k' = 8q + s (any
// order of k' is the same sum), so a lane's eight A values per stage are contiguous and
// come in 16-byte loads.  The next stage's A values and C tile are loaded into registers
// while the matrix cores work on the current one.
template <typename T>
__global__ void __launch_bounds__(256) k_rotate(const T *Ap, const T *Aq, int ld, const cplx<T> *Cm, int nr, int n,
                                                long K, cplx<T> *Out, double br, double bi, int tiles_p, int tiles_i) {
  using M = mfma_t<T>;
  using C = cplx<T>;
  using V = typename M::v16;
  using P2 = typename M::pair;  // one complex value as a vector (loads stay in registers)
  constexpr int PV = 16 / sizeof(T);  // T per 16-byte load
  constexpr int NV = 8 / PV;          // 16-byte loads per lane per operand and stage
  __shared__ P2 Ct[kRotJC][64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // Work-groups are dealt round-robin to the 8 XCDs; renumber so that each XCD runs a
  // contiguous range, in which the tiles_i row tiles of one column block p are
  // neighbours and read that block of C through the same L2.
  const int G = gridDim.x, b = blockIdx.x, xq = G / 8, xr = G % 8, x = b % 8;
  const int logical = (x < xr ? x * (xq + 1) : xr * (xq + 1) + (x - xr) * xq) + b / 8;
  const int ti = logical % tiles_i, tp = logical / tiles_i;
  const long p0 = (long)tp * 64;
  const int i0 = ti * 64 + wave * 16;  // this wave's 16 rows
  typename M::acc re[4], im[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    re[c] = typename M::acc{0, 0, 0, 0};
    im[c] = typename M::acc{0, 0, 0, 0};
  }
  const int q = lane >> 4;
  // rows are padded to the tile and ld covers every stage, so these loads need no check
  const GAS V *apv = (const GAS V *)(lgp(Ap) + (long)(i0 + (lane & 15)) * ld + 8 * q);
  const GAS V *aqv = (const GAS V *)(lgp(Aq) + (long)(i0 + (lane & 15)) * ld + 8 * q);
  const GAS P2 *Cg = (const GAS P2 *)lgp(Cm);
  V an[NV], bn[NV];
  P2 cn[4];
  auto load_stage = [&](int jb) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      an[v] = apv[(2 * jb) / PV + v];
      bn[v] = aqv[(2 * jb) / PV + v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, r = e >> 6, c = e & 63;
      const int j = jb + r;
      const long p = p0 + c;
      cn[u] = (j < n && p < K) ? Cg[(long)j * K + p] : P2{0, 0};
    }
  };
  load_stage(0);
  for (int jb = 0; jb < n; jb += kRotJC) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      Ct[e >> 6][e & 63] = cn[u];
    }
    V ac[NV], bc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      ac[v] = an[v];
      bc[v] = bn[v];
    }
    __syncthreads();
    if (jb + kRotJC < n) load_stage(jb + kRotJC);
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const T ap = ac[s / PV][s % PV], aq = bc[s / PV][s % PV];
      const int jl = 4 * q + (s >> 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const P2 v = Ct[jl][16 * c + (lane & 15)];
        const T bv = (s & 1) ? v.y : v.x;
        re[c] = M::run(ap, bv, re[c]);
        im[c] = M::run(aq, bv, im[c]);
      }
    }
    __syncthreads();
  }
  GAS C *Og = lgw(Out);
  const bool beta = (br != 0.0 || bi != 0.0);
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + M::row(lane, r);
      const long p = p0 + 16 * c + (lane & 15);
      if (i < nr && p < K) {
        C o{re[c][r], im[c][r]};
        if (beta) {
          const C qv = Og[(long)i * K + p];
          o.x += (T)(br * (double)qv.x - bi * (double)qv.y);
          o.y += (T)(br * (double)qv.y + bi * (double)qv.x);
        }
        Og[(long)i * K + p] = o;
      }
    }
}

// ------------------------------------------------------------------------ Cholesky / inverse (FP64)
// Blocked in 32-wide panels, one launch per step so that every read of a value
// another work-group wrote happens in a later launch.
//
// Cholesky S = L L^H in place (lower; the strict upper triangle is zeroed), per panel c0:
//   k_chol_diag    factor the diagonal block (already updated by earlier panels) in LDS
//   k_chol_panel   rows below: L[i][c0+j] = (S[i][c0+j] - sum_{k<j} L[i][c0+k] conj(L_dd[j][k])) / L_dd[j][j]
//   k_chol_update  S[i][j] -= sum_k L[i][c0+k] conj(L[j][c0+k]) on the lower trailing matrix
// info[0] = 0 or the first non-positive leading minor (LAPACK convention); the first
// diagonal launch writes it, and every later launch returns at once when it is set.
constexpr int kChB = 32;

// right-looking in LDS with one barrier per column: every lane reads the same pivot d,
// column j (scaled by 1/sqrt(d)) goes straight to S, and the lanes update the lower
// block right of column j from the unscaled column (column j is not written in step j)
__global__ void __launch_bounds__(256) k_chol_diag(double *S, int n, int c0, int *info) {
  __shared__ double Dr[kChB][kChB + 1], Di[kChB][kChB + 1];
  const int tid = threadIdx.x;
  if (c0 > 0 && info[0] != 0) return;
  const int b = min(kChB, n - c0);
  for (int e = tid; e < kChB * kChB; e += 256) {
    const int i = e / kChB, j = e - i * kChB;
    double vr = 0.0, vi = 0.0;
    if (i < b && j <= i) {
      vr = S[2 * ((long)(c0 + i) * n + c0 + j)];
      vi = S[2 * ((long)(c0 + i) * n + c0 + j) + 1];
    }
    Dr[i][j] = vr;
    Di[i][j] = vi;
  }
  __syncthreads();
  int bad = 0;
  for (int j = 0; j < b; ++j) {
    const double d = Dr[j][j];
    if (!(d > 0.0)) {
      bad = c0 + j + 1;
      break;
    }
    const double ljj = sqrt(d), rl = 1.0 / ljj;
    if (tid < b - j) {
      const int i = j + tid;
      const long o = 2 * ((long)(c0 + i) * n + c0 + j);
      S[o] = tid == 0 ? ljj : Dr[i][j] * rl;
      S[o + 1] = tid == 0 ? 0.0 : Di[i][j] * rl;
    }
    const double rd = rl * rl;  // L[i][j] conj(L[k][j]) = D[i][j] conj(D[k][j]) / d
#pragma unroll
    for (int e = tid; e < kChB * kChB; e += 256) {
      const int i = e / kChB, k = e - i * kChB;
      if (k > j && k <= i && i < b) {
        const double ar = Dr[i][j] * rd, ai = Di[i][j] * rd, br = Dr[k][j], bi = Di[k][j];
        Dr[i][k] -= ar * br + ai * bi;
        Di[i][k] -= ai * br - ar * bi;
      }
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0) info[0] = bad;
    return;
  }
  if (c0 == 0 && tid == 0) info[0] = 0;
  for (int e = tid; e < b * b; e += 256) {  // upper part of the diagonal block
    const int i = e / b, j = e - i * b;
    if (j > i) {
      S[2 * ((long)(c0 + i) * n + c0 + j)] = 0.0;
      S[2 * ((long)(c0 + i) * n + c0 + j) + 1] = 0.0;
    }
  }
  // rows of this panel right of the diagonal block are upper triangle: zero
  const int w = n - c0 - b;
  for (int e = tid; e < b * w; e += 256) {
    const int i = e / w, j = e - i * w;
    S[2 * ((long)(c0 + i) * n + c0 + b + j)] = 0.0;
    S[2 * ((long)(c0 + i) * n + c0 + b + j) + 1] = 0.0;
  }
}

// n <= kCholSmall: the whole factorisation in one work-group, the lower triangle
// packed in LDS (row i at i(i+1)/2, n(n+1)/2 x 16 bytes <= 132 KB).  Right-looking
// with one barrier per column: the lanes update the trailing lower triangle from
// the unscaled column j, then, after the barrier, scale column j while the next
// column's update (which reads column j + 1 and writes columns > j + 1) runs.
constexpr int kCholSmall = 128;
__global__ void __launch_bounds__(1024) k_chol_small(double *S, int n, int *info) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2 *T = reinterpret_cast<double2 *>(smem_raw);
  const int tid = threadIdx.x, ty = tid >> 5, tx = tid & 31;
  for (int i = ty; i < n; i += 32)
    for (int k = tx; k <= i; k += 32) T[i * (i + 1) / 2 + k] = double2{S[2 * ((long)i * n + k)], S[2 * ((long)i * n + k) + 1]};
  __syncthreads();
  int bad = 0;
  for (int j = 0; j < n; ++j) {
    const double d = T[j * (j + 1) / 2 + j].x;
    if (!(d > 0.0)) {
      bad = j + 1;
      break;
    }
    const double rd = 1.0 / d;
    for (int i = j + 1 + ty; i < n; i += 32) {
      const double2 a = T[i * (i + 1) / 2 + j];
      const double ar = a.x * rd, ai = a.y * rd;
      for (int k = j + 1 + tx; k <= i; k += 32) {  // L[i][j] conj(L[k][j]) = D[i][j] conj(D[k][j]) / d
        const double2 b = T[k * (k + 1) / 2 + j];
        double2 &t = T[i * (i + 1) / 2 + k];
        t.x -= ar * b.x + ai * b.y;
        t.y -= ai * b.x - ar * b.y;
      }
    }
    __syncthreads();
    if (tid < n - j) {
      const int i = j + tid;
      double2 &c = T[i * (i + 1) / 2 + j];
      const double ljj = sqrt(d);
      if (tid == 0) {
        c = double2{ljj, 0.0};
      } else {
        const double rl = 1.0 / ljj;
        c.x *= rl;
        c.y *= rl;
      }
    }
  }
  __syncthreads();
  if (tid == 0) info[0] = bad;
  if (bad) return;
  for (int e = tid; e < n * n; e += 1024) {
    const int i = e / n, k = e - i * n;
    const double2 v = k <= i ? T[i * (i + 1) / 2 + k] : double2{0.0, 0.0};
    S[2 * (long)e] = v.x;
    S[2 * (long)e + 1] = v.y;
  }
}

// one lane per row i >= c0 + 32 (launched only when such rows exist, so the panel is 32 wide)
__global__ void __launch_bounds__(256) k_chol_panel(double *S, int n, int c0, const int *info) {
  __shared__ double Lr[kChB][kChB + 1], Li[kChB][kChB + 1], rdiag[kChB];
  if (info[0] != 0) return;
  for (int e = threadIdx.x; e < kChB * kChB; e += 256) {
    const int i = e / kChB, j = e - i * kChB;
    Lr[i][j] = j <= i ? S[2 * ((long)(c0 + i) * n + c0 + j)] : 0.0;
    Li[i][j] = j <= i ? S[2 * ((long)(c0 + i) * n + c0 + j) + 1] : 0.0;
  }
  if (threadIdx.x < kChB) rdiag[threadIdx.x] = 1.0 / S[2 * ((long)(c0 + threadIdx.x) * n + c0 + threadIdx.x)];
  __syncthreads();
  const int i = c0 + kChB + blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double *row = S + 2 * ((long)i * n + c0);
  double xr[kChB], xi[kChB];
#pragma unroll
  for (int j = 0; j < kChB; ++j) {
    xr[j] = row[2 * j];
    xi[j] = row[2 * j + 1];
  }
  // right-looking: once x_j is final, x_m -= x_j conj(L_dd[m][j]) for every m > j
  // (independent updates, short dependency chains)
#pragma unroll
  for (int j = 0; j < kChB; ++j) {
    xr[j] *= rdiag[j];
    xi[j] *= rdiag[j];
#pragma unroll
    for (int m = j + 1; m < kChB; ++m) {
      xr[m] -= xr[j] * Lr[m][j] + xi[j] * Li[m][j];
      xi[m] -= xi[j] * Lr[m][j] - xr[j] * Li[m][j];
    }
  }
#pragma unroll
  for (int j = 0; j < kChB; ++j) {
    row[2 * j] = xr[j];
    row[2 * j + 1] = xi[j];
  }
}

// 32x32 output tile (ti, tj), tj <= ti, of the trailing matrix starting at c0 + 32;
// lane: row tid / 8, columns tid % 8 + 8 q
__global__ void __launch_bounds__(256) k_chol_update(double *S, int n, int c0, const int *info) {
  __shared__ double Ar[kChB][kChB + 1], Ai[kChB][kChB + 1], Br[kChB][kChB + 1], Bi[kChB][kChB + 1];
  const int ti = blockIdx.y, tj = blockIdx.x;
  if (tj > ti || info[0] != 0) return;
  const int r0 = c0 + kChB, i0 = r0 + ti * kChB, j0 = r0 + tj * kChB;
  for (int e = threadIdx.x; e < kChB * kChB; e += 256) {
    const int r = e / kChB, k = e - r * kChB;
    const bool ia = i0 + r < n, jb = j0 + r < n;
    Ar[r][k] = ia ? S[2 * ((long)(i0 + r) * n + c0 + k)] : 0.0;
    Ai[r][k] = ia ? S[2 * ((long)(i0 + r) * n + c0 + k) + 1] : 0.0;
    Br[r][k] = jb ? S[2 * ((long)(j0 + r) * n + c0 + k)] : 0.0;
    Bi[r][k] = jb ? S[2 * ((long)(j0 + r) * n + c0 + k) + 1] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x >> 3, cb = threadIdx.x & 7, i = i0 + r;
  double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
#pragma unroll 8
  for (int k = 0; k < kChB; ++k) {  // A[r][k] conj(B[c][k]); out-of-range rows are zero
    const double ar = Ar[r][k], ai = Ai[r][k];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cb + 8 * q;
      sr[q] += ar * Br[c][k] + ai * Bi[c][k];
      si[q] += ai * Br[c][k] - ar * Bi[c][k];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = j0 + cb + 8 * q;
    if (i < n && j <= i) {
      S[2 * ((long)i * n + j)] -= sr[q];
      S[2 * ((long)i * n + j) + 1] -= si[q];
    }
  }
}

// X = L^{-1} for lower-triangular L (complex double), blocked in 32x32 blocks:
//   k_trinv_diag  X_bb = L_bb^{-1} (one lane per column, forward substitution in LDS),
//                 blocks above the diagonal zeroed
//   k_trinv_row   block row r, one work-group per block column jb < r:
//                 X_r,jb = -X_rr sum_{k=jb}^{r-1} L_rk X_k,jb  (every X read was written
//                 by an earlier launch)
__global__ void __launch_bounds__(256) k_trinv_diag(const double *L, int n, double *X) {
  __shared__ double Lr[kChB][kChB + 1], Li[kChB][kChB + 1], Xr[kChB][kChB + 1], Xi[kChB][kChB + 1];
  const int bi = blockIdx.y, bj = blockIdx.x, tid = threadIdx.x;
  const int r0 = bi * kChB, q0 = bj * kChB;
  const int b = min(kChB, n - r0), bc = min(kChB, n - q0);
  if (bj > bi) {
    for (int e = tid; e < b * bc; e += 256) {
      const int i = e / bc, j = e - i * bc;
      X[2 * ((long)(r0 + i) * n + q0 + j)] = 0.0;
      X[2 * ((long)(r0 + i) * n + q0 + j) + 1] = 0.0;
    }
    return;
  }
  if (bj < bi) return;
  for (int e = tid; e < kChB * kChB; e += 256) {
    const int i = e / kChB, j = e - i * kChB;
    const bool in = i < b && j <= i;
    Lr[i][j] = in ? L[2 * ((long)(r0 + i) * n + r0 + j)] : 0.0;
    Li[i][j] = in ? L[2 * ((long)(r0 + i) * n + r0 + j) + 1] : 0.0;
    Xr[i][j] = 0.0;
    Xi[i][j] = 0.0;
  }
  __shared__ double Ir[kChB], Ii[kChB];  // 1 / L[i][i]
  if (tid < b) {
    const double lr = L[2 * ((long)(r0 + tid) * n + r0 + tid)], li = L[2 * ((long)(r0 + tid) * n + r0 + tid) + 1];
    const double rd = 1.0 / (lr * lr + li * li);
    Ir[tid] = lr * rd;
    Ii[tid] = -li * rd;
  }
  __syncthreads();
  const int c = tid;
  if (c < b) {
    Xr[c][c] = Ir[c];
    Xi[c][c] = Ii[c];
    for (int i = c + 1; i < b; ++i) {
      double sr = 0.0, si = 0.0;
      for (int k = c; k < i; ++k) {
        sr += Lr[i][k] * Xr[k][c] - Li[i][k] * Xi[k][c];
        si += Lr[i][k] * Xi[k][c] + Li[i][k] * Xr[k][c];
      }
      Xr[i][c] = -(sr * Ir[i] - si * Ii[i]);  // -(s / l)
      Xi[i][c] = -(sr * Ii[i] + si * Ir[i]);
    }
  }
  __syncthreads();
  for (int e = tid; e < b * b; e += 256) {
    const int i = e / b, j = e - i * b;
    X[2 * ((long)(r0 + i) * n + r0 + j)] = Xr[i][j];
    X[2 * ((long)(r0 + i) * n + r0 + j) + 1] = Xi[i][j];
  }
}

__global__ void __launch_bounds__(256) k_trinv_row(const double *L, int n, int r, double *X) {
  __shared__ double Ar[kChB][kChB + 1], Ai[kChB][kChB + 1], Br[kChB][kChB + 1], Bi[kChB][kChB + 1];
  const int jb = blockIdx.x, tid = threadIdx.x;
  const int r0 = r * kChB, q0 = jb * kChB;
  const int row = tid >> 3, cb = tid & 7;
  double tr[4] = {0, 0, 0, 0}, tim[4] = {0, 0, 0, 0};
  // the next k block's L and X values are loaded into registers during the products
  double par[4], pai[4], pbr[4], pbi[4];
  auto fetch = [&](int k) {
    const int k0 = k * kChB;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, a = e / kChB, m = e - a * kChB;
      const bool la = r0 + a < n && k0 + m < n, xb = k0 + a < n && q0 + m < n;
      par[u] = la ? L[2 * ((long)(r0 + a) * n + k0 + m)] : 0.0;
      pai[u] = la ? L[2 * ((long)(r0 + a) * n + k0 + m) + 1] : 0.0;
      pbr[u] = xb ? X[2 * ((long)(k0 + a) * n + q0 + m)] : 0.0;
      pbi[u] = xb ? X[2 * ((long)(k0 + a) * n + q0 + m) + 1] : 0.0;
    }
  };
  fetch(jb);
  for (int k = jb; k < r; ++k) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u, a = e / kChB, m = e - a * kChB;
      Ar[a][m] = par[u];
      Ai[a][m] = pai[u];
      Br[a][m] = pbr[u];
      Bi[a][m] = pbi[u];
    }
    __syncthreads();
    if (k + 1 < r) fetch(k + 1);
#pragma unroll 8
    for (int m = 0; m < kChB; ++m) {
      const double ar = Ar[row][m], ai = Ai[row][m];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = cb + 8 * q;
        tr[q] += ar * Br[m][c] - ai * Bi[m][c];
        tim[q] += ar * Bi[m][c] + ai * Br[m][c];
      }
    }
    __syncthreads();
  }
  // Y = -X_rr T: T to LDS (B), X_rr to LDS (A)
  for (int e = tid; e < kChB * kChB; e += 256) {
    const int a = e / kChB, m = e - a * kChB;
    const bool xa = r0 + a < n && r0 + m < n;
    Ar[a][m] = xa ? X[2 * ((long)(r0 + a) * n + r0 + m)] : 0.0;
    Ai[a][m] = xa ? X[2 * ((long)(r0 + a) * n + r0 + m) + 1] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    Br[row][cb + 8 * q] = tr[q];
    Bi[row][cb + 8 * q] = tim[q];
  }
  __syncthreads();
  double yr[4] = {0, 0, 0, 0}, yi[4] = {0, 0, 0, 0};
#pragma unroll 8
  for (int m = 0; m < kChB; ++m) {
    const double ar = Ar[row][m], ai = Ai[row][m];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cb + 8 * q;
      yr[q] += ar * Br[m][c] - ai * Bi[m][c];
      yi[q] += ar * Bi[m][c] + ai * Br[m][c];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = cb + 8 * q;
    if (r0 + row < n && q0 + c < n) {
      X[2 * ((long)(r0 + row) * n + q0 + c)] = -yr[q];
      X[2 * ((long)(r0 + row) * n + q0 + c) + 1] = -yi[q];
    }
  }
}

// ------------------------------------------------------------------------ HBM copy (calibration)
// plain 16-byte-per-lane device copy: the STREAM-style bandwidth reference and
// the known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters
// one 16-byte element per thread and no loop: the fastest plain copy form
// measured on MI355X (6.1 TB/s for 2 GiB; a grid-stride loop over 8192
// work-groups reached 4.6 TB/s, several elements per thread less)
__global__ void __launch_bounds__(256) k_stream_copy(const V4f *src, V4f *dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------------ host launchers
void launch_overlap(Ctx &cx, int prec, int n1, int n2, long K, const void *A, const void *B, double vol, double *S,
                    DevMem &partial) {
  const long K2 = 2 * K;
  const int ti = (n1 + kOvTile - 1) / kOvTile, tj = (n2 + kOvTile - 1) / kOvTile;
  const size_t es = prec == INQX_FP64 ? 8 : 4;
  const long kc = overlap_chunk(K2, ti * tj, es);
  INQX_CHECK(kc <= INT32_MAX, INQX_ERR_ARG, "overlap: K too long");
  const int kchunk = (int)kc;
  const int nchunk = (int)((K2 + kchunk - 1) / kchunk);
  partial.ensure(es * (size_t)nchunk * ti * tj * kOvTile * kOvTile * 2);
  const int grid = nchunk * ti * tj;
  const int rgrid = ti * tj * kOvTile;
  if (prec == INQX_FP64) {
    {
      ProfScope ps(cx, "overlap_mfma_f64");
      hipLaunchKernelGGL(k_overlap_partial<double>, dim3(grid), dim3(256), 0, cx.stream, (const double *)A,
                         (const double *)B, n1, n2, K2, ti, tj, kchunk, partial.as<double>());
    }
    INQX_HIP(hipGetLastError());
    ProfScope ps(cx, "overlap_reduce_f64");
    hipLaunchKernelGGL(k_overlap_reduce<double>, dim3(rgrid), dim3(256), 0, cx.stream, partial.as<double>(), n1, n2,
                       ti, tj, nchunk, vol, S);
  } else {
    {
      ProfScope ps(cx, "overlap_mfma_f32");
      hipLaunchKernelGGL(k_overlap_partial<float>, dim3(grid), dim3(256), 0, cx.stream, (const float *)A,
                         (const float *)B, n1, n2, K2, ti, tj, kchunk, partial.as<float>());
    }
    INQX_HIP(hipGetLastError());
    ProfScope ps(cx, "overlap_reduce_f32");
    hipLaunchKernelGGL(k_overlap_reduce<float>, dim3(rgrid), dim3(256), 0, cx.stream, partial.as<float>(), n1, n2, ti,
                       tj, nchunk, vol, S);
  }
  INQX_HIP(hipGetLastError());
}

void launch_rotate(Ctx &cx, int prec, int nr, int nc, long K, const double *R, int mode, double ar, double ai,
                   const void *Cm, void *Out, double br, double bi, DevMem &ops) {
  INQX_CHECK(mode == 1 || nr == nc, INQX_ERR_ARG, "rotate: modes 0 and 2 need a square matrix");
  const size_t es = prec == INQX_FP64 ? 8 : 4;
  const int tiles_p = (int)((K + 63) / 64), tiles_i = (nr + 63) / 64;
  const int nr_pad = tiles_i * 64, ld = (2 * nc + 2 * kRotJC - 1) / (2 * kRotJC) * (2 * kRotJC);
  const size_t mat = (size_t)nr_pad * ld;
  ops.ensure(es * mat * 2);
  const int gprep = (int)(((size_t)nr_pad * (ld / 2) + 255) / 256);
  const int grid = tiles_p * tiles_i;
  if (prec == INQX_FP64) {
    double *Ap = ops.as<double>(), *Aq = Ap + mat;
    hipLaunchKernelGGL(k_rot_prep<double>, dim3(gprep), dim3(256), 0, cx.stream, R, nr, nc, mode, ar, ai, nr_pad, ld,
                       Ap, Aq);
    INQX_HIP(hipGetLastError());
    ProfScope ps(cx, "rotate_mfma_f64");
    hipLaunchKernelGGL(k_rotate<double>, dim3(grid), dim3(256), 0, cx.stream, Ap, Aq, ld, (const cplx<double> *)Cm, nr,
                       nc, K, (cplx<double> *)Out, br, bi, tiles_p, tiles_i);
  } else {
    float *Ap = ops.as<float>(), *Aq = Ap + mat;
    hipLaunchKernelGGL(k_rot_prep<float>, dim3(gprep), dim3(256), 0, cx.stream, R, nr, nc, mode, ar, ai, nr_pad, ld,
                       Ap, Aq);
    INQX_HIP(hipGetLastError());
    ProfScope ps(cx, "rotate_mfma_f32");
    hipLaunchKernelGGL(k_rotate<float>, dim3(grid), dim3(256), 0, cx.stream, Ap, Aq, ld, (const cplx<float> *)Cm, nr, nc,
                       K, (cplx<float> *)Out, br, bi, tiles_p, tiles_i);
  }
  INQX_HIP(hipGetLastError());
}

// zero fill with vector stores (no memset node when a caller captures the stream)
__global__ void __launch_bounds__(256) k_fill_zero(V4f *p, long n16, unsigned char *tail, int ntail) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = V4f{0.f, 0.f, 0.f, 0.f};
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// ------------------------------------------------------------------------ overlap_diagonal
// per state: sum conj(phi1) phi2 and sum |phi2|^2 in FP64 (one work-group per
// state, fixed reduction order); out = vol * value, or value / norm
template <typename T>
__global__ void __launch_bounds__(256) k_overlap_diag(const cplx<T> *A, const cplx<T> *B, long K, double vol,
                                                     int normalized, double *out) {
  __shared__ double red[3][256];
  const int tid = threadIdx.x, s = blockIdx.x;
  const GAS cplx<T> *a = lgp(A) + (size_t)s * K;
  const GAS cplx<T> *b = lgp(B) + (size_t)s * K;
  double vr = 0.0, vi = 0.0, nn = 0.0;
  for (long p = tid; p < K; p += 256) {
    const cplx<T> x = a[p], y = b[p];
    vr += (double)x.x * y.x + (double)x.y * y.y;
    vi += (double)x.x * y.y - (double)x.y * y.x;
    nn += (double)y.x * y.x + (double)y.y * y.y;
  }
  red[0][tid] = vr;
  red[1][tid] = vi;
  red[2][tid] = nn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int c = 0; c < 3; ++c) red[c][tid] += red[c][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    GAS double *o = lgw(out) + 2 * (size_t)s;
    if (normalized) {
      o[0] = red[0][0] / red[2][0];
      o[1] = red[1][0] / red[2][0];
    } else {
      o[0] = vol * red[0][0];
      o[1] = vol * red[1][0];
    }
  }
}

void launch_overlap_diagonal(Ctx &cx, int prec, int n, long K, double vol, const void *A, const void *B,
                             int normalized, double *out) {
  ProfScope ps(cx, "overlap_diagonal");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_overlap_diag<double>, dim3(n), dim3(256), 0, cx.stream, (const cplx<double> *)A,
                       (const cplx<double> *)B, K, vol, normalized, out);
  else
    hipLaunchKernelGGL(k_overlap_diag<float>, dim3(n), dim3(256), 0, cx.stream, (const cplx<float> *)A,
                       (const cplx<float> *)B, K, vol, normalized, out);
  INQX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------ density
// density[p][ncomp] (component ispin) += sum_i occ_i |phi_i(p)|^2, states
// state-major [nst][N]; FP64, states summed in order
template <typename T>
__global__ void __launch_bounds__(256) k_density_add(const cplx<T> *phi, const double *occ, int nst, long N,
                                                    int ncomp, int ispin, double *den) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int i = 0; i < nst; ++i) {
      const cplx<T> v = lgp(phi)[(size_t)i * N + p];
      acc += lgp(occ)[i] * ((double)v.x * v.x + (double)v.y * v.y);
    }
    lgw(den)[p * ncomp + ispin] += acc;
  }
}

void launch_density_add(Ctx &cx, int prec, const void *phi, const double *occ, int nst, long N, int ncomp,
                        int ispin, double *den) {
  const int grid = (int)std::max<long>(1, std::min<long>((N + 255) / 256, (long)cx.num_cu * 16));
  ProfScope ps(cx, "density_add");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_density_add<double>, dim3(grid), dim3(256), 0, cx.stream, (const cplx<double> *)phi, occ,
                       nst, N, ncomp, ispin, den);
  else
    hipLaunchKernelGGL(k_density_add<float>, dim3(grid), dim3(256), 0, cx.stream, (const cplx<float> *)phi, occ,
                       nst, N, ncomp, ispin, den);
  INQX_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------ randomize
// operations::randomize (src/operations/randomize.hpp:48-70): value (ip, ist)
// comes from pcg32 (64-bit LCG, XSH-RR output of the pre-step state) seeded
// with bump(seed + inc) and advanced by (ip nst + ist) samples (jump-ahead,
// Brown 1994); U = 2 r / 2^32 - 1, real part first
constexpr uint64_t kPcgMult = 6364136223846793005ULL;
constexpr uint64_t kPcgInc = 1442695040888963407ULL;

__device__ __forceinline__ uint64_t pcg_jump(uint64_t state, uint64_t delta) {
  uint64_t cm = kPcgMult, cp = kPcgInc, am = 1, ap = 0;
  while (delta > 0) {
    if (delta & 1) {
      am *= cm;
      ap = ap * cm + cp;
    }
    cp = (cm + 1) * cp;
    cm *= cm;
    delta >>= 1;
  }
  return am * state + ap;
}

__device__ __forceinline__ double pcg_uniform(uint64_t s) {
  const uint32_t rot = (uint32_t)(s >> 59);
  const uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
  const uint32_t r = (x >> rot) | (x << ((32u - rot) & 31u));
  return 2.0 * (double)r / 4294967296.0 - 1.0;
}

__global__ void __launch_bounds__(256) k_randomize(long npoints, int nst, uint64_t start, int type, int layout,
                                                   void *out) {
  const long total = npoints * nst;
  const int per = type == 2 ? 1 : 2;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const long ip = idx / nst;
    const int ist = (int)(idx - ip * nst);
    uint64_t st = pcg_jump(start, (uint64_t)idx * per);
    const double re = pcg_uniform(st);
    double im = 0.0;
    if (per == 2) im = pcg_uniform(st * kPcgMult + kPcgInc);
    const long o = layout == 0 ? (long)ist * npoints + ip : idx;
    if (type == INQX_FP64) {
      lgw((double *)out)[2 * o] = re;
      lgw((double *)out)[2 * o + 1] = im;
    } else if (type == INQX_FP32) {
      lgw((float *)out)[2 * o] = (float)re;
      lgw((float *)out)[2 * o + 1] = (float)im;
    } else {
      lgw((double *)out)[o] = re;
    }
  }
}

void launch_randomize(Ctx &cx, int type, long npoints, int nst, int index, int layout, void *out) {
  const uint64_t seed = (uint64_t)npoints * (uint64_t)nst + (uint64_t)index;
  const uint64_t start = (seed + kPcgInc) * kPcgMult + kPcgInc;
  const long total = npoints * nst;
  const int grid = (int)std::max<long>(1, std::min<long>((total + 255) / 256, (long)cx.num_cu * 32));
  ProfScope ps(cx, "randomize");
  hipLaunchKernelGGL(k_randomize, dim3(grid), dim3(256), 0, cx.stream, npoints, nst, start, type, layout, out);
  INQX_HIP(hipGetLastError());
}

void launch_fill_zero(Ctx &cx, void *p, size_t bytes) {
  if (bytes == 0) return;
  const long n16 = (long)(bytes / 16);
  const int ntail = (int)(bytes - 16 * (size_t)n16);
  const int grid = (int)std::max<long>(1, std::min<long>((n16 + 255) / 256, (long)cx.num_cu * 16));
  hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, cx.stream, (V4f *)p, n16,
                     (unsigned char *)p + 16 * (size_t)n16, ntail);
  INQX_HIP(hipGetLastError());
}

void launch_cholesky(Ctx &cx, int n, double *S, int *info) {
  ProfScope ps(cx, "cholesky_f64");
  if (n <= kCholSmall) {
    const size_t lds = sizeof(double2) * (size_t)n * (n + 1) / 2;
    static DeviceOnce attr;
    if (!attr.test_and_set(cx.device))
      INQX_HIP(hipFuncSetAttribute((const void *)k_chol_small, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(sizeof(double2) * kCholSmall * (kCholSmall + 1) / 2)));
    hipLaunchKernelGGL(k_chol_small, dim3(1), dim3(1024), lds, cx.stream, S, n, info);
    INQX_HIP(hipGetLastError());
    return;
  }
  for (int c0 = 0; c0 < n; c0 += kChB) {
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, cx.stream, S, n, c0, info);
    INQX_HIP(hipGetLastError());
    const int rest = n - c0 - kChB;
    if (rest <= 0) break;
    hipLaunchKernelGGL(k_chol_panel, dim3((rest + 255) / 256), dim3(256), 0, cx.stream, S, n, c0, (const int *)info);
    INQX_HIP(hipGetLastError());
    const int m = (rest + kChB - 1) / kChB;
    hipLaunchKernelGGL(k_chol_update, dim3(m, m), dim3(256), 0, cx.stream, S, n, c0, (const int *)info);
    INQX_HIP(hipGetLastError());
  }
}

void launch_stream_copy(Ctx &cx, const void *src, void *dst, size_t bytes) {
  const long n = (long)(bytes / 16);
  INQX_CHECK((n + 255) / 256 <= INT32_MAX, INQX_ERR_ARG, "stream copy too long");
  const int grid = (int)((n + 255) / 256);
  if (grid == 0) return;
  ProfScope ps(cx, "stream_copy");
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, cx.stream, (const V4f *)src, (V4f *)dst, n);
  INQX_HIP(hipGetLastError());
}

void launch_trinv(Ctx &cx, int n, const double *L, double *X) {
  ProfScope ps(cx, "trinv_f64");
  const int nb = (n + kChB - 1) / kChB;
  hipLaunchKernelGGL(k_trinv_diag, dim3(nb, nb), dim3(256), 0, cx.stream, L, n, X);
  INQX_HIP(hipGetLastError());
  for (int r = 1; r < nb; ++r) {
    hipLaunchKernelGGL(k_trinv_row, dim3(r), dim3(256), 0, cx.stream, L, n, r, X);
    INQX_HIP(hipGetLastError());
  }
}

}  // namespace inqx
