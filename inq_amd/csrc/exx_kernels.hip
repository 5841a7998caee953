// exx_kernels.hip -- batched pair-density Poisson solves for exact exchange, gfx950.
//
// The reference's exchange_operator::block_exchange
// (src/hamiltonian/exchange_operator.hpp:124-161) loops over the occupied
// states j and, for all states i at once, forms rho_ij = conj(hf_j) phi_i,
// solves Poisson (solvers::poisson::in_place, src/solvers/poisson.hpp:71-110,
// 199-213: forward FFT, v_G = -4 pi/N * (-1/|G+q|^2) rho_G, zeroterm at
// |G+q|^2 < 1e-6, unnormalised inverse FFT) and accumulates
// exxphi_i += factor occ_j hf_j v_ij.
//
// Here a launch covers a chunk of j's times all i, as kernels over
// real-space arrays in state-major layout [state][i0][i1][i2]:
//
//   pair_plane_fwd   per (j, plane i0, i): forms rho on the fly from the two
//                    planes, 2-D forward FFT along i1 (registers, a column
//                    per lane pair) then i2 (LDS rows), stores the plane of t
//   pair_xline       per (j, i, 128 lines (i1,i2)): forward FFT along i0 in
//                    registers, Poisson kernel, inverse FFT along i0, in place
//   pair_plane_inv   per (j, plane i0, i): 2-D inverse FFT of the plane of t
//   pair_reduce      w_i += sum_j coef_j hf_j t_ji in j order (deterministic;
//                    chunks of j give the same bits); pair_reduce_sym adds the
//                    conjugate pairs' terms when hf = phi
//
// Planes with n1 != n2 run the 2-D transforms as a column kernel and a row
// kernel in place in t (pair_cols, pair_rows) instead of pair_plane_fwd/inv.
// Without hf the same kernels solve Poisson in place on a batch of densities.
#include "plane_common.hpp"

#include <algorithm>
#include <cmath>

namespace inqx {

namespace {

template <typename T> struct PairDev {
  const cplx<T> *phi;  // [nst][N]
  const cplx<T> *hf;   // [nhf][N] or null
  cplx<T> *t;          // [nj][nst][N]
  cplx<T> *w;          // [nst][N]
  const int *jlist;    // [nj]
  const double *coef;  // [nhf]
  int nst, nj;
  int n0, n1, n2;      // grid (the plane passes take n0 at run time, the line pass n1 and n2)
  // symmetric pairs (hf = phi, no shift): pair (j, i) with i occupied and
  // i > j is not solved; its potential is conj of pair (i, j)'s, added into
  // wb by k_pair_reduce_sym.  occ: [nst] 1 for an occupied state, or null.
  const int *occ;
  cplx<T> *wb;         // [nst][N]
  T b[3][3];           // reciprocal lattice vectors (rows, cartesian, 2 pi included)
  T q[3];              // cartesian shift added to G
  T scal;              // 4 pi / N
  T zero;              // value of the kernel at |G + q|^2 < 1e-6: zeroterm / N
};

// a plane pass works on one N1 x N2 plane in LDS (rows padded to N2 + 1); every
// row or column of the plane is split over LANES lanes of a wave (a lane pair
// (l, l + 32), or the four lanes l + 16 q for complex128 planes from 72
// points, whose planes fill the LDS of a CU: four lanes per line give that
// one work-group twice the waves), so a wave covers 64 / LANES lines
template <typename T, int N1, int N2> struct PairPlaneCfg {
  static constexpr int STRIDE = N2 + 1;
  static constexpr int NL = N1 > N2 ? N1 : N2;
  static constexpr int LANES = sizeof(T) == 8 && N1 % 4 == 0 && N2 % 4 == 0 && NL >= 72 ? 4 : 2;
  static constexpr int LPW = 64 / LANES;
  static constexpr int THREADS = (NL + LPW - 1) / LPW * 64;
  static_assert(THREADS <= 512, "plane pass sized for at most 8 waves");
  // the launch bound: 256 keeps up to 512 VGPRs per lane-pair thread
  static constexpr int MAXT = THREADS <= 256 ? 256 : 512;
  static constexpr int LDS = N1 * STRIDE * (int)sizeof(cplx<T>);
};
template <typename T, int N1, int N2> constexpr int pair_plane_maxt() { return PairPlaneCfg<T, N1, N2>::MAXT; }

// symmetric mode: the pair (hf index j of chunk entry jj, state i) is skipped
// when both are occupied and i > j (the potential of pair (i, j) stands for it)
template <typename T> __device__ __forceinline__ bool pair_skipped(const PairDev<T> &d, int jj, int i) {
  return d.occ && i > gp(d.jlist)[jj] && gp(d.occ)[i] != 0;
}

// FFT along i2 of row r of the LDS plane, split over S lanes (part q of the
// row); rows >= N1 run on a clamped row and do not store
template <typename C, int N1, int N2, int SIGN, int S>
__device__ __forceinline__ void lds_row_fft_split(C *lds, int r, int q) {
  constexpr int M = N2 / S;
  const int rc = r < N1 ? r : N1 - 1;
  C *row = lds + rc * (N2 + 1);
  C v[M];
#pragma unroll
  for (int n = 0; n < M; ++n) v[n] = row[S * n + q];
  line_dit<N2, SIGN, C, S>(v, q);
  __builtin_amdgcn_wave_barrier();
  if (r < N1) {
    const int o = M * line_part_out<S>(q);
#pragma unroll
    for (int k = 0; k < M; ++k) row[k + o] = v[k];
  }
}

template <int N> __host__ __device__ constexpr int symi(int i) { return i >= (N + 1) / 2 ? i - N : i; }

// reciprocal: the hardware approximation (1 ulp) for the complex64 path, an
// exact division for complex128
__device__ __forceinline__ float rcp_t(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_t(double x) { return 1.0 / x; }

template <typename T, int N1, int N2>
__global__ void __launch_bounds__((pair_plane_maxt<T, N1, N2>())) k_pair_plane_fwd(PairDev<T> d, int total) {
  using C = cplx<T>;
  using Cfg = PairPlaneCfg<T, N1, N2>;
  constexpr int NP = N1 * N2;
  constexpr int S = Cfg::LANES, M1 = N1 / S;
  const long N = (long)d.n0 * NP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x;
  const int line = (tid >> 6) * Cfg::LPW + (tid & 63) % Cfg::LPW;
  const int q = (tid & 63) / Cfg::LPW, qo = line_part_out<S>(q);  // part of the line, its output position
  // one work item per work-group, the grid being the work count: a grid-stride loop
  // here made the complex128 96-point kernels spill to scratch
  {
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int i = wg % d.nst;
    const int r = wg / d.nst;
    const int i0 = r % d.n0;
    const int jj = r / d.n0;
    if (pair_skipped(d, jj, i)) return;  // uniform over the work-group
    const size_t plane = (size_t)i0 * NP;
    const GAS C *ph = gp(d.phi) + (size_t)i * N + plane;
    const GAS C *hf = nullptr;
    if (d.hf) hf = gp(d.hf) + (size_t)gp(d.jlist)[jj] * N + plane;
    {
      // column z = line along i1: part q holds y = S n + q
      const int z = line < N2 ? line : N2 - 1;
      C v[M1];
      if (hf) {
#pragma unroll
        for (int n = 0; n < M1; ++n) {
          const int y = S * n + q;
          const C a = ph[y * N2 + z], h = hf[y * N2 + z];
          v[n] = C{h.x * a.x + h.y * a.y, h.x * a.y - h.y * a.x};  // conj(h) a
        }
      } else {
#pragma unroll
        for (int n = 0; n < M1; ++n) v[n] = ph[(S * n + q) * N2 + z];
      }
      line_dit<N1, -1, C, S>(v, q);
      if (line < N2) {
#pragma unroll
        for (int k = 0; k < M1; ++k) lds[(k + qo * M1) * Cfg::STRIDE + z] = v[k];
      }
    }
    __syncthreads();
    lds_row_fft_split<C, N1, N2, -1, S>(lds, line, q);
    __syncthreads();
    GAS C *tp = gw(d.t) + ((size_t)jj * d.nst + i) * N + plane;
    for (int e = tid; e < NP; e += Cfg::THREADS) {
      const int y = e / N2, z = e - y * N2;
      tp[e] = lds[y * Cfg::STRIDE + z];
    }
    __syncthreads();
  }
}

// lanes per line of the pair line pass: a quarter line per lane for
// complex128 lines of 72, 80 and above 96 points, else a lane pair.  Measured
// per 16 complex128 Poisson solves, lane pair vs four lanes: 72^3 53.9 vs
// 48.3 us, 80^3 79.2 vs 61.9 us, 96^3 122.9 vs 135.1 us, 32-64 points within
// 1 us; above 96 points the lane-pair form spills 13-308 VGPRs to scratch
template <typename T, int N0> constexpr int xline_lanes() {
  return sizeof(T) == 8 && N0 % 4 == 0 && N0 >= 72 && N0 != 96 ? 4 : 2;
}

// each line is split over S lanes of a wave, a lane pair (l, l + 32) or the
// four lanes l + 16 q (q < 4), so a wave covers 64 / S lines
// CUBE: n1 = n2 = N0 known at compile time (the row strides of the loads
// become constants; 143 instead of 198 VGPRs at 96^3 complex64), otherwise
// n1 and n2 are read at run time
template <typename T, int N0, bool CUBE>
__global__ void __launch_bounds__(256) k_pair_xline(PairDev<T> d, int total) {
  using C = cplx<T>;
  constexpr int S = xline_lanes<T, N0>(), LPW = 64 / S, M = N0 / S;
  const int N1 = CUBE ? N0 : d.n1, N2 = CUBE ? N0 : d.n2;
  const int NP = N1 * N2;
  const int NTILE = (NP + LPW - 1) / LPW;
  const int NGRP = (NTILE + 3) / 4;
  const long N = (long)N0 * NP;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int q = lane / LPW, qo = line_part_out<S>(q);  // part of the line, its output position
  // one work item per work-group, the grid being the work count: a grid-stride loop
  // here made the complex128 96-point kernels spill to scratch
  {
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int pair = wg / NGRP;
    const int tile = (wg - pair * NGRP) * 4 + wave;
    if (tile >= NTILE) return;  // wave-uniform; no barrier follows
    if (pair_skipped(d, pair / d.nst, pair % d.nst)) return;
    const int L = tile * LPW + lane % LPW;
    const int Lc = L < NP ? L : NP - 1;
    const auto rs = make_rsrc(d.t + (size_t)pair * N, (unsigned)(N * sizeof(C)));
    const unsigned voff = (unsigned)((Lc + q * NP) * sizeof(C));  // i0 = S n + q
    C v[M];
#pragma unroll
    for (int n = 0; n < M; ++n) bload(rs, voff, (unsigned)(S * n * NP * sizeof(C)), v[n]);
    line_dit<N0, -1, C, S>(v, q);
    __builtin_amdgcn_sched_barrier(0);  // keep the kernel factors out of the FFT's live range
    // frequency f = k + qo N0/S; G + q = m0 b0 + (m1 b1 + m2 b2 + q)
    const int i1 = Lc / N2, i2 = Lc - (Lc / N2) * N2;
    const T m1 = (T)(i1 >= (N1 + 1) / 2 ? i1 - N1 : i1), m2 = (T)(i2 >= (N2 + 1) / 2 ? i2 - N2 : i2);
    T u[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[c] = m1 * d.b[1][c] + m2 * d.b[2][c] + d.q[c];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int f = k + qo * M;
      const T m0 = (T)(f >= (N0 + 1) / 2 ? f - N0 : f);
      const T g0 = m0 * d.b[0][0] + u[0], g1 = m0 * d.b[0][1] + u[1], g2 = m0 * d.b[0][2] + u[2];
      const T gg = g0 * g0 + g1 * g1 + g2 * g2;
      const T fct = gg < (T)1e-6 ? d.zero : d.scal * rcp_t(gg);
      v[k].x *= fct;
      v[k].y *= fct;
    }
    __builtin_amdgcn_sched_barrier(0);
    line_dif<N0, +1, C, S>(v, q);  // v[n]: i0 = S n + q
    if (L < NP) {
#pragma unroll
      for (int n = 0; n < M; ++n) bstore(rs, voff, (unsigned)(S * n * NP * sizeof(C)), v[n]);
    }
  }
}

// inverse 2D transform of one pair plane in place, the same decomposition as
// the forward pass: the column transform reads t straight from memory, the
// row transform runs from LDS.  k_pair_reduce then forms sum_j c_j hf_j t_j.
template <typename T, int N1, int N2>
__global__ void __launch_bounds__((pair_plane_maxt<T, N1, N2>())) k_pair_plane_inv(PairDev<T> d, int total) {
  using C = cplx<T>;
  using Cfg = PairPlaneCfg<T, N1, N2>;
  constexpr int NP = N1 * N2;
  constexpr int S = Cfg::LANES, M1 = N1 / S;
  const long N = (long)d.n0 * NP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x;
  const int line = (tid >> 6) * Cfg::LPW + (tid & 63) % Cfg::LPW;
  const int q = (tid & 63) / Cfg::LPW, qo = line_part_out<S>(q);
  // one work item per work-group, the grid being the work count: a grid-stride loop
  // here made the complex128 96-point kernels spill to scratch
  {
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int i = wg % d.nst;
    const int r = wg / d.nst;
    const int i0 = r % d.n0;
    const int jj = r / d.n0;
    if (pair_skipped(d, jj, i)) return;  // uniform over the work-group
    const size_t plane = (size_t)i0 * NP;
    GAS C *tp = gw(d.t) + ((size_t)jj * d.nst + i) * N + plane;
    {
      const int z = line < N2 ? line : N2 - 1;
      C v[M1];
#pragma unroll
      for (int n = 0; n < M1; ++n) v[n] = tp[(S * n + q) * N2 + z];
      line_dit<N1, +1, C, S>(v, q);
      if (line < N2) {
#pragma unroll
        for (int k = 0; k < M1; ++k) lds[(k + qo * M1) * Cfg::STRIDE + z] = v[k];
      }
    }
    __syncthreads();
    lds_row_fft_split<C, N1, N2, +1, S>(lds, line, q);
    __syncthreads();
    for (int e = tid; e < NP; e += Cfg::THREADS) {
      const int y = e / N2, z = e - y * N2;
      tp[e] = lds[y * Cfg::STRIDE + z];
    }
    __syncthreads();
  }
}

// ---- planes with n1 != n2: the 2-D transforms of a pair plane as a column
// kernel (built for N1, n2 at run time) and a row kernel (built for N2, n1 at
// run time), both in place in t, a lane pair per line and 32 lines per wave
// (a kernel per (n1, n2) pair would square the instantiations)

// pair plane (chunk entry jj, plane i0, state i) of work-group wg
__device__ __forceinline__ void pair_plane_item(int wg, int nst, int n0, int &jj, int &i0, int &i) {
  i = wg % nst;
  const int r = wg / nst;
  i0 = r % n0;
  jj = r / n0;
}

// FFT along i1 of every column of one pair plane; PROD: the input is
// conj(hf_j) phi_i (phi_i alone without hf), formed on the fly as in
// pair_plane_fwd, else the plane of t itself
template <typename T, int N1, int SIGN, bool PROD>
__global__ void __launch_bounds__(256) k_pair_cols(PairDev<T> d, int total) {
  using C = cplx<T>;
  constexpr int H1 = N1 / 2;
  const int n2 = d.n2, NP = N1 * n2;
  const long N = (long)d.n0 * NP;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int half = lane >= 32 ? 1 : 0;
  const int wg = blockIdx.x;
  if (wg >= total) return;
  int jj, i0, i;
  pair_plane_item(wg, d.nst, d.n0, jj, i0, i);
  if (pair_skipped(d, jj, i)) return;  // uniform over the work-group
  const size_t plane = (size_t)i0 * NP;
  GAS C *tp = gw(d.t) + ((size_t)jj * d.nst + i) * N + plane;
  const GAS C *ph = gp(d.phi) + (size_t)i * N + plane;
  const GAS C *hf = (PROD && d.hf) ? gp(d.hf) + (size_t)gp(d.jlist)[jj] * N + plane : nullptr;
  // the lanes past the last column repeat it and store nothing; they are in
  // the same wave as its owner, whose loads all precede its stores
  for (int g = wave; g * 32 < n2; g += (int)(blockDim.x >> 6)) {
    const int zl = g * 32 + (lane & 31);
    const bool ok = zl < n2;
    const int z = ok ? zl : n2 - 1;
    C v[H1];
#pragma unroll
    for (int n = 0; n < H1; ++n) {
      const int e = (2 * n + half) * n2 + z;
      if (PROD) {
        const C a = ph[e];
        if (hf) {
          const C h = hf[e];
          v[n] = C{h.x * a.x + h.y * a.y, h.x * a.y - h.y * a.x};  // conj(h) a
        } else {
          v[n] = a;
        }
      } else {
        v[n] = tp[e];
      }
    }
    fft_split<N1, SIGN, C>::dit(v, half != 0);
    if (ok) {
#pragma unroll
      for (int k = 0; k < H1; ++k) tp[(k + half * H1) * n2 + z] = v[k];
    }
  }
}

// FFT along i2 of every row of one pair plane of t, in place
template <typename T, int N2, int SIGN>
__global__ void __launch_bounds__(256) k_pair_rows(PairDev<T> d, int total) {
  using C = cplx<T>;
  constexpr int H2 = N2 / 2;
  const int n1 = d.n1, NP = n1 * N2;
  const long N = (long)d.n0 * NP;
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int half = lane >= 32 ? 1 : 0;
  const int wg = blockIdx.x;
  if (wg >= total) return;
  int jj, i0, i;
  pair_plane_item(wg, d.nst, d.n0, jj, i0, i);
  if (pair_skipped(d, jj, i)) return;
  GAS C *tp = gw(d.t) + ((size_t)jj * d.nst + i) * N + (size_t)i0 * NP;
  for (int g = wave; g * 32 < n1; g += (int)(blockDim.x >> 6)) {
    const int yl = g * 32 + (lane & 31);
    const bool ok = yl < n1;
    GAS C *row = tp + (size_t)(ok ? yl : n1 - 1) * N2;
    C v[H2];
#pragma unroll
    for (int n = 0; n < H2; ++n) v[n] = row[2 * n + half];
    fft_split<N2, SIGN, C>::dit(v, half != 0);
    if (ok) {
#pragma unroll
      for (int k = 0; k < H2; ++k) row[k + half * H2] = v[k];
    }
  }
}

// the column then the row transforms of every pair plane (n1 != n2)
template <typename T, int SIGN, bool PROD>
void pair_planes_rect(Ctx &cx, const PairDev<T> &d, int total) {
  switch (d.n1) {
#define CASE(S)                                                                                          \
  case S:                                                                                                \
    hipLaunchKernelGGL((k_pair_cols<T, S, SIGN, PROD>), dim3(total), dim3(256), 0, cx.stream, d, total); \
    break;
    INQX_SIZES(CASE)
#undef CASE
  }
  INQX_HIP(hipGetLastError());
  switch (d.n2) {
#define CASE(S)                                                                                  \
  case S:                                                                                        \
    hipLaunchKernelGGL((k_pair_rows<T, S, SIGN>), dim3(total), dim3(256), 0, cx.stream, d, total); \
    break;
    INQX_SIZES(CASE)
#undef CASE
  }
  INQX_HIP(hipGetLastError());
}

// w_i += sum_jj c_j hf_j t[jj][i] over the chunk, added in jj order so that
// splitting the j range into chunks gives the same bits.  16-byte vectors
// (two complex64 or one complex128), 4 per thread; the work-groups of one
// point range run over consecutive i so that hf_j is read once from memory.
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

__device__ __forceinline__ void cmac(float4 &a, float c, const float4 &h, const float4 &t) {
  a.x += c * (h.x * t.x - h.y * t.y);
  a.y += c * (h.x * t.y + h.y * t.x);
  a.z += c * (h.z * t.z - h.w * t.w);
  a.w += c * (h.z * t.w + h.w * t.z);
}
__device__ __forceinline__ void cmac(double2 &a, double c, const double2 &h, const double2 &t) {
  a.x += c * (h.x * t.x - h.y * t.y);
  a.y += c * (h.x * t.y + h.y * t.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_pair_reduce(PairDev<T> d, long nvec) {
  using V = typename Vec16<T>::type;
  constexpr int U = 4;
  const int i = (int)(blockIdx.x % (unsigned)d.nst);
  const long v0 = (long)(blockIdx.x / (unsigned)d.nst) * 256 * U + threadIdx.x;
  GAS V *wv = gw(reinterpret_cast<V *>(d.w)) + (size_t)i * nvec;
  const GAS V *tv = gp(reinterpret_cast<const V *>(d.t)) + (size_t)i * nvec;
  const GAS V *hv = gp(reinterpret_cast<const V *>(d.hf));
  const size_t jstride = (size_t)d.nst * nvec;
  V acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long v = v0 + u * 256;
    if (v < nvec) acc[u] = wv[v];
  }
  for (int jj = 0; jj < d.nj; ++jj) {
    const int j = gp(d.jlist)[jj];
    if (d.occ && i > j && gp(d.occ)[i] != 0) continue;  // pair not solved: k_pair_reduce_sym adds it
    const T c = (T)gp(d.coef)[j];
    V x[U], h[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long v = v0 + u * 256;
      if (v < nvec) {
        x[u] = tv[jj * jstride + v];
        h[u] = hv[(size_t)j * nvec + v];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) cmac(acc[u], c, h[u], x[u]);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long v = v0 + u * 256;
    if (v < nvec) wv[v] = acc[u];
  }
}

// acc += c h conj(t)
__device__ __forceinline__ void cmac_conj(float4 &a, float c, const float4 &h, const float4 &t) {
  a.x += c * (h.x * t.x + h.y * t.y);
  a.y += c * (h.y * t.x - h.x * t.y);
  a.z += c * (h.z * t.z + h.w * t.w);
  a.w += c * (h.w * t.z - h.z * t.w);
}
__device__ __forceinline__ void cmac_conj(double2 &a, double c, const double2 &h, const double2 &t) {
  a.x += c * (h.x * t.x + h.y * t.y);
  a.y += c * (h.y * t.x - h.x * t.y);
}

// symmetric mode: for every chunk entry jj (an occupied state j),
// wb_j = sum over occupied states s < j, in s order, of c_s phi_s conj(t[jj][s])
// -- the terms of w_j whose pair (j, s) was not solved (its potential is the
// conjugate of pair (s, j)'s, exchange_operator.hpp:144-160 with real kernels).
// Every w_j's terms land in one chunk, so wb needs no accumulation across chunks.
template <typename T>
__global__ void __launch_bounds__(256) k_pair_reduce_sym(PairDev<T> d, long nvec) {
  using V = typename Vec16<T>::type;
  constexpr int U = 4;
  const int jj = (int)(blockIdx.x % (unsigned)d.nj);
  const long v0 = (long)(blockIdx.x / (unsigned)d.nj) * 256 * U + threadIdx.x;
  const int j = gp(d.jlist)[jj];
  const GAS V *tv = gp(reinterpret_cast<const V *>(d.t)) + (size_t)jj * d.nst * nvec;
  const GAS V *hv = gp(reinterpret_cast<const V *>(d.hf));
  V acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) acc[u] = V{};
  for (int s = 0; s < j; ++s) {
    if (gp(d.occ)[s] == 0) continue;
    const T c = (T)gp(d.coef)[s];
    V x[U], h[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long v = v0 + u * 256;
      if (v < nvec) {
        x[u] = tv[(size_t)s * nvec + v];
        h[u] = hv[(size_t)s * nvec + v];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) cmac_conj(acc[u], c, h[u], x[u]);
  }
  GAS V *bv = gw(reinterpret_cast<V *>(d.wb)) + (size_t)j * nvec;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long v = v0 + u * 256;
    if (v < nvec) bv[v] = acc[u];
  }
}

// S0: line length along i0; S > 0: square planes S x S (pair_plane_fwd /
// pair_plane_inv); S == 0: planes n1 x n2 with n1 != n2 (column and row kernels)
template <typename T, int S0, int S> void run_pairs(inqx_ham_s *h, const PairArgs &a) {
  using Cfg = PairPlaneCfg<T, (S > 0 ? S : 16), (S > 0 ? S : 16)>;
  Ctx &cx = h->ctx->c;
  const Grid &g = h->g;
  PairDev<T> d;
  d.phi = reinterpret_cast<const cplx<T> *>(a.phi);
  d.hf = reinterpret_cast<const cplx<T> *>(a.hf);
  d.t = reinterpret_cast<cplx<T> *>(a.t);
  d.w = reinterpret_cast<cplx<T> *>(a.w);
  d.jlist = a.jlist;
  d.coef = a.coef;
  d.nst = a.nst;
  d.nj = a.hf ? a.nj : 1;
  d.occ = a.hf ? a.sym_occ : nullptr;
  d.wb = reinterpret_cast<cplx<T> *>(a.wb);
  d.n0 = g.n[0];
  d.n1 = g.n[1];
  d.n2 = g.n[2];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) d.b[r][c] = (T)g.b[r][c];
  for (int c = 0; c < 3; ++c) d.q[c] = (T)a.gshift[c];
  d.scal = (T)(4.0 * M_PI / (double)g.N);
  d.zero = (T)(a.zeroterm / (double)g.N);
  if constexpr (S > 0) {
    static DeviceOnce attr;
    if (Cfg::LDS > 65536 && !attr.test_and_set(cx.device)) {
      INQX_HIP(hipFuncSetAttribute((const void *)k_pair_plane_fwd<T, S, S>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, Cfg::LDS));
      INQX_HIP(hipFuncSetAttribute((const void *)k_pair_plane_inv<T, S, S>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, Cfg::LDS));
    }
  }
  const bool f64 = std::is_same<T, double>::value;
  const int tot1 = d.nj * S0 * d.nst;
  {
    ProfScope ps(cx, f64 ? "pair_plane_fwd_f64" : "pair_plane_fwd_f32");
    if constexpr (S > 0)
      hipLaunchKernelGGL((k_pair_plane_fwd<T, S, S>), dim3(tot1), dim3(Cfg::THREADS), Cfg::LDS,
                         cx.stream, d, tot1);
    else
      pair_planes_rect<T, -1, true>(cx, d, tot1);
  }
  INQX_HIP(hipGetLastError());
  constexpr int xlpw = 64 / xline_lanes<T, S0>();  // lines per wave of the line pass
  const int ngrp = ((d.n1 * d.n2 + xlpw - 1) / xlpw + 3) / 4;
  const int tot2 = d.nj * d.nst * ngrp;
  {
    ProfScope ps(cx, f64 ? "pair_xline_f64" : "pair_xline_f32");
    if (S0 == S) hipLaunchKernelGGL((k_pair_xline<T, S0, true>), dim3(tot2), dim3(256), 0, cx.stream, d, tot2);
    else hipLaunchKernelGGL((k_pair_xline<T, S0, false>), dim3(tot2), dim3(256), 0, cx.stream, d, tot2);
  }
  INQX_HIP(hipGetLastError());
  {
    ProfScope ps(cx, f64 ? "pair_plane_inv_f64" : "pair_plane_inv_f32");
    if constexpr (S > 0)
      hipLaunchKernelGGL((k_pair_plane_inv<T, S, S>), dim3(tot1), dim3(Cfg::THREADS), Cfg::LDS,
                         cx.stream, d, tot1);
    else
      pair_planes_rect<T, +1, false>(cx, d, tot1);
  }
  INQX_HIP(hipGetLastError());
  if (d.hf) {
    // 16-byte vectors per state: N complex values of 2 sizeof(T) bytes
    const long nvec = (long)g.N * 2 * (long)sizeof(T) / 16;
    const long per = (nvec + 1023) / 1024;
    const long tot4 = per * d.nst;
    {
      ProfScope ps(cx, f64 ? "pair_reduce_f64" : "pair_reduce_f32");
      hipLaunchKernelGGL((k_pair_reduce<T>), dim3((unsigned)tot4), dim3(256), 0, cx.stream, d, nvec);
    }
    INQX_HIP(hipGetLastError());
    if (d.occ) {
      const long tot5 = per * d.nj;
      ProfScope ps(cx, f64 ? "pair_reduce_sym_f64" : "pair_reduce_sym_f32");
      hipLaunchKernelGGL((k_pair_reduce_sym<T>), dim3((unsigned)tot5), dim3(256), 0, cx.stream, d, nvec);
      INQX_HIP(hipGetLastError());
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- Hartree
namespace {
constexpr int kHChunk = 8192;

__global__ void __launch_bounds__(256) k_real_to_cplx(const double *r, double2 *c, long N) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p < N) gw(c)[p] = double2{gp(r)[p], 0.0};
}

// vh = Re(c); part[chunk] = sum over the chunk of rho vh (fixed tree order)
__global__ void __launch_bounds__(256) k_hartree_out(const double2 *c, const double *rho, double *vh, long N,
                                                     double *part) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const long p0 = (long)blockIdx.x * kHChunk;
  double acc = 0.0;
  for (long p = p0 + tid; p < p0 + kHChunk && p < N; p += 256) {
    const double v = gp(c)[p].x;
    if (vh) gw(vh)[p] = v;
    acc += gp(rho)[p] * v;
  }
  red[tid] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) gw(part)[blockIdx.x] = red[0];
}
}  // namespace

int hartree_chunks(long N) { return (int)((N + kHChunk - 1) / kHChunk); }

// ---------------------------------------------------------------- allocation-free Hartree
// The scratch of inqx_hartree_noalloc: the complex128 work grid, the energy
// partial sums and their total, in one file-scope device array that every
// device holds from module load until the process ends.  Host code never
// passes the array's name to a kernel (that is a host-side stand-in, not a
// device address): it uses the per-device address from hipGetSymbolAddress.
namespace {
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void *g_scratch_addr[64] = {};
DeviceOnce g_scratch_once;

// sum of the partials in chunk order (the order of inqx_hartree's host sum)
__global__ void __launch_bounds__(64) k_sum_partials(const double *part, int n, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double e = 0.0;
    for (int c = 0; c < n; ++c) e += gp(part)[c];
    gw(out)[0] = e;
  }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

void noalloc_scratch_init(int device) {
  INQX_CHECK(device >= 0 && device < 64, INQX_ERR_ARG, "device index beyond 63");
  if (g_scratch_once.test_and_set(device)) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    g_scratch_once.bits.fetch_and(~(1ull << device));  // retry on the next context of this device
    throw Error{INQX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  }
  g_scratch_addr[device] = p;
}

size_t hartree_noalloc_bytes(long N) {
  return align256(16 * (size_t)N) + align256(sizeof(double) * (size_t)hartree_chunks(N)) + 256;
}

double *launch_hartree_noalloc(inqx_ham_s *h, const double *rho, double *vh) {
  Ctx &cx = h->ctx->c;
  const long N = h->g.N;
  INQX_CHECK(hartree_noalloc_bytes(N) <= kNoAllocScratchBytes, INQX_ERR_ARG,
             "grid of " + std::to_string(N) + " points exceeds the allocation-free Hartree scratch (" +
                 std::to_string(kNoAllocMaxPoints) + " points at most)");
  void *base = g_scratch_addr[cx.device & 63];
  INQX_CHECK(base != nullptr, INQX_ERR_STATE, "allocation-free scratch not set up on this device");
  unsigned char *p = static_cast<unsigned char *>(base);
  void *work = p;
  double *part = reinterpret_cast<double *>(p + align256(16 * (size_t)N));
  double *total = reinterpret_cast<double *>(p + align256(16 * (size_t)N) + align256(sizeof(double) * hartree_chunks(N)));
  launch_hartree(h, rho, vh, work, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, cx.stream, part, hartree_chunks(N), total);
  INQX_HIP(hipGetLastError());
  return total;
}

void launch_hartree(inqx_ham_s *h, const double *rho, double *vh, void *work, double *part) {
  Ctx &cx = h->ctx->c;
  const long N = h->g.N;
  {
    ProfScope ps(cx, "hartree_in");
    hipLaunchKernelGGL(k_real_to_cplx, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, cx.stream, rho,
                       reinterpret_cast<double2 *>(work), N);
  }
  INQX_HIP(hipGetLastError());
  PairArgs a;
  a.phi = work;
  a.t = work;
  a.w = work;
  a.nst = 1;
  a.nj = 1;
  launch_pairs(h, INQX_FP64, a);
  {
    ProfScope ps(cx, "hartree_out");
    hipLaunchKernelGGL(k_hartree_out, dim3((unsigned)hartree_chunks(N)), dim3(256), 0, cx.stream,
                       reinterpret_cast<const double2 *>(work), rho, vh, N, part);
  }
  INQX_HIP(hipGetLastError());
}

namespace {
__global__ void __launch_bounds__(256) k_add_vec4(float4 *w, const float4 *b, long n4) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n4) {
    float4 x = gp(w)[i];
    const float4 y = gp(b)[i];
    x.x += y.x;
    x.y += y.y;
    x.z += y.z;
    x.w += y.w;
    gw(w)[i] = x;
  }
}
__global__ void __launch_bounds__(256) k_add_dvec2(double2 *w, const double2 *b, long n2) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n2) {
    double2 x = gp(w)[i];
    const double2 y = gp(b)[i];
    x.x += y.x;
    x.y += y.y;
    gw(w)[i] = x;
  }
}
}  // namespace

// w += b over bytes (a multiple of 16) of complex values of the precision
void launch_add_inplace(Ctx &cx, int prec, void *w, const void *b, size_t bytes) {
  const long n = (long)(bytes / 16);
  if (n == 0) return;
  INQX_CHECK((n + 255) / 256 <= INT32_MAX, INQX_ERR_ARG, "add: array too long");
  ProfScope ps(cx, "pair_add_sym");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_add_dvec2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.stream, (double2 *)w,
                       (const double2 *)b, n);
  else
    hipLaunchKernelGGL(k_add_vec4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.stream, (float4 *)w,
                       (const float4 *)b, n);
  INQX_HIP(hipGetLastError());
}

bool pair_grid_supported(const int *n) {
  auto ok = [](int v) {
    switch (v) {
#define CASE(S) case S:
      INQX_SIZES(CASE)
#undef CASE
      return true;
      default: return false;
    }
  };
  return ok(n[0]) && ok(n[1]) && ok(n[2]);
}

namespace {
// square planes whose LDS plane passes do not fit the CU's 160 KB (complex128
// from 100 points) run the column and row kernels instead
template <typename T, int S0, int S> void run_pairs_square(inqx_ham_s *h, const PairArgs &a) {
  if constexpr (PairPlaneCfg<T, S, S>::LDS > 160 * 1024) run_pairs<T, S0, 0>(h, a);
  else run_pairs<T, S0, S>(h, a);
}

template <typename T, int S0> void pairs_n12(inqx_ham_s *h, const PairArgs &a) {
  if (h->g.n[1] != h->g.n[2]) {
    run_pairs<T, S0, 0>(h, a);
    return;
  }
  switch (h->g.n[1]) {
#define CASE(S)                          \
  case S:                                \
    run_pairs_square<T, S0, S>(h, a);    \
    break;
    INQX_SIZES(CASE)
#undef CASE
    default: break;
  }
}
}  // namespace

void launch_pairs(inqx_ham_s *h, int prec, const PairArgs &a) {
  const int *n = h->g.n;
  INQX_CHECK(pair_grid_supported(n), INQX_ERR_UNSUPPORTED,
             "pair Poisson solves need built grid sizes");
  INQX_CHECK(a.nst > 0 && a.nj > 0 && a.phi && a.t && a.w, INQX_ERR_ARG, "bad pair arguments");
  INQX_CHECK(!a.hf || (a.jlist && a.coef), INQX_ERR_ARG, "pair chunk without j list");
  INQX_CHECK(a.hf || a.t == a.w, INQX_ERR_ARG, "a plain Poisson solve works in place");
  switch (n[0]) {
#define CASE(S)                                       \
  case S:                                             \
    if (prec == INQX_FP64) pairs_n12<double, S>(h, a); \
    else pairs_n12<float, S>(h, a);                   \
    break;
    INQX_SIZES(CASE)
#undef CASE
    default: break;
  }
}

}  // namespace inqx
