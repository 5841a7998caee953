// ham_kernels.hip -- H|psi> on G-sphere blocks, hand-written for gfx950.
//
// One H application on a block of nst states stored on their G sphere
// ([nst][npw], band-major) runs as five stream-ordered kernels with no host
// synchronisation:
//
//   plane_inverse  sphere -> zero-padded (i1,i2) planes of every plane index
//                  i0 the sphere touches, inverse FFT along i2 (rows that hold
//                  sphere points only) then along i1; writes t1[slot][ist][i1][i2]
//   xpass_project  inverse FFT along i0 of each line (i1,i2) in registers,
//                  gathers the line's KB sphere points into LDS and reduces
//                  beta * exp(+ik.r) * psi per (tile, atom) slot
//   kb_reduce      sums the slots of each atom, scales by dV, multiplies by
//                  the KB coefficient matrix D
//   xpass_apply    inverse FFT along i0 again, multiplies by V(r)/N, adds the
//                  KB term exp(-ik.r) beta D P at the sphere points, forward
//                  FFT along i0, keeps the plane indices of the sphere (in
//                  place in t1)
//   plane_forward  forward FFT along i1 then along i2 (sphere rows only),
//                  gathers the sphere and adds 1/2|G+k|^2 psi
//
// This is the reference's G-space operator ks_hamiltonian::operator()
// (src/hamiltonian/ks_hamiltonian.hpp:160-180) restricted to the sphere:
// to_real (transform.hpp:113-182), KB projections (projector_all.hpp:113-211),
// scalar_potential (scalar_potential.hpp:60-63), KB apply
// (projector_all.hpp:215-288), to_fourier (transform.hpp:52-109) and
// laplacian_add (laplacian.hpp:24-37) with the 1/2|k|^2 shift of
// ks_hamiltonian.hpp:149 folded into the kinetic factor.
#include "internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace inqx {

void invalidate_launch_cache(inqx_ham_s *h);

struct HamDev {
  int n0, n1, n2, n12;
  long N;
  int natoms, maxsph, has_kb, ntile;
  const uint32_t *line_mask;
  const int *tile_info;
  const int *slot_atom, *slot_ent, *ent_c, *ent_ap;
  const int *atom_slots, *atom_slot_list;
  const int *pt_mem, *mem_ap;
  const float *pos;
  const void *beta;
  const double *coeff;
  double kb_scale;  // dV / N
};

__device__ __forceinline__ int find_block(const int *prefix, int nblocks, int wg) {
  int lo = 0, hi = nblocks;  // prefix[lo] <= wg < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= wg) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <typename T> __device__ __forceinline__ void sincos_t(T a, T *s, T *c);
template <> __device__ __forceinline__ void sincos_t<float>(float a, float *s, float *c) { sincosf(a, s, c); }
template <> __device__ __forceinline__ void sincos_t<double>(double a, double *s, double *c) { sincos(a, s, c); }

// -------------------------------------------------------------------------------------------
// plane passes
// -------------------------------------------------------------------------------------------
template <typename T, int N1, int N2> struct PlaneCfg {
  static constexpr int NL = (N1 > N2) ? N1 : N2;
  static constexpr int PLANE = N1 * (N2 + 1);  // padded row stride N2+1
  static constexpr int PLANE_BYTES = PLANE * (int)sizeof(cplx<T>);
  static constexpr int P0 = 256 / NL;
  static constexpr int PL = (65536 / PLANE_BYTES) > 0 ? (65536 / PLANE_BYTES) : 1;
  static constexpr int P = (P0 < PL ? P0 : PL) > 0 ? (P0 < PL ? P0 : PL) : 1;
  static constexpr int THREADS = ((P * NL + 63) / 64) * 64;
  static constexpr int LDS = P * PLANE_BYTES;
};

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_plane_inverse(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                       int total) {
  using C = cplx<T>;
  using Cfg = PlaneCfg<T, N1, N2>;
  constexpr int P = Cfg::P;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> &B = blocks[b];
    const int local = wg - prefix[b];
    const int ngrp = (B.nst + P - 1) / P;
    const int slot = local / ngrp;
    const int s0 = (local - slot * ngrp) * P;
    const int np = min(P, B.nst - s0);
    for (int i = tid; i < np * Cfg::PLANE; i += nth) lds[i] = C{0, 0};
    __syncthreads();
    const int seg0 = B.seg_off[slot], nseg = B.seg_off[slot + 1] - seg0;
    for (int i = tid; i < np * nseg; i += nth) {
      const int p = i / nseg, j = i - p * nseg;
      const int yz = B.pw_yz[seg0 + j];
      const int y = yz / N2, z = yz - y * N2;
      lds[p * Cfg::PLANE + y * (N2 + 1) + z] = B.psi[(size_t)(s0 + p) * B.npw + seg0 + j];
    }
    __syncthreads();
    const int r0 = B.row_off[slot], nrow = B.row_off[slot + 1] - r0;
    for (int i = tid; i < np * nrow; i += nth) {
      const int p = i / nrow;
      const int y = B.rows[r0 + i - p * nrow];
      C *row = lds + p * Cfg::PLANE + y * (N2 + 1);
      C x[N2];
#pragma unroll
      for (int z = 0; z < N2; ++z) x[z] = row[z];
      fft_line<N2, +1, C>::run(x);
#pragma unroll
      for (int z = 0; z < N2; ++z) row[z] = x[z];
    }
    __syncthreads();
    for (int i = tid; i < np * N2; i += nth) {
      const int p = i / N2, z = i - p * N2;
      const C *col = lds + p * Cfg::PLANE + z;
      C x[N1];
#pragma unroll
      for (int y = 0; y < N1; ++y) x[y] = col[y * (N2 + 1)];
      fft_line<N1, +1, C>::run(x);
      C *out = B.t1 + ((size_t)slot * B.nst + s0 + p) * (N1 * N2) + z;
#pragma unroll
      for (int y = 0; y < N1; ++y) out[y * N2] = x[y];
    }
    __syncthreads();
  }
}

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_plane_forward(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                       int total) {
  using C = cplx<T>;
  using Cfg = PlaneCfg<T, N1, N2>;
  constexpr int P = Cfg::P;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> &B = blocks[b];
    const int local = wg - prefix[b];
    const int ngrp = (B.nst + P - 1) / P;
    const int slot = local / ngrp;
    const int s0 = (local - slot * ngrp) * P;
    const int np = min(P, B.nst - s0);
    for (int i = tid; i < np * N2; i += nth) {
      const int p = i / N2, z = i - p * N2;
      const C *in = B.t1 + ((size_t)slot * B.nst + s0 + p) * (N1 * N2) + z;
      C x[N1];
#pragma unroll
      for (int y = 0; y < N1; ++y) x[y] = in[y * N2];
      fft_line<N1, -1, C>::run(x);
      C *col = lds + p * Cfg::PLANE + z;
#pragma unroll
      for (int y = 0; y < N1; ++y) col[y * (N2 + 1)] = x[y];
    }
    __syncthreads();
    const int r0 = B.row_off[slot], nrow = B.row_off[slot + 1] - r0;
    for (int i = tid; i < np * nrow; i += nth) {
      const int p = i / nrow;
      const int y = B.rows[r0 + i - p * nrow];
      C *row = lds + p * Cfg::PLANE + y * (N2 + 1);
      C x[N2];
#pragma unroll
      for (int z = 0; z < N2; ++z) x[z] = row[z];
      fft_line<N2, -1, C>::run(x);
#pragma unroll
      for (int z = 0; z < N2; ++z) row[z] = x[z];
    }
    __syncthreads();
    const int seg0 = B.seg_off[slot], nseg = B.seg_off[slot + 1] - seg0;
    for (int i = tid; i < np * nseg; i += nth) {
      const int p = i / nseg, j = i - p * nseg;
      const int yz = B.pw_yz[seg0 + j];
      const int y = yz / N2, z = yz - y * N2;
      const size_t o = (size_t)(s0 + p) * B.npw + seg0 + j;
      const C v = lds[p * Cfg::PLANE + y * (N2 + 1) + z];
      const C c = B.psi[o];
      const T kf = B.kin[seg0 + j];
      B.hpsi[o] = C{v.x + kf * c.x, v.y + kf * c.y};
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------------------------------
// x passes: one lane per line (i1,i2), 64 consecutive lines per wave, one state per wave
// -------------------------------------------------------------------------------------------
template <typename T, int N0>
__device__ __forceinline__ void load_line_inverse(const SphereBlock<T> &B, int n12, int s, int L, bool valid,
                                                  cplx<T> (&x)[N0]) {
  using C = cplx<T>;
#pragma unroll
  for (int i0 = 0; i0 < N0; ++i0) {
    const int j = B.slot_of[i0];  // uniform
    x[i0] = (j >= 0 && valid) ? B.t1[((size_t)j * B.nst + s) * n12 + L] : C{0, 0};
  }
  fft_line<N0, +1, C>::run(x);
}

template <typename T, int N0>
__global__ void __launch_bounds__(256) k_xpass_project(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                       int total, HamDev hd, int lds_per_wave) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  C *lds = reinterpret_cast<C *>(smem_raw) + wave * lds_per_wave;
  const T *beta = reinterpret_cast<const T *>(hd.beta);
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> &B = blocks[b];
    const int local = wg - prefix[b];
    const int ngrp = (B.nst + nwave - 1) / nwave;
    const int tile = local / ngrp;
    const int s = (local - tile * ngrp) * nwave + wave;
    const int L = tile * kTile + lane;
    const int npts = hd.tile_info[4 * tile + 0];
    const bool active = (s < B.nst) && npts > 0;
    if (active) {
      const bool valid = L < hd.n12;
      C x[N0];
      load_line_inverse<T, N0>(B, hd.n12, s, L, valid, x);
      uint32_t m[3] = {0, 0, 0};
      int cnt = 0;
      if (valid) {
        const uint4 w = reinterpret_cast<const uint4 *>(hd.line_mask)[L];
        m[0] = w.x;
        m[1] = w.y;
        m[2] = w.z;
        cnt = (int)w.w;
      }
#pragma unroll
      for (int i0 = 0; i0 < N0; ++i0) {
        if ((m[i0 >> 5] >> (i0 & 31)) & 1u) {
          lds[cnt] = x[i0];
          ++cnt;
        }
      }
    }
    __syncthreads();
    if (active) {
      const int sl0 = hd.tile_info[4 * tile + 1], sl1 = hd.tile_info[4 * tile + 2];
      for (int sl = sl0; sl < sl1; ++sl) {
        const int a = hd.slot_atom[sl];
        const int e0 = hd.slot_ent[sl], e1 = hd.slot_ent[sl + 1];
        C acc[kMaxNlm];
#pragma unroll
        for (int l = 0; l < kMaxNlm; ++l) acc[l] = C{0, 0};
        for (int e = e0 + lane; e < e1; e += 64) {
          const int ap = hd.ent_ap[e];
          C v = lds[hd.ent_c[e]];
          if (!B.gamma) {
            const float *ps = hd.pos + 3 * (size_t)ap;
            const T arg = (T)(B.kcov[0] * (double)ps[0] + B.kcov[1] * (double)ps[1] + B.kcov[2] * (double)ps[2]);
            T sn, cs;
            sincos_t<T>(arg, &sn, &cs);
            v = cmul(v, C{cs, sn});
          }
          const int p = ap - a * hd.maxsph;
#pragma unroll
          for (int l = 0; l < kMaxNlm; ++l) {
            const T bt = beta[((size_t)a * kMaxNlm + l) * hd.maxsph + p];
            acc[l].x += bt * v.x;
            acc[l].y += bt * v.y;
          }
        }
#pragma unroll
        for (int l = 0; l < kMaxNlm; ++l) {
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) {
            acc[l].x += __shfl_xor(acc[l].x, off, 64);
            acc[l].y += __shfl_xor(acc[l].y, off, 64);
          }
        }
        if (lane == 0) {
#pragma unroll
          for (int l = 0; l < kMaxNlm; ++l) B.partial[((size_t)sl * kMaxNlm + l) * B.nst + s] = acc[l];
        }
      }
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_kb_reduce(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                   int total, HamDev hd) {
  using C = cplx<T>;
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> &B = blocks[b];
    const int idx = (wg - prefix[b]) * blockDim.x + threadIdx.x;
    if (idx >= hd.natoms * B.nst) continue;
    const int a = idx / B.nst, s = idx - a * B.nst;
    double pr[kMaxNlm], pi[kMaxNlm];
#pragma unroll
    for (int l = 0; l < kMaxNlm; ++l) pr[l] = pi[l] = 0.0;
    const int q0 = hd.atom_slots[a], q1 = hd.atom_slots[a + 1];
    for (int q = q0; q < q1; ++q) {
      const int sl = hd.atom_slot_list[q];
#pragma unroll
      for (int l = 0; l < kMaxNlm; ++l) {
        const C v = B.partial[((size_t)sl * kMaxNlm + l) * B.nst + s];
        pr[l] += (double)v.x;
        pi[l] += (double)v.y;
      }
    }
    const double *D = hd.coeff + (size_t)a * kMaxNlm * kMaxNlm;
#pragma unroll
    for (int l = 0; l < kMaxNlm; ++l) {
      double dr = 0.0, di = 0.0;
#pragma unroll
      for (int m = 0; m < kMaxNlm; ++m) {
        dr += D[l * kMaxNlm + m] * pr[m];
        di += D[l * kMaxNlm + m] * pi[m];
      }
      B.dp[((size_t)a * kMaxNlm + l) * B.nst + s] = C{(T)(dr * hd.kb_scale), (T)(di * hd.kb_scale)};
    }
  }
}

template <typename T, int N0>
__global__ void __launch_bounds__(256) k_xpass_apply(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                     int total, HamDev hd, int lds_per_wave) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  C *lds = reinterpret_cast<C *>(smem_raw) + wave * lds_per_wave;
  const T *beta = reinterpret_cast<const T *>(hd.beta);
  for (int wg = blockIdx.x; wg < total; wg += gridDim.x) {
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> &B = blocks[b];
    const int local = wg - prefix[b];
    const int ngrp = (B.nst + nwave - 1) / nwave;
    const int tile = local / ngrp;
    const int s = (local - tile * ngrp) * nwave + wave;
    const int L = tile * kTile + lane;
    const bool valid = (L < hd.n12) && (s < B.nst);
    const int npts = hd.has_kb ? hd.tile_info[4 * tile + 0] : 0;
    C x[N0];
    if (s < B.nst) {
      load_line_inverse<T, N0>(B, hd.n12, s, L, valid, x);
      if (valid) {
#pragma unroll
        for (int i0 = 0; i0 < N0; ++i0) {
          const T v = B.vloc[(size_t)i0 * hd.n12 + L];
          x[i0].x *= v;
          x[i0].y *= v;
        }
      }
      if (npts > 0) {
        // KB values of the tile's sphere points, one point per lane
        const int ptb = hd.tile_info[4 * tile + 3];
        for (int c = lane; c < npts; c += 64) {
          const int q0 = hd.pt_mem[ptb + c], q1 = hd.pt_mem[ptb + c + 1];
          C val = C{0, 0};
          for (int q = q0; q < q1; ++q) {
            const int ap = hd.mem_ap[q];
            const int a = ap / hd.maxsph, p = ap - a * hd.maxsph;
            C t = C{0, 0};
#pragma unroll
            for (int l = 0; l < kMaxNlm; ++l) {
              const T bt = beta[((size_t)a * kMaxNlm + l) * hd.maxsph + p];
              const C d = B.dp[((size_t)a * kMaxNlm + l) * B.nst + s];
              t.x += bt * d.x;
              t.y += bt * d.y;
            }
            if (!B.gamma) {
              const float *ps = hd.pos + 3 * (size_t)ap;
              const T arg = (T)(-(B.kcov[0] * (double)ps[0] + B.kcov[1] * (double)ps[1] + B.kcov[2] * (double)ps[2]));
              T sn, cs;
              sincos_t<T>(arg, &sn, &cs);
              t = cmul(t, C{cs, sn});
            }
            val = cadd(val, t);
          }
          lds[c] = val;
        }
      }
    }
    __syncthreads();
    if (s < B.nst) {
      if (npts > 0 && valid) {
        const uint4 w = reinterpret_cast<const uint4 *>(hd.line_mask)[L];
        const uint32_t m[3] = {w.x, w.y, w.z};
        int cnt = (int)w.w;
#pragma unroll
        for (int i0 = 0; i0 < N0; ++i0) {
          if ((m[i0 >> 5] >> (i0 & 31)) & 1u) {
            x[i0] = cadd(x[i0], lds[cnt]);
            ++cnt;
          }
        }
      }
      fft_line<N0, -1, C>::run(x);
      if (valid) {
#pragma unroll
        for (int i0 = 0; i0 < N0; ++i0) {
          const int j = B.slot_of[i0];
          if (j >= 0) B.t1[((size_t)j * B.nst + s) * hd.n12 + L] = x[i0];
        }
      }
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------------------------------
// host side: tables and launches
// -------------------------------------------------------------------------------------------
#define INQX_SIZES(X) X(12) X(16) X(18) X(20) X(24) X(28) X(30) X(32) X(36) X(40) X(48) X(54) X(60) X(64) X(72) X(80) X(96)

bool grid_supported(const int *n) {
  if (n[1] != n[2]) return false;
  auto ok = [](int v) {
    switch (v) {
#define CASE(S) case S:
      INQX_SIZES(CASE)
#undef CASE
      return true;
      default: return false;
    }
  };
  return ok(n[0]) && ok(n[1]) && n[0] <= 96;
}

void build_kb_tables(inqx_ham_s *h) {
  invalidate_launch_cache(h);
  KBTables &kb = h->kb;
  const Grid &g = h->g;
  hipStream_t st = h->ctx->c.stream;
  const bool dev = !h->ctx->c.host_only;
  const int n12 = g.n[1] * g.n[2];
  kb.natoms = (int)h->projs.size();
  kb.maxsph = 1;
  for (auto &p : h->projs) kb.maxsph = std::max<int>(kb.maxsph, (int)p.npts());
  kb.ntile = (n12 + kTile - 1) / kTile;
  // memberships (a, p) per grid point, grouped by tile / line / i0
  struct Mem {
    int tile, lane, i0, a, p;
  };
  std::vector<Mem> mem;
  for (int a = 0; a < kb.natoms; ++a) {
    const Projector &P = h->projs[a];
    for (long ip = 0; ip < P.npts(); ++ip) {
      const int *c = &P.coords[3 * ip];
      const int L = c[1] * g.n[2] + c[2];
      mem.push_back(Mem{L / kTile, L % kTile, c[0], a, (int)ip});
    }
  }
  // unique points in each tile sorted by (lane, i0): compact index
  std::vector<uint32_t> mask((size_t)n12 * 4, 0u);
  std::vector<std::vector<std::pair<int, int>>> tile_pts(kb.ntile);  // (lane, i0)
  for (auto &m : mem) tile_pts[m.tile].push_back({m.lane, m.i0});
  std::vector<int> tile_info((size_t)kb.ntile * 4, 0);
  std::vector<int> pt_mem, mem_ap;
  std::vector<std::vector<int>> cidx_of(kb.ntile);
  kb.maxpts = 1;
  long ptbase = 0;
  std::vector<int> ptcount(kb.ntile);
  for (int t = 0; t < kb.ntile; ++t) {
    auto &v = tile_pts[t];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    ptcount[t] = (int)v.size();
    kb.maxpts = std::max(kb.maxpts, (int)v.size());
    // line bitmasks and bases
    int c = 0;
    for (size_t k = 0; k < v.size(); ++k) {
      const int L = t * kTile + v[k].first;
      if (k == 0 || v[k].first != v[k - 1].first) mask[(size_t)L * 4 + 3] = (uint32_t)c;
      mask[(size_t)L * 4 + (v[k].second >> 5)] |= 1u << (v[k].second & 31);
      ++c;
    }
    tile_info[4 * t + 0] = (int)v.size();
    tile_info[4 * t + 3] = (int)ptbase;
    ptbase += (long)v.size();
  }
  auto compact = [&](int t, int lane, int i0) {
    auto &v = tile_pts[t];
    return (int)(std::lower_bound(v.begin(), v.end(), std::make_pair(lane, i0)) - v.begin());
  };
  // K4 memberships per unique point, in projector order
  std::vector<std::vector<int>> per_pt(ptbase);
  for (auto &m : mem) {
    long gi = tile_info[4 * m.tile + 3] + compact(m.tile, m.lane, m.i0);
    per_pt[gi].push_back(m.a * kb.maxsph + m.p);
  }
  pt_mem.push_back(0);
  for (auto &v : per_pt) {
    for (int x : v) mem_ap.push_back(x);
    pt_mem.push_back((int)mem_ap.size());
  }
  // K2 slots: (tile, atom) pairs with their entries
  std::vector<Mem> sorted = mem;
  std::stable_sort(sorted.begin(), sorted.end(), [](const Mem &x, const Mem &y) {
    if (x.tile != y.tile) return x.tile < y.tile;
    return x.a < y.a;
  });
  std::vector<int> slot_atom, slot_ent, ent_c, ent_ap;
  std::vector<std::vector<int>> atom_slot(kb.natoms);
  for (size_t k = 0; k < sorted.size(); ++k) {
    const Mem &m = sorted[k];
    if (k == 0 || m.tile != sorted[k - 1].tile || m.a != sorted[k - 1].a) {
      if (k == 0 || m.tile != sorted[k - 1].tile) tile_info[4 * m.tile + 1] = (int)slot_atom.size();
      atom_slot[m.a].push_back((int)slot_atom.size());
      slot_atom.push_back(m.a);
      slot_ent.push_back((int)ent_c.size());
    }
    ent_c.push_back(compact(m.tile, m.lane, m.i0));
    ent_ap.push_back(m.a * kb.maxsph + m.p);
    tile_info[4 * m.tile + 2] = (int)slot_atom.size();
  }
  slot_ent.push_back((int)ent_c.size());
  kb.nslot = (int)slot_atom.size();
  kb.nent = (long)ent_c.size();
  kb.npt = ptbase;
  kb.nmem = (long)mem_ap.size();
  std::vector<int> atom_slots(kb.natoms + 1, 0), atom_slot_list;
  for (int a = 0; a < kb.natoms; ++a) {
    atom_slots[a] = (int)atom_slot_list.size();
    for (int x : atom_slot[a]) atom_slot_list.push_back(x);
  }
  atom_slots[kb.natoms] = (int)atom_slot_list.size();
  // per-atom tables
  std::vector<float> pos((size_t)kb.natoms * kb.maxsph * 3, 0.0f);
  std::vector<double> b64((size_t)kb.natoms * kMaxNlm * kb.maxsph, 0.0);
  std::vector<float> b32(b64.size(), 0.0f);
  std::vector<double> coeff((size_t)kb.natoms * kMaxNlm * kMaxNlm, 0.0);
  for (int a = 0; a < kb.natoms; ++a) {
    const Projector &P = h->projs[a];
    for (long ip = 0; ip < P.npts(); ++ip)
      for (int d = 0; d < 3; ++d) pos[((size_t)a * kb.maxsph + ip) * 3 + d] = P.relpos[3 * ip + d];
    for (int l = 0; l < P.nlm; ++l)
      for (long ip = 0; ip < P.npts(); ++ip) {
        double v = P.matrix[(size_t)l * P.npts() + ip];
        b64[((size_t)a * kMaxNlm + l) * kb.maxsph + ip] = v;
        b32[((size_t)a * kMaxNlm + l) * kb.maxsph + ip] = (float)v;
      }
    for (int l = 0; l < P.nlm; ++l)
      for (int m = 0; m < P.nlm; ++m) coeff[((size_t)a * kMaxNlm + l) * kMaxNlm + m] = P.coeff[l * P.nlm + m];
  }
  if (!dev) {
    kb.built = true;
    return;
  }
  upload(kb.line_mask, mask, st);
  upload(kb.tile_info, tile_info, st);
  upload(kb.slot_atom, slot_atom, st);
  upload(kb.slot_ent, slot_ent, st);
  upload(kb.ent_c, ent_c, st);
  upload(kb.ent_ap, ent_ap, st);
  upload(kb.atom_slots, atom_slots, st);
  upload(kb.atom_slot_list, atom_slot_list, st);
  upload(kb.pt_mem, pt_mem, st);
  upload(kb.mem_ap, mem_ap, st);
  upload(kb.pos, pos, st);
  upload(kb.beta64, b64, st);
  upload(kb.beta32, b32, st);
  upload(kb.coeff, coeff, st);
  INQX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  kb.built = true;
}

void ensure_basis_device(inqx_ham_s *h, Basis &b) {
  hipStream_t st = h->ctx->c.stream;
  if (!b.d_pw_yz.p) {
    upload(b.d_slot_i0, b.slot_i0, st);
    upload(b.d_slot_of, b.slot_of, st);
    upload(b.d_seg_off, b.seg_off, st);
    upload(b.d_row_off, b.row_off, st);
    upload(b.d_rows, b.rows, st);
    upload(b.d_pw_yz, b.pw_yz, st);
    INQX_HIP(hipStreamSynchronize(st));
  }
  if (b.kin_A[0] != h->A[0] || b.kin_A[1] != h->A[1] || b.kin_A[2] != h->A[2]) {
    std::vector<double> k64(b.npw);
    std::vector<float> k32(b.npw);
    const Grid &g = h->g;
    for (long i = 0; i < b.npw; ++i) {
      const int f = b.gidx[i];
      const int i2 = f % g.n[2], i1 = (f / g.n[2]) % g.n[1], i0 = f / (g.n[2] * g.n[1]);
      double v[3] = {2.0 * M_PI * Grid::sym(i0, g.n[0]) + b.kcov[0] + h->A[0],
                     2.0 * M_PI * Grid::sym(i1, g.n[1]) + b.kcov[1] + h->A[1],
                     2.0 * M_PI * Grid::sym(i2, g.n[2]) + b.kcov[2] + h->A[2]};
      double c[3];
      g.cov_to_cart(v, c);
      k64[i] = 0.5 * (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      k32[i] = (float)k64[i];
    }
    upload(b.d_kin64, k64, st);
    upload(b.d_kin32, k32, st);
    INQX_HIP(hipStreamSynchronize(st));
    for (int d = 0; d < 3; ++d) b.kin_A[d] = h->A[d];
  }
}

void ensure_workspace(inqx_ham_s *h, Basis &b, int nst, int prec) {
  const size_t cs = prec == INQX_FP64 ? 16 : 8;
  if (b.ws_nst[prec] >= nst && b.t1[prec].p && b.ws_kb_nslot[prec] == h->kb.nslot &&
      b.ws_natoms[prec] == h->kb.natoms)
    return;
  const int n12 = h->g.n[1] * h->g.n[2];
  const int ns = std::max(nst, b.ws_nst[prec]);
  b.t1[prec].ensure(cs * (size_t)b.slot_i0.size() * ns * n12);
  b.partial[prec].ensure(cs * (size_t)std::max(1, h->kb.nslot) * kMaxNlm * ns);
  b.dp[prec].ensure(cs * (size_t)std::max(1, h->kb.natoms) * kMaxNlm * ns);
  b.ws_nst[prec] = ns;
  b.ws_kb_nslot[prec] = h->kb.nslot;
  b.ws_natoms[prec] = h->kb.natoms;
  invalidate_launch_cache(h);
}

// ---- per-launch descriptor cache
struct LaunchKey {
  std::vector<uintptr_t> k;
  bool operator<(const LaunchKey &o) const { return k < o.k; }
};
struct LaunchData {
  DevMem desc;  // SphereBlock array followed by 4 prefix arrays
  int total_plane = 0, total_x = 0, total_red = 0;
  size_t off_pref[3] = {0, 0, 0};
  int nwave = 4;
  int lds_per_wave = 0;
};
static std::map<std::pair<inqx_ham_s *, LaunchKey>, std::unique_ptr<LaunchData>> g_launch_cache;

// descriptors hold raw pointers to workspaces, tables and bases: drop every
// cached descriptor of a Hamiltonian whenever one of those may change
void invalidate_launch_cache(inqx_ham_s *h) {
  for (auto it = g_launch_cache.begin(); it != g_launch_cache.end();) {
    if (it->first.first == h) it = g_launch_cache.erase(it);
    else ++it;
  }
}

template <typename T, int N1, int N2> static void set_plane_attr() {
  static bool done = false;
  if (done) return;
  using Cfg = PlaneCfg<T, N1, N2>;
  if (Cfg::LDS > 65536) {
    INQX_HIP(hipFuncSetAttribute((const void *)k_plane_inverse<T, N1, N2>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, Cfg::LDS));
    INQX_HIP(hipFuncSetAttribute((const void *)k_plane_forward<T, N1, N2>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, Cfg::LDS));
  }
  done = true;
}

template <typename T, int N0> static void set_x_attr(int lds) {
  static int done = 65536;
  if (lds <= done) return;
  INQX_HIP(hipFuncSetAttribute((const void *)k_xpass_project<T, N0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  INQX_HIP(hipFuncSetAttribute((const void *)k_xpass_apply<T, N0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  done = lds;
}

template <typename T, int N0, int N12>
static void run_passes(inqx_ham_s *h, LaunchData &ld, int nblocks, const HamDev &hd) {
  using Cfg = PlaneCfg<T, N12, N12>;
  hipStream_t st = h->ctx->c.stream;
  const SphereBlock<T> *blk = ld.desc.as<SphereBlock<T>>();
  const int *pref_plane = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[0]);
  const int *pref_x = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[1]);
  const int *pref_red = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[2]);
  set_plane_attr<T, N12, N12>();
  const int cu = h->ctx->c.num_cu;
  const int grid_plane = std::min(ld.total_plane, cu * 16);
  Ctx &cx = h->ctx->c;
  const bool f64 = std::is_same<T, double>::value;
  {
  ProfScope ps(cx, f64 ? "plane_inverse_f64" : "plane_inverse_f32");
  hipLaunchKernelGGL((k_plane_inverse<T, N12, N12>), dim3(grid_plane), dim3(Cfg::THREADS), Cfg::LDS, st, blk,
                     pref_plane, nblocks, ld.total_plane);
  }
  INQX_HIP(hipGetLastError());
  const int xthreads = 64 * ld.nwave;
  const int xlds = ld.nwave * ld.lds_per_wave * (int)sizeof(cplx<T>);
  set_x_attr<T, N0>(xlds);
  const int grid_x = std::min(ld.total_x, cu * 16);
  if (hd.has_kb) {
    {
      ProfScope ps(cx, f64 ? "xpass_project_f64" : "xpass_project_f32");
      hipLaunchKernelGGL((k_xpass_project<T, N0>), dim3(grid_x), dim3(xthreads), xlds, st, blk, pref_x, nblocks,
                         ld.total_x, hd, ld.lds_per_wave);
    }
    INQX_HIP(hipGetLastError());
    {
      ProfScope ps(cx, f64 ? "kb_reduce_f64" : "kb_reduce_f32");
      hipLaunchKernelGGL((k_kb_reduce<T>), dim3(std::min(ld.total_red, cu * 16)), dim3(256), 0, st, blk, pref_red,
                         nblocks, ld.total_red, hd);
    }
    INQX_HIP(hipGetLastError());
  }
  {
    ProfScope ps(cx, f64 ? "xpass_apply_f64" : "xpass_apply_f32");
    hipLaunchKernelGGL((k_xpass_apply<T, N0>), dim3(grid_x), dim3(xthreads), xlds, st, blk, pref_x, nblocks,
                       ld.total_x, hd, ld.lds_per_wave);
  }
  INQX_HIP(hipGetLastError());
  {
    ProfScope ps(cx, f64 ? "plane_forward_f64" : "plane_forward_f32");
    hipLaunchKernelGGL((k_plane_forward<T, N12, N12>), dim3(grid_plane), dim3(Cfg::THREADS), Cfg::LDS, st, blk,
                       pref_plane, nblocks, ld.total_plane);
  }
  INQX_HIP(hipGetLastError());
}

template <typename T>
void launch_apply(inqx_ham_s *h, int nblocks, Basis *const *bases, const int *nst, const void *const *psi,
                  void *const *hpsi) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const Grid &g = h->g;
  INQX_CHECK(h->has_v, INQX_ERR_STATE, "potential not set");
  INQX_CHECK(grid_supported(g.n), INQX_ERR_UNSUPPORTED, "grid size not built for the sphere path");
  if (!h->kb.built) build_kb_tables(h);
  const bool has_kb = h->kb.natoms > 0;
  if (prec == INQX_FP32 && !h->v32.p) INQX_CHECK(false, INQX_ERR_STATE, "potential not uploaded");
  LaunchKey key;
  key.k.push_back((uintptr_t)prec);
  for (int i = 0; i < nblocks; ++i) {
    ensure_basis_device(h, *bases[i]);
    ensure_workspace(h, *bases[i], nst[i], prec);
    key.k.push_back((uintptr_t)bases[i]);
    key.k.push_back((uintptr_t)nst[i]);
    key.k.push_back((uintptr_t)psi[i]);
    key.k.push_back((uintptr_t)hpsi[i]);
    key.k.push_back((uintptr_t)bases[i]->t1[prec].p);
  }
  for (int d = 0; d < 3; ++d) {
    uint64_t bits;
    std::memcpy(&bits, &h->A[d], sizeof bits);
    key.k.push_back((uintptr_t)bits);
  }
  auto ck = std::make_pair(h, key);
  auto it = g_launch_cache.find(ck);
  if (it == g_launch_cache.end()) {
    if (g_launch_cache.size() > 256) g_launch_cache.clear();
    auto ld = std::make_unique<LaunchData>();
    std::vector<SphereBlock<T>> blk(nblocks);
    const int P = PlaneCfg<T, 16, 16>::P;  // placeholder, replaced below by the real plane config
    (void)P;
    std::vector<int> pp(nblocks + 1, 0), px(nblocks + 1, 0), pr(nblocks + 1, 0);
    int plane_p = 1;
    switch (g.n[1]) {
#define CASE(S)                            \
  case S:                                  \
    plane_p = PlaneCfg<T, S, S>::P;        \
    break;
      INQX_SIZES(CASE)
#undef CASE
      default: break;
    }
    // waves (states) per x-pass workgroup within the LDS budget
    const int cs = (int)sizeof(cplx<T>);
    ld->lds_per_wave = h->kb.maxpts;
    int nw = 4;
    while (nw > 1 && nw * ld->lds_per_wave * cs > 160 * 1024) --nw;
    INQX_CHECK(ld->lds_per_wave * cs <= 160 * 1024, INQX_ERR_UNSUPPORTED, "KB tile too dense for LDS");
    ld->nwave = nw;
    const double *Afull = h->A;
    for (int i = 0; i < nblocks; ++i) {
      Basis &B = *bases[i];
      SphereBlock<T> &d = blk[i];
      d.psi = reinterpret_cast<const cplx<T> *>(psi[i]);
      d.hpsi = reinterpret_cast<cplx<T> *>(hpsi[i]);
      d.t1 = B.t1[prec].as<cplx<T>>();
      d.partial = B.partial[prec].as<cplx<T>>();
      d.dp = B.dp[prec].as<cplx<T>>();
      d.slot_i0 = B.d_slot_i0.as<int>();
      d.slot_of = B.d_slot_of.as<int>();
      d.seg_off = B.d_seg_off.as<int>();
      d.row_off = B.d_row_off.as<int>();
      d.rows = B.d_rows.as<int>();
      d.pw_yz = B.d_pw_yz.as<int>();
      d.kin = prec == INQX_FP64 ? (const T *)B.d_kin64.p : (const T *)B.d_kin32.p;
      const DevMem &vm = prec == INQX_FP64 ? h->v64 : h->v32;
      d.vloc = reinterpret_cast<const T *>(vm.as<char>() + (size_t)B.spin * g.N * sizeof(T));
      d.nst = nst[i];
      d.npw = (int)B.npw;
      d.nslot = (int)B.slot_i0.size();
      for (int k = 0; k < 3; ++k) d.kcov[k] = B.kcov[k] + Afull[k];
      d.gamma = (d.kcov[0] == 0.0 && d.kcov[1] == 0.0 && d.kcov[2] == 0.0) ? 1 : 0;
      pp[i + 1] = pp[i] + d.nslot * ((nst[i] + plane_p - 1) / plane_p);
      px[i + 1] = px[i] + h->kb.ntile * ((nst[i] + nw - 1) / nw);
      pr[i + 1] = pr[i] + (std::max(1, h->kb.natoms) * nst[i] + 255) / 256;
    }
    ld->total_plane = pp[nblocks];
    ld->total_x = px[nblocks];
    ld->total_red = pr[nblocks];
    size_t bytes = sizeof(SphereBlock<T>) * nblocks;
    bytes = (bytes + 255) & ~size_t(255);
    ld->off_pref[0] = bytes;
    ld->off_pref[1] = bytes + sizeof(int) * (nblocks + 1);
    ld->off_pref[2] = bytes + 2 * sizeof(int) * (nblocks + 1);
    size_t tot = bytes + 3 * sizeof(int) * (nblocks + 1);
    std::vector<char> host(tot);
    std::memcpy(host.data(), blk.data(), sizeof(SphereBlock<T>) * nblocks);
    std::memcpy(host.data() + ld->off_pref[0], pp.data(), sizeof(int) * (nblocks + 1));
    std::memcpy(host.data() + ld->off_pref[1], px.data(), sizeof(int) * (nblocks + 1));
    std::memcpy(host.data() + ld->off_pref[2], pr.data(), sizeof(int) * (nblocks + 1));
    ld->desc.ensure(tot);
    INQX_HIP(hipMemcpyAsync(ld->desc.p, host.data(), tot, hipMemcpyHostToDevice, h->ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(h->ctx->c.stream));
    it = g_launch_cache.emplace(ck, std::move(ld)).first;
  }
  LaunchData &ld = *it->second;
  HamDev hd;
  hd.n0 = g.n[0];
  hd.n1 = g.n[1];
  hd.n2 = g.n[2];
  hd.n12 = g.n[1] * g.n[2];
  hd.N = g.N;
  hd.natoms = h->kb.natoms;
  hd.maxsph = h->kb.maxsph;
  hd.has_kb = has_kb ? 1 : 0;
  hd.ntile = h->kb.ntile;
  hd.line_mask = h->kb.line_mask.as<uint32_t>();
  hd.tile_info = h->kb.tile_info.as<int>();
  hd.slot_atom = h->kb.slot_atom.as<int>();
  hd.slot_ent = h->kb.slot_ent.as<int>();
  hd.ent_c = h->kb.ent_c.as<int>();
  hd.ent_ap = h->kb.ent_ap.as<int>();
  hd.atom_slots = h->kb.atom_slots.as<int>();
  hd.atom_slot_list = h->kb.atom_slot_list.as<int>();
  hd.pt_mem = h->kb.pt_mem.as<int>();
  hd.mem_ap = h->kb.mem_ap.as<int>();
  hd.pos = h->kb.pos.as<float>();
  hd.beta = prec == INQX_FP64 ? h->kb.beta64.p : h->kb.beta32.p;
  hd.coeff = h->kb.coeff.as<double>();
  hd.kb_scale = (g.vol / (double)g.N) / (double)g.N;
  switch (g.n[0] * 1000 + g.n[1]) {
#define CASE(S)                                            \
  case S * 1000 + S:                                       \
    run_passes<T, S, S>(h, ld, nblocks, hd);               \
    break;
    INQX_SIZES(CASE)
#undef CASE
    default:
      INQX_CHECK(false, INQX_ERR_UNSUPPORTED, "only cubic grid shapes are built for the sphere path");
  }
}

template void launch_apply<double>(inqx_ham_s *, int, Basis *const *, const int *, const void *const *, void *const *);
template void launch_apply<float>(inqx_ham_s *, int, Basis *const *, const int *, const void *const *, void *const *);

}  // namespace inqx
