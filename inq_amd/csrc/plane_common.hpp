// plane_common.hpp -- device helpers shared by the plane passes of the sphere
// path: global-address-space pointers, raw buffer loads and stores, the
// plane layout in LDS, the scatter round and the plane work items.
#pragma once
#include "internal.hpp"

// global address space in the device pass; the host pass of the same
// templates sees plain pointers
#if defined(__HIP_DEVICE_COMPILE__)
#define GAS __attribute__((address_space(1)))
#else
#define GAS
#endif

namespace inqx {

template <typename T> __device__ __forceinline__ const GAS T *gp(const T *p) { return (const GAS T *)p; }
template <typename T> __device__ __forceinline__ GAS T *gw(T *p) { return (GAS T *)p; }

// 4- and 2-wide loads of plain records from global memory
template <typename S> struct V4 {
  S x, y, z, w;
};
template <typename S> struct V2 {
  S x, y;
};
template <typename S> __device__ __forceinline__ V4<S> ld4(const void *p, size_t i) {
  const GAS S *q = (const GAS S *)__builtin_assume_aligned(p, 4 * sizeof(S)) + 4 * i;
  return V4<S>{q[0], q[1], q[2], q[3]};
}
template <typename S> __device__ __forceinline__ V2<S> ld2(const void *p, size_t i) {
  const GAS S *q = (const GAS S *)__builtin_assume_aligned(p, 2 * sizeof(S)) + 2 * i;
  return V2<S>{q[0], q[1]};
}

__device__ __forceinline__ int find_block(const int *prefix, int nblocks, int wg) {
  const GAS int *pf = gp(prefix);
  int lo = 0, hi = nblocks;  // prefix[lo] <= wg < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pf[mid] <= wg) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- raw buffer access: per-lane byte offset in a VGPR, wave-uniform part in an SGPR
// the base and size of a buffer resource are wave-uniform by construction;
// readfirstlane makes that explicit so the resource stays in SGPRs
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, unsigned bytes) {
  const unsigned long long a = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void *p = (void *)(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)__builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, cplx<float> &v) {
  const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  v.x = __int_as_float(q[0]);
  v.y = __int_as_float(q[1]);
}
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, cplx<double> &v) {
  const auto q = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  v.x = __longlong_as_double(((long long)(unsigned)q[1] << 32) | (unsigned)q[0]);
  v.y = __longlong_as_double(((long long)(unsigned)q[3] << 32) | (unsigned)q[2]);
}
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float &v) {
  v = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double &v) {
  const auto q = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  v = __longlong_as_double(((long long)(unsigned)q[1] << 32) | (unsigned)q[0]);
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, cplx<float> v) {
  typedef int i2 __attribute__((ext_vector_type(2)));
  i2 q = {__float_as_int(v.x), __float_as_int(v.y)};
  __builtin_amdgcn_raw_buffer_store_b64(q, r, voff, soff, 0);
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, cplx<double> v) {
  typedef int i4 __attribute__((ext_vector_type(4)));
  const long long a = __double_as_longlong(v.x), b = __double_as_longlong(v.y);
  i4 q = {(int)a, (int)(a >> 32), (int)b, (int)(b >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(q, r, voff, soff, 0);
}

// a per-lane buffer offset beyond every record: loads return 0, stores are dropped
constexpr unsigned kOOB = 0x80000000u;

// Plane layout in LDS: a plane holds only its row window (the symmetric range
// of i1 that holds sphere rows, nwin rows from wlo) at LDS rows 1 .. nwin, so
// more work-groups fit per CU (51 of 72 rows at the 72^3 sphere: complex64
// planes of 29 KB, complex128 planes of 62 KB, two work-groups per CU).
// plane_row(y) is the LDS row of a row y that holds sphere points;
// row_in_plane(y) whether row y has an LDS row at all.
// (plane_row_n and row_in_plane_n take the plane's row count n at run time)
__device__ __forceinline__ int plane_row_n(int y, int wlo, int n) { return (y >= (n + 1) / 2 ? y - n : y) - wlo + 1; }
__device__ __forceinline__ bool row_in_plane_n(int y, int wlo, int nwin, int n) {
  return (unsigned)((y >= (n + 1) / 2 ? y - n : y) - wlo) < (unsigned)nwin;
}
template <int N> __device__ __forceinline__ int plane_row(int y, int wlo) { return plane_row_n(y, wlo, N); }
template <int N> __device__ __forceinline__ bool row_in_plane(int y, int wlo, int nwin) {
  return row_in_plane_n(y, wlo, nwin, N);
}

// i / d for 0 <= i < 2^24 and d >= 1 from the float reciprocal rd = 1/d: the
// float quotient is off by at most one, corrected from the remainder
__device__ __forceinline__ int div_rcp(int i, int d, float rd) {
  int q = (int)((float)i * rd);
  const int r = i - q * d;
  q += r >= d ? 1 : 0;
  q -= r < 0 ? 1 : 0;
  return q;
}

// FFT of one LDS row of N values split over the lane pair; lanes with ok ==
// false transform a copy of a valid row and store nothing
template <typename C, int N, int SIGN>
__device__ __forceinline__ void plane_row_split(C *row, bool upper, bool ok) {
  constexpr int H = N / 2;
  C v[H];
#pragma unroll
  for (int n = 0; n < H; ++n) v[n] = row[2 * n + (upper ? 1 : 0)];
  fft_split<N, SIGN, C>::dit(v, upper);
  __builtin_amdgcn_wave_barrier();
  if (ok) {
#pragma unroll
    for (int k = 0; k < H; ++k) row[k + (upper ? H : 0)] = v[k];
  }
}

// the transform of a line split over S lanes (part q of the line): input
// x[S n + q], output X[k + (N/S) line_part_out<S>(q)]
template <int N, int SIGN, typename C, int S>
__device__ __forceinline__ void line_dit(C (&v)[N / S], int q) {
  if constexpr (S == 2) fft_split<N, SIGN, C>::dit(v, q != 0);
  else fft_quad<N, SIGN, C>::dit(v, q);
}
template <int S> __device__ __forceinline__ int line_part_out(int q) { return S == 2 ? q : quad_rev(q); }
// the mirror: input X[k + (N/S) line_part_out<S>(q)] (line_dit's output order),
// output x[S k + q] (line_dit's input order)
template <int N, int SIGN, typename C, int S>
__device__ __forceinline__ void line_dif(C (&v)[N / S], int q) {
  if constexpr (S == 2) fft_split<N, SIGN, C>::dif(v, q != 0);
  else fft_quad<N, SIGN, C>::dif(v, q);
}

// FFT of one LDS row of N values split over S lanes; lanes with ok == false
// transform a copy of a valid row and store nothing
template <typename C, int N, int SIGN, int S>
__device__ __forceinline__ void plane_row_lanes(C *row, int q, bool ok) {
  if constexpr (S == 2) {
    plane_row_split<C, N, SIGN>(row, q != 0, ok);
  } else {
    constexpr int M = N / S;
    C v[M];
#pragma unroll
    for (int n = 0; n < M; ++n) v[n] = row[S * n + q];
    line_dit<N, SIGN, C, S>(v, q);
    __builtin_amdgcn_wave_barrier();
    if (ok) {
      const int o = M * line_part_out<S>(q);
#pragma unroll
      for (int k = 0; k < M; ++k) row[k + o] = v[k];
    }
  }
}

// In both plane passes every row and column of a plane is split over LANES
// lanes of a wave; P planes share a work-group of at most 256 lanes.
// complex64: a lane pair (lane, lane + 32), 32 lines per wave.  complex128, when
// both plane axes divide by 4: four lanes (lane + 16 q, part q = 0..3), 16 lines
// per wave; a quarter line per lane keeps the passes near 170 VGPRs, so two
// work-groups share a CU (a lane pair per line needs ~300 VGPRs, one wave per
// SIMD: plane_forward 337 -> 251 us, plane_inverse 265 -> 221 us at Si64).
template <typename T, int N1, int N2> struct PlaneCfg {
  static_assert(N1 % 2 == 0 && N2 % 2 == 0, "split lines need even lengths");
  static constexpr int LANES = (sizeof(T) == 8 && N1 % 4 == 0 && N2 % 4 == 0) ? 4 : 2;
  static constexpr int LPW = 64 / LANES;  // lines per wave
  static constexpr int NL = (N1 > N2) ? N1 : N2;
  // a plane in LDS (padded row stride N2+1): the rows of its row window (at
  // most N1) and two more
  static constexpr int PLANE = (N1 + 2) * (N2 + 1);
  static constexpr int PLANE_BYTES = PLANE * (int)sizeof(cplx<T>);
  static constexpr int P0 = 128 / NL > 0 ? 128 / NL : 1;  // about 128 lines per work-group
  static constexpr int PL = (65536 / PLANE_BYTES) > 0 ? (65536 / PLANE_BYTES) : 1;
  static constexpr int P = (P0 < PL ? P0 : PL) > 0 ? (P0 < PL ? P0 : PL) : 1;
  static constexpr int WAVES = (P * NL + LPW - 1) / LPW < 4 ? (P * NL + LPW - 1) / LPW : 4;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int LDS = P * PLANE_BYTES;
};

// sphere points per thread in one round of the plane passes' scatter and
// gather: all their loads are in flight together (fewer for complex128, whose
// transforms leave no registers for more without spilling)
template <typename T> constexpr int gather_round() { return sizeof(T) == 4 ? 16 : 8; }


// loads of one scatter round: KG sphere points per thread from point index
// i0 in steps of nth (past the end: the last point again, not stored later).
// Point i is point j of plane p (i = p nseg + j); pwl[j] is its LDS offset in
// the plane.  (Real and imaginary parts in separate arrays: an array of
// 16-byte complex structs here was kept in scratch memory instead of registers.)
template <typename T, int KG>
__device__ __forceinline__ void scatter_round(__amdgpu_buffer_rsrc_t rpsi, const GAS int *pwl, int npw, int nseg,
                                              float rseg, int tot, int i0, int nth, int (&ldx)[KG],
                                              T (&vr)[KG], T (&vi)[KG]) {
#pragma unroll
  for (int u = 0; u < KG; ++u) {
    const int i = min(i0 + u * nth, tot - 1);
    const int p = div_rcp(i, nseg, rseg), j = i - p * nseg;
    ldx[u] = pwl[j];  // the plane's offset p pstr is added when the value is stored
    cplx<T> v;
    bload(rpsi, (unsigned)((p * npw + j) * (int)sizeof(cplx<T>)), 0u, v);
    vr[u] = v.x;
    vi[u] = v.y;
  }
}

// work item of a plane-pass work-group: block, slot and first state (the
// block's work-groups take consecutive slots of the same states)
struct PlaneItem {
  int b, slot, s0, np;
};
template <typename T>
__device__ __forceinline__ PlaneItem plane_item(const SphereBlock<T> *blocks, const int *prefix, int nblocks, int wg,
                                                int P, SphereBlock<T> &B) {
  PlaneItem it;
  it.b = find_block(prefix, nblocks, wg);
  B = blocks[it.b];
  const int local = wg - gp(prefix)[it.b];
  const int sgrp = local / B.nslot;
  it.slot = local - sgrp * B.nslot;
  it.s0 = sgrp * P;
  it.np = min(P, B.nst - it.s0);
  return it;
}

// ---- the forward pass's gather: sphere values out of the transformed planes
// in LDS, plus 1/2|G+k|^2 psi, into H psi and the Taylor accumulators

// buffer resources of the gather based at point seg0 of state s0: one 32-bit
// offset per point; points past the end use an offset beyond the buffer (loads
// give zero, stores are dropped) instead of a branch
struct GatherBufs {
  __amdgpu_buffer_rsrc_t psi, kin, h, a1, a2;
};
template <typename T>
__device__ __forceinline__ GatherBufs gather_bufs(const SphereBlock<T> &B, int s0, int seg0, int np, int nseg) {
  using C = cplx<T>;
  GatherBufs R;
  const size_t base = (size_t)s0 * B.npw + seg0;
  const unsigned nb = (unsigned)(((size_t)(np - 1) * B.npw + nseg) * sizeof(C));
  R.psi = make_rsrc(B.psi + base, nb);
  R.kin = make_rsrc(B.kin ? (const void *)(B.kin + seg0) : (const void *)B.psi, B.kin ? (unsigned)(nseg * sizeof(T)) : 0u);
  R.h = make_rsrc(B.hpsi + base, nb);
  // accumulators in the working precision, or complex128 (acc_wide: an FP64
  // state accumulating FP32 H applications; offsets then scale by 2)
  const unsigned wsh = B.acc_wide ? 1u : 0u;
  R.a1 = make_rsrc(B.acc1 ? (const void *)((const char *)B.acc1 + (base * sizeof(C) << wsh)) : (const void *)(B.hpsi + base),
                   B.acc1 ? nb << wsh : 0u);
  R.a2 = make_rsrc(B.acc2 ? (const void *)((const char *)B.acc2 + (base * sizeof(C) << wsh)) : (const void *)(B.hpsi + base),
                   B.acc2 ? nb << wsh : 0u);
  return R;
}

// LDS offsets of one gather round: KG sphere points per thread from point
// index i0 in steps of nth (past the end: the last point again)
template <int KG>
__device__ __forceinline__ void gather_offsets(int (&lx)[KG], const GAS int *pwl, int i0, int nth, int tot, int nseg,
                                               float rseg, int pstr) {
#pragma unroll
  for (int u = 0; u < KG; ++u) {
    const int i = i0 + u * nth;
    const int ic = i < tot ? i : tot - 1;
    const int p = div_rcp(ic, nseg, rseg), j = ic - p * nseg;
    lx[u] = p * pstr + pwl[j];
  }
}

// a + c out for complex a in the working precision or complex128 (wide)
// (wide only for T = float: complex128 states accumulate in complex128 anyway)
template <typename T>
__device__ __forceinline__ void acc_update(__amdgpu_buffer_rsrc_t r, unsigned o, bool wide, T cr, T ci, cplx<T> out) {
  if (sizeof(T) == 4 && wide) {
    cplx<double> q;
    const unsigned ow = o == kOOB ? kOOB : o << 1;
    bload(r, ow, 0u, q);
    q.x += (double)cr * (double)out.x - (double)ci * (double)out.y;
    q.y += (double)cr * (double)out.y + (double)ci * (double)out.x;
    bstore(r, ow, 0u, q);
  } else {
    cplx<T> q;
    bload(r, o, 0u, q);
    q.x += cr * out.x - ci * out.y;
    q.y += cr * out.y + ci * out.x;
    bstore(r, o, 0u, q);
  }
}

// one gather round at the offsets lx: the sphere values and kinetic factors
// are loaded first, then the LDS values, so that no memory load waits behind
// another; then H psi (stored or added) and the accumulators
template <typename T, int KG>
__device__ __forceinline__ void gather_update(const SphereBlock<T> &B, const GatherBufs &R, const cplx<T> *lds,
                                              const int (&lx)[KG], int i0, int nth, int tot, int nseg, float rseg,
                                              T c1r, T c1i, T c2r, T c2i) {
  using C = cplx<T>;
  unsigned oo[KG];
  C outv[KG], cv[KG];
  T kv[KG];
#pragma unroll
  for (int u = 0; u < KG; ++u) {
    const int i = i0 + u * nth;
    const int ic = i < tot ? i : tot - 1;
    const int p = div_rcp(ic, nseg, rseg), j = ic - p * nseg;
    oo[u] = i < tot ? (unsigned)((p * B.npw + j) * (int)sizeof(C)) : kOOB;
    if (B.kin) {
      bload(R.psi, oo[u], 0u, cv[u]);
      bload(R.kin, (unsigned)(j * (int)sizeof(T)), 0u, kv[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < KG; ++u) outv[u] = lds[lx[u]];
#pragma unroll
  for (int u = 0; u < KG; ++u) {
    C out = outv[u];
    if (B.kin) {  // H application: + 1/2|G+k|^2 psi; without it a plain masked forward FFT
      out.x += kv[u] * cv[u].x;
      out.y += kv[u] * cv[u].y;
    }
    if (B.write_h) {
      if (B.accumulate) {
        C q;
        bload(R.h, oo[u], 0u, q);
        q.x += out.x;
        q.y += out.y;
        bstore(R.h, oo[u], 0u, q);
      } else {
        bstore(R.h, oo[u], 0u, out);
      }
    }
    if (B.acc1) acc_update<T>(R.a1, oo[u], B.acc_wide != 0, c1r, c1i, out);
    if (B.acc2) acc_update<T>(R.a2, oo[u], B.acc_wide != 0, c2r, c2i, out);
  }
}

}  // namespace inqx
