// plane_rect.hip -- the plane passes of the sphere path on grids whose plane
// is not square (n1 != n2).
//
// A square plane runs each pass as one kernel per plane size that keeps the
// plane in LDS between its row and column transforms (ham_kernels.hip).  A
// kernel per (n1, n2) pair would square the number of instantiations, so a
// non-square plane splits each pass at the t1 planes instead, every kernel
// built for one line length:
//
//   inverse  rows_inverse<N2>: sphere -> LDS plane (row window), FFT along i2
//            of the rows that hold sphere points, whole plane to t1 (zeros
//            outside the window); cols<N1, +1>: FFT along i1 in place in t1
//   forward  cols<N1, -1> in place in t1; rows_forward<N2>: the sphere rows
//            of t1 -> LDS, FFT along i2, gather of the sphere (+ 1/2|G+k|^2 psi
//            and the Taylor accumulators, as the square forward pass)
//
// The same transforms as the square passes (transform.hpp:52-182 of the
// reference, restricted to the sphere), with one more round trip of the
// planes through t1 per pass.
#include "plane_common.hpp"

namespace inqx {
namespace {

template <typename T, int N2>
__global__ void __launch_bounds__(256) k_rows_inverse(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                      int total, int pb, int pstr, int n1) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = nth >> 6;
  const bool upper = lane >= 32;
  const int wg = blockIdx.x;  // one work item per work-group
  if (wg >= total) return;
  SphereBlock<T> B;
  const PlaneItem it = plane_item(blocks, prefix, nblocks, wg, pb, B);
  const int slot = it.slot, s0 = it.s0, np = it.np;
  const V4<int> sg = ld4<int>(B.slot_seg, 2 * slot);
  const int seg0 = sg.x, nseg = sg.y, r0 = sg.z, nrow = sg.w;
  const V4<int> sw = ld4<int>(B.slot_seg, 2 * slot + 1);
  const int wlo = sw.x, nwin = sw.y;  // row window of the plane in LDS
  const float rseg = 1.0f / (float)max(nseg, 1), rrow = 1.0f / (float)max(nrow, 1);
  const GAS int *pwl = gp(B.pw_lds) + seg0;
  const GAS int *rows = gp(B.rows) + r0;
  constexpr int KG = gather_round<T>();
  const int tot = np * nseg, nr = np * nrow;
  // the sphere points into the cleared planes
  for (int i = tid; i < np * pstr; i += nth) lds[i] = C{0, 0};
  __syncthreads();
  if (tot > 0) {
    const auto rpsi = make_rsrc(B.psi + (size_t)s0 * B.npw + seg0,
                                (unsigned)(((size_t)(np - 1) * B.npw + nseg) * sizeof(C)));
    int ldx[KG];
    T vr[KG], vi[KG];
    for (int i0 = tid; i0 < tot; i0 += KG * nth) {
      scatter_round<T, KG>(rpsi, pwl, B.npw, nseg, rseg, tot, i0, nth, ldx, vr, vi);
#pragma unroll
      for (int u = 0; u < KG; ++u) {
        const int i = i0 + u * nth;
        if (i < tot) lds[div_rcp(i, nseg, rseg) * pstr + ldx[u]] = C{vr[u], vi[u]};
      }
    }
  }
  __syncthreads();
  // rows along i2 (sphere rows only), a lane pair per row
  for (int g = wave; g * 32 < nr; g += nwave) {
    const int li = g * 32 + (lane & 31);
    const bool ok = li < nr;
    const int lc = ok ? li : nr - 1;
    const int p = div_rcp(lc, nrow, rrow);
    const int y = rows[lc - p * nrow];
    plane_row_split<C, N2, +1>(lds + p * pstr + plane_row_n(y, wlo, n1) * (N2 + 1), upper, ok);
  }
  __syncthreads();
  // the whole planes to t1 (slot `slot` of states s0 .. s0+np-1), rows outside
  // the row window zero
  const int n12 = n1 * N2;
  const size_t pstride = (size_t)B.nslot * n12;  // between the states' planes
  GAS C *dst = gw(B.t1) + (size_t)s0 * pstride + (size_t)slot * n12;
  const float r12 = 1.0f / (float)n12;
  for (int e = tid; e < np * n12; e += nth) {
    const int p = div_rcp(e, n12, r12), r = e - p * n12;
    const int y = r / N2, z = r - y * N2;
    C v = C{0, 0};
    if (row_in_plane_n(y, wlo, nwin, n1)) v = lds[p * pstr + plane_row_n(y, wlo, n1) * (N2 + 1) + z];
    dst[(size_t)p * pstride + r] = v;
  }
}

// FFT along i1 of every column of the np planes, in place in t1, a lane pair
// per column (no LDS: the column's values stay in the pair's registers)
template <typename T, int N1, int SIGN>
__global__ void __launch_bounds__(256) k_plane_cols(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                    int total, int pb, int n2) {
  using C = cplx<T>;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = nth >> 6;
  const bool upper = lane >= 32;
  const int half = upper ? 1 : 0;
  const int wg = blockIdx.x;
  if (wg >= total) return;
  SphereBlock<T> B;
  const PlaneItem it = plane_item(blocks, prefix, nblocks, wg, pb, B);
  const int slot = it.slot, s0 = it.s0, np = it.np;
  // a plane with no sphere points is zero (inverse) or not gathered (forward)
  if (ld4<int>(B.slot_seg, 2 * slot).y == 0) return;
  const int n12 = N1 * n2;
  const size_t pstride = (size_t)B.nslot * n12;
  const auto rt = make_rsrc(B.t1 + (size_t)s0 * pstride + (size_t)slot * n12,
                            (unsigned)(((size_t)(np - 1) * pstride + n12) * sizeof(C)));
  const float rn2 = 1.0f / (float)n2;
  for (int g = wave; g * 32 < np * n2; g += nwave) {
    const int li = g * 32 + (lane & 31);
    const bool ok = li < np * n2;
    const int lc = ok ? li : np * n2 - 1;
    const int p = div_rcp(lc, n2, rn2), z = lc - p * n2;
    // rows y = 2n + half in, rows y = k + half N1/2 out
    const unsigned vi = (unsigned)(((size_t)p * pstride + z + half * n2) * sizeof(C));
    C v[N1 / 2];
#pragma unroll
    for (int n = 0; n < N1 / 2; ++n) bload(rt, vi, (unsigned)(2 * n * n2 * (int)sizeof(C)), v[n]);
    fft_split<N1, SIGN, C>::dit(v, upper);
    // lanes past the last column store out of range (dropped)
    const unsigned vo = ok ? (unsigned)(((size_t)p * pstride + z + half * (N1 / 2) * n2) * sizeof(C)) : kOOB;
#pragma unroll
    for (int k = 0; k < N1 / 2; ++k) bstore(rt, vo, (unsigned)(k * n2 * (int)sizeof(C)), v[k]);
  }
}

template <typename T, int N2>
__global__ void __launch_bounds__(256) k_rows_forward(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                      int total, int pb, int pstr, int n1, T c1r, T c1i, T c2r,
                                                      T c2i) {
  using C = cplx<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = nth >> 6;
  const bool upper = lane >= 32;
  const int wg = blockIdx.x;
  if (wg >= total) return;
  SphereBlock<T> B;
  const PlaneItem it = plane_item(blocks, prefix, nblocks, wg, pb, B);
  const int slot = it.slot, s0 = it.s0, np = it.np;
  const V4<int> sg = ld4<int>(B.slot_seg, 2 * slot);
  const int seg0 = sg.x, nseg = sg.y, r0 = sg.z, nrow = sg.w;
  if (nseg == 0) return;  // uniform: nothing of this plane is on the sphere
  const V4<int> sw = ld4<int>(B.slot_seg, 2 * slot + 1);
  const int wlo = sw.x;
  const float rseg = 1.0f / (float)nseg, rrow = 1.0f / (float)nrow;
  const GAS int *pwl = gp(B.pw_lds) + seg0;
  const GAS int *rows = gp(B.rows) + r0;
  const int tot = np * nseg, nr = np * nrow;
  // the sphere rows of the np planes (columns already transformed) into LDS
  const int n12 = n1 * N2;
  const size_t pstride = (size_t)B.nslot * n12;
  const GAS C *src = gp(B.t1) + (size_t)s0 * pstride + (size_t)slot * n12;
  for (int e = tid; e < nr * N2; e += nth) {
    const int lr = e / N2, z = e - lr * N2;
    const int p = div_rcp(lr, nrow, rrow);
    const int y = rows[lr - p * nrow];
    lds[p * pstr + plane_row_n(y, wlo, n1) * (N2 + 1) + z] = src[(size_t)p * pstride + y * N2 + z];
  }
  __syncthreads();
  for (int g = wave; g * 32 < nr; g += nwave) {
    const int li = g * 32 + (lane & 31);
    const bool ok = li < nr;
    const int lc = ok ? li : nr - 1;
    const int p = div_rcp(lc, nrow, rrow);
    const int y = rows[lc - p * nrow];
    plane_row_split<C, N2, -1>(lds + p * pstr + plane_row_n(y, wlo, n1) * (N2 + 1), upper, ok);
  }
  __syncthreads();
  const GatherBufs R = gather_bufs(B, s0, seg0, np, nseg);
  constexpr int KG = gather_round<T>();
  int lx[KG];
  for (int i0 = tid; i0 < tot; i0 += KG * nth) {
    gather_offsets<KG>(lx, pwl, i0, nth, tot, nseg, rseg, pstr);
    gather_update<T, KG>(B, R, lds, lx, i0, nth, tot, nseg, rseg, c1r, c1i, c2r, c2i);
  }
}

constexpr int kRectThreads = 256;

// dynamic LDS up to the CU's 160 KB (a complex128 plane of 98 x 97 rows x columns is 152 KB)
template <typename K> void allow_lds(K kernel, int device, DeviceOnce &once) {
  if (once.test_and_set(device)) return;
  INQX_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
}

template <typename T, int N2>
void rows_inverse(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb, int pstr,
                  int n1, int lds) {
  static DeviceOnce once;
  allow_lds(k_rows_inverse<T, N2>, cx.device, once);
  hipLaunchKernelGGL((k_rows_inverse<T, N2>), dim3(total), dim3(kRectThreads), lds, cx.stream, blk, pref, nblocks,
                     total, pb, pstr, n1);
}

template <typename T, int N2>
void rows_forward(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb, int pstr,
                  int n1, int lds, const double *cf) {
  static DeviceOnce once;
  allow_lds(k_rows_forward<T, N2>, cx.device, once);
  hipLaunchKernelGGL((k_rows_forward<T, N2>), dim3(total), dim3(kRectThreads), lds, cx.stream, blk, pref, nblocks,
                     total, pb, pstr, n1, (T)cf[0], (T)cf[1], (T)cf[2], (T)cf[3]);
}

template <typename T, int N1, int SIGN>
void cols(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb, int n2) {
  hipLaunchKernelGGL((k_plane_cols<T, N1, SIGN>), dim3(total), dim3(kRectThreads), 0, cx.stream, blk, pref, nblocks,
                     total, pb, n2);
}

}  // namespace

bool rect_plane_supported(int n1, int n2) {
  auto ok = [](int v) {
    switch (v) {
#define CASE(S) case S:
      INQX_SIZES(CASE)
#undef CASE
      return true;
      default: return false;
    }
  };
  return ok(n1) && ok(n2);
}

int rect_plane_batch(int prec, int n1, int n2) {
  const int cs = prec == INQX_FP64 ? 16 : 8;
  const int plane = (n1 + 2) * (n2 + 1) * cs;  // the row window (at most n1 rows) and two more
  const int p0 = std::max(1, 128 / std::max(n1, n2));  // about 128 lines per work-group
  return std::max(1, std::min(p0, 65536 / plane));
}

template <typename T>
void launch_rect_inverse(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb,
                         int pstr, int n1, int n2) {
  INQX_CHECK(rect_plane_supported(n1, n2), INQX_ERR_UNSUPPORTED, "plane size not built for the sphere path");
  const bool f64 = std::is_same<T, double>::value;
  const int lds = pb * pstr * (int)sizeof(cplx<T>);
  INQX_CHECK(lds <= 160 * 1024, INQX_ERR_UNSUPPORTED, "the plane passes' row window does not fit the 160 KB of LDS");
  {
    ProfScope ps(cx, f64 ? "plane_rows_inverse_f64" : "plane_rows_inverse_f32");
    switch (n2) {
#define CASE(S)                                                          \
  case S:                                                                \
    rows_inverse<T, S>(cx, blk, pref, nblocks, total, pb, pstr, n1, lds); \
    break;
      INQX_SIZES(CASE)
#undef CASE
    }
    INQX_HIP(hipGetLastError());
  }
  {
    ProfScope ps(cx, f64 ? "plane_cols_inverse_f64" : "plane_cols_inverse_f32");
    switch (n1) {
#define CASE(S)                                                 \
  case S:                                                       \
    cols<T, S, +1>(cx, blk, pref, nblocks, total, pb, n2);      \
    break;
      INQX_SIZES(CASE)
#undef CASE
    }
    INQX_HIP(hipGetLastError());
  }
}

template <typename T>
void launch_rect_forward(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb,
                         int pstr, int n1, int n2, const double *cf) {
  INQX_CHECK(rect_plane_supported(n1, n2), INQX_ERR_UNSUPPORTED, "plane size not built for the sphere path");
  const bool f64 = std::is_same<T, double>::value;
  const int lds = pb * pstr * (int)sizeof(cplx<T>);
  INQX_CHECK(lds <= 160 * 1024, INQX_ERR_UNSUPPORTED, "the plane passes' row window does not fit the 160 KB of LDS");
  {
    ProfScope ps(cx, f64 ? "plane_cols_forward_f64" : "plane_cols_forward_f32");
    switch (n1) {
#define CASE(S)                                                 \
  case S:                                                       \
    cols<T, S, -1>(cx, blk, pref, nblocks, total, pb, n2);      \
    break;
      INQX_SIZES(CASE)
#undef CASE
    }
    INQX_HIP(hipGetLastError());
  }
  {
    ProfScope ps(cx, f64 ? "plane_rows_forward_f64" : "plane_rows_forward_f32");
    switch (n2) {
#define CASE(S)                                                              \
  case S:                                                                    \
    rows_forward<T, S>(cx, blk, pref, nblocks, total, pb, pstr, n1, lds, cf); \
    break;
      INQX_SIZES(CASE)
#undef CASE
    }
    INQX_HIP(hipGetLastError());
  }
}

template void launch_rect_inverse<float>(Ctx &, const SphereBlock<float> *, const int *, int, int, int, int, int, int);
template void launch_rect_inverse<double>(Ctx &, const SphereBlock<double> *, const int *, int, int, int, int, int,
                                          int);
template void launch_rect_forward<float>(Ctx &, const SphereBlock<float> *, const int *, int, int, int, int, int, int,
                                         const double *);
template void launch_rect_forward<double>(Ctx &, const SphereBlock<double> *, const int *, int, int, int, int, int,
                                          int, const double *);

}  // namespace inqx
