// layout.hip -- Fourier-space states in the reference's layouts.
//
// The reference keeps Fourier-space orbitals point-major: the full cube
// [Gx][Gy][Gz][ist] of an orbital_set<fourier_space> (states/orbital_set.hpp,
// basis/fourier_space.hpp), or [ipw][ist] for a point list.  The sphere path
// works band-major on the sphere, [ist][ipw].  These kernels move a block
// between the two through LDS tiles of 64 points x 16 states, so both the
// point-major rows (states contiguous) and the band-major rows (points
// contiguous) are read and written in coalesced runs.
#include "internal.hpp"

namespace inqx {
namespace {

constexpr int kTP = 64;  // points per tile
constexpr int kTS = 16;  // states per tile

// band[s][i] = src[row(i) nst + s], row(i) = rows ? rows[i] : i
template <typename C>
__global__ void __launch_bounds__(256) k_point_to_band(const C *__restrict__ src, const int *__restrict__ rows,
                                                       int npw, int nst, int ntp, C *__restrict__ band) {
  __shared__ C tile[kTP][kTS + 1];
  const int tp = blockIdx.x % ntp, ts = blockIdx.x / ntp;
  const int i0 = tp * kTP, s0 = ts * kTS;
  for (int e = threadIdx.x; e < kTP * kTS; e += blockDim.x) {
    const int il = e / kTS, sl = e - il * kTS;
    const int i = i0 + il, s = s0 + sl;
    if (i < npw && s < nst) tile[il][sl] = src[(size_t)(rows ? rows[i] : i) * nst + s];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTP * kTS; e += blockDim.x) {
    const int sl = e / kTP, il = e - sl * kTP;
    const int i = i0 + il, s = s0 + sl;
    if (i < npw && s < nst) band[(size_t)s * npw + i] = tile[il][sl];
  }
}

// dst[row(i) nst + s] = band[s][i]
template <typename C>
__global__ void __launch_bounds__(256) k_band_to_point(const C *__restrict__ band, const int *__restrict__ rows,
                                                       int npw, int nst, int ntp, C *__restrict__ dst) {
  __shared__ C tile[kTP][kTS + 1];
  const int tp = blockIdx.x % ntp, ts = blockIdx.x / ntp;
  const int i0 = tp * kTP, s0 = ts * kTS;
  for (int e = threadIdx.x; e < kTP * kTS; e += blockDim.x) {
    const int sl = e / kTP, il = e - sl * kTP;
    const int i = i0 + il, s = s0 + sl;
    if (i < npw && s < nst) tile[il][sl] = band[(size_t)s * npw + i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTP * kTS; e += blockDim.x) {
    const int il = e / kTS, sl = e - il * kTS;
    const int i = i0 + il, s = s0 + sl;
    if (i < npw && s < nst) dst[(size_t)(rows ? rows[i] : i) * nst + s] = tile[il][sl];
  }
}

template <typename C>
void launch_layout(Ctx &cx, bool to_band, const void *src, const int *rows, long npw, int nst, void *dst) {
  const int ntp = (int)((npw + kTP - 1) / kTP), nts = (nst + kTS - 1) / kTS;
  const long grid = (long)ntp * nts;
  INQX_CHECK(grid <= INT32_MAX && npw <= INT32_MAX, INQX_ERR_ARG, "state block too large for the layout kernels");
  if (grid == 0) return;
  ProfScope ps(cx, to_band ? "layout_to_band" : "layout_from_band");
  if (to_band)
    hipLaunchKernelGGL(k_point_to_band<C>, dim3((unsigned)grid), dim3(256), 0, cx.stream, (const C *)src, rows,
                       (int)npw, nst, ntp, (C *)dst);
  else
    hipLaunchKernelGGL(k_band_to_point<C>, dim3((unsigned)grid), dim3(256), 0, cx.stream, (const C *)src, rows,
                       (int)npw, nst, ntp, (C *)dst);
  INQX_HIP(hipGetLastError());
}

}  // namespace

void launch_point_to_band(Ctx &cx, int prec, const void *src, const int *rows, long npw, int nst, void *band) {
  if (prec == INQX_FP64) launch_layout<cplx<double>>(cx, true, src, rows, npw, nst, band);
  else launch_layout<cplx<float>>(cx, true, src, rows, npw, nst, band);
}

void launch_band_to_point(Ctx &cx, int prec, const void *band, const int *rows, long npw, int nst, void *dst) {
  if (prec == INQX_FP64) launch_layout<cplx<double>>(cx, false, band, rows, npw, nst, dst);
  else launch_layout<cplx<float>>(cx, false, band, rows, npw, nst, dst);
}

}  // namespace inqx
