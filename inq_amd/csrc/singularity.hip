// singularity.hip -- the exchange singularity correction of Carrier et al.,
// PRB 75, 205126 (2007), as hamiltonian::singularity_correction computes it
// (src/hamiltonian/singularity_correction.hpp:19-103).
//
// f(q) = 4 pi^2 / sum_j [4 |b_j sin(a_j.q/2)|^2 + 2 (b_j sin(a_j.q)).(b_{j+1} sin(a_{j+1}.q))]
// (the auxiliary function of Eq. 16), summed over a k-point set for f_k and
// integrated over a nested cubic grid of q for f_0.  The f_0 grid (7 scales x
// 61 x 121 x 121 points, the inner third of each scale left out) is summed
// on the device: one work-group per (scale, ikx) plane with a fixed-order
// reduction, then the 427 plane sums in order on the host, so the result does
// not depend on timing.  The value the Poisson solves of the exchange use at
// G + q = 0 is sing(ik) = -nk V (f_k - f_0).
#include "internal.hpp"

#include <cmath>
#include <vector>

namespace inqx {
namespace {

constexpr int kSingSteps = 7;
constexpr int kSingNk = 60;
constexpr int kSingLine = 2 * kSingNk + 1;  // 121 points per axis of a scale

struct SingCell {
  double a[3][3];  // lattice vectors (rows, cartesian)
  double b[3][3];  // reciprocal vectors (rows, cartesian, 2 pi included)
};

__host__ __device__ inline double sing_aux(const SingCell &c, const double qc[3]) {
  double val = 0.0;
  for (int j = 0; j < 3; ++j) {
    const int jp = j == 2 ? 0 : j + 1;
    const double aq = c.a[j][0] * qc[0] + c.a[j][1] * qc[1] + c.a[j][2] * qc[2];
    const double apq = c.a[jp][0] * qc[0] + c.a[jp][1] * qc[1] + c.a[jp][2] * qc[2];
    const double s1 = sin(0.5 * aq), s2 = sin(aq), s3 = sin(apq);
    const double bb = c.b[j][0] * c.b[j][0] + c.b[j][1] * c.b[j][1] + c.b[j][2] * c.b[j][2];
    const double bbp = c.b[j][0] * c.b[jp][0] + c.b[j][1] * c.b[jp][1] + c.b[j][2] * c.b[jp][2];
    val += 4.0 * bb * s1 * s1 + 2.0 * bbp * s2 * s3;
  }
  return 4.0 * M_PI * M_PI / val;
}

// covariant q -> cartesian: q_cart = sum_k q_k b_k / (2 pi)
__host__ __device__ inline void sing_cov_to_cart(const SingCell &c, const double q[3], double qc[3]) {
  for (int d = 0; d < 3; ++d)
    qc[d] = (q[0] * c.b[0][d] + q[1] * c.b[1][d] + q[2] * c.b[2][d]) / (2.0 * M_PI);
}

// one work-group per (scale, ikx): sum of f over (iky, ikz) of that plane,
// times the scale's k-volume element, outside the inner third of the scale
__global__ void __launch_bounds__(256) k_sing_fzero(SingCell c, double vol, double *part) {
  __shared__ double red[256];
  const int step = blockIdx.x / (kSingNk + 1);
  const int ikx = blockIdx.x - step * (kSingNk + 1);
  double length = 1.0, kvol = pow(2.0 * M_PI / (2.0 * kSingNk + 1.0), 3) / vol;
  for (int s = 0; s < step; ++s) {
    length /= 3.0;
    kvol /= 27.0;
  }
  double acc = 0.0;
  for (int e = threadIdx.x; e < kSingLine * kSingLine; e += 256) {
    const int iky = e / kSingLine - kSingNk, ikz = e % kSingLine - kSingNk;
    if (ikx <= kSingNk / 3 && abs(iky) <= kSingNk / 3 && abs(ikz) <= kSingNk / 3) continue;
    const double q[3] = {2.0 * M_PI * ikx * length / (2.0 * kSingNk), 2.0 * M_PI * iky * length / (2.0 * kSingNk),
                         2.0 * M_PI * ikz * length / (2.0 * kSingNk)};
    double qc[3];
    sing_cov_to_cart(c, q, qc);
    acc += kvol * sing_aux(c, qc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace
}  // namespace inqx

using namespace inqx;

extern "C" {

int inqx_singularity_correction(inqx_ctx ctx, const double lattice[9], int nk, const double *kcov,
                                const double *kweight, double *fk, double *fzero) {
  API_BEGIN
  INQX_CHECK(ctx && lattice && fzero && nk >= 0 && (nk == 0 || (kcov && kweight && fk)), INQX_ERR_ARG,
             "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  Grid g;
  const int n1[3] = {1, 1, 1};
  g.init(lattice, n1);
  INQX_CHECK(g.vol > 1e-10, INQX_ERR_ARG, "the lattice volume is not positive");
  SingCell c;
  for (int r = 0; r < 3; ++r)
    for (int d = 0; d < 3; ++d) {
      c.a[r][d] = g.a[r][d];
      c.b[r][d] = g.b[r][d];
    }
  // f_k: sum over the other k-points of weight * f(k - k') (singularity_correction.hpp:50-60)
  for (int ik = 0; ik < nk; ++ik) {
    double s = 0.0;
    for (int ik2 = 0; ik2 < nk; ++ik2) {
      const double q[3] = {kcov[3 * ik] - kcov[3 * ik2], kcov[3 * ik + 1] - kcov[3 * ik2 + 1],
                           kcov[3 * ik + 2] - kcov[3 * ik2 + 2]};
      double qc[3];
      sing_cov_to_cart(c, q, qc);
      if (qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2] < 1e-6) continue;  // metric norm (squared)
      s += kweight[ik2] * sing_aux(c, qc);
    }
    fk[ik] = s * 4.0 * M_PI / g.vol;
  }
  // f_0 (singularity_correction.hpp:62-84)
  const int nblk = kSingSteps * (kSingNk + 1);
  DevMem part;
  part.ensure(sizeof(double) * nblk);
  hipLaunchKernelGGL(k_sing_fzero, dim3(nblk), dim3(256), 0, ctx->c.stream, c, g.vol, part.as<double>());
  INQX_HIP(hipGetLastError());
  std::vector<double> h(nblk);
  INQX_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * nblk, hipMemcpyDeviceToHost, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  double fz = 0.0;
  for (int i = 0; i < nblk; ++i) fz += h[i];
  double length = 1.0;
  for (int s = 0; s < kSingSteps - 1; ++s) length /= 3.0;
  fz *= 8.0 * M_PI / std::pow(2.0 * M_PI, 3);
  fz += 4.0 * M_PI * std::pow(3.0 / (4.0 * M_PI), 1.0 / 3.0) * std::pow(g.vol, 2.0 / 3.0) / M_PI / g.vol * length;
  *fzero = fz;
  API_END(ctx)
}

}  // extern "C"
