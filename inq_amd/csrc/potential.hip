// potential.hip -- atomic-sphere fields on the real-space grid.
//
// field[p] += sum_a sum_T f_{s(a)}(|r_p - R_a - T|) over the lattice images T
// with |r_p - R_a - T| <= radius_a: the per-atom sphere sums of
// atomic_potential::local_potential (short-range radial tables added on each
// atom's sphere, src/hamiltonian/atomic_potential.hpp:116-165) and
// atomic_potential::ionic_density (valence charge x long-range density on the
// sphere of the range separation, atomic_potential.hpp:170-200).  The reference
// scatters each sphere point with an atomic add; here one lane owns one grid
// point and adds its atoms and images in a fixed order (atoms ascending, images
// in loop order), so the result is reproducible bit for bit.
//
// Grid point p = (i0 n1 + i1) n2 + i2 sits at contravariant (i0/n0, i1/n1, i2/n2).
// The images needed along axis k satisfy |x_k + s_k| <= radius |b_k| / 2 pi, since
// the contravariant component of a cartesian vector d is (b_k / 2 pi) . d.
// Radial tables are uniform, r = j dr (j < nr), interpolated linearly.
#include "internal.hpp"

namespace inqx {

struct AtomFieldArgs {
  int n0, n1, n2;
  double a[3][3];  // lattice vectors, rows
  double rln[3];   // |b_k| / 2 pi
  int natoms, nr;
  double inv_dr;
  const double *atoms;   // [natoms][kAtomRec]
  const double *tables;  // [nspecies][nr]
  double *field;         // [N]
};

__global__ void __launch_bounds__(256) k_atomic_field_add(AtomFieldArgs d) {
  const long N = (long)d.n0 * d.n1 * d.n2;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  const int i2 = (int)(p % d.n2), i1 = (int)((p / d.n2) % d.n1), i0 = (int)(p / ((long)d.n1 * d.n2));
  const double xp[3] = {(double)i0 / d.n0, (double)i1 / d.n1, (double)i2 / d.n2};
  double acc = 0.0;
  for (int at = 0; at < d.natoms; ++at) {
    const double *rec = d.atoms + (long)at * kAtomRec;
    const double rad = rec[3];
    const double *tab = d.tables + (long)rec[4] * d.nr;
    double f[3];
    int lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      f[k] = xp[k] - rec[k];
      const double w = rad * d.rln[k];
      lo[k] = (int)ceil(-w - f[k]);
      hi[k] = (int)floor(w - f[k]);
    }
    for (int s0 = lo[0]; s0 <= hi[0]; ++s0)
      for (int s1 = lo[1]; s1 <= hi[1]; ++s1)
        for (int s2 = lo[2]; s2 <= hi[2]; ++s2) {
          const double c0 = f[0] + s0, c1 = f[1] + s1, c2 = f[2] + s2;
          const double x = c0 * d.a[0][0] + c1 * d.a[1][0] + c2 * d.a[2][0];
          const double y = c0 * d.a[0][1] + c1 * d.a[1][1] + c2 * d.a[2][1];
          const double z = c0 * d.a[0][2] + c1 * d.a[1][2] + c2 * d.a[2][2];
          const double r = sqrt(x * x + y * y + z * z);
          if (r > rad) continue;
          const double u = r * d.inv_dr;
          int j = (int)u;
          if (j > d.nr - 2) j = d.nr - 2;  // r <= radius <= (nr - 1) dr
          const double t = u - j;
          acc += (1.0 - t) * tab[j] + t * tab[j + 1];
        }
  }
  d.field[p] += acc;
}

void launch_atomic_field_add(inqx_ham_s *h, int natoms, const double *atoms_dev, int nr, double dr,
                             const double *tables_dev, double *field) {
  const Grid &g = h->g;
  AtomFieldArgs d;
  d.n0 = g.n[0];
  d.n1 = g.n[1];
  d.n2 = g.n[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) d.a[i][j] = g.a[i][j];
  for (int k = 0; k < 3; ++k) d.rln[k] = std::sqrt(g.rl[k][0] * g.rl[k][0] + g.rl[k][1] * g.rl[k][1] + g.rl[k][2] * g.rl[k][2]);
  d.natoms = natoms;
  d.nr = nr;
  d.inv_dr = 1.0 / dr;
  d.atoms = atoms_dev;
  d.tables = tables_dev;
  d.field = field;
  Ctx &cx = h->ctx->c;
  ProfScope ps(cx, "atomic_field_add");
  hipLaunchKernelGGL(k_atomic_field_add, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, cx.stream, d);
  INQX_HIP(hipGetLastError());
}

}  // namespace inqx
