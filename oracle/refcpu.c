/* refcpu.c -- the reference's CPU data flow for H|psi>, overlap + Cholesky and
 * the exchange pair solves, on the FFTW3 / BLAS / LAPACK interfaces.
 *
 * TEST INFRASTRUCTURE: this is the timed CPU baseline of the benchmark (and is
 * checked against the oracle in the tests); the product never loads it.
 *
 * It follows the reference's CPU path, not the GPU build's design:
 *   - orbitals on the full grid, point-major [x][y][z][ist] (field_set.hpp:52,62);
 *   - H on fourier_space orbitals (ks_hamiltonian.hpp:160-180): to_real as a
 *     batched 3-D FFT over the state axis followed by a separate x 1/N pass
 *     (transform.hpp:113-182), V psi (scalar_potential.hpp:60-63), the KB term
 *     as gather / DGEMM on interleaved re,im / coefficients / DGEMM / scatter-add
 *     (projector_all.hpp:113-288), to_fourier (transform.hpp:52-109) and
 *     laplacian_add (laplacian.hpp:24-37);
 *   - overlap as ZGEMM over the full grid (overlap.hpp:26-95) and ZPOTRF
 *     (cholesky.hpp:28-100);
 *   - exchange pair solves rho = conj(phi_j) phi_i, forward FFT, -4 pi/N *
 *     (-1/|G|^2), inverse FFT, accumulation (exchange_operator.hpp:124-161,
 *     poisson.hpp:71-110).
 * The reference runs one MPI rank per core over states; here OpenMP threads
 * own contiguous chunks of states (their own FFT plans) and the library runs
 * sequentially inside each thread.
 *
 * The FFTW3 / BLAS / LAPACK symbols are taken at run time from Intel MKL's
 * single runtime library (libmkl_rt) when it can be loaded, with MKL's threading
 * layer set to sequential; otherwise the oracle's own FFT and plain loops are
 * used, and rc_backend() says which. */
#define _GNU_SOURCE
#include <complex.h>
#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double _Complex cplx;
typedef void *fftw_plan_t;

/* FFTW3 and BLAS/LAPACK entry points (public interface signatures) */
static fftw_plan_t (*x_plan_many)(int rank, const int *n, int howmany, void *in, const int *inembed, int istride,
                                  int idist, void *out, const int *onembed, int ostride, int odist, int sign,
                                  unsigned flags);
static void (*x_execute_dft)(const fftw_plan_t p, void *in, void *out);
static void (*x_destroy_plan)(fftw_plan_t p);
static void (*x_dgemm)(const char *ta, const char *tb, const int *m, const int *n, const int *k, const double *alpha,
                       const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
                       double *c, const int *ldc);
static void (*x_zgemm)(const char *ta, const char *tb, const int *m, const int *n, const int *k, const cplx *alpha,
                       const cplx *a, const int *lda, const cplx *b, const int *ldb, const cplx *beta, cplx *c,
                       const int *ldc);
static void (*x_zpotrf)(const char *uplo, const int *n, cplx *a, const int *lda, int *info);
static int (*x_set_threading_layer)(int);

#define FFTW_FORWARD_SIGN (-1)
#define FFTW_BACKWARD_SIGN (+1)
#define FFTW_ESTIMATE_FLAG (1U << 6)
#define MKL_THREADING_SEQUENTIAL 1

static int g_backend = -1; /* 1: MKL FFTW3/BLAS/LAPACK, 0: own FFT and loops */

int rc_init(void) {
  if (g_backend >= 0) return g_backend;
  const char *names[] = {"libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1", NULL};
  void *lib = NULL;
  for (int i = 0; names[i] && !lib; ++i) lib = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
  g_backend = 0;
  if (!lib) return 0;
  *(void **)&x_set_threading_layer = dlsym(lib, "MKL_Set_Threading_Layer");
  *(void **)&x_plan_many = dlsym(lib, "fftw_plan_many_dft");
  *(void **)&x_execute_dft = dlsym(lib, "fftw_execute_dft");
  *(void **)&x_destroy_plan = dlsym(lib, "fftw_destroy_plan");
  *(void **)&x_dgemm = dlsym(lib, "dgemm_");
  *(void **)&x_zgemm = dlsym(lib, "zgemm_");
  *(void **)&x_zpotrf = dlsym(lib, "zpotrf_");
  if (x_set_threading_layer && x_plan_many && x_execute_dft && x_destroy_plan && x_dgemm && x_zgemm && x_zpotrf) {
    x_set_threading_layer(MKL_THREADING_SEQUENTIAL);
    g_backend = 1;
  }
  return g_backend;
}

int rc_backend(void) { return g_backend; }

/* ----------------------------------------------------------------- thread chunks of states */
typedef struct {
  int nthreads, nst;
  int *s0, *ns;
  fftw_plan_t *fwd, *bwd; /* per thread: howmany = ns[t], stride nst, in place on the chunk's first column */
} chunks_t;

static void chunks_init(chunks_t *c, const int n[3], int nst, int nthreads, cplx *layout_buf) {
  c->nthreads = nthreads < nst ? nthreads : nst;
  c->nst = nst;
  c->s0 = (int *)calloc(c->nthreads, sizeof(int));
  c->ns = (int *)calloc(c->nthreads, sizeof(int));
  c->fwd = (fftw_plan_t *)calloc(c->nthreads, sizeof(fftw_plan_t));
  c->bwd = (fftw_plan_t *)calloc(c->nthreads, sizeof(fftw_plan_t));
  for (int t = 0; t < c->nthreads; ++t) {
    c->s0[t] = (int)((long)nst * t / c->nthreads);
    c->ns[t] = (int)((long)nst * (t + 1) / c->nthreads) - c->s0[t];
    if (g_backend == 1) {
      cplx *p = layout_buf + c->s0[t];
      c->fwd[t] = x_plan_many(3, n, c->ns[t], p, NULL, nst, 1, p, NULL, nst, 1, FFTW_FORWARD_SIGN, FFTW_ESTIMATE_FLAG);
      c->bwd[t] = x_plan_many(3, n, c->ns[t], p, NULL, nst, 1, p, NULL, nst, 1, FFTW_BACKWARD_SIGN, FFTW_ESTIMATE_FLAG);
    }
  }
}

static void chunks_free(chunks_t *c) {
  for (int t = 0; t < c->nthreads; ++t) {
    if (c->fwd[t]) x_destroy_plan(c->fwd[t]);
    if (c->bwd[t]) x_destroy_plan(c->bwd[t]);
  }
  free(c->s0);
  free(c->ns);
  free(c->fwd);
  free(c->bwd);
}

/* unnormalised 3-D FFT of the thread's chunk of columns of data [N][nst], in place */
static void chunk_fft(const chunks_t *c, int t, const int n[3], long N, cplx *data, int sign) {
  if (g_backend == 1) {
    cplx *p = data + c->s0[t];
    x_execute_dft(sign < 0 ? c->fwd[t] : c->bwd[t], p, p);
    return;
  }
  /* own FFT on a packed copy of the chunk */
  const int ns = c->ns[t], nst = c->nst, s0 = c->s0[t];
  cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)N * ns);
  for (long ip = 0; ip < N; ++ip)
    for (int s = 0; s < ns; ++s) tmp[ip * ns + s] = data[ip * nst + s0 + s];
  orc_fft3d(n, ns, tmp, sign);
  for (long ip = 0; ip < N; ++ip)
    for (int s = 0; s < ns; ++s) data[ip * nst + s0 + s] = tmp[ip * ns + s];
  free(tmp);
}

/* ----------------------------------------------------------------- Hamiltonian on the full cube */
typedef struct {
  int n[3];
  long N;
  int nst, natoms, nlm;
  double dv;
  double *v;      /* [N] */
  double *kin;    /* [N] 1/2 |G|^2 of every Fourier point (Gamma) */
  long *off;      /* [natoms + 1] offsets into pts */
  int *pts;       /* sphere point flat indices */
  double *beta;   /* per atom [nlm][npts] */
  double *coeff;  /* per atom [nlm][nlm] */
  cplx *work;     /* [N][nst] real-space orbitals */
  chunks_t ch;
} rc_ham;

static int symi(int i, int n) { return i >= (n + 1) / 2 ? i - n : i; }

static double cell_volume(const double *a) {
  return fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
              a[2] * (a[3] * a[7] - a[4] * a[6]));
}

/* 1/2 |G|^2 of every Fourier point, G = sum_k m_k b_k with m in the symmetric
 * range and b_k = 2 pi (a_l x a_m) / V */
static void kinetic_table(const double *a, const int n[3], double *kin) {
  const double vol = cell_volume(a);
  double b[3][3];
  for (int k = 0; k < 3; ++k) {
    const double *u = a + 3 * ((k + 1) % 3), *w = a + 3 * ((k + 2) % 3);
    b[k][0] = 2 * M_PI * (u[1] * w[2] - u[2] * w[1]) / vol;
    b[k][1] = 2 * M_PI * (u[2] * w[0] - u[0] * w[2]) / vol;
    b[k][2] = 2 * M_PI * (u[0] * w[1] - u[1] * w[0]) / vol;
  }
  for (int i0 = 0; i0 < n[0]; ++i0)
    for (int i1 = 0; i1 < n[1]; ++i1)
      for (int i2 = 0; i2 < n[2]; ++i2) {
        const int m[3] = {symi(i0, n[0]), symi(i1, n[1]), symi(i2, n[2])};
        double g[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 3; ++c) g[c] += m[k] * b[k][c];
        kin[((long)i0 * n[1] + i1) * n[2] + i2] = 0.5 * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      }
}

rc_ham *rc_ham_create(const double lattice[9], const int n[3], int nst, const double *v, int natoms, const int *npts,
                      const int *pts, int nlm, const double *beta, const double *coeff, int nthreads) {
  rc_init();
  rc_ham *h = (rc_ham *)calloc(1, sizeof(rc_ham));
  for (int d = 0; d < 3; ++d) h->n[d] = n[d];
  h->N = (long)n[0] * n[1] * n[2];
  h->nst = nst;
  h->natoms = natoms;
  h->nlm = nlm;
  h->dv = cell_volume(lattice) / (double)h->N;
  h->v = (double *)malloc(sizeof(double) * h->N);
  memcpy(h->v, v, sizeof(double) * h->N);
  h->kin = (double *)malloc(sizeof(double) * h->N);
  kinetic_table(lattice, n, h->kin);
  h->off = (long *)malloc(sizeof(long) * (natoms + 1));
  h->off[0] = 0;
  for (int at = 0; at < natoms; ++at) h->off[at + 1] = h->off[at] + npts[at];
  const long tot = h->off[natoms];
  h->pts = (int *)malloc(sizeof(int) * (tot ? tot : 1));
  memcpy(h->pts, pts, sizeof(int) * tot);
  h->beta = (double *)malloc(sizeof(double) * (nlm * tot > 0 ? nlm * tot : 1));
  memcpy(h->beta, beta, sizeof(double) * nlm * tot);
  h->coeff = (double *)malloc(sizeof(double) * (size_t)natoms * nlm * nlm + 1);
  memcpy(h->coeff, coeff, sizeof(double) * (size_t)natoms * nlm * nlm);
  h->work = (cplx *)malloc(sizeof(cplx) * (size_t)h->N * nst);
  chunks_init(&h->ch, n, nst, nthreads, h->work);
  return h;
}

void rc_ham_destroy(rc_ham *h) {
  if (!h) return;
  chunks_free(&h->ch);
  free(h->v);
  free(h->kin);
  free(h->off);
  free(h->pts);
  free(h->beta);
  free(h->coeff);
  free(h->work);
  free(h);
}

/* KB term of one thread's states: gather, P = dV beta S (DGEMM on interleaved
 * re/im), P <- D P, S' = beta^T P (DGEMM), scatter-add (projector_all.hpp:113-288) */
static void kb_chunk(const rc_ham *h, int t, const cplx *psi_rs, cplx *hpsi_rs, double *sbuf, double *pbuf) {
  const int nst = h->nst, s0 = h->ch.s0[t], ns = h->ch.ns[t], nlm = h->nlm;
  const int m2 = 2 * ns;
  for (int at = 0; at < h->natoms; ++at) {
    const long p0 = h->off[at];
    const int np = (int)(h->off[at + 1] - p0);
    if (np == 0) continue;
    const double *beta = h->beta + (size_t)nlm * p0;
    for (int p = 0; p < np; ++p) {
      const cplx *src = psi_rs + (size_t)h->pts[p0 + p] * nst + s0;
      memcpy(sbuf + (size_t)p * m2, src, sizeof(cplx) * ns);
    }
    if (g_backend == 1) {
      const double zero = 0.0;
      x_dgemm("N", "N", &m2, &nlm, &np, &h->dv, sbuf, &m2, beta, &np, &zero, pbuf, &m2);
    } else {
      for (int l = 0; l < nlm; ++l)
        for (int c = 0; c < m2; ++c) {
          double acc = 0.0;
          for (int p = 0; p < np; ++p) acc += beta[(size_t)l * np + p] * sbuf[(size_t)p * m2 + c];
          pbuf[(size_t)l * m2 + c] = h->dv * acc;
        }
    }
    /* P <- D P (multiply_by_coefficients) */
    const double *D = h->coeff + (size_t)at * nlm * nlm;
    double *q = pbuf + (size_t)nlm * m2;
    for (int l = 0; l < nlm; ++l)
      for (int c = 0; c < m2; ++c) {
        double acc = 0.0;
        for (int k = 0; k < nlm; ++k) acc += D[l * nlm + k] * pbuf[(size_t)k * m2 + c];
        q[(size_t)l * m2 + c] = acc;
      }
    if (g_backend == 1) {
      const double one = 1.0, zero = 0.0;
      x_dgemm("N", "T", &m2, &np, &nlm, &one, q, &m2, beta, &np, &zero, sbuf, &m2);
    } else {
      for (int p = 0; p < np; ++p)
        for (int c = 0; c < m2; ++c) {
          double acc = 0.0;
          for (int l = 0; l < nlm; ++l) acc += beta[(size_t)l * np + p] * q[(size_t)l * m2 + c];
          sbuf[(size_t)p * m2 + c] = acc;
        }
    }
    for (int p = 0; p < np; ++p) {
      cplx *dst = hpsi_rs + (size_t)h->pts[p0 + p] * nst + s0;
      const cplx *sv = (const cplx *)(sbuf + (size_t)p * m2);
      for (int s = 0; s < ns; ++s) dst[s] += sv[s];
    }
  }
}

/* hpsi_fs = H psi_fs on the full cube, both [N][nst] (Gamma point) */
void rc_ham_apply(rc_ham *h, const cplx *psi_fs, cplx *hpsi_fs) {
  const long N = h->N;
  const int nst = h->nst;
  long maxnp = 1;
  for (int at = 0; at < h->natoms; ++at)
    if (h->off[at + 1] - h->off[at] > maxnp) maxnp = h->off[at + 1] - h->off[at];
  const double inv = 1.0 / (double)N;
#pragma omp parallel num_threads(h->ch.nthreads)
  {
    const int t = omp_get_thread_num();
    /* to_real: copy, batched inverse FFT of the thread's states, x 1/N in a separate pass */
#pragma omp for schedule(static)
    for (long ip = 0; ip < N; ++ip) memcpy(h->work + ip * nst, psi_fs + ip * nst, sizeof(cplx) * nst);
    chunk_fft(&h->ch, t, h->n, N, h->work, +1);
#pragma omp barrier
#pragma omp for schedule(static)
    for (long ip = 0; ip < N; ++ip)
      for (int s = 0; s < nst; ++s) h->work[ip * nst + s] *= inv;
    /* scalar potential: hpsi_rs = V psi_rs (into the output array) */
#pragma omp for schedule(static)
    for (long ip = 0; ip < N; ++ip)
      for (int s = 0; s < nst; ++s) hpsi_fs[ip * nst + s] = h->v[ip] * h->work[ip * nst + s];
    /* KB projectors on the thread's states */
    double *sbuf = (double *)malloc(sizeof(double) * (size_t)maxnp * 2 * h->ch.ns[t]);
    double *pbuf = (double *)malloc(sizeof(double) * (size_t)2 * h->nlm * 2 * h->ch.ns[t]);
    kb_chunk(h, t, h->work, hpsi_fs, sbuf, pbuf);
    free(sbuf);
    free(pbuf);
    /* to_fourier of the thread's states, then laplacian_add */
    chunk_fft(&h->ch, t, h->n, N, hpsi_fs, -1);
#pragma omp barrier
#pragma omp for schedule(static)
    for (long ip = 0; ip < N; ++ip)
      for (int s = 0; s < nst; ++s) hpsi_fs[ip * nst + s] += h->kin[ip] * psi_fs[ip * nst + s];
  }
}

/* S = vol_el Phi^H Phi over the full cube (ZGEMM; Phi [N][nst]) and its Cholesky
 * factor (ZPOTRF); returns the LAPACK info */
int rc_overlap_cholesky(long N, int nst, double vol_el, const cplx *phi, cplx *S) {
  rc_init();
  if (g_backend == 1) {
    const int m = nst, k = (int)N;
    const cplx alpha = vol_el, beta = 0.0;
    x_zgemm("N", "C", &m, &m, &k, &alpha, phi, &m, phi, &m, &beta, S, &m);
    int info = 0;
    x_zpotrf("L", &m, S, &m, &info);
    return info;
  }
  for (int i = 0; i < nst; ++i)
    for (int j = 0; j < nst; ++j) {
      cplx acc = 0.0;
      for (long g = 0; g < N; ++g) acc += phi[g * nst + i] * conj(phi[g * nst + j]);
      S[(size_t)j * nst + i] = vol_el * acc;
    }
  /* column-major lower Cholesky */
  for (int j = 0; j < nst; ++j) {
    double d = creal(S[(size_t)j * nst + j]);
    for (int k = 0; k < j; ++k) d -= creal(S[(size_t)k * nst + j] * conj(S[(size_t)k * nst + j]));
    if (d <= 0.0) return j + 1;
    d = sqrt(d);
    S[(size_t)j * nst + j] = d;
    for (int i = j + 1; i < nst; ++i) {
      cplx acc = S[(size_t)j * nst + i];
      for (int k = 0; k < j; ++k) acc -= S[(size_t)k * nst + i] * conj(S[(size_t)k * nst + j]);
      S[(size_t)j * nst + i] = acc / d;
    }
  }
  return 0;
}

/* Exchange pair solves at the Gamma point (exchange_operator.hpp:124-161):
 * for each occupied j, rho_i = conj(hf_j) phi_i, Poisson (G = 0 dropped,
 * poisson.hpp:71-110), kphi_i += coef_j hf_j v_i.  phi, kphi [N][nst]; hf [N][nhf]. */
void rc_exchange_pairs(const double lattice[9], const int n[3], int nst, const cplx *phi, int nhf, const cplx *hf,
                       const double *coef, cplx *kphi, int nthreads) {
  rc_init();
  const long N = (long)n[0] * n[1] * n[2];
  cplx *rho = (cplx *)malloc(sizeof(cplx) * (size_t)N * nst);
  chunks_t ch;
  chunks_init(&ch, n, nst, nthreads, rho);
  /* kernel -4 pi/N * (-1/|G|^2) at every Fourier point, 0 at G = 0 */
  double *kern = (double *)malloc(sizeof(double) * N);
  kinetic_table(lattice, n, kern);
  for (long ip = 0; ip < N; ++ip) {
    const double g2 = 2.0 * kern[ip];
    kern[ip] = g2 < 1e-6 ? 0.0 : (4.0 * M_PI / (double)N) / g2;
  }
  for (int j = 0; j < nhf; ++j) {
#pragma omp parallel num_threads(ch.nthreads)
    {
      const int t = omp_get_thread_num();
#pragma omp for schedule(static)
      for (long ip = 0; ip < N; ++ip)
        for (int s = 0; s < nst; ++s) rho[ip * nst + s] = conj(hf[ip * nhf + j]) * phi[ip * nst + s];
      chunk_fft(&ch, t, n, N, rho, -1);
#pragma omp barrier
#pragma omp for schedule(static)
      for (long ip = 0; ip < N; ++ip)
        for (int s = 0; s < nst; ++s) rho[ip * nst + s] *= kern[ip];
      chunk_fft(&ch, t, n, N, rho, +1);
#pragma omp barrier
#pragma omp for schedule(static)
      for (long ip = 0; ip < N; ++ip)
        for (int s = 0; s < nst; ++s) kphi[ip * nst + s] += coef[j] * hf[ip * nhf + j] * rho[ip * nst + s];
    }
  }
  chunks_free(&ch);
  free(kern);
  free(rho);
}
