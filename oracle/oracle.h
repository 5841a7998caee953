/*
 * oracle.h -- FP64 CPU restatement of LLNL/inq's Kohn-Sham H|psi> hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle checks results: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (inq_amd/libinqx.so) never links, loads or calls it.
 *
 * Conventions (all restated from the reference, file:line cited at each
 * function in oracle.c):
 *   - grid index i in [0,n) maps to the symmetric integer sym(i) = i - n for
 *     i >= (n+1)/2                                   (src/basis/grid.hpp:78-84)
 *   - real-space orbitals: layout [i0][i1][i2][ist], states fastest
 *                                                   (src/basis/field_set.hpp:52,62,232-238)
 *   - forward FFT: unnormalised, exp(-i G.r); inverse FFT: exp(+i G.r) then
 *     x 1/N                                           (src/operations/transform.hpp:79,141,175-181)
 *   - G-sphere coefficients ("band-major" sphere layout [ist][ipw]) use the
 *     same unnormalised forward convention; their volume element is V/N^2
 *                                                   (src/basis/fourier_space.hpp:48-50)
 * Parity status of each function is stated in oracle.c's header.
 */
#ifndef INQX_ORACLE_H
#define INQX_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_cplx;

typedef struct orc_ham orc_ham;

/* ---- grid / cell ---- */
int  orc_version(void);
void orc_set_threads(int nthreads);
int  orc_good_size(int n);

/* batched 3-D FFT of nst states in [i0][i1][i2][ist] layout; sign -1 forward
 * (unnormalised), +1 backward (unnormalised, caller scales) */
void orc_fft3d(const int n[3], int nst, orc_cplx *data, int sign);

/* ---- Hamiltonian object ---- */
orc_ham *orc_ham_create(const double lattice[9], const int n[3]);
void     orc_ham_destroy(orc_ham *h);
double   orc_ham_volume(const orc_ham *h);
void     orc_ham_set_potential(orc_ham *h, const double *v);          /* [N] */
/* spherical grid of the reference (src/basis/spherical_grid.hpp:59-158) */
long     orc_sphere_build(const orc_ham *h, const double center[3], double radius,
                          int *coords, float *dist, float *relpos, long cap);
/* projector with a Gaussian radial exp(-r^2/(2 w^2)) x real Y_lm (s, p_y, p_z, p_x) */
int      orc_ham_add_projector_gauss(orc_ham *h, const double center[3], double radius,
                                     int nlm, double width, const double *kb_coeff);
/* projector with a caller-supplied matrix [nlm][npts] on the reference sphere */
int      orc_ham_add_projector_matrix(orc_ham *h, const double center[3], double radius,
                                      int nlm, const double *matrix, const double *kb_coeff);
int      orc_ham_num_projectors(const orc_ham *h);
long     orc_ham_projector_size(const orc_ham *h, int iproj);
void     orc_ham_projector_points(const orc_ham *h, int iproj, int *coords, float *dist,
                                  float *relpos, double *matrix);

/* ---- H|psi> ---- */
/* a1: real-space input, [N][nst] layout (ks_hamiltonian.hpp:137-156) */
void orc_ham_apply_rs(const orc_ham *h, int nst, const double kcov[3],
                      const orc_cplx *phi, orc_cplx *hphi);
/* a2: Fourier-space input on the full cube (ks_hamiltonian.hpp:160-180) */
void orc_ham_apply_fs(const orc_ham *h, int nst, const double kcov[3],
                      const orc_cplx *phi, orc_cplx *hphi);
/* sphere form: gather o a2 o scatter, band-major [nst][npw] */
void orc_ham_apply_sphere(const orc_ham *h, int nst, const double kcov[3], long npw,
                          const int *gidx, const orc_cplx *phi, orc_cplx *hphi);
/* projections P[a][l][ist] (already multiplied by D when apply_coeff != 0) */
void orc_ham_projections(const orc_ham *h, int nst, const double kcov[3],
                         const orc_cplx *phi_rs, int apply_coeff, orc_cplx *proj);
/* non-local energy sum_i occ_i sum_a P^H D P (projector_all.hpp:306-327) */
double orc_ham_nonlocal_energy(const orc_ham *h, int nst, const double kcov[3],
                               const orc_cplx *phi_rs, const double *occ);

/* ---- G sphere ---- */
long orc_gsphere(const orc_ham *h, const double kcov[3], double ecut, int *gidx, long cap);
void orc_gsphere_kinetic(const orc_ham *h, const double kcov[3], long npw, const int *gidx,
                         double *kin);
void orc_sphere_to_cube(const orc_ham *h, int nst, long npw, const int *gidx,
                        const orc_cplx *phi_sph, orc_cplx *cube);
void orc_cube_to_sphere(const orc_ham *h, int nst, long npw, const int *gidx,
                        const orc_cplx *cube, orc_cplx *phi_sph);
void orc_to_real(const orc_ham *h, int nst, const orc_cplx *fs, orc_cplx *rs, int normalize);
void orc_to_fourier(const orc_ham *h, int nst, const orc_cplx *rs, orc_cplx *fs);

/* ---- dense linear algebra on nst x nst row-major matrices ---- */
void orc_overlap(long np, int nst1, int nst2, double vol_el, int band_major,
                 const orc_cplx *phi1, const orc_cplx *phi2, orc_cplx *olap);
int  orc_cholesky(int n, orc_cplx *mat);
void orc_invert_lower(int n, orc_cplx *mat);
void orc_rotate_h(long np, int nst, int band_major, const orc_cplx *rot, orc_cplx *phi);
void orc_rotate5(long np, int nst, int band_major, const orc_cplx *rot, const orc_cplx *phi,
                 orc_cplx *rotphi, orc_cplx alpha, orc_cplx beta);
int  orc_orthogonalize(long np, int nst, double vol_el, int band_major, orc_cplx *phi);
int  orc_eigh(int n, orc_cplx *mat, double *evals);
int  orc_subspace_diag_sphere(const orc_ham *h, int nst, const double kcov[3], long npw,
                              const int *gidx, orc_cplx *phi, double *evals);

/* ---- propagation (sphere form, band-major) ---- */
void orc_exponential_in_place(const orc_ham *h, int nst, const double kcov[3], long npw,
                              const int *gidx, orc_cplx factor, int order, orc_cplx *phi);
void orc_exponential_2_for_1(const orc_ham *h, int nst, const double kcov[3], long npw,
                             const int *gidx, orc_cplx f1, orc_cplx f2, int order,
                             orc_cplx *phi, orc_cplx *expphi);
void orc_etrs_step(const orc_ham *h, int nst, const double kcov[3], long npw, const int *gidx,
                   double dt, orc_cplx *phi);
/* self-consistent ETRS step (etrs.hpp:22-79) with update_hamiltonian =
 * vion + V_Hartree[density] (self_consistency.hpp:100-125, no XC): one block,
 * one spin component.  rho_out [N] receives the last density; returns the
 * number of SCF passes, *delta the last integral |drho|. */
int orc_etrs_sc_step(orc_ham *h, int nst, const double kcov[3], long npw, const int *gidx, double dt,
                     const double *vion, const double *occ, int nscf, double threshold, orc_cplx *phi,
                     double *rho_out, double *delta);
void orc_energy_sphere(const orc_ham *h, int nst, const double kcov[3], long npw,
                       const int *gidx, const orc_cplx *phi, const double *occ, double out[4]);
void orc_density_dipole_sphere(const orc_ham *h, int nst, long npw, const int *gidx,
                               const orc_cplx *phi, const double *occ, double *density,
                               double dipole[3]);

/* real spherical harmonic Y_lm(v / |v|), l <= 2, m = -l..l (projector channels) */
double orc_ylm_real(int l, int m, const double v[3]);

/* ---- Poisson / exchange ---- */
void orc_poisson_in_place(const orc_ham *h, int nst, orc_cplx *rho, const double gshift_cart[3],
                          double zeroterm);
double orc_sing_auxiliary(const double lattice[9], const double q_cov[3]);
void   orc_sing_correction(const double lattice[9], int nk, const double *kcov,
                           const double *kweight, double *fk, double *fzero);
double orc_ace_build(const orc_ham *h, int nst, const orc_cplx *phi_rs, const double *occ,
                     double exx_coeff, orc_cplx *xi);
void   orc_exchange_direct(const orc_ham *h, int nst, const orc_cplx *phi_rs, int nhf,
                           const orc_cplx *hf_rs, const double *occ, double exx_coeff,
                           double scale, orc_cplx *exxphi);
void   orc_exchange_direct_shift(const orc_ham *h, int nst, const orc_cplx *phi_rs, int nhf,
                                 const orc_cplx *hf_rs, const double *occ, double exx_coeff,
                                 double scale, const double gshift_cart[3], double zeroterm,
                                 orc_cplx *exxphi);
void   orc_ace_apply(const orc_ham *h, int nxi, const orc_cplx *xi, int nst, const orc_cplx *phi_rs,
                     orc_cplx *exxphi);

/* ---- operations::randomize (pcg32), values [npoints][nst] ---- */
void orc_randomize(long npoints, int nst, int index, int is_complex, double *out);

#ifdef __cplusplus
}
#endif
#endif
