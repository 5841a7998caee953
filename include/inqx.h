/*
 * inqx.h -- C ABI of the MI355X-native Kohn-Sham Hamiltonian engine.
 *
 * Drop-in boundary for LLNL/inq's hot path (paths relative to the reference
 * tree).  The reference exposes this path as header-only C++ with duck-typed
 * operators; each entry point below names the member or free function it
 * replaces.  Plain C types only: device pointers are `void *` to complex
 * arrays (complex128 for INQX_FP64, complex64 for INQX_FP32 precision), host
 * pointers are plain `double *`/`int *`.
 *
 * Error convention: every call returns an int status (INQX_OK == 0, negative
 * on error) and never aborts or throws across the boundary;
 * inqx_last_error() gives the message.  The reference aborts on GPU errors
 * (external_libs/gpurun/include/gpu/host.hpp:42-59) and on Cholesky
 * info < 0 (src/matrix/cholesky.hpp:77-82); here both become status codes.
 *
 * Streams: every compute call is enqueued on the context's stream with no
 * hidden host synchronisation (the reference synchronises after every
 * kernel, external_libs/gpurun/include/gpu/run.hpp:77,117,154,201), so
 * apply/exponential/etrs sequences can be captured in a hipGraph once their
 * workspaces exist (inqx_ham_prepare).  A context is not thread-safe.
 *
 * State layouts:
 *   sphere ("band-major"): [nst][npw] G-sphere coefficients of one block,
 *     in the reference's unnormalised forward-FFT convention
 *     (src/operations/transform.hpp:79); sphere order is the ascending
 *     flattened FFT index (i0*n1 + i1)*n2 + i2 returned by
 *     inqx_basis_gindex.
 *   real space / full Fourier cube: [i0][i1][i2][ist] with states fastest,
 *     the reference's field_set layout (src/basis/field_set.hpp:52,62,232-238).
 */
#ifndef INQX_H
#define INQX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct inqx_ctx_s *inqx_ctx;
typedef struct inqx_ham_s *inqx_ham;
typedef struct inqx_basis_s *inqx_basis;
typedef struct inqx_prop_s *inqx_prop;

enum {
  INQX_OK = 0,
  INQX_ERR_ARG = -1,         /* bad argument / shape */
  INQX_ERR_HIP = -2,         /* HIP runtime failure */
  INQX_ERR_UNSUPPORTED = -3, /* grid size or option not built */
  INQX_ERR_NOT_PD = -4,      /* Cholesky: matrix not positive definite */
  INQX_ERR_STATE = -5        /* call order (e.g. no potential set) */
};

enum { INQX_FP64 = 0, INQX_FP32 = 1 };
/* value type of inqx_randomize beyond the two state precisions: real double */
enum { INQX_REAL64 = 2 };

/* ------------------------------------------------------------ context */
/* Environment/device selection the reference leaves to input::environment
 * (src/input/environment.hpp:27-92) and scripts/gpubind. */
int inqx_ctx_create(int device, void *stream, inqx_ctx *out);
int inqx_ctx_destroy(inqx_ctx ctx);
int inqx_ctx_set_stream(inqx_ctx ctx, void *stream);
int inqx_sync(inqx_ctx ctx);
const char *inqx_last_error(inqx_ctx ctx);
int inqx_version(void);

/* per-kernel timing with HIP events on the context stream (replaces the
 * reference's caliper / kalimotxo scope timers, src/utils/profiling.hpp,
 * external_libs/kalimotxo/cali.h:74-168).  inqx_prof_get enumerates the
 * accumulated (kernel name, launch count, total milliseconds) after
 * synchronising the stream. */
int inqx_prof_enable(inqx_ctx ctx, int on);
int inqx_prof_reset(inqx_ctx ctx);
int inqx_prof_count(inqx_ctx ctx, int *nkernels);
int inqx_prof_get(inqx_ctx ctx, int index, char *name, int name_cap, long *launches, double *total_ms);

/* device memory helpers, for callers without their own allocator
 * (replaces the gpu::array caching pool, external_libs/gpurun/include/gpu/array.hpp:65-119) */
int inqx_malloc(inqx_ctx ctx, size_t bytes, void **ptr);
int inqx_free(inqx_ctx ctx, void *ptr);
int inqx_memcpy_h2d(inqx_ctx ctx, void *dst, const void *src, size_t bytes);
int inqx_memcpy_d2h(inqx_ctx ctx, void *dst, const void *src, size_t bytes);
int inqx_memcpy_d2d(inqx_ctx ctx, void *dst, const void *src, size_t bytes);

/* device-to-device copy with a plain streaming kernel (bytes a multiple of
 * 16): the HBM bandwidth reference the roofline figures are checked against
 * (the role of speed_tests/copies.cpp in the reference) */
int inqx_stream_copy(inqx_ctx ctx, const void *src, void *dst, size_t bytes);

/* ------------------------------------------------------------ Hamiltonian */
/* hamiltonian::ks_hamiltonian constructor (src/hamiltonian/ks_hamiltonian.hpp:92-102):
 * lattice rows are the cell vectors in bohr (src/systems/cell.hpp:36-52), n the
 * real-space grid (src/basis/real_space.hpp:39-50).  Each axis takes any of the
 * lengths 12, 16, 18, 20, 24, 28, 30, 32, 36, 40, 48, 54, 60, 64, 72, 80, 96, 100,
 * 108, 120 or 128, independently of the others (other lengths return
 * INQX_ERR_UNSUPPORTED from the calls that transform).  A complex128 apply
 * whose sphere's plane row window (rows x (n2 + 1) x 16 bytes) exceeds the
 * 160 KB of LDS also returns INQX_ERR_UNSUPPORTED (a whole-cube basis from 100
 * points; spheres of a plane-wave cutoff fit to 128^3). */
int inqx_ham_create(inqx_ctx ctx, const double lattice[9], const int n[3], inqx_ham *out);
int inqx_ham_destroy(inqx_ham ham);
/* ks_hamiltonian::scalar_potential() (ks_hamiltonian.hpp:241-243): host array
 * [N][ncomp] in Hartree; spin index `spin` of a basis selects the column
 * (src/hamiltonian/scalar_potential.hpp:31-34).  INQX_ERR_ARG when a basis of
 * this Hamiltonian has a spin index >= ncomp (an apply on such a basis is
 * rejected the same way); INQX_ERR_STATE while a propagator with a captured
 * graph uses the Hamiltonian. */
int inqx_ham_set_potential(inqx_ham ham, const double *v, int ncomp);
/* ks_hamiltonian::uniform_vector_potential() (ks_hamiltonian.hpp:257-262):
 * covariant A added to every block's k-point (velocity-gauge kick,
 * src/perturbations/kick.hpp:36-48).  INQX_ERR_STATE while a propagator with
 * a captured graph uses the Hamiltonian. */
int inqx_ham_set_vector_potential(inqx_ham ham, const double a_cov[3]);
/* one projector of projector_all (src/hamiltonian/projector.hpp:102-113,
 * projector_all.hpp:40-109).  The sphere is built as
 * basis::spherical_grid(grid, center, radius) (src/basis/spherical_grid.hpp:59-158).
 * nlm = 1..9 channels (l up to 2).  matrix: host [nlm][npts] projector
 * values on that sphere, or NULL for the Gaussian form exp(-r^2/(2 width^2))
 * Y_lm in the channel order of projector.hpp:63 (l = 0; l = 1 with m = -1, 0,
 * 1 = y, z, x; l = 2 with m = -2..2 = xy, yz, 3z^2 - r^2, xz, x^2 - y^2, the
 * standard real harmonics).  kb_coeff: host [nlm][nlm] (projector.hpp:85-93),
 * any coupling between channels.  Returns the projector index (>= 0) or a
 * negative status. */
int inqx_ham_add_projector(inqx_ham ham, const double center[3], double radius, int nlm,
                           const double *matrix, const double *kb_coeff, double width);
int inqx_ham_projector_size(inqx_ham ham, int iproj, long *npts);
/* sphere points of a projector: coords [npts][3] grid indices, dist [npts]
 * (float, spherical_grid.hpp:40-43), relpos [npts][3] contravariant (float) */
int inqx_ham_projector_points(inqx_ham ham, int iproj, int *coords, float *dist, float *relpos);

/* ------------------------------------------------------------ basis (one k-point/spin block) */
/* G sphere 1/2|G+k|^2 < ecut of orbital_set(k, spin) (src/states/orbital_set.hpp:18-192) */
int inqx_basis_create(inqx_ham ham, const double kcov[3], double ecut, int spin, inqx_basis *out);
int inqx_basis_destroy(inqx_basis basis);
int inqx_basis_size(inqx_basis basis, long *npw);
int inqx_basis_gindex(inqx_basis basis, int *gidx);
/* 1/2 |G+k+A|^2 per sphere point (host, FP64) */
int inqx_basis_kinetic(inqx_basis basis, double *kin);
/* allocate the workspaces apply needs for nst states at a precision, so that
 * later calls allocate nothing (required before hipGraph capture) */
int inqx_ham_prepare(inqx_ham ham, inqx_basis basis, int nst, int precision);

/* launch-shape autotuner: times the H application of the given blocks (one
 * launch sequence, as inqx_ham_apply_batch runs them) for every launch shape
 * (x-pass states per work-group 1, 2, 4 or 8, times plane-pass planes per
 * work-group from the largest the grid's LDS budget allows down to 1) and
 * keeps the fastest for this Hamiltonian and precision.  cache_path: a text
 * file of earlier results, one line per (device, precision, grid, KB size,
 * blocks) key; a hit applies the stored shape without timing, a miss times and
 * appends (NULL: no cache).  shape[2] (out, may be NULL): x-pass states,
 * plane batch.  Allocates scratch states; do not call during graph capture. */
int inqx_ham_autotune_batch(inqx_ham ham, int nblocks, const inqx_basis *bases, const int *nst, int precision,
                            const char *cache_path, int *shape, double *best_ms);
/* sets the plane-pass batch directly (planes per work-group; 0: the largest the
 * grid's LDS budget allows, larger values are capped to it) */
int inqx_ham_set_plane_batch(inqx_ham ham, int precision, int planes);
/* the single-block form (no cache); best_xwaves / best_ms may be NULL */
int inqx_ham_autotune(inqx_ham ham, inqx_basis basis, int nst, int precision, int *best_xwaves, double *best_ms);
/* sets the x-pass shape directly (1, 2, 4 or 8 states per work-group) */
int inqx_ham_set_launch_shape(inqx_ham ham, int precision, int xwaves);

/* ------------------------------------------------------------ H |psi> */
/* ks_hamiltonian::operator()(orbital_set<fourier_space>) (ks_hamiltonian.hpp:160-180)
 * restricted to the G sphere: hpsi = P H P psi, psi/hpsi band-major device arrays. */
int inqx_ham_apply(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *psi, void *hpsi);
/* the same for many blocks (k-points x spins) in one launch sequence */
int inqx_ham_apply_batch(inqx_ham ham, int nblocks, const inqx_basis *bases, const int *nst, int precision,
                         const void *const *psi, void *const *hpsi);

/* Fourier-space layouts of a block of sphere states */
enum { INQX_SPHERE_BAND = 0, INQX_SPHERE_POINT = 1, INQX_CUBE_POINT = 2 };
/* states of basis in a layout -> band-major sphere [nst][npw] (device arrays).
 * INQX_SPHERE_POINT: src [npw][nst] (the reference's point-major [ipw][ist]);
 * INQX_CUBE_POINT: src [N][nst], the full Fourier cube [Gx][Gy][Gz][ist] of an
 * orbital_set<fourier_space> (states/orbital_set.hpp:19-161; points off the
 * sphere are dropped); INQX_SPHERE_BAND: a copy. */
int inqx_states_to_sphere(inqx_ham ham, inqx_basis basis, int nst, int precision, int layout, const void *src,
                          void *sphere);
/* band-major sphere -> layout (INQX_CUBE_POINT: points off the sphere are set to zero) */
int inqx_states_from_sphere(inqx_ham ham, inqx_basis basis, int nst, int precision, int layout, const void *sphere,
                            void *dst);
/* H|psi> with psi and hpsi in one layout.  INQX_CUBE_POINT on a basis that
 * holds the whole cube (an ecut above every |G+k|^2/2) is the reference's
 * Fourier-space operator on its own layout, ks_hamiltonian::operator()
 * (orbital_set<fourier_space>) (ks_hamiltonian.hpp:160-180); on a sphere basis
 * it is mask H mask.  Converts through a band-major workspace of the
 * Hamiltonian; stream-ordered, no host synchronisation. */
int inqx_ham_apply_layout(inqx_ham ham, inqx_basis basis, int nst, int precision, int layout, const void *psi,
                          void *hpsi);

/* ------------------------------------------------------------ overlap / orthogonalisation */
/* All matrices are device arrays of complex128, row-major.  States are
 * band-major device arrays [n][npw] of the given precision.  vol_el is the
 * volume element of the basis (V/N^2 on the G sphere,
 * src/basis/fourier_space.hpp:48-50). */
/* operations::overlap (src/operations/overlap.hpp:26-95): S = vol_el phi1^H phi2 */
int inqx_overlap(inqx_ctx ctx, int precision, int n1, int n2, long npw, double vol_el, const void *phi1,
                 const void *phi2, void *S);
/* operations::overlap_diagonal (src/operations/overlap_diagonal.hpp:25-60):
 * out[i] = vol_el sum conj(phi1_i) phi2_i; with normalized != 0,
 * overlap_diagonal_normalized (:100-153): out[i] = <phi1_i|phi2_i> / <phi2_i|phi2_i>.
 * out is a device complex128 array [n]; sums in FP64. */
int inqx_overlap_diagonal(inqx_ctx ctx, int precision, int n, long npw, double vol_el, const void *phi1,
                          const void *phi2, int normalized, void *out);
/* matrix::cholesky (src/matrix/cholesky.hpp:28-100): in place S = L L^H, L lower,
 * upper part zeroed; INQX_ERR_NOT_PD when a leading minor is not positive
 * (the reference prints a warning or aborts, cholesky.hpp:77-82) */
int inqx_cholesky(inqx_ctx ctx, int n, void *S);
/* matrix::invert_triangular (src/matrix/invert_triangular.hpp:20-50): Linv = L^{-1} */
int inqx_invert_triangular(inqx_ctx ctx, int n, const void *L, void *Linv);
/* operations::rotate, 2-argument form (src/operations/rotate.hpp:26-80): phi <- phi rot^H */
int inqx_rotate(inqx_ctx ctx, int precision, int n, long npw, const void *rot, void *phi);
/* operations::rotate, 5-argument form (rotate.hpp:84-140): rotphi <- alpha phi rot + beta rotphi */
int inqx_rotate_add(inqx_ctx ctx, int precision, int n, long npw, const void *rot, const void *phi, void *rotphi,
                    double alpha_re, double alpha_im, double beta_re, double beta_im);
/* matrix::diagonalize (src/matrix/diagonalize.hpp:167-189): matrix is a device
 * complex128 [n][n] Hermitian matrix, replaced by its eigenvectors in the
 * reference's orientation (row k = the conjugate of the k-th eigenvector);
 * eigenvalues (host [n]) ascending.  FP64 on the device (Householder
 * tridiagonalisation, bisection, inverse iteration, back-transformation);
 * n <= 1024.  Synchronises the context's stream. */
int inqx_diagonalize(inqx_ctx ctx, int n, void *matrix, double *eigenvalues);
/* operations::orthogonalize (src/operations/orthogonalize.hpp:25-32): overlap ->
 * Cholesky -> rotate_trs, in place */
int inqx_orthogonalize(inqx_ctx ctx, int precision, int n, long npw, double vol_el, void *phi);
/* ground_state::subspace_diagonalization (src/ground_state/subspace_diagonalization.hpp:21-29):
 * H_sub = phi^H H phi, FP64 eigendecomposition on the device (eigenvalues
 * ascending, copied to the host array), phi <- phi U; nst <= 1024 */
int inqx_subspace_diag(inqx_ham ham, inqx_basis basis, int nst, int precision, void *phi, double *eigenvalues);
/* operations::preconditioner on Fourier-space states (src/operations/preconditioner.hpp:26-58),
 * Teter-Payne-Allan: <T>_i = Re(<(1/2)|G|^2 phi_i|phi_i> / <phi_i|phi_i>), then
 * phi_i(G) *= K((1/2)|G|^2 / <T>_i), K(x) = (27+18x+12x^2+8x^3)/(27+18x+12x^2+8x^3+16x^4);
 * |G| without the k shift (point_op.g2).  In place on the sphere states. */
int inqx_precondition(inqx_ham ham, inqx_basis basis, int nst, int precision, void *phi);
/* eigensolvers::steepest_descent (src/eigensolvers/steepest_descent.hpp:23-86):
 * num_steps (the reference uses 5) line minimisations of each state along its
 * (with precondition != 0 TPA-preconditioned) residual H phi - <H>phi, then
 * operations::orthogonalize (complex64 states: two Cholesky passes, so that
 * the FP32 rotation keeps them orthonormal); in place.  Overlaps and the step lengths are FP64
 * and stay on the device; the only host synchronisation is the Cholesky status
 * of the final orthogonalisation (INQX_ERR_NOT_PD if it fails). */
int inqx_steepest_descent(inqx_ham ham, inqx_basis basis, int nst, int precision, int num_steps, int precondition,
                          void *phi);

/* The Hartree step of hamiltonian::self_consistency::update_hamiltonian
 * (src/hamiltonian/self_consistency.hpp:115-120): vh = poisson::solve(density)
 * for a real field (src/solvers/poisson.hpp:217-223; periodic cell, G = 0 term
 * dropped) and energy = 1/2 sum density vh dV with dV = V/N
 * (operations::integral_product).  density, vh: device double arrays [N], point
 * order (i0, i1, i2); vh or energy may be NULL; energy is a host double (the
 * call then synchronises).  FP64 throughout. */
int inqx_hartree(inqx_ham ham, const double *density, double *vh, double *energy);
/* inqx_hartree without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as inqx_hartree (bit for bit); its work grid and
 * partial sums live in a 16 MiB device array of the library that each GPU
 * holds until the process ends, so grids are limited to 1 048 320 points
 * (INQX_ERR_ARG beyond).  Every call on a GPU uses that GPU's one array: calls
 * that overlap on one GPU (other threads or streams) race. */
int inqx_hartree_noalloc(inqx_ham ham, const double *density, double *vh, double *energy);

/* Atomic-sphere fields: the per-atom sphere sums of
 * atomic_potential::local_potential (short-range radial functions added on each
 * atom's sphere, src/hamiltonian/atomic_potential.hpp:116-165) and
 * atomic_potential::ionic_density (valence charge x long-range density,
 * atomic_potential.hpp:170-200):
 *   field[p] += sum_a sum_T f_{species[a]}(|r_p - R_a - T|)
 * over the periodic images T with |r_p - R_a - T| <= radius[a].  f_s is given at
 * r = j dr, j < nr (tables [nspecies][nr], host) and interpolated linearly;
 * radius[a] <= (nr - 1) dr.  positions: host cartesian [natoms][3]; radius,
 * species: host [natoms]; field: device double [N] in point order (i0, i1, i2).
 * Each grid point adds its atoms in a fixed order (no atomics), so results are
 * reproducible bit for bit.  The call synchronises the context's stream. */
int inqx_atomic_field_add(inqx_ham ham, int natoms, const double *positions, const double *radius,
                          const int *species, int nspecies, int nr, double dr, const double *tables, double *field);

/* ------------------------------------------------------------ restart files */
/* The reference's restart format (src/operations/io.hpp:41-211,
 * src/systems/electrons.hpp:323-381): raw native-endian files without headers,
 * named by utils::num_to_str ("%010ld", src/utils/num_str.hpp:20-24).
 * io::save / io::load of an orbital_set (io.hpp:168-211): file
 * dirname/<first_state+ist>_0000000000 per state, N complex128 real-space values
 * in point order (i0, i1, i2).  Save goes through to_real (x 1/N) from the
 * sphere; load through the masked forward transform onto the sphere.
 * complex64 states are widened to complex128 in the files.  Load returns
 * INQX_ERR_ARG when a file is missing or short (the reference returns false). */
int inqx_states_save(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *psi, const char *dirname,
                     int first_state);
int inqx_states_load(inqx_ham ham, inqx_basis basis, int nst, int precision, const char *dirname, int first_state,
                     void *psi);
/* io::save / io::load of a real field_set (io.hpp:124-162), e.g. the spin density:
 * dirname/<icomp> with npoints doubles each; v is a host array [npoints][ncomp] */
int inqx_field_save(inqx_ctx ctx, const char *dirname, long npoints, int ncomp, const double *v);
int inqx_field_load(inqx_ctx ctx, const char *dirname, long npoints, int ncomp, double *v);
/* io::save / io::load of a plain array (io.hpp:104-120), e.g. the occupations: dirname/array */
int inqx_array_save(inqx_ctx ctx, const char *dirname, long n, const double *v);
int inqx_array_load(inqx_ctx ctx, const char *dirname, long n, double *v);

/* ------------------------------------------------------------ real-time propagation */
/* operations::exponential_in_place (src/operations/exponential.hpp:21-34):
 * phi <- sum_{n<=order} (f H)^n / n! phi, f = f_re + i f_im, in place */
int inqx_exponential(inqx_ham ham, inqx_basis basis, int nst, int precision, double f_re, double f_im, int order,
                     void *phi);
/* operations::exponential_2_for_1 (exponential.hpp:49-69): expphi <- exp_T(f1 H) phi and
 * phi <- exp_T(f2 H) phi sharing the H^n phi */
int inqx_exponential_2_for_1(inqx_ham ham, inqx_basis basis, int nst, int precision, double f1_re, double f1_im,
                             double f2_re, double f2_im, int order, void *phi, void *expphi);
/* real_time::propagate's step loop (src/real_time/propagate.hpp:87-119) for a
 * Hamiltonian independent of the density: each step is the ETRS step
 * (src/real_time/etrs.hpp:22-79, Taylor order 4) followed by
 * energy::calculate (src/hamiltonian/energy.hpp:61-113) and the electronic
 * dipole (src/observables/dipole.hpp:24-59).  Every step appends one record
 * {kinetic, sum occ eigenvalue, non-local, sum of norms, dipole x, y, z,
 * step} (FP64) to a device ring of ring_capacity records.  phi are the
 * caller's band-major device states (updated in place), occ host arrays.
 * With use_graph the step is captured once in a hipGraph and replayed. */
int inqx_prop_create(inqx_ham ham, int nblocks, const inqx_basis *bases, const int *nst, int precision,
                     void *const *phi, const double *const *occ, double dt, int ring_capacity, inqx_prop *out);
int inqx_prop_step(inqx_prop prop, int nsteps, int use_graph);
/* observables of the current states (no propagation), appended to the ring */
int inqx_prop_energy(inqx_prop prop);
/* energy::calculate for a ground-state or any other caller, without a propagator
 * (src/hamiltonian/energy.hpp:61-113 with ks_hamiltonian::non_local_energy,
 * ks_hamiltonian.hpp:124-133, and observables::dipole, dipole.hpp:24-59): one H
 * application of the given band-major device states of every block, reduced in
 * FP64 into out[8] (host) = {kinetic, sum occ eigenvalue, non-local, sum of
 * norms, dipole x, y, z, 0}, the record layout of the propagator.  occ: host
 * arrays [nst[i]].  Synchronises the context's stream. */
int inqx_energy(inqx_ham ham, int nblocks, const inqx_basis *bases, const int *nst, int precision,
                void *const *phi, const double *const *occ, double *out);
/* number of records written so far and the most recent min(count, cap) records [k][8] */
int inqx_prop_records(inqx_prop prop, int *count, double *records, int cap);
/* Self-consistent steps: real_time::etrs for a density-dependent Hamiltonian
 * (src/real_time/etrs.hpp:22-79).  Each step takes the half and full steps
 * with H(t) (2-for-1), the spin density of the full-step prediction, updates
 * the potential to vion + V_Hartree[density] (inqx_ks_potential), restores the
 * half step, then runs up to max_passes (the reference: 5) second half steps
 * with H(t + dt), each followed by the density, the change integral |drho| dV
 * and the potential update, stopping once the change is below threshold (the
 * reference: 5e-5).  vion: device double [N] (ionic plus external potential),
 * kept by the propagator.  The caller first sets the potential of the initial
 * density (inqx_ks_potential).  These steps run eagerly (use_graph is
 * rejected with INQX_ERR_UNSUPPORTED): the pass count depends on the density
 * change, read after each pass.  max_passes = 0 returns to fixed-H steps.
 * Grids of built axis lengths only. */
int inqx_prop_set_scf(inqx_prop prop, const double *vion, int max_passes, double threshold);
/* State precision of an FP32 propagator: INQX_FP64 keeps a complex128 copy of
 * the states that the Taylor terms of the FP32 H applications accumulate into
 * (the rounding of the states to complex64 between steps is what makes FP32
 * propagation drift); the caller's complex64 arrays receive the state rounded
 * after every half step and feed the first H application of the next one.
 * INQX_FP32 (the default) accumulates in complex64.  Set before the first
 * graph step; fixed-H steps only. */
int inqx_prop_set_state_precision(inqx_prop prop, int precision);
/* SCF passes and the last integral |drho| dV of the most recent self-consistent step */
int inqx_prop_scf_info(inqx_prop prop, int *passes, double *delta);
/* hamiltonian::self_consistency::update_hamiltonian at the Hartree level
 * (src/hamiltonian/self_consistency.hpp:100-125, no XC): the Hamiltonian's
 * potential becomes vion + V_H[total density] in every spin component, on the
 * device, in place (vion: device double [N]; density: device double
 * [N][ncomp] spin density, ncomp that of the Hamiltonian's potential, which
 * must have been set once).  energies (host, optional) = {external = integral
 * density vion, Hartree = 1/2 integral density V_H}; the call synchronises. */
int inqx_ks_potential(inqx_ham ham, const double *vion, const double *density, double *energies);
/* a propagator that captured a graph pins its Hamiltonian and bases: they
 * cannot be destroyed (INQX_ERR_STATE) or have their workspaces grown until
 * the propagator is destroyed */
int inqx_prop_destroy(inqx_prop prop);

/* operations::randomize (src/operations/randomize.hpp:48-70): fills a field of
 * npoints x nst values with the reference's pcg32 sequence (seed npoints*nst +
 * index, U[-1,1), complex values real part first).  type: INQX_FP64
 * (complex128), INQX_FP32 (complex64, the double values rounded) or
 * INQX_REAL64; layout INQX_POINT_MAJOR ([npoints][nst], the reference's) or
 * INQX_STATE_MAJOR ([nst][npoints]); out is a device array. */
int inqx_randomize(inqx_ctx ctx, int type, long npoints, int nst, int index, int layout, void *out);

/* ------------------------------------------------------------ sphere <-> real space */
/* real-space layouts: state-major [nst][N] (the layout of the exchange
 * entry points), or point-major [N][nst], the reference's field_set layout
 * (src/basis/field_set.hpp:52,62); N = n0 n1 n2, point index (i0 n1 + i1) n2 + i2 */
enum { INQX_STATE_MAJOR = 0, INQX_POINT_MAJOR = 1 };
/* operations::transform::to_real (src/operations/transform.hpp:113-182) of
 * sphere states: psi_rs = (1/N) IFFT(scatter(psi)) */
int inqx_sphere_to_real(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *psi, int layout,
                        void *psi_rs);
/* operations::transform::to_fourier (transform.hpp:52-109) restricted to the
 * sphere: psi = mask(FFT(psi_rs)) (unnormalised); accumulate != 0 adds into psi */
int inqx_real_to_sphere(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *psi_rs, int layout,
                        int accumulate, void *psi);

/* ks_hamiltonian::operator()(orbital_set<real_space>) (ks_hamiltonian.hpp:137-156)
 * on arbitrary real-space states: hpsi_rs = to_real(H to_fourier(psi_rs)).
 * The basis must hold the whole cube (inqx_basis_create with an ecut above
 * every |G+k|^2/2, e.g. 1e30), so that no component is masked. */
int inqx_ham_apply_real(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *psi_rs, int layout,
                        void *hpsi_rs);
/* observables::density::calculate_add (src/observables/density.hpp:24-38),
 * non-spinor: density[p][ncomp] (component ispin) += sum_i occ_i |phi_i(p)|^2
 * with phi_i = to_real of the sphere states; density is a device double
 * array [N][ncomp], occ a host array [nst]; FP64 sums. */
int inqx_density_add(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *phi, const double *occ,
                     int ncomp, int ispin, void *density);

/* ------------------------------------------------------------ exact exchange */
/* solvers::poisson::in_place (src/solvers/poisson.hpp:71-110,199-213) on n
 * real-space densities, state-major [n][N], in place: rho_G <- -4 pi/N *
 * (-1/|G+q|^2) rho_G, zeroterm at |G+q|^2 < 1e-6, unnormalised inverse FFT.
 * gshift_cart = q (cartesian). */
int inqx_poisson(inqx_ham ham, int n, int precision, void *rho, const double gshift_cart[3], double zeroterm);
/* exchange_operator::direct (src/hamiltonian/exchange_operator.hpp:124-186)
 * in sphere form, the hf states at the basis' own k-point: exxphi +=
 * mask(FFT(sum_j f occ_j hf_j Poisson(conj(hf_j) phi_i))), f = -exx_coeff
 * scale / 2, phi_i = to_real of the sphere states (the periodic part on the
 * grid; j with |occ_j| < 1e-10 skipped).  phi [nst][npw], hf [nhf][npw] and
 * exxphi [nst][npw] on the basis' sphere; occ host [nhf].  complex128
 * (INQX_FP64) is offered up to a 64^3 grid; larger grids return
 * INQX_ERR_UNSUPPORTED for it (the same holds for every exchange entry point). */
int inqx_exchange_direct(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *phi, int nhf,
                         const void *hf, const double *occ, double exx_coeff, double scale, double zeroterm,
                         void *exxphi);
/* the same with the occupied states of several k-points, as exchange_operator
 * keeps them (exchange_operator.hpp:76-91): block b holds nhf[b] states hf[b]
 * on the sphere of hf_bases[b] (k-point q_b, a basis of the same
 * Hamiltonian), occupations occ[b] (host) and the singularity value of its
 * k-point zeroterms[b]; each pair solve is poisson::in_place(rho_ij,
 * -k + q_b, zeroterms[b]) (exchange_operator.hpp:151) with k the k-point of
 * basis. */
int inqx_exchange_direct_kpts(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *phi, int nblocks,
                              const inqx_basis *hf_bases, const int *nhf, const void *const *hf,
                              const double *const *occ, const double *zeroterms, double exx_coeff, double scale,
                              void *exxphi);
/* exchange_operator::update's ACE build (exchange_operator.hpp:93-117):
 * W = direct(phi, -1), M = vol_el W^H phi (FP64), E_x = -1/2 sum occ_i
 * Re M_ii, M = L L^H, xi = W L^{-H}.  xi [nst][npw] on the sphere (the
 * sphere projection of the reference's real-space xi); *energy = E_x. */
int inqx_ace_build(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *phi, const double *occ,
                   double exx_coeff, double zeroterm, void *xi, double *energy);
/* the ACE build of one (k-point, spin) block against the occupied states of
 * every k-point (exchange_operator.hpp:97-114: W = direct(phi, -1) over all
 * stored orbitals), hf blocks as in inqx_exchange_direct_kpts; occ: host
 * [nst], the block's own occupations for E_x */
int inqx_ace_build_kpts(inqx_ham ham, inqx_basis basis, int nst, int precision, const void *phi, const double *occ,
                        int nblocks, const inqx_basis *hf_bases, const int *nhf, const void *const *hf,
                        const double *const *hf_occ, const double *zeroterms, double exx_coeff, void *xi,
                        double *energy);
/* exchange_operator::ace (exchange_operator.hpp:217-231):
 * exxphi <- exxphi - xi (vol_el xi^H psi), xi [nxi][npw], psi/exxphi [nst][npw] */
int inqx_ace_apply(inqx_ctx ctx, int precision, int nxi, int nst, long npw, double vol_el, const void *xi,
                   const void *psi, void *exxphi);
/* hamiltonian::singularity_correction (src/hamiltonian/singularity_correction.hpp:
 * 19-103, Carrier et al. PRB 75, 205126): fk[ik] = 4 pi / V sum_{k' != k}
 * w_k' f(k - k') over the k-points (covariant kcov [nk][3], weights
 * kweight [nk]) and the integral fzero, with f the auxiliary function of Eq.
 * 16.  The value an exchange Poisson solve uses at G + q = 0 for the hf states
 * of k-point ik is zeroterm = -nk V (fk[ik] - fzero) (singularity_correction.hpp:
 * 100-102); with the Gamma point alone, zeroterm = V fzero.  The fzero grid
 * (7 x 61 x 121 x 121 points) is summed on the device. */
int inqx_singularity_correction(inqx_ctx ctx, const double lattice[9], int nk, const double *kcov,
                                const double *kweight, double *fk, double *fzero);
/* upper bound on the pair workspace of the exchange entry points (bytes;
 * 0, the default, takes a quarter of the free device memory, at most 64 GiB;
 * at least one j times all states is always used).  Larger chunks amortise
 * the accumulation into the exchange result over more j. */
int inqx_ham_set_exchange_workspace(inqx_ham ham, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
