// internal.hpp -- host-side objects behind the inqx C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/inqx.h"
#include "fft_reg.hpp"

namespace inqx {

// ------------------------------------------------------------------ errors
struct Error {
  int code;
  std::string msg;
};

#define INQX_HIP(call)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw ::inqx::Error{INQX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};   \
  } while (0)

#define INQX_CHECK(cond, code, msg)                           \
  do {                                                        \
    if (!(cond)) throw ::inqx::Error{(code), std::string(msg)}; \
  } while (0)

// ------------------------------------------------------------------ device buffers
// one-time per-device settings (function attributes are set for the current
// device): test_and_set(dev) is false the first time for a device
struct DeviceOnce {
  std::atomic<unsigned long long> bits{0};
  bool test_and_set(int dev) {
    const unsigned long long m = 1ull << (dev & 63);
    return (bits.fetch_or(m) & m) != 0;
  }
};

struct DevMem {
  void *p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t nb) {
    if (nb <= bytes && p) return;
    release();
    if (nb == 0) nb = 16;
    INQX_HIP(hipMalloc(&p, nb));
    bytes = nb;
  }
  template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

template <typename T> void upload(DevMem &m, const std::vector<T> &v, hipStream_t s) {
  m.ensure(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (!v.empty()) INQX_HIP(hipMemcpyAsync(m.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
}

// ------------------------------------------------------------------ grid / cell metric
// systems::cell and its cell_metric (src/systems/cell.hpp:36-52,146-224)
struct Grid {
  int n[3];
  long N;
  double a[3][3];   // lattice vectors, rows, cartesian
  double b[3][3];   // reciprocal vectors 2 pi / V (a_j x a_k)
  double rl[3][3];  // b / 2 pi
  double vol;
  void init(const double *lat, const int *nn);
  void cov_to_cart(const double *v, double *c) const;
  void contra_to_cart(const double *v, double *c) const;
  void cart_to_contra(const double *c, double *v) const;
  static int sym(int i, int n) { return (i >= (n + 1) / 2) ? i - n : i; }
};

// per-kernel timing with HIP events on the context's stream (the role of the
// reference's CALI_CXX_MARK_* timers, src/utils/profiling.hpp)
struct KStat {
  long count = 0;
  double total_ms = 0.0;
};
struct Ctx {
  int device = 0;
  bool host_only = false;  // device < 0: set-up queries only, no HIP calls
  hipStream_t stream = nullptr;
  std::string err;
  int num_cu = 256;
  bool prof = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, KStat> stats;
  // workspaces of the linear-algebra entry points
  DevMem ws_partial, ws_ops, ws_tmp, ws_tmp2, ws_S, ws_X, ws_info, ws_eig, ws_lam;
  hipEvent_t get_event();
  void flush_prof();
};

// brackets one kernel launch with events when profiling is on
struct ProfScope {
  Ctx &c;
  std::string name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ProfScope(Ctx &cc, const char *n) : c(cc) {
    if (c.prof) {
      name = n;
      e0 = c.get_event();
      e1 = c.get_event();
      (void)hipEventRecord(e0, c.stream);
    }
  }
  ~ProfScope() {
    if (c.prof) {
      (void)hipEventRecord(e1, c.stream);
      c.pending.push_back({name, {e0, e1}});
    }
  }
};

// ------------------------------------------------------------------ projectors
struct Projector {
  double center[3];
  double radius;
  int nlm;
  std::vector<int> coords;    // [npts][3]
  std::vector<float> dist;    // [npts]
  std::vector<float> relpos;  // [npts][3] contravariant
  std::vector<double> matrix; // [nlm][npts]
  std::vector<double> coeff;  // [nlm][nlm]
  long npts() const { return (long)dist.size(); }
};

// line lengths with compiled register FFTs (radix-4/2/3/5/7 stages)
#define INQX_SIZES(X) X(12) X(16) X(18) X(20) X(24) X(28) X(30) X(32) X(36) X(40) X(48) X(54) X(60) X(64) X(72) X(80) X(96) X(100) X(108) X(120) X(128)

constexpr int kMaxNlm = 4;   // channels of one KB group (s + 3 p, padded)
// A projector of nlm channels (up to l = 2: s + 3 p + 5 d = 9) is split into
// ceil(nlm / 4) KB groups of four channels over the same sphere; the x passes
// treat every group as an atom, and kb_reduce couples a projector's groups
// through its full nlm x nlm coefficient matrix.
constexpr int kMaxNlmAtom = 9;
constexpr int kMaxGroups = (kMaxNlmAtom + kMaxNlm - 1) / kMaxNlm;  // groups of one projector
constexpr int kCoeffCols = kMaxGroups * kMaxNlm;                   // coefficient row length per group
constexpr int kExxFp64MaxGrid = 64;  // largest grid side of the complex128 exchange
constexpr int kTile = 32;    // x-pencils (lines) per tile: one wavefront, a lane pair per line

// tables of the real-space KB pass, shared by all blocks of a Hamiltonian
// A tile is 64 consecutive lines L = i1*n2 + i2.  Every (atom, sphere point)
// pair becomes one entry of the tile that holds the point; a slot is one
// (tile, atom) pair.  Entries carry everything a lane needs (compact point
// index, tile-local slot, projector values, relative position), so the
// x-pass kernels read them with independent, coalesced loads.
struct KBTables {
  bool built = false;
  int gen = 0;  // incremented by every build: the per-basis phase tables follow it
  int natoms = 0;  // KB groups (see kMaxGroups): the x passes' atoms
  int maxsph = 0, ntile = 0, nslot = 0, maxpts = 0, maxslot = 0;
  int maxchunk = 0;  // max over tiles of ceil(entries / 64): 64-entry rows of a tile's records
  long nent = 0;
  // per line: bitmask of sphere points along i0 (3 words) and the line's
  // first compact point index inside its tile
  DevMem line_mask;   // [n12] uint4 m0, m1, m2, base (n0 <= 96) or 2 x uint4 m0-m3, base, 0, 0, 0
  DevMem tile_info;   // [ntile] int4: npts, nslot_t, slot_base, ent_base ; [ntile] ent count in a 2nd int4
  DevMem lane_info;   // [ntile][64] int: project-pass run length | (slot + 1) << 8 | (first lane of slot) << 16
  DevMem slot_atom;   // [nslot] atom of each global slot
  DevMem ent_cs;      // [nent] int2: compact point index, tile-local slot
  DevMem ent_b32;     // [nent] float4: projector values l = 0..3
  DevMem ent_b64;     // [nent] double4
  DevMem atom_slots;  // [natoms+1] offsets into atom_slot_list
  DevMem atom_slot_list;
  DevMem coeff;       // [natoms][kMaxNlm][kCoeffCols] double: group rows of the projector's D
  DevMem group_sib;   // [natoms] int2: first group of the group's projector, number of groups
  std::vector<float> pos_host;  // [nent][4] the entry positions, for the per-basis phase tables
};

// ------------------------------------------------------------------ device block descriptor
template <typename T> struct SphereBlock {
  using C = cplx<T>;
  const C *psi;        // [nst][npw]
  C *hpsi;             // [nst][npw]
  C *t1;               // [nslot][nst][n1*n2] plane intermediates (in place for the x pass)
  C *partial;          // [kb nslot][kMaxNlm][nst]
  C *dp;               // [natoms][kMaxNlm][nst]
  const int *slot_seg; // [nslot] 2 x int4: sphere segment start, length, row list start, row count;
                       // row window start (symmetric i1), window length, 0, 0
  const int *rows;     // rows (i1) holding sphere points, per slot
  const int *pw_lds;   // [npw] LDS offset of each sphere point in its plane (plane-pass layout)
  const T *kin;        // [npw] 1/2 |G+k|^2
  const T *vloc;       // [N] V/N of the block's spin
  C *acc1, *acc2;      // Taylor accumulators acc += c (H psi), or null
  int acc_wide;        // the accumulators hold complex128 values (T = float)
  C *proj;             // KB projections P [natoms][kMaxNlm][nst], or null
  double *dip;         // dipole partials [ntile][nst][4], or null
  C *dps;              // D P per (KB slot, state): [kb nslot][nst][kMaxNlm]
  int nst, npw;
  int nslot, m_lo;     // slot j holds plane index m = m_lo + j (symmetric range)
  int gamma;           // k+A == 0: no KB phases
  int write_h;         // store H psi into hpsi
  double kcov[3];      // k + A, covariant
  // sphere <-> real-space transforms (kin == nullptr in plane_forward: plain masked FFT)
  const C *kbph;       // [nent] exp(+i (k+A).r) of every KB entry, or null at k + A = 0
  C *rs;               // real-space states, [nst][N] (rs_layout 0) or [N][nst] (1)
  int rs_layout;
  int accumulate;      // plane_forward adds into hpsi instead of storing
};

// optional work fused into one H application
struct ApplyOpts {
  void *const *acc1 = nullptr;  // per block, acc1 += c1 (H psi)
  void *const *acc2 = nullptr;  // per block, acc2 += c2 (H psi)
  double c1r = 0.0, c1i = 0.0, c2r = 0.0, c2i = 0.0;
  bool acc_wide = false;        // acc1/acc2 are complex128 arrays for FP32 applications
  bool write_h = true;
  bool observables = false;     // projections and dipole partials into the basis workspaces
  // mode 0: H application; 1: sphere -> real space (psi -> rs, x 1/N);
  // 2: real space -> sphere (rs -> hpsi, masked forward FFT)
  int mode = 0;
  void *const *rs = nullptr;    // per block real-space arrays for modes 1 and 2
  int rs_layout = 0;            // 0: [nst][N], 1: [N][nst]
  bool accumulate = false;      // mode 2: hpsi += result
};

// one G-sphere block
struct Basis {
  struct inqx_ham_s *ham = nullptr;
  double kcov[3];
  double ecut;
  int spin;
  long npw = 0;
  std::vector<int> gidx;      // flattened FFT index
  // plane slots: slot j <-> symmetric plane index m_lo + j, contiguous range
  int m_lo = 0, nslot = 0;
  std::vector<int> slot_seg;  // [nslot][8] segment start, length, row start, row count, row window start
                              // (symmetric i1 of the plane's first row), window length, 0, 0
  int nwin_max = 0;           // longest row window of a plane
  std::vector<int> rows, pw_yz;
  DevMem d_slot_seg, d_rows, d_pw_lw, d_kin64, d_kin32;
  DevMem d_g2h;  // (1/2)|G|^2 per sphere point, FP64, no k shift (preconditioner)
  DevMem d_gidx; // flattened FFT index per sphere point (layout conversions), on first use
  double kin_A[3] = {1e300, 0, 0};  // A for which d_kin* were built
  // KB phases exp(+i (k+A).r_e) per KB entry e (FP64 and FP32 copies), built for
  // the vector potential ph_A and the KB table generation ph_gen
  DevMem d_ph64, d_ph32;
  double ph_A[3] = {1e300, 0, 0};
  int ph_gen = -1;
  // workspaces per precision
  DevMem t1[2], partial[2], dp[2], dps[2], proj[2], dip, dots, hn[2][2], save[2];
  int ws_nst[2] = {0, 0};
  int ws_kb_nslot[2] = {-1, -1}, ws_natoms[2] = {-1, -1};
  int pinned = 0;  // > 0 while a captured graph uses the workspaces
};

}  // namespace inqx

// C entry points: every inqx::Error becomes its status code and the context's
// last-error message; nothing propagates across the boundary
#define API_BEGIN try {
#define API_END(ctxp)                                    \
  }                                                      \
  catch (const inqx::Error &e) {                         \
    if (ctxp) (ctxp)->c.err = e.msg;                     \
    return e.code;                                       \
  }                                                      \
  catch (const std::exception &e) {                      \
    if (ctxp) (ctxp)->c.err = e.what();                  \
    return INQX_ERR_ARG;                                 \
  }                                                      \
  return INQX_OK;

struct inqx_ctx_s {
  inqx::Ctx c;
};

struct inqx_ham_s {
  inqx_ctx ctx;
  inqx::Grid g;
  int ncomp = 1;
  bool has_v = false;
  std::vector<double> v;  // [N][ncomp]
  inqx::DevMem v64, v32;  // [ncomp][N] of V/N
  double A[3] = {0, 0, 0};
  std::vector<inqx::Projector> projs;
  inqx::KBTables kb;
  std::vector<inqx_basis> bases;
  int pinned = 0;  // > 0 while a captured graph uses the projector tables
  // launch shape of the x-pass kernels per precision: states (waves) per work-group
  int xwaves[2] = {4, 4};  // states per x-pass work-group (complex128, complex64)
  // plane-pass planes (states) per work-group per precision; 0: the largest the grid allows
  int plane_batch[2] = {0, 0};
  // exact-exchange workspaces: real-space states, hf states, accumulator, pair planes
  inqx::DevMem exx_phi, exx_hf, exx_w, exx_t, exx_jlist, exx_coef;
  inqx::DevMem exx_wb, exx_occ;  // symmetric pairs: the conjugate-pair accumulator and occupied flags
  // real-space apply and density workspaces (sphere arrays of a full-cube basis, occupations)
  inqx::DevMem rs_a, rs_b, dens_occ;
  inqx::DevMem lay_a, lay_b;  // band-major copies for inqx_ham_apply_layout
  // steepest-descent workspace: H phi, residual, H residual, dot partials and results
  inqx::DevMem sd_states, sd_scal;
  inqx::DevMem hartree_work, hartree_part;  // Hartree solve: complex128 [N] and the energy partials
  inqx::DevMem atom_work;                   // atomic-sphere fields: atom records and radial tables
  inqx::DevMem en_ring, en_counter, en_occ, en_desc;  // inqx_energy: one observables record
  std::vector<char> en_host;
  size_t exx_pair_bytes = 0;  // budget of the pair workspace (one j-chunk); 0: a quarter of free memory, <= 64 GiB
};

struct inqx_basis_s {
  inqx::Basis b;
};

namespace inqx {
// ham_kernels.hip
void build_kb_tables(inqx_ham_s *h);
void invalidate_launch_cache(inqx_ham_s *h);
void set_capture_keep(std::vector<std::shared_ptr<void>> *keep);
template <typename T>
void launch_apply(inqx_ham_s *h, int nblocks, Basis *const *bases, const int *nst, const void *const *psi,
                  void *const *hpsi, const ApplyOpts *opts = nullptr);
void ensure_basis_device(inqx_ham_s *h, Basis &b);
void ensure_workspace(inqx_ham_s *h, Basis &b, int nst, int prec);
bool grid_supported(const int *n);
int max_plane_batch(const int *n, int prec);
// plane passes of a grid with n1 != n2 (plane_rect.hip): rows and columns in
// separate kernels, each built for one line length, the planes passing through t1
bool rect_plane_supported(int n1, int n2);
int rect_plane_batch(int prec, int n1, int n2);
template <typename T>
void launch_rect_inverse(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb,
                         int pstr, int n1, int n2);
template <typename T>
void launch_rect_forward(Ctx &cx, const SphereBlock<T> *blk, const int *pref, int nblocks, int total, int pb,
                         int pstr, int n1, int n2, const double *cf);
// linalg.hip
void launch_overlap(Ctx &cx, int prec, int n1, int n2, long K, const void *A, const void *B, double vol, double *S,
                    DevMem &partial);
// Out[nr][K] = M C (+ beta Out), C [nc][K]; M from R by mode (0: conj(R), 2: R, both square;
// 1: alpha R^T with R [nc][nr])
void launch_rotate(Ctx &cx, int prec, int nr, int nc, long K, const double *R, int mode, double ar, double ai,
                   const void *Cm, void *Out, double br, double bi, DevMem &ops);
void launch_cholesky(Ctx &cx, int n, double *S, int *info);
void launch_trinv(Ctx &cx, int n, const double *L, double *X);
void launch_stream_copy(Ctx &cx, const void *src, void *dst, size_t bytes);
void launch_fill_zero(Ctx &cx, void *p, size_t bytes);
void launch_randomize(Ctx &cx, int type, long npoints, int nst, int index, int layout, void *out);
void launch_overlap_diagonal(Ctx &cx, int prec, int n, long K, double vol, const void *A, const void *B,
                             int normalized, double *out);
void launch_density_add(Ctx &cx, int prec, const void *phi, const double *occ, int nst, long N, int ncomp,
                        int ispin, double *den);
// eigensolver.hip: per-state passes of steepest descent and the TPA
// preconditioner over band-major sphere states [nst][K].  Dots: out[s][k] =
// vol sum_G w conj(a_k) b_k as (re, im), n <= 4 pairs, w applied to the pairs
// in wmask; part holds sd_dots_partial_bytes(nst, K).
size_t sd_dots_partial_bytes(int nst, long K);
void launch_sd_dots(Ctx &cx, int prec, int nst, long K, int n, const void *const *a, const void *const *b,
                    const double *w, int wmask, double vol, double *part, double *out);
void launch_sd_residual(Ctx &cx, int prec, int nst, long K, const void *phi, const void *hphi, void *res,
                        const double *en);
void launch_tpa_scale(Ctx &cx, int prec, int nst, long K, void *res, const double *g2h, const double *kn);
void launch_sd_lambda(Ctx &cx, int nst, const double *m03, const double *m45, double *lam);
void launch_sd_update(Ctx &cx, int prec, int nst, long K, void *phi, void *hphi, const void *res, const void *hres,
                      const double *lam);
// exx_kernels.hip: batched pair-density Poisson solves on real-space arrays in
// state-major layout [n][N].  With hf: for every chunk entry jj (hf index
// j = jlist[jj]) and state i, rho = conj(hf_j) phi_i -> Poisson -> w_i +=
// coef[j] hf_j v.  Without hf (nj = 1): w_i = Poisson(phi_i).
struct PairArgs {
  const void *phi = nullptr;   // [nst][N]
  const void *hf = nullptr;    // [nhf][N] or null
  void *t = nullptr;           // [nj][nst][N] pair workspace (may be phi when hf is null)
  void *w = nullptr;           // [nst][N]
  const int *jlist = nullptr;  // device [nj]: hf index of each chunk entry
  const double *coef = nullptr;  // device [nhf]: coefficient per hf index
  int nst = 0, nj = 1;
  double gshift[3] = {0, 0, 0};  // cartesian, added to G
  double zeroterm = 0.0;         // kernel value at |G + q|^2 < 1e-6 (times -4 pi)
  // symmetric pairs (hf = phi, gshift = 0): device [nst] occupied flags; pairs
  // (j, i) with both occupied and i > j are skipped and their terms of w_i land
  // in wb [nst][N] (every occupied i whose index is in this chunk's j list)
  const int *sym_occ = nullptr;
  void *wb = nullptr;
};
void launch_pairs(inqx_ham_s *h, int prec, const PairArgs &a);
// w += b, bytes a multiple of 16 (complex values of the precision)
void launch_add_inplace(Ctx &cx, int prec, void *w, const void *b, size_t bytes);
bool pair_grid_supported(const int *n);
// Hartree potential of a real density [N] (FP64): work holds N complex128,
// part hartree_chunks(N) doubles of sum rho vh partials (fixed order)
int hartree_chunks(long N);
// atomic-sphere fields (potential.hip): per atom record = contravariant position (3),
// radius, radial table index
constexpr int kAtomRec = 5;
void launch_atomic_field_add(inqx_ham_s *h, int natoms, const double *atoms_dev, int nr, double dr,
                             const double *tables_dev, double *field);
void launch_hartree(inqx_ham_s *h, const double *rho, double *vh, void *work, double *part);
// eigh.hip: FP64 Hermitian eigensolver on the device.  A [n][n] (destroyed)
// -> lam [n] ascending, U [n][n] rows = eigenvectors; ws of eigh_workspace_bytes(n)
constexpr int kMaxEigN = 1024;
size_t eigh_workspace_bytes(int n);
void launch_eigh(Ctx &cx, int n, void *A, double *lam, void *U, void *ws);
void launch_conj(Ctx &cx, void *a, long n);
// layout.hip: band[s][i] <-> point-major rows (rows[i], or i when rows is null) of [row][nst]
void launch_point_to_band(Ctx &cx, int prec, const void *src, const int *rows, long npw, int nst, void *band);
void launch_band_to_point(Ctx &cx, int prec, const void *band, const int *rows, long npw, int nst, void *dst);
// propagate.cpp: workspaces of inqx_ks_potential, dropped with the Hamiltonian
void release_ks_work(inqx_ham_s *h);
// allocation-free Hartree (inqx_hartree_noalloc): its scratch is a file-scope
// device array of kNoAllocScratchBytes; noalloc_scratch_init looks up its
// address once per device (at context creation, with the device current).
// The launcher returns the device address of the energy sum (sum rho vh).
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr long kNoAllocMaxPoints = 1048320;  // largest grid that fits (2^20 - 256 points)
void noalloc_scratch_init(int device);
size_t hartree_noalloc_bytes(long N);
double *launch_hartree_noalloc(inqx_ham_s *h, const double *rho, double *vh);
}  // namespace inqx
