// ham_kernels.hip -- H|psi> on G-sphere blocks, hand-written for gfx950.
//
// One H application on a block of nst states stored on their G sphere
// ([nst][npw], band-major) runs as five stream-ordered kernels with no host
// synchronisation:
//
//   plane_inverse  sphere -> zero-padded (i1,i2) planes of every plane index
//                  i0 of the sphere, inverse FFT along i2 (rows that hold
//                  sphere points only) then along i1; writes t1[slot][ist][i1][i2]
//   xpass_project  inverse FFT along i0 of each line (i1,i2) in registers,
//                  keeps the line's KB sphere points in LDS and reduces
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
//
// Block descriptors are copied into scalar registers at the top of each work
// item and every array access goes through an explicit global-address-space
// pointer, so the compiler emits scalar loads for uniform data and
// global_load/global_store (not flat) for the streams.
#include "plane_common.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace inqx {

void invalidate_launch_cache(inqx_ham_s *h);

// x-pass work-groups per group of states: the tiles padded to a multiple of 8
// (work-groups are dealt to the 8 XCDs round robin)
static inline int x_tiles_padded(int ntile) { return (ntile + 7) / 8 * 8; }

struct HamDev {
  int n0, n1, n2, n12;
  long N;
  int natoms, has_kb, ntile, maxpts, maxslot;
  int ntg;  // x-pass work-groups per group of states: the tiles padded to a multiple of 8
  const uint32_t *line_mask;
  const int *tile_info;
  const int *lane_info;  // [ntile][64] project-pass runs (see build_kb_tables)
  const int *slot_atom;
  const int *ent_cs;
  const void *ent_b;
  const int *atom_slots, *atom_slot_list;
  const double *coeff;    // [natoms][kMaxNlm][kCoeffCols]
  const int *group_sib;   // [natoms] first sibling group, sibling count
  double kb_scale;  // dV / N
  double lat[9];    // lattice vectors (rows), for positions r of grid points
};

// x-pass work-group size limit: up to 8 states for complex64 lines up to 96
// points; complex128 and longer lines need a wave per SIMD (VGPRs plus AGPRs)
// and take at most 4
template <typename T> constexpr int xpass_threads(int n0) { return sizeof(T) == 4 && n0 <= 96 ? 512 : 256; }

// -------------------------------------------------------------------------------------------
// plane passes (shared helpers in plane_common.hpp)
// -------------------------------------------------------------------------------------------
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256) k_plane_inverse(const SphereBlock<T> *blocks,
                                                                          const int *prefix, int nblocks,
                                                                          int total, int pb, int pstr) {
  using C = cplx<T>;
  using Cfg = PlaneCfg<T, N1, N2>;
  constexpr int S = Cfg::LANES, LPW = Cfg::LPW;  // lanes per line, lines per wave
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane / LPW, ll = lane % LPW;  // part of the line, line of the wave
  const int qo = line_part_out<S>(q);
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
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
    // the first row-pass line's row index, loaded now so that its latency
    // overlaps the scatter
    const int nr = np * nrow;
    const int lr0 = min(wave * LPW + ll, max(nr - 1, 0));
    const int pr0 = div_rcp(lr0, nrow, rrow);
    const int yr0 = nr > 0 ? rows[lr0 - pr0 * nrow] : 0;
    // the planes' sphere points, KG per thread in one round (a whole
    // plane at the configurations' sizes): the first round's loads go out
    // before the plane is cleared, so their latency overlaps the clearing
    const int tot = np * nseg;
    const auto rpsi = make_rsrc(B.psi + (size_t)s0 * B.npw + seg0,
                                (unsigned)(((size_t)(np - 1) * B.npw + nseg) * sizeof(C)));
    int ldx[KG];
    T vr[KG], vi[KG];
    if (tot > 0) scatter_round<T, KG>(rpsi, pwl, B.npw, nseg, rseg, tot, tid, nth, ldx, vr, vi);
    for (int i = tid; i < np * pstr; i += nth) lds[i] = C{0, 0};
    __syncthreads();
    for (int i0 = tid; i0 < tot; i0 += KG * nth) {
      if (i0 != tid) scatter_round<T, KG>(rpsi, pwl, B.npw, nseg, rseg, tot, i0, nth, ldx, vr, vi);
#pragma unroll
      for (int u = 0; u < KG; ++u) {
        const int i = i0 + u * nth;
        if (i < tot) lds[div_rcp(i, nseg, rseg) * pstr + ldx[u]] = C{vr[u], vi[u]};
      }
    }
    __syncthreads();
    // rows along i2 (sphere rows only), S lanes per row
    for (int g = wave; g * LPW < nr; g += Cfg::WAVES) {
      const int li = g * LPW + ll;
      const bool ok = li < nr;
      const int lc = ok ? li : nr - 1;
      const int p = div_rcp(lc, nrow, rrow);
      const int y = g == wave ? yr0 : rows[lc - p * nrow];
      plane_row_lanes<C, N2, +1, S>(lds + p * pstr + plane_row<N1>(y, wlo) * (N2 + 1), q, ok);
    }
    __syncthreads();
    // columns along i1 to the plane slots, S lanes per column; the np
    // planes of this work-group are slot `slot` of states s0 .. s0+np-1
    const size_t pstride = (size_t)B.nslot * (N1 * N2);  // between the states' planes
    const auto rt = make_rsrc(B.t1 + (size_t)s0 * pstride + (size_t)slot * (N1 * N2),
                              (unsigned)(((size_t)(np - 1) * pstride + N1 * N2) * sizeof(C)));
    for (int g = wave; g * LPW < np * N2; g += Cfg::WAVES) {
      const int li = g * LPW + ll;
      const bool ok = li < np * N2;
      const int lc = ok ? li : np * N2 - 1;
      const int p = lc / N2, z = lc - p * N2;
      // rows y = S n + q; in the window layout rows outside the window are zero
      // (row(y) = y + const within each half of the FFT order: one base per half)
      const int bA = p * pstr + z + plane_row<N1>(q, wlo) * (N2 + 1);  // rows y < N1/2
      const int bB = bA - N1 * (N2 + 1);                                 // rows y >= N1/2
      C v[N1 / S];
#pragma unroll
      for (int n = 0; n < N1 / S; ++n) {
        const int y = S * n + q;
        v[n] = row_in_plane<N1>(y, wlo, nwin) ? lds[(y < N1 / 2 ? bA : bB) + S * n * (N2 + 1)] : C{0, 0};
      }
      line_dit<N1, +1, C, S>(v, q);
      // v[k] is row k + qo N1/S; lanes past the last column store out of range
      // (dropped), no branch
      const unsigned vo = ok ? (unsigned)(((size_t)p * pstride + z + qo * (N1 / S) * N2) * sizeof(C)) : kOOB;
#pragma unroll
      for (int k = 0; k < N1 / S; ++k) bstore(rt, vo, (unsigned)(k * N2 * sizeof(C)), v[k]);
    }
    __syncthreads();
  }
}

// complex64 up to 72-point lines: at most 168 VGPRs, three waves per SIMD
// (longer lines would spill)
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && N1 <= 72 ? 3 : 1)))
k_plane_forward(const SphereBlock<T> *blocks, const int *prefix, int nblocks, int total, int pb, int pstr, T c1r,
                T c1i, T c2r, T c2i) {
  using C = cplx<T>;
  using Cfg = PlaneCfg<T, N1, N2>;
  constexpr int S = Cfg::LANES, LPW = Cfg::LPW;  // lanes per line, lines per wave
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *lds = reinterpret_cast<C *>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane / LPW, ll = lane % LPW;  // part of the line, line of the wave
  const int qo = line_part_out<S>(q);
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    SphereBlock<T> B;
    const PlaneItem it = plane_item(blocks, prefix, nblocks, wg, pb, B);
    const int slot = it.slot, s0 = it.s0, np = it.np;
    const V4<int> sg = ld4<int>(B.slot_seg, 2 * slot);
    const int seg0 = sg.x, nseg = sg.y, r0 = sg.z, nrow = sg.w;
    if (nseg == 0) return;  // uniform: nothing of this plane is on the sphere
    const V4<int> sw = ld4<int>(B.slot_seg, 2 * slot + 1);
    const int wlo = sw.x, nwin = sw.y;  // row window of the plane in LDS
    const float rseg = 1.0f / (float)nseg, rrow = 1.0f / (float)nrow;
    const GAS int *pwl = gp(B.pw_lds) + seg0;
    const GAS int *rows = gp(B.rows) + r0;
    const int tot = np * nseg, nr = np * nrow;
    // the first row-pass line's row index, loaded now
    const int lr0 = min(wave * LPW + ll, nr - 1);
    const int pr0 = div_rcp(lr0, nrow, rrow);
    const int yr0 = rows[lr0 - pr0 * nrow];
    // columns along i1 from the plane slots, then the sphere rows along i2,
    // S lanes per line; the np planes are slot `slot` of states s0..s0+np-1
    const size_t pstride = (size_t)B.nslot * (N1 * N2);
    const auto rt = make_rsrc(B.t1 + (size_t)s0 * pstride + (size_t)slot * (N1 * N2),
                              (unsigned)(((size_t)(np - 1) * pstride + N1 * N2) * sizeof(C)));
    for (int g = wave; g * LPW < np * N2; g += Cfg::WAVES) {
      const int li = g * LPW + ll;
      const bool ok = li < np * N2;
      const int lc = ok ? li : np * N2 - 1;
      const int p = lc / N2, z = lc - p * N2;
      const unsigned vo = (unsigned)(((size_t)p * pstride + z + q * N2) * sizeof(C));  // rows S n + q
      C v[N1 / S];
#pragma unroll
      for (int n = 0; n < N1 / S; ++n) bload(rt, vo, (unsigned)(S * n * N2 * sizeof(C)), v[n]);
      line_dit<N1, -1, C, S>(v, q);
      // v[k] is row y = k + qo N1/S (complex64: the lower lanes hold y < N1/2,
      // the upper lanes y >= N1/2); in the window layout rows outside the window
      // are not stored
      const int sb = p * pstr + z + plane_row<N1>(qo * (N1 / S), wlo) * (N2 + 1);
      if (ok) {
#pragma unroll
        for (int k = 0; k < N1 / S; ++k) {
          const int y = k + qo * (N1 / S);
          if (row_in_plane<N1>(y, wlo, nwin)) lds[sb + k * (N2 + 1)] = v[k];
        }
      }
    }
    __syncthreads();
    // the sphere values of this work-group's states s0 .. s0+np-1 through
    // buffer resources based at point seg0 of state s0 (past the end: an
    // offset beyond the buffer, loads give zero and stores are dropped)
    const GatherBufs R = gather_bufs(B, s0, seg0, np, nseg);
    // complex128: the LDS offsets of the gather's first round are loaded before
    // the row transforms, so that their latency overlaps them (420 -> 358 us at
    // Si64; complex64 has no registers to spare under its 168-VGPR limit and
    // loads them in the gather)
    constexpr int KG = gather_round<T>();
    constexpr bool kEarly = sizeof(T) == 8;
    int lx[KG];
    if constexpr (kEarly) gather_offsets<KG>(lx, pwl, tid, nth, tot, nseg, rseg, pstr);
    for (int g = wave; g * LPW < nr; g += Cfg::WAVES) {
      const int li = g * LPW + ll;
      const bool ok = li < nr;
      const int lc = ok ? li : nr - 1;
      const int p = div_rcp(lc, nrow, rrow);
      const int y = g == wave ? yr0 : rows[lc - p * nrow];
      plane_row_lanes<C, N2, -1, S>(lds + p * pstr + plane_row<N1>(y, wlo) * (N2 + 1), q, ok);
    }
    __syncthreads();
    for (int i0 = tid; i0 < tot; i0 += KG * nth) {
      if (!kEarly || i0 != tid) gather_offsets<KG>(lx, pwl, i0, nth, tot, nseg, rseg, pstr);
      gather_update<T, KG>(B, R, lds, lx, i0, nth, tot, nseg, rseg, c1r, c1i, c2r, c2i);
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------------------------------------
// x passes: a lane pair per line (i1,i2), 32 consecutive lines (a tile) per wave, one state per wave
// -------------------------------------------------------------------------------------------

// symmetric plane index of FFT index i (compile time)
template <int N> constexpr int sym_of(int i) { return i >= (N + 1) / 2 ? i - N : i; }


// A line (i1,i2) along i0 is shared by a lane pair (lane, lane + 32): the
// lower lane holds i0 in [0, N0/2), the upper lane [N0/2, N0) after an
// inverse transform, and the even / odd i0 before it.  t1 is state-major
// [state][slot][i1][i2]; slot j holds the symmetric plane index m_lo + j.
//
// loads the plane slots i0 = 2n + half of the pair's line (zero outside the
// slot range) and transforms along i0: afterwards v[k] = x[k + half N0/2]
template <typename T, int N0>
__device__ __forceinline__ void load_line_slots_split(const SphereBlock<T> &B, int n12, int s, int Lc, bool upper,
                                                      cplx<T> (&v)[N0 / 2]) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  const unsigned pl = (unsigned)(n12 * sizeof(C));
  const auto rs = make_rsrc(B.t1 + (size_t)s * B.nslot * n12, (unsigned)B.nslot * pl);
  const unsigned vl = (unsigned)(Lc * sizeof(C));
#pragma unroll
  for (int n = 0; n < H; ++n) {
    const int jl = sym_of<N0>(2 * n) - B.m_lo, jh = sym_of<N0>(2 * n + 1) - B.m_lo;  // uniform
    const bool vlo = (unsigned)jl < (unsigned)B.nslot, vhi = (unsigned)jh < (unsigned)B.nslot;
    C x{0, 0};
    if (vlo && vhi && jh == jl + 1) {
      bload(rs, vl + (upper ? pl : 0u), (unsigned)jl * pl, x);
    } else if (vlo || vhi) {  // no divergent branch: the invalid half reads out of range (zero)
      const int j = upper ? jh : jl;
      bload(rs, (unsigned)j < (unsigned)B.nslot ? vl + (unsigned)j * pl : kOOB, 0u, x);
    }
    v[n] = x;
  }
}
template <typename T, int N0>
__device__ __forceinline__ void load_line_inverse_split(const SphereBlock<T> &B, int n12, int s, int Lc, bool upper,
                                                        cplx<T> (&v)[N0 / 2]) {
  load_line_slots_split<T, N0>(B, n12, s, Lc, upper, v);
  fft_split<N0, +1, cplx<T>>::dit(v, upper);
}

// stores v[k] = X[2k + half] (after the forward split transform) into the
// plane slots of the line; lanes with ok == false store nothing
template <typename T, int N0>
__device__ __forceinline__ void store_line_slots_split(const SphereBlock<T> &B, int n12, int s, int L, bool upper,
                                                       bool ok, const cplx<T> (&v)[N0 / 2]) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  const unsigned pl = (unsigned)(n12 * sizeof(C));
  const auto rs = make_rsrc(B.t1 + (size_t)s * B.nslot * n12, (unsigned)B.nslot * pl);
  const unsigned vl = ok ? (unsigned)(L * sizeof(C)) : kOOB;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int jl = sym_of<N0>(2 * k) - B.m_lo, jh = sym_of<N0>(2 * k + 1) - B.m_lo;
    const bool vlo = (unsigned)jl < (unsigned)B.nslot, vhi = (unsigned)jh < (unsigned)B.nslot;
    if (vlo && vhi && jh == jl + 1) {
      bstore(rs, vl + (upper ? pl : 0u), (unsigned)jl * pl, v[k]);
    } else if (vlo || vhi) {
      const int j = upper ? jh : jl;
      bstore(rs, (unsigned)j < (unsigned)B.nslot ? vl + (unsigned)j * pl : kOOB, 0u, v[k]);
    }
  }
}

// a line's point mask (bit i0: the line holds a sphere point of the tile's KB
// records at i0) and the number of the tile's points on its earlier lines.
// The table holds one 16-byte record per line (mask words 0-2, count) for lines
// up to 96 points, two (mask words 0-3, count) for longer lines.
template <int N0> constexpr int mask_words() { return N0 > 96 ? 4 : 3; }
template <int N0> struct LineMask {
  uint32_t m[mask_words<N0>()];
  int base;
};
template <int N0> __device__ __forceinline__ LineMask<N0> load_line_mask(const uint32_t *tab, int L) {
  LineMask<N0> r;
  if constexpr (N0 > 96) {
    const V4<uint32_t> a = ld4<uint32_t>(tab, 2 * (size_t)L), b = ld4<uint32_t>(tab, 2 * (size_t)L + 1);
    r.m[0] = a.x;
    r.m[1] = a.y;
    r.m[2] = a.z;
    r.m[3] = a.w;
    r.base = (int)b.x;
  } else {
    const V4<uint32_t> a = ld4<uint32_t>(tab, L);
    r.m[0] = a.x;
    r.m[1] = a.y;
    r.m[2] = a.z;
    r.base = (int)a.w;
  }
  return r;
}

// the lane's half of a line's point mask as bits 0..N0/2-1 (bit k: i0 = k +
// half N0/2 holds a sphere point), and the number of points of the line
// before that half plus the tile's points on earlier lines
template <int N0>
__device__ __forceinline__ unsigned long long half_mask(const LineMask<N0> &lm, bool upper, int &cnt0) {
  constexpr int H = N0 / 2;
  static_assert(H <= 64, "lines up to 128 points");
  const unsigned long long lo = (unsigned long long)lm.m[0] | ((unsigned long long)lm.m[1] << 32);  // bits 0..63
  unsigned long long hi = lm.m[2];                                                                 // bits 64..
  if constexpr (N0 > 96) hi |= (unsigned long long)lm.m[3] << 32;
  const unsigned long long keep = H >= 64 ? ~0ull : (1ull << (H & 63)) - 1ull;
  const unsigned long long ml = lo & keep;
  const unsigned long long mu = H >= 64 ? hi : ((lo >> (H & 63)) | (hi << ((64 - H) & 63))) & keep;
  cnt0 = lm.base + (upper ? __popcll(ml) : 0);
  return upper ? mu : ml;
}

__device__ __forceinline__ void lds_add(float *p, float v) { atomicAdd(p, v); }
__device__ __forceinline__ void lds_add(double *p, double v) { atomicAdd(p, v); }

// one 16-byte record per lane from global memory straight into LDS (no
// registers): lane l writes dst + 16 l, dst uniform over the wave; the wait
// for it is the vmcnt wait in front of the next barrier
__device__ __forceinline__ void dma16(const void *src, void *dst) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}
// nbytes (a multiple of 16) of contiguous global memory into LDS, shared out
// over the ns waves of the work-group
__device__ __forceinline__ void dma_copy(const void *src, void *dst, int nbytes, int wave, int ns, int lane) {
  const int nch = nbytes >> 4;
  for (int i = wave; i * 64 < nch; i += ns) {
    const int c = i * 64 + lane;
    if (c < nch) dma16(static_cast<const char *>(src) + (size_t)c * 16, static_cast<char *>(dst) + i * 1024);
  }
}

// LDS layout of the x passes (byte offsets, computed on the host): the
// work-group's shared part (V tile of the apply pass, then the tile's KB
// records: point/slot ints, projector values, phases), then one region per
// wave (state): compact sphere points, then the slot values (projection
// accumulators of the project pass, D P of the apply pass)
struct XLds {
  int v, cs, bt, ph, li, wave0, wave_bytes, aux;
};

// Inverse FFT along i0 of the 32 lines of a tile, then the KB projections
// P[slot][l] = sum_entries beta_l exp(+ik.pos) psi.  A work-group takes ns
// states (one per wave) of one tile and loads the tile's entry records and
// lane table into LDS once, before the line loads.  Each lane walks a run of
// records of one slot from LDS and sums it in registers; the lanes' partial
// sums of a slot are then added in lane order (no atomics, the sums do not
// depend on timing).
template <typename T, int N0>
__global__ void __launch_bounds__(xpass_threads<T>(N0)) k_xpass_project(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                       int total, HamDev hd, XLds xl) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int ns = blockDim.x >> 6, ptid = lane;
  const bool upper = lane >= 32;
  const int2 *scs = reinterpret_cast<const int2 *>(smem_raw + xl.cs);
  const V4<T> *sbt = reinterpret_cast<const V4<T> *>(smem_raw + xl.bt);
  const C *sph = reinterpret_cast<const C *>(smem_raw + xl.ph);
  const int *sli = reinterpret_cast<const int *>(smem_raw + xl.li);
  C *pts = reinterpret_cast<C *>(smem_raw + xl.wave0 + wave * xl.wave_bytes);
  C *part = reinterpret_cast<C *>(smem_raw + xl.wave0 + wave * xl.wave_bytes + xl.aux);  // [64 lanes][kMaxNlm]
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> B = blocks[b];
    // work order: consecutive work-groups take consecutive tiles of the same
    // states; the tiles of one state group are padded to hd.ntg, a multiple of
    // 8, so a tile's work-groups always go to the same XCD
    const int local = wg - gp(prefix)[b];
    const int sgrp = local / hd.ntg;
    const int tile = local - sgrp * hd.ntg;
    if (tile >= hd.ntile) return;  // uniform: a padding work-group
    const int s = sgrp * ns + wave;
    const int L = tile * kTile + (lane & 31);
    const V4<int> ti = ld4<int>(hd.tile_info, 2 * tile);
    const V4<int> tj = ld4<int>(hd.tile_info, 2 * tile + 1);
    const int npts = ti.x, nsl = ti.y, slot_base = ti.z, ent_base = ti.w, nent = tj.x;
    if (npts == 0) return;  // uniform: no sphere point in the tile
    const int rows = tj.y;
    dma_copy(hd.ent_cs + 2 * (size_t)ent_base, smem_raw + xl.cs, 64 * rows * 8, wave, ns, lane);
    dma_copy(static_cast<const char *>(hd.ent_b) + (size_t)ent_base * 4 * sizeof(T), smem_raw + xl.bt,
             64 * rows * 4 * (int)sizeof(T), wave, ns, lane);
    if (!B.gamma) dma_copy(B.kbph + ent_base, smem_raw + xl.ph, 64 * rows * (int)sizeof(C), wave, ns, lane);
    dma_copy(hd.lane_info + (size_t)tile * 64, smem_raw + xl.li, 64 * 4, wave, ns, lane);
    const bool active = s < B.nst;
    const bool valid = L < hd.n12;
    C v[H];
    if (active) {
      // line loads and the line's point mask in flight together
      load_line_slots_split<T, N0>(B, hd.n12, s, valid ? L : hd.n12 - 1, upper, v);
      const LineMask<N0> lm = load_line_mask<N0>(hd.line_mask, valid ? L : hd.n12 - 1);
      fft_split<N0, +1, C>::dit(v, upper);
      // every value goes out: the line's sphere points to their compact
      // slots, the others (and lanes past the last line) to the trash slot
      // pts[npts + 1], so the stores need no branches
      int cnt;
      const unsigned long long mh = valid ? half_mask<N0>(lm, upper, cnt) : 0ull;
      if (!valid) cnt = 0;
      const int trash = npts + 1;
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const int bit = (int)((mh >> k) & 1ull);
        pts[bit ? cnt : trash] = v[k];
        cnt += bit;
      }
    }
    __syncthreads();  // entry records in LDS, sphere points staged
    if (active) {
      // the lane's run (records of one slot): record j at 64 j + lane
      const int cnt = sli[ptid] & 255;
      C a0{0, 0}, a1{0, 0}, a2{0, 0}, a3{0, 0};
      for (int j = 0; j < cnt; j += 2) {
        // two records: all their LDS reads before the accumulation
        int2 cs[2];
        V4<T> bt[2];
        C u[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) cs[q] = scs[64 * min(j + q, cnt - 1) + ptid];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int e = 64 * min(j + q, cnt - 1) + ptid;
          bt[q] = sbt[e];
          u[q] = pts[cs[q].x];
          if (!B.gamma) u[q] = cmul(u[q], sph[e]);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (j + q < cnt) {
            const V4<T> b = bt[q];
            const C x = u[q];
            a0.x += b.x * x.x; a0.y += b.x * x.y;
            a1.x += b.y * x.x; a1.y += b.y * x.y;
            a2.x += b.z * x.x; a2.y += b.z * x.y;
            a3.x += b.w * x.x; a3.y += b.w * x.y;
          }
        }
      }
      part[ptid * kMaxNlm + 0] = a0;
      part[ptid * kMaxNlm + 1] = a1;
      part[ptid * kMaxNlm + 2] = a2;
      part[ptid * kMaxNlm + 3] = a3;
    }
    __syncthreads();
    if (active) {
      // slot q's sum over its lanes [first(q), first(q + 1)), in lane order
      GAS C *pt = gw(B.partial);
      for (int i = ptid; i < nsl * kMaxNlm; i += 64) {
        const int q = i / kMaxNlm, l = i - q * kMaxNlm;
        const int f0 = sli[q] >> 16, f1 = sli[q + 1] >> 16;
        C sum = part[f0 * kMaxNlm + l];
        for (int k = f0 + 1; k < f1; ++k) sum = cadd(sum, part[k * kMaxNlm + l]);
        pt[((size_t)(slot_base + q) * kMaxNlm + l) * B.nst + s] = sum;
      }
    }
  }
}

// FP64 sum over a KB group's slots, x dV, x D.  A projector with more than
// four channels spans several groups: every group's thread sums the slots of
// all groups of its projector and multiplies by its rows of the projector's
// coefficient matrix (projector_all.hpp:136-189, multiply_by_coefficients
// over all lm of an atom).  Writes D P per (group, l, state) for the energy
// record and per (slot, state, l) -- every slot of the group -- so that the
// apply pass loads a tile's values as contiguous records.
template <typename T>
__global__ void __launch_bounds__(256) k_kb_reduce(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                   int total, HamDev hd) {
  using C = cplx<T>;
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> B = blocks[b];
    // four lanes per (group, state): lane quarter qr sums the slots qr, qr + 4,
    // ... of every sibling group; the quarters are combined by two butterfly
    // exchanges (every lane ends with the same sum, in a fixed order), and lane
    // qr then writes channel qr
    const int item = (wg - gp(prefix)[b]) * (blockDim.x >> 2) + (threadIdx.x >> 2);
    const int qr = threadIdx.x & 3;
    const bool live = item < hd.natoms * B.nst;  // uniform over the four lanes of an item
    const int ic = live ? item : 0;
    const int a = ic / B.nst, s = ic - a * B.nst;
    // the projections of every group of this group's projector (siblings),
    // each the FP64 sum over its slots
    const V2<int> sb = ld2<int>(hd.group_sib, a);
    const int g0 = sb.x, ng = sb.y;
    double pr[kMaxGroups][kMaxNlm], pi[kMaxGroups][kMaxNlm];
#pragma unroll
    for (int q = 0; q < kMaxGroups; ++q)
#pragma unroll
      for (int l = 0; l < kMaxNlm; ++l) pr[q][l] = pi[q][l] = 0.0;
    const GAS C *part = gp(B.partial);
#pragma unroll
    for (int gq = 0; gq < kMaxGroups; ++gq) {
      if (gq < ng) {
        const int q0 = gp(hd.atom_slots)[g0 + gq], q1 = gp(hd.atom_slots)[g0 + gq + 1];
        // two of this lane's slots per step, their loads in flight together
        for (int q = q0 + qr; q < q1; q += 8) {
          const int sl0 = gp(hd.atom_slot_list)[q];
          const int sl1 = gp(hd.atom_slot_list)[min(q + 4, q1 - 1)];
          C v0[kMaxNlm], v1[kMaxNlm];
#pragma unroll
          for (int l = 0; l < kMaxNlm; ++l) {
            v0[l] = part[((size_t)sl0 * kMaxNlm + l) * B.nst + s];
            v1[l] = part[((size_t)sl1 * kMaxNlm + l) * B.nst + s];
          }
#pragma unroll
          for (int l = 0; l < kMaxNlm; ++l) {
            pr[gq][l] += (double)v0[l].x;
            pi[gq][l] += (double)v0[l].y;
          }
          if (q + 4 < q1) {
#pragma unroll
            for (int l = 0; l < kMaxNlm; ++l) {
              pr[gq][l] += (double)v1[l].x;
              pi[gq][l] += (double)v1[l].y;
            }
          }
        }
      }
    }
    // (q0 + q1) + (q2 + q3) on every lane of the item
#pragma unroll
    for (int q = 0; q < kMaxGroups; ++q)
#pragma unroll
      for (int l = 0; l < kMaxNlm; ++l) {
        pr[q][l] += __shfl_xor(pr[q][l], 1, 64);
        pi[q][l] += __shfl_xor(pi[q][l], 1, 64);
        pr[q][l] += __shfl_xor(pr[q][l], 2, 64);
        pi[q][l] += __shfl_xor(pi[q][l], 2, 64);
      }
    if (!live) return;
    const int me = a - g0;  // this group among its siblings
    const int l = qr;       // the channel this lane writes
    if (B.proj) {
      double xr = 0.0, xi = 0.0;
#pragma unroll
      for (int q = 0; q < kMaxGroups; ++q)
#pragma unroll
        for (int m = 0; m < kMaxNlm; ++m)
          if (q == me && m == l) {
            xr = pr[q][m];
            xi = pi[q][m];
          }
      gw(B.proj)[((size_t)a * kMaxNlm + l) * B.nst + s] = C{(T)(xr * hd.kb_scale), (T)(xi * hd.kb_scale)};
    }
    const GAS double *D = gp(hd.coeff) + ((size_t)a * kMaxNlm + l) * kCoeffCols;
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxGroups; ++q)
#pragma unroll
      for (int m = 0; m < kMaxNlm; ++m) {
        dr += D[q * kMaxNlm + m] * pr[q][m];
        di += D[q * kMaxNlm + m] * pi[q][m];
      }
    const C r{(T)(dr * hd.kb_scale), (T)(di * hd.kb_scale)};
    gw(B.dp)[((size_t)a * kMaxNlm + l) * B.nst + s] = r;
    GAS C *dps = gw(B.dps);
    const int q0 = gp(hd.atom_slots)[a], q1 = gp(hd.atom_slots)[a + 1];
    for (int q = q0; q < q1; ++q) dps[((size_t)gp(hd.atom_slot_list)[q] * B.nst + s) * kMaxNlm + l] = r;
  }
}

// Inverse FFT along i0, times V/N, plus the KB term exp(-ik.r) beta D P at
// the sphere points, forward FFT along i0, stored back into the sphere's
// plane slots; optionally the dipole partial sums of the tile.  A work-group
// takes ns states (one per wave) of one tile.  Before the line loads, the
// work-group loads into LDS everything else it will read: the tile's V/N
// (shared), its KB entry records (shared) and each state's D P of the tile's
// slots, so no memory round trip sits between the two line transforms.
template <typename T, int N0>
__global__ void __launch_bounds__(xpass_threads<T>(N0)) k_xpass_apply(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                     int total, HamDev hd, XLds xl) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const int ns = blockDim.x >> 6, ptid = lane;
  const bool upper = lane >= 32;
  const T *sv = reinterpret_cast<const T *>(smem_raw + xl.v);  // [N0][32] V/N of the tile's lines
  const int2 *scs = reinterpret_cast<const int2 *>(smem_raw + xl.cs);
  const V4<T> *sbt = reinterpret_cast<const V4<T> *>(smem_raw + xl.bt);
  const C *sph = reinterpret_cast<const C *>(smem_raw + xl.ph);
  C *pts = reinterpret_cast<C *>(smem_raw + xl.wave0 + wave * xl.wave_bytes);
  C *dpl = reinterpret_cast<C *>(smem_raw + xl.wave0 + wave * xl.wave_bytes + xl.aux);
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> B = blocks[b];
    // work order as in k_xpass_project
    const int local = wg - gp(prefix)[b];
    const int sgrp = local / hd.ntg;
    const int tile = local - sgrp * hd.ntg;
    if (tile >= hd.ntile) return;  // uniform: a padding work-group
    const int s = sgrp * ns + wave;
    const int L = tile * kTile + (lane & 31);
    const bool valid = L < hd.n12;
    const int Lc = valid ? L : hd.n12 - 1;
    const bool act = s < B.nst;
    int npts = 0, nsl = 0, slot_base = 0, ent_base = 0, nent = 0;
    if (hd.has_kb) {
      const V4<int> ti = ld4<int>(hd.tile_info, 2 * tile);
      const V4<int> tj = ld4<int>(hd.tile_info, 2 * tile + 1);
      npts = ti.x;
      nsl = ti.y;
      slot_base = ti.z;
      ent_base = ti.w;
      nent = tj.x;
    }
    int rows = 0;
    if (hd.has_kb) rows = ld4<int>(hd.tile_info, 2 * tile + 1).y;
    const int ncap = 64 * rows;  // the tile's records including padding
    {
      // V/N rows of the tile: N0 rows of 32 lines; a 16-byte record past the
      // last line (partial last tile) loads the row's first record instead
      constexpr int RPR = 32 * (int)sizeof(T) / 16;  // records per row
      constexpr int LPR = 16 / (int)sizeof(T);       // lines per record
      const char *vg = reinterpret_cast<const char *>(B.vloc);
      for (int i = wave; i * 64 < N0 * RPR; i += ns) {
        const int c = i * 64 + lane;
        if (c < N0 * RPR) {
          const int r = c / RPR, cc = c - r * RPR;
          const int line = tile * kTile + cc * LPR;
          const int l0 = line < hd.n12 ? line : tile * kTile;
          dma16(vg + ((size_t)r * hd.n12 + l0) * sizeof(T), smem_raw + xl.v + i * 1024);
        }
      }
    }
    if (npts > 0) {
      dma_copy(hd.ent_cs + 2 * (size_t)ent_base, smem_raw + xl.cs, ncap * 8, wave, ns, lane);
      dma_copy(static_cast<const char *>(hd.ent_b) + (size_t)ent_base * 4 * sizeof(T), smem_raw + xl.bt,
               ncap * 4 * (int)sizeof(T), wave, ns, lane);
      if (!B.gamma) dma_copy(B.kbph + ent_base, smem_raw + xl.ph, ncap * (int)sizeof(C), wave, ns, lane);
      if (act) {
        // this state's D P of the tile's slots: kMaxNlm values per slot
        constexpr int RPS = kMaxNlm * (int)sizeof(C) / 16;  // records per slot
        const char *dg = reinterpret_cast<const char *>(B.dps);
        for (int i = 0; i * 64 < nsl * RPS; ++i) {
          const int c = i * 64 + lane;
          if (c < nsl * RPS) {
            const int q = c / RPS, cc = c - q * RPS;
            dma16(dg + ((size_t)(slot_base + q) * B.nst + s) * kMaxNlm * sizeof(C) + cc * 16,
                  reinterpret_cast<char *>(dpl) + i * 1024);
          }
        }
      }
    }
    C v[H];
    LineMask<N0> lm{};
    if (act) {
      load_line_slots_split<T, N0>(B, hd.n12, s, Lc, upper, v);
      if (npts > 0) {
        lm = load_line_mask<N0>(hd.line_mask, Lc);
        // the zero slot pts[npts] and the trash slot pts[npts + 1]; the compact
        // points too unless every point has exactly one record (then the KB
        // loop below stores each point once)
        for (int i = (nent == npts ? npts : 0) + ptid; i < npts + 2; i += 64) pts[i] = C{0, 0};
      }
      fft_split<N0, +1, C>::dit(v, upper);
      if (B.dip) {
        // sum |u|^2 and sum f0 |u|^2 along the lane's half line, f0 = sym(i0)/n0
        double s0 = 0.0, s1 = 0.0;
        if (valid) {
#pragma unroll
          for (int k = 0; k < H; ++k) {
            const double x2 = (double)v[k].x * (double)v[k].x + (double)v[k].y * (double)v[k].y;
            const double f0 = upper ? (double)sym_of<N0>(k + H) / (double)N0 : (double)sym_of<N0>(k) / (double)N0;
            s0 += x2;
            s1 += x2 * f0;
          }
        }
        const int i1 = Lc / hd.n2, i2 = Lc - i1 * hd.n2;
        const double f1 = (double)(i1 >= (hd.n1 + 1) / 2 ? i1 - hd.n1 : i1) / hd.n1;
        const double f2 = (double)(i2 >= (hd.n2 + 1) / 2 ? i2 - hd.n2 : i2) / hd.n2;
        double d[4];
        d[0] = s0;
#pragma unroll
        for (int c = 0; c < 3; ++c) d[c + 1] = s1 * hd.lat[c] + s0 * (f1 * hd.lat[3 + c] + f2 * hd.lat[6 + c]);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) d[c] += __shfl_xor(d[c], off, 64);
        if (lane == 0) {
          GAS double *dq = gw(B.dip) + ((size_t)tile * B.nst + s) * 4;
#pragma unroll
          for (int c = 0; c < 4; ++c) dq[c] = d[c];
        }
      }
    }
    __syncthreads();  // V tile, entry records and D P in LDS
    if (act) {
      // V/N at i0 = k + half N0/2
      const T *vrow = sv + (upper ? H * kTile : 0) + (lane & 31);
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const T vv = vrow[k * kTile];
        v[k].x *= vv;
        v[k].y *= vv;
      }
      if (npts > 0) {
        // the tile's records in LDS order, two rows per step with all reads
        // first; padding records (point -1, zero projector values) go to the
        // trash slot.  Where no two records share a point (no overlapping
        // projector spheres in the tile: as many records as points) each
        // point is stored once; otherwise the records add with LDS atomics.
        const int trash = npts + 1;
        const bool once = nent == npts;  // uniform
        for (int e0 = 0; e0 < ncap; e0 += 128) {
          const int nr = min(2, (ncap - e0) >> 6);  // uniform
          int2 cs[2];
          V4<T> bt[2];
          C dd[2][kMaxNlm];
          C ph[2];
#pragma unroll
          for (int q = 0; q < 2; ++q) cs[q] = scs[e0 + 64 * min(q, nr - 1) + ptid];
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int e = e0 + 64 * min(q, nr - 1) + ptid;
            bt[q] = sbt[e];
            const C *d = dpl + max(cs[q].y, 0) * kMaxNlm;
#pragma unroll
            for (int l = 0; l < kMaxNlm; ++l) dd[q][l] = d[l];
            if (!B.gamma) ph[q] = sph[e];
          }
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            if (q < nr) {
              C t;
              t.x = bt[q].x * dd[q][0].x + bt[q].y * dd[q][1].x + bt[q].z * dd[q][2].x + bt[q].w * dd[q][3].x;
              t.y = bt[q].x * dd[q][0].y + bt[q].y * dd[q][1].y + bt[q].z * dd[q][2].y + bt[q].w * dd[q][3].y;
              if (!B.gamma) t = cmul(t, C{ph[q].x, -ph[q].y});  // exp(-i (k+A).r)
              C *r = pts + (cs[q].x < 0 ? trash : cs[q].x);
              if (once) {
                *r = t;
              } else {
                lds_add(&r->x, t.x);
                lds_add(&r->y, t.y);
              }
            }
          }
        }
      }
    }
    __syncthreads();
    if (act) {
      if (npts > 0) {
        // every value reads: its compact point, or the zero slot pts[npts]
        // (no branches, so the reads are in flight together)
        int cnt;
        const unsigned long long mh = valid ? half_mask<N0>(lm, upper, cnt) : 0ull;
        if (!valid) cnt = 0;
#pragma unroll
        for (int k = 0; k < H; ++k) {
          const int bit = (int)((mh >> k) & 1ull);
          v[k] = cadd(v[k], pts[bit ? cnt : npts]);
          cnt += bit;
        }
      }
      fft_split<N0, -1, C>::dif(v, upper);
      store_line_slots_split<T, N0>(B, hd.n12, s, L, upper, valid, v);
    }
  }
}


// -------------------------------------------------------------------------------------------
// observables of a block of sphere states (FP64 accumulation, fixed reduction order)
// -------------------------------------------------------------------------------------------
constexpr int kDotChunk = 4096;

// per (state, chunk): sum kin |c|^2, sum |c|^2, sum conj(c) h (re, im)
template <typename T>
__global__ void __launch_bounds__(256) k_sphere_dots(const cplx<T> *psi, const cplx<T> *hpsi, const T *kin, int nst,
                                                     int npw, int nch, double *partial) {
  using C = cplx<T>;
  __shared__ double red[4][256];
  const int s = blockIdx.x / nch, ch = blockIdx.x - s * nch;
  const int tid = threadIdx.x;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  const GAS C *ps = gp(psi) + (size_t)s * npw;
  const GAS C *hs = gp(hpsi) + (size_t)s * npw;
  const GAS T *kn = gp(kin);
  const int e0 = ch * kDotChunk, e1 = min(e0 + kDotChunk, npw);
  for (int i = e0 + tid; i < e1; i += 256) {
    const C c = ps[i], h = hs[i];
    const double n2 = (double)c.x * c.x + (double)c.y * c.y;
    a[0] += (double)kn[i] * n2;
    a[1] += n2;
    a[2] += (double)c.x * h.x + (double)c.y * h.y;
    a[3] += (double)c.x * h.y - (double)c.y * h.x;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[c][tid] = a[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int c = 0; c < 4; ++c) red[c][tid] += red[c][tid + w];
    __syncthreads();
  }
  if (tid < 4) partial[((size_t)s * nch + ch) * 4 + tid] = red[tid][0];
}

struct ObsBlock {
  const double *dots;   // [nst][nch][4]
  const void *proj;     // [natoms][kMaxNlm][nst] complex T
  const void *dp;       // [natoms][kMaxNlm][nst] complex T
  const double *dip;    // [ntile][nst][4]
  const double *occ;    // [nst]
  int nst, nch;
};

// one work-group: combines every block into record[slot] =
// {kinetic, eigenvalue sum, non-local, norm sum, dipole x, y, z, step}
// (energy::calculate, src/hamiltonian/energy.hpp:61-113; dipole, src/observables/dipole.hpp:24-59)
template <typename T>
__global__ void __launch_bounds__(256) k_obs_combine(const ObsBlock *ob, int nblocks, int natoms, int ntile,
                                                     double vol_g, double dip_scale, double *ring, int ring_cap,
                                                     int *counter) {
  using C = cplx<T>;
  __shared__ double red[7][256];
  const int tid = threadIdx.x;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nblocks; ++b) {
    const ObsBlock B = ob[b];
    for (int s = tid; s < B.nst; s += 256) {
      double d[4] = {0, 0, 0, 0};
      // unrolled so that several loads of each chain are in flight (one lane per state)
#pragma unroll 8
      for (int ch = 0; ch < B.nch; ++ch)
        for (int c = 0; c < 4; ++c) d[c] += B.dots[((size_t)s * B.nch + ch) * 4 + c];
      const double occ = B.occ[s];
      acc[0] += occ * d[0] / d[1];
      acc[1] += occ * d[2] / d[1];
      acc[3] += d[1] * vol_g;
      double enl = 0.0;
      if (B.proj) {
        const C *pj = reinterpret_cast<const C *>(B.proj), *dq = reinterpret_cast<const C *>(B.dp);
#pragma unroll 2
        for (int a = 0; a < natoms; ++a)
#pragma unroll
          for (int l = 0; l < kMaxNlm; ++l) {
            const C p = pj[((size_t)a * kMaxNlm + l) * B.nst + s], q = dq[((size_t)a * kMaxNlm + l) * B.nst + s];
            enl += (double)p.x * q.x + (double)p.y * q.y;
          }
      }
      acc[2] += occ * enl;
      if (B.dip) {
        double dx = 0, dy = 0, dz = 0;
#pragma unroll 8
        for (int t = 0; t < ntile; ++t) {
          const double *q = B.dip + ((size_t)t * B.nst + s) * 4;
          dx += q[1];
          dy += q[2];
          dz += q[3];
        }
        acc[4] += occ * dx * dip_scale;
        acc[5] += occ * dy * dip_scale;
        acc[6] += occ * dz * dip_scale;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 7; ++c) red[c][tid] = acc[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int c = 0; c < 7; ++c) red[c][tid] += red[c][tid + w];
    __syncthreads();
  }
  if (tid == 0) {
    const int step = counter[0];
    double *rec = ring + (size_t)(step % ring_cap) * 8;
    for (int c = 0; c < 7; ++c) rec[c] = red[c][0];
    rec[7] = (double)step;
    counter[0] = step + 1;
  }
}

// -------------------------------------------------------------------------------------------
// host side: tables and launches
// -------------------------------------------------------------------------------------------

bool grid_supported(const int *n) {
  auto ok = [](int v) {
    switch (v) {
#define CASE(S) case S:
      INQX_SIZES(CASE)
#undef CASE
      return true;
      default: return false;
    }
  };
  return ok(n[0]) && ok(n[1]) && ok(n[2]);
}

void build_kb_tables(inqx_ham_s *h) {
  INQX_CHECK(h->pinned == 0, INQX_ERR_STATE, "projector tables are pinned by a captured propagation graph");
  invalidate_launch_cache(h);
  KBTables &kb = h->kb;
  const Grid &g = h->g;
  hipStream_t st = h->ctx->c.stream;
  const bool dev = !h->ctx->c.host_only;
  const int n12 = g.n[1] * g.n[2];
  // KB groups: projector j of nlm channels -> groups gfirst[j] .. + ng(j) - 1
  std::vector<int> gproj, gindex, sib;  // per group: projector, group index in it; sibling (first, count)
  for (int j = 0; j < (int)h->projs.size(); ++j) {
    const int ng = (h->projs[j].nlm + kMaxNlm - 1) / kMaxNlm;
    const int first = (int)gproj.size();
    for (int q = 0; q < ng; ++q) {
      gproj.push_back(j);
      gindex.push_back(q);
      sib.push_back(first);
      sib.push_back(ng);
    }
  }
  kb.natoms = (int)gproj.size();
  kb.maxsph = 1;
  kb.maxchunk = 0;
  for (auto &p : h->projs) kb.maxsph = std::max<int>(kb.maxsph, (int)p.npts());
  kb.ntile = (n12 + kTile - 1) / kTile;
  struct Mem {
    int tile, lane, i0, a, p;
  };
  std::vector<Mem> mem;
  for (int a = 0; a < kb.natoms; ++a) {
    const Projector &P = h->projs[gproj[a]];
    for (long ip = 0; ip < P.npts(); ++ip) {
      const int *c = &P.coords[3 * ip];
      const int L = c[1] * g.n[2] + c[2];
      mem.push_back(Mem{L / kTile, L % kTile, c[0], a, (int)ip});
    }
  }
  // unique points of each tile sorted by (lane, i0) give the compact index
  // line mask records: 4 words (mask 0-2, count) up to 96 points per line, 8
  // (mask 0-3, count, padding) above
  const int mw = g.n[0] > 96 ? 8 : 4, mcount = g.n[0] > 96 ? 4 : 3;
  std::vector<uint32_t> mask((size_t)n12 * mw, 0u);
  std::vector<std::vector<std::pair<int, int>>> tile_pts(kb.ntile);
  for (auto &m : mem) tile_pts[m.tile].push_back({m.lane, m.i0});
  std::vector<int> tile_info((size_t)kb.ntile * 8, 0);
  kb.maxpts = 1;
  for (int t = 0; t < kb.ntile; ++t) {
    auto &v = tile_pts[t];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    kb.maxpts = std::max(kb.maxpts, (int)v.size());
    int c = 0;
    for (size_t k = 0; k < v.size(); ++k) {
      const int L = t * kTile + v[k].first;
      if (k == 0 || v[k].first != v[k - 1].first) mask[(size_t)L * mw + mcount] = (uint32_t)c;
      mask[(size_t)L * mw + (v[k].second >> 5)] |= 1u << (v[k].second & 31);
      ++c;
    }
    tile_info[8 * t + 0] = (int)v.size();
  }
  auto compact = [&](int t, int lane, int i0) {
    auto &v = tile_pts[t];
    return (int)(std::lower_bound(v.begin(), v.end(), std::make_pair(lane, i0)) - v.begin());
  };
  // entries sorted by (tile, atom, point) -> slots = (tile, atom) runs
  std::stable_sort(mem.begin(), mem.end(), [](const Mem &x, const Mem &y) {
    if (x.tile != y.tile) return x.tile < y.tile;
    return x.a < y.a;
  });
  std::vector<int> slot_atom, ent_cs;
  std::vector<float> eb32, epos;
  std::vector<double> eb64;
  std::vector<std::vector<int>> atom_slot(kb.natoms);
  kb.maxslot = 1;
  for (size_t k = 0; k < mem.size(); ++k) {
    const Mem &m = mem[k];
    const bool new_tile = (k == 0 || m.tile != mem[k - 1].tile);
    if (new_tile) {
      tile_info[8 * m.tile + 2] = (int)slot_atom.size();
      tile_info[8 * m.tile + 3] = (int)k;
    }
    if (new_tile || m.a != mem[k - 1].a) {
      atom_slot[m.a].push_back((int)slot_atom.size());
      slot_atom.push_back(m.a);
    }
    tile_info[8 * m.tile + 1] = (int)slot_atom.size() - tile_info[8 * m.tile + 2];
    tile_info[8 * m.tile + 4] = (int)k + 1 - tile_info[8 * m.tile + 3];
    kb.maxslot = std::max(kb.maxslot, tile_info[8 * m.tile + 1]);
    ent_cs.push_back(compact(m.tile, m.lane, m.i0));
    ent_cs.push_back((int)slot_atom.size() - 1 - tile_info[8 * m.tile + 2]);
    const Projector &P = h->projs[gproj[m.a]];
    for (int l = 0; l < kMaxNlm; ++l) {
      const int ch = kMaxNlm * gindex[m.a] + l;  // channel of the projector
      const double v = ch < P.nlm ? P.matrix[(size_t)ch * P.npts() + m.p] : 0.0;
      eb64.push_back(v);
      eb32.push_back((float)v);
    }
    for (int d = 0; d < 3; ++d) epos.push_back(P.relpos[3 * (size_t)m.p + d]);
    epos.push_back(0.0f);
  }
  kb.nslot = (int)slot_atom.size();
  // Record order per tile, for the project pass: every lane of a wave walks a
  // run of records of ONE slot, so it sums in registers and its partial
  // needs no atomics.  A slot's records are split into contiguous runs over
  // a contiguous range of lanes, the lanes dealt out so that the longest run
  // (the tile's rows) is as short as possible; record j of lane l's run is
  // stored at the tile base + 64 j + l (the 64 lanes' j-th loads are
  // consecutive), runs shorter than the rows padded with empty records
  // (point and slot -1, zero projector values).  lane_info[tile][l] = run
  // length | (slot + 1) << 8 | (first lane of slot l, l <= slots) << 16.
  std::vector<int> lane_info((size_t)kb.ntile * 64, 0);
  {
    std::vector<int> cs2;
    std::vector<float> b32, pos2;
    std::vector<double> b64;
    long base = 0;
    for (int t = 0; t < kb.ntile; ++t) {
      const int e0 = tile_info[8 * t + 3], ne = tile_info[8 * t + 4], nsl_t = tile_info[8 * t + 1];
      INQX_CHECK(nsl_t <= 63, INQX_ERR_UNSUPPORTED, "more than 63 projector atoms meet one 32-line tile");
      std::vector<int> off(nsl_t + 1, 0), cnt(nsl_t, 0), nl(nsl_t, 1);
      for (int k = 0; k < ne; ++k) ++cnt[ent_cs[2 * ((size_t)e0 + k) + 1]];
      for (int q = 0; q < nsl_t; ++q) off[q + 1] = off[q] + cnt[q];
      auto rows_of = [&](int q) { return (cnt[q] + nl[q] - 1) / nl[q]; };
      for (int used = nsl_t; used < 64; ++used) {  // one more lane to the slot with the longest run
        int best = -1;
        for (int q = 0; q < nsl_t; ++q)
          if (rows_of(q) > 1 && (best < 0 || rows_of(q) > rows_of(best))) best = q;
        if (best < 0) break;
        ++nl[best];
      }
      int rows = 0;
      for (int q = 0; q < nsl_t; ++q) rows = std::max(rows, rows_of(q));
      INQX_CHECK(rows <= 255, INQX_ERR_UNSUPPORTED, "KB tile too dense (a lane run above 255 records)");
      std::vector<int> lane_slot(64, -1), lane_first(64, 0), lane_len(64, 0);
      int l = 0;
      for (int q = 0; q < nsl_t; ++q) {
        const int c = rows_of(q);
        lane_info[(size_t)t * 64 + q] |= l << 16;  // first lane of slot q
        for (int part = 0; part < nl[q]; ++part, ++l) {
          const int b0 = std::min(cnt[q], part * c), b1 = std::min(cnt[q], (part + 1) * c);
          lane_slot[l] = q;
          lane_first[l] = off[q] + b0;
          lane_len[l] = b1 - b0;
        }
      }
      lane_info[(size_t)t * 64 + nsl_t] |= l << 16;  // one past the last lane used
      for (int k = 0; k < 64; ++k) lane_info[(size_t)t * 64 + k] |= lane_len[k] | ((lane_slot[k] + 1) << 8);
      for (int j = 0; j < rows; ++j)
        for (int k = 0; k < 64; ++k) {
          const bool ok = j < lane_len[k];
          const size_t o = (size_t)e0 + (size_t)(ok ? lane_first[k] + j : 0);
          cs2.push_back(ok ? ent_cs[2 * o] : -1);
          cs2.push_back(ok ? ent_cs[2 * o + 1] : -1);
          for (int q = 0; q < kMaxNlm; ++q) {
            b32.push_back(ok ? eb32[kMaxNlm * o + q] : 0.0f);
            b64.push_back(ok ? eb64[kMaxNlm * o + q] : 0.0);
          }
          for (int q = 0; q < 4; ++q) pos2.push_back(ok ? epos[4 * o + q] : 0.0f);
        }
      tile_info[8 * t + 3] = (int)base;
      tile_info[8 * t + 5] = rows;
      base += 64L * rows;
      kb.maxchunk = std::max(kb.maxchunk, rows);
    }
    ent_cs.swap(cs2);
    eb32.swap(b32);
    eb64.swap(b64);
    epos.swap(pos2);
  }
  kb.nent = (long)ent_cs.size() / 2;
  ++kb.gen;
  kb.pos_host = epos;
  std::vector<int> atom_slots(kb.natoms + 1, 0), atom_slot_list;
  for (int a = 0; a < kb.natoms; ++a) {
    atom_slots[a] = (int)atom_slot_list.size();
    for (int x : atom_slot[a]) atom_slot_list.push_back(x);
  }
  atom_slots[kb.natoms] = (int)atom_slot_list.size();
  // coefficient rows of each group: rows 4q .. 4q+3 of its projector's D,
  // columns over all channels of the projector (group by group)
  std::vector<double> coeff((size_t)std::max(1, kb.natoms) * kMaxNlm * kCoeffCols, 0.0);
  for (int a = 0; a < kb.natoms; ++a) {
    const Projector &P = h->projs[gproj[a]];
    for (int l = 0; l < kMaxNlm; ++l) {
      const int row = kMaxNlm * gindex[a] + l;
      if (row >= P.nlm) continue;
      for (int m = 0; m < P.nlm; ++m) coeff[((size_t)a * kMaxNlm + l) * kCoeffCols + m] = P.coeff[row * P.nlm + m];
    }
  }
  if (sib.empty()) sib.assign(2, 0);
  if (!dev) {
    kb.built = true;
    return;
  }
  upload(kb.line_mask, mask, st);
  upload(kb.tile_info, tile_info, st);
  upload(kb.lane_info, lane_info, st);
  upload(kb.slot_atom, slot_atom, st);
  upload(kb.ent_cs, ent_cs, st);
  upload(kb.ent_b32, eb32, st);
  upload(kb.ent_b64, eb64, st);
  upload(kb.atom_slots, atom_slots, st);
  upload(kb.atom_slot_list, atom_slot_list, st);
  upload(kb.coeff, coeff, st);
  upload(kb.group_sib, sib, st);
  INQX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
  kb.built = true;
}

void ensure_basis_device(inqx_ham_s *h, Basis &b) {
  hipStream_t st = h->ctx->c.stream;
  if (!b.d_pw_lw.p) {
    // LDS offset of every sphere point in its plane (row window layout)
    const Grid &g = h->g;
    const int n2p = g.n[2] + 1;
    std::vector<int> lw(b.npw);
    for (int sl = 0; sl < b.nslot; ++sl) {
      const int *ss = &b.slot_seg[8 * sl];
      for (int j = 0; j < ss[1]; ++j) {
        const long q = ss[0] + j;
        const int y = b.pw_yz[q] / g.n[2], z = b.pw_yz[q] - y * g.n[2];
        lw[q] = (Grid::sym(y, g.n[1]) - ss[4] + 1) * n2p + z;
      }
    }
    upload(b.d_slot_seg, b.slot_seg, st);
    upload(b.d_rows, b.rows, st);
    upload(b.d_pw_lw, lw, st);
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

// exp(+i (k+A).r_e) of every KB entry e for this basis' k and the current A,
// with r_e the entry's float contravariant position as the reference keeps it
// (projector_all.hpp:123,284): the phase depends on (k + A, entry) only, so it
// is tabulated once instead of evaluated per (entry, state) in the x passes
static void ensure_phase_table(inqx_ham_s *h, Basis &b) {
  const KBTables &kb = h->kb;
  if (b.ph_gen == kb.gen && b.ph_A[0] == h->A[0] && b.ph_A[1] == h->A[1] && b.ph_A[2] == h->A[2]) return;
  INQX_CHECK(b.pinned == 0, INQX_ERR_STATE, "basis tables are pinned by a captured propagation graph");
  const long nent = kb.nent;
  std::vector<double> p64(2 * (size_t)std::max(1L, nent));
  std::vector<float> p32(2 * (size_t)std::max(1L, nent));
  const double kk[3] = {b.kcov[0] + h->A[0], b.kcov[1] + h->A[1], b.kcov[2] + h->A[2]};
  for (long e = 0; e < nent; ++e) {
    const float *r = &kb.pos_host[4 * (size_t)e];
    const double arg = kk[0] * (double)r[0] + kk[1] * (double)r[1] + kk[2] * (double)r[2];
    p64[2 * e] = std::cos(arg);
    p64[2 * e + 1] = std::sin(arg);
    p32[2 * e] = (float)p64[2 * e];
    p32[2 * e + 1] = (float)p64[2 * e + 1];
  }
  hipStream_t st = h->ctx->c.stream;
  upload(b.d_ph64, p64, st);
  upload(b.d_ph32, p32, st);
  INQX_HIP(hipStreamSynchronize(st));
  b.ph_gen = kb.gen;
  for (int d = 0; d < 3; ++d) b.ph_A[d] = h->A[d];
}

void ensure_workspace(inqx_ham_s *h, Basis &b, int nst, int prec) {
  const size_t cs = prec == INQX_FP64 ? 16 : 8;
  if (b.ws_nst[prec] >= nst && b.t1[prec].p && b.ws_kb_nslot[prec] == h->kb.nslot &&
      b.ws_natoms[prec] == h->kb.natoms)
    return;
  INQX_CHECK(b.pinned == 0, INQX_ERR_STATE, "basis workspaces are pinned by a captured propagation graph");
  const int n12 = h->g.n[1] * h->g.n[2];
  const int ns = std::max(nst, b.ws_nst[prec]);
  b.t1[prec].ensure(cs * (size_t)std::max(1, b.nslot) * ns * n12);
  b.partial[prec].ensure(cs * (size_t)std::max(1, h->kb.nslot) * kMaxNlm * ns);
  b.dp[prec].ensure(cs * (size_t)std::max(1, h->kb.natoms) * kMaxNlm * ns);
  b.dps[prec].ensure(cs * (size_t)std::max(1, h->kb.nslot) * kMaxNlm * ns);
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
  int mode = 0;                      // ApplyOpts::mode
  int total_line = 0, ntile_line = 0;  // line passes of the transform modes
  unsigned long long last_use = 0;     // lookup counter at the most recent use (eviction order)
  int plane_batch = 1;                 // planes (states) per plane-pass work-group
  int plane_pstr = 0;                  // LDS elements per plane: rows (the longest row window + 2, or n1) x (n2 + 1)
  bool observables = false;            // the blocks point at the bases' projection and dipole buffers
  XLds xl_apply{}, xl_project{};       // x-pass LDS layouts (byte offsets)
  int xlds_apply = 0, xlds_project = 0;  // x-pass LDS bytes per work-group
};
static std::map<std::pair<inqx_ham_s *, LaunchKey>, std::shared_ptr<LaunchData>> g_launch_cache;
static unsigned long long g_launch_clock = 0;
constexpr size_t kLaunchCacheCap = 512;

// drops the least recently used half of the descriptors: those of the current
// step were just used and stay (a graph capture that follows an eager step
// finds every configuration the step needs)
static void evict_launch_cache() {
  std::vector<unsigned long long> uses;
  uses.reserve(g_launch_cache.size());
  for (auto &kv : g_launch_cache) uses.push_back(kv.second->last_use);
  std::nth_element(uses.begin(), uses.begin() + uses.size() / 2, uses.end());
  const unsigned long long cut = uses[uses.size() / 2];
  for (auto it = g_launch_cache.begin(); it != g_launch_cache.end();) {
    if (it->second->last_use < cut) it = g_launch_cache.erase(it);
    else ++it;
  }
}

// while a hipGraph is being captured, every descriptor it uses is kept alive
// by the capturing object, whatever happens to the cache later
static thread_local std::vector<std::shared_ptr<void>> *g_capture_keep = nullptr;
void set_capture_keep(std::vector<std::shared_ptr<void>> *keep) { g_capture_keep = keep; }

// descriptors hold raw pointers to workspaces, tables and bases: drop every
// cached descriptor of a Hamiltonian whenever one of those may change
void invalidate_launch_cache(inqx_ham_s *h) {
  for (auto it = g_launch_cache.begin(); it != g_launch_cache.end();) {
    if (it->first.first == h) it = g_launch_cache.erase(it);
    else ++it;
  }
}
// only the descriptors that point at the projection and dipole buffers: the
// plain applications of the same step stay cached (a graph captured after an
// eager step whose energy record grew those buffers still finds them)
static void invalidate_observable_launches(inqx_ham_s *h) {
  for (auto it = g_launch_cache.begin(); it != g_launch_cache.end();) {
    if (it->first.first == h && it->second->observables) it = g_launch_cache.erase(it);
    else ++it;
  }
}

// -------------------------------------------------------------------------------------------
// sphere <-> real-space line passes: one lane per line, 64 lines x 4 states per work-group
// -------------------------------------------------------------------------------------------

// inverse FFT along i0 of the plane slots written by k_plane_inverse, times
// 1/N, into the full real-space grid (operations::transform::to_real,
// src/operations/transform.hpp:113-182); 32 lines (lane pairs) per wave,
// one state per wave
template <typename T, int N0>
__global__ void __launch_bounds__(256) k_line_to_real(const SphereBlock<T> *blocks, const int *prefix, int nblocks,
                                                      int total, int n12, int ntile, T scale) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const bool upper = lane >= 32;
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> B = blocks[b];
    const int local = wg - gp(prefix)[b];
    const int ngrp = (B.nst + 3) >> 2;
    const int tile = local / ngrp;
    const int s = (local - tile * ngrp) * 4 + wave;
    if (s >= B.nst) return;  // wave-uniform
    const int L = tile * 32 + (lane & 31);
    const int Lc = L < n12 ? L : n12 - 1;
    C v[H];
    load_line_inverse_split<T, N0>(B, n12, s, Lc, upper, v);
    if (B.rs_layout == 0) {
      const unsigned pl = (unsigned)(n12 * sizeof(C));
      const auto rs = make_rsrc(B.rs + (size_t)s * N0 * n12, (unsigned)N0 * pl);
      const unsigned vo = L < n12 ? (unsigned)(L * sizeof(C)) + (upper ? (unsigned)H * pl : 0u) : kOOB;
#pragma unroll
      for (int k = 0; k < H; ++k) bstore(rs, vo, (unsigned)k * pl, C{v[k].x * scale, v[k].y * scale});
    } else if (L < n12) {
      {
        GAS C *out = gw(B.rs) + ((size_t)L + (upper ? (size_t)H * n12 : 0)) * B.nst + s;
#pragma unroll
        for (int k = 0; k < H; ++k) out[(size_t)k * n12 * B.nst] = C{v[k].x * scale, v[k].y * scale};
      }
    }
  }
}

// forward FFT along i0 of full real-space lines, keeping the plane indices of
// the sphere for k_plane_forward (operations::transform::to_fourier,
// src/operations/transform.hpp:52-109, restricted to the sphere)
template <typename T, int N0>
__global__ void __launch_bounds__(256) k_line_from_real(const SphereBlock<T> *blocks, const int *prefix,
                                                        int nblocks, int total, int n12, int ntile) {
  using C = cplx<T>;
  constexpr int H = N0 / 2;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
  const bool upper = lane >= 32;
  {  // one work item per work-group: the grid is the work count
    const int wg = blockIdx.x;
    if (wg >= total) return;
    const int b = find_block(prefix, nblocks, wg);
    const SphereBlock<T> B = blocks[b];
    const int local = wg - gp(prefix)[b];
    const int ngrp = (B.nst + 3) >> 2;
    const int tile = local / ngrp;
    const int s = (local - tile * ngrp) * 4 + wave;
    if (s >= B.nst) return;
    const int L = tile * 32 + (lane & 31);
    const int Lc = L < n12 ? L : n12 - 1;
    C v[H];
    if (B.rs_layout == 0) {
      const unsigned pl = (unsigned)(n12 * sizeof(C));
      const auto rs = make_rsrc(B.rs + (size_t)s * N0 * n12, (unsigned)N0 * pl);
      const unsigned vo = (unsigned)(Lc * sizeof(C)) + (upper ? (unsigned)H * pl : 0u);
#pragma unroll
      for (int k = 0; k < H; ++k) bload(rs, vo, (unsigned)k * pl, v[k]);
    } else {
      const GAS C *in = gp(B.rs) + ((size_t)Lc + (upper ? (size_t)H * n12 : 0)) * B.nst + s;
#pragma unroll
      for (int k = 0; k < H; ++k) v[k] = in[(size_t)k * n12 * B.nst];
    }
    fft_split<N0, -1, C>::dif(v, upper);
    store_line_slots_split<T, N0>(B, n12, s, L, upper, L < n12, v);
  }
}

// dynamic LDS of the plane passes: up to the largest plane window of the
// configuration (checked against the CU's 160 KB at each launch)
template <typename T, int N1, int N2> static void set_plane_attr(int device) {
  static DeviceOnce once;
  using Cfg = PlaneCfg<T, N1, N2>;
  if (Cfg::LDS <= 65536 || once.test_and_set(device)) return;
  const int lds = std::min(Cfg::LDS, 160 * 1024);
  INQX_HIP(hipFuncSetAttribute((const void *)k_plane_inverse<T, N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  INQX_HIP(hipFuncSetAttribute((const void *)k_plane_forward<T, N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
}

// the x-pass LDS depends on the KB tables: the attribute grows to the largest
// size requested on each device
template <typename T, int N0> static void set_x_attr(int device, int lds) {
  static int done[64];
  static DeviceOnce init;
  int &d = done[device & 63];
  if (!init.test_and_set(device)) d = 65536;
  if (lds <= d) return;
  INQX_HIP(hipFuncSetAttribute((const void *)k_xpass_project<T, N0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  INQX_HIP(hipFuncSetAttribute((const void *)k_xpass_apply<T, N0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  d = lds;
}

// N12 > 0: square planes N12 x N12 (one kernel per plane pass); N12 == 0:
// planes n1 x n2 with n1 != n2 (row and column kernels, plane_rect.hip)
template <typename T, int N0, int N12>
static void run_passes(inqx_ham_s *h, LaunchData &ld, int nblocks, const HamDev &hd, const double *cf) {
  constexpr int NP = N12 > 0 ? N12 : 16;  // (square configuration; unused when N12 == 0)
  using Cfg = PlaneCfg<T, NP, NP>;
  hipStream_t st = h->ctx->c.stream;
  const SphereBlock<T> *blk = ld.desc.as<SphereBlock<T>>();
  const int *pref_plane = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[0]);
  const int *pref_x = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[1]);
  const int *pref_red = reinterpret_cast<const int *>(ld.desc.as<char>() + ld.off_pref[2]);
  if constexpr (N12 > 0) set_plane_attr<T, N12, N12>(h->ctx->c.device);
  const int cu = h->ctx->c.num_cu;
  const int grid_plane = ld.total_plane;  // one item per work-group
  const int plane_lds = ld.plane_batch * ld.plane_pstr * (int)sizeof(cplx<T>);
  INQX_CHECK(plane_lds <= 160 * 1024, INQX_ERR_UNSUPPORTED,
             "the plane passes' row window does not fit the 160 KB of LDS in this precision (" +
                 std::to_string(plane_lds) + " bytes)");
  Ctx &cx = h->ctx->c;
  const bool f64 = std::is_same<T, double>::value;
  auto plane_inv = [&]() {
    if constexpr (N12 > 0) {
      ProfScope ps(cx, f64 ? "plane_inverse_f64" : "plane_inverse_f32");
      hipLaunchKernelGGL((k_plane_inverse<T, N12, N12>), dim3(grid_plane), dim3(Cfg::THREADS), plane_lds, st, blk,
                         pref_plane, nblocks, ld.total_plane, ld.plane_batch, ld.plane_pstr);
      INQX_HIP(hipGetLastError());
    } else {
      launch_rect_inverse<T>(cx, blk, pref_plane, nblocks, ld.total_plane, ld.plane_batch, ld.plane_pstr, hd.n1,
                             hd.n2);
    }
  };
  auto plane_fwd = [&](const double *c) {
    if constexpr (N12 > 0) {
      ProfScope ps(cx, f64 ? "plane_forward_f64" : "plane_forward_f32");
      hipLaunchKernelGGL((k_plane_forward<T, N12, N12>), dim3(grid_plane), dim3(Cfg::THREADS), plane_lds, st, blk,
                         pref_plane, nblocks, ld.total_plane, ld.plane_batch, ld.plane_pstr, (T)c[0], (T)c[1],
                         (T)c[2], (T)c[3]);
      INQX_HIP(hipGetLastError());
    } else {
      launch_rect_forward<T>(cx, blk, pref_plane, nblocks, ld.total_plane, ld.plane_batch, ld.plane_pstr, hd.n1,
                             hd.n2, c);
    }
  };
  if (ld.mode == 2) {  // real space -> sphere
    const int grid_l = ld.total_line;
    {
      ProfScope ps(cx, f64 ? "line_from_real_f64" : "line_from_real_f32");
      hipLaunchKernelGGL((k_line_from_real<T, N0>), dim3(grid_l), dim3(256), 0, st, blk, pref_x, nblocks,
                         ld.total_line, hd.n12, ld.ntile_line);
    }
    INQX_HIP(hipGetLastError());
    const double zero[4] = {0.0, 0.0, 0.0, 0.0};
    plane_fwd(zero);
    return;
  }
  plane_inv();
  if (ld.mode == 1) {  // sphere -> real space
    const int grid_l = ld.total_line;
    ProfScope ps(cx, f64 ? "line_to_real_f64" : "line_to_real_f32");
    hipLaunchKernelGGL((k_line_to_real<T, N0>), dim3(grid_l), dim3(256), 0, st, blk, pref_x, nblocks, ld.total_line,
                       hd.n12, ld.ntile_line, (T)(1.0 / (double)hd.N));
    INQX_HIP(hipGetLastError());
    return;
  }
  const int xthreads = 64 * ld.nwave;  // one wave per (tile, state)
  set_x_attr<T, N0>(h->ctx->c.device, std::max(ld.xlds_apply, ld.xlds_project));
  const int grid_x = ld.total_x;
  if (hd.has_kb) {
    {
      ProfScope ps(cx, f64 ? "xpass_project_f64" : "xpass_project_f32");
      hipLaunchKernelGGL((k_xpass_project<T, N0>), dim3(grid_x), dim3(xthreads), ld.xlds_project, st, blk, pref_x,
                         nblocks, ld.total_x, hd, ld.xl_project);
    }
    INQX_HIP(hipGetLastError());
    {
      ProfScope ps(cx, f64 ? "kb_reduce_f64" : "kb_reduce_f32");
      hipLaunchKernelGGL((k_kb_reduce<T>), dim3(ld.total_red), dim3(256), 0, st, blk, pref_red,
                         nblocks, ld.total_red, hd);
    }
    INQX_HIP(hipGetLastError());
  }
  {
    ProfScope ps(cx, f64 ? "xpass_apply_f64" : "xpass_apply_f32");
    hipLaunchKernelGGL((k_xpass_apply<T, N0>), dim3(grid_x), dim3(xthreads), ld.xlds_apply, st, blk, pref_x, nblocks,
                       ld.total_x, hd, ld.xl_apply);
  }
  INQX_HIP(hipGetLastError());
  plane_fwd(cf);
}

// the pass templates for a grid n0 x n1 x n2: x-pass line length N0 and the
// plane size (N12 x N12 when n1 = n2, else the row and column kernels of
// plane_rect.hip) are independent compile-time sizes (host code only: every
// kernel is instantiated once per size)
template <typename T, int N0>
static void dispatch_n12(int n1, int n2, inqx_ham_s *h, LaunchData &ld, int nblocks, const HamDev &hd,
                         const double *cf) {
  if (n1 != n2) {
    run_passes<T, N0, 0>(h, ld, nblocks, hd, cf);
    return;
  }
  switch (n1) {
#define CASE(S)                                     \
  case S:                                           \
    run_passes<T, N0, S>(h, ld, nblocks, hd, cf);   \
    break;
    INQX_SIZES(CASE)
#undef CASE
    default:
      INQX_CHECK(false, INQX_ERR_UNSUPPORTED, "grid size not built for the sphere path");
  }
}
template <typename T>
static void dispatch_passes(int n0, int n1, int n2, inqx_ham_s *h, LaunchData &ld, int nblocks, const HamDev &hd,
                            const double *cf) {
  switch (n0) {
#define CASE(S)                                             \
  case S:                                                   \
    dispatch_n12<T, S>(n1, n2, h, ld, nblocks, hd, cf);     \
    break;
    INQX_SIZES(CASE)
#undef CASE
    default:
      INQX_CHECK(false, INQX_ERR_UNSUPPORTED, "grid size not built for the sphere path");
  }
}

template <typename T>
void launch_apply(inqx_ham_s *h, int nblocks, Basis *const *bases, const int *nst, const void *const *psi,
                  void *const *hpsi, const ApplyOpts *opts) {
  static const ApplyOpts kNone;
  if (!opts) opts = &kNone;
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const Grid &g = h->g;
  const int mode = opts->mode;
  INQX_CHECK(grid_supported(g.n), INQX_ERR_UNSUPPORTED, "grid size not built for the sphere path");
  if (mode == 0) {
    INQX_CHECK(h->has_v, INQX_ERR_STATE, "potential not set");
    for (int i = 0; i < nblocks; ++i)
      INQX_CHECK(bases[i]->spin < h->ncomp, INQX_ERR_ARG,
                 "basis spin index " + std::to_string(bases[i]->spin) + " has no potential component (ncomp " +
                     std::to_string(h->ncomp) + ")");
    if (!h->kb.built) build_kb_tables(h);
    if (prec == INQX_FP32 && !h->v32.p) INQX_CHECK(false, INQX_ERR_STATE, "potential not uploaded");
  } else {
    INQX_CHECK(opts->rs != nullptr, INQX_ERR_ARG, "transform without real-space arrays");
  }
  const bool has_kb = mode == 0 && h->kb.natoms > 0;
  LaunchKey key;
  key.k.push_back((uintptr_t)prec);
  key.k.push_back((uintptr_t)mode);
  key.k.push_back((uintptr_t)opts->rs_layout);
  key.k.push_back((uintptr_t)(opts->accumulate ? 1 : 0));
  key.k.push_back((uintptr_t)h->xwaves[prec]);
  key.k.push_back((uintptr_t)h->plane_batch[prec]);
  if (mode != 0)
    for (int i = 0; i < nblocks; ++i) key.k.push_back((uintptr_t)opts->rs[i]);
  for (int i = 0; i < nblocks; ++i) {
    ensure_basis_device(h, *bases[i]);
    ensure_workspace(h, *bases[i], nst[i], prec);
    const double *kc = bases[i]->kcov;
    const bool gamma_i = kc[0] + h->A[0] == 0.0 && kc[1] + h->A[1] == 0.0 && kc[2] + h->A[2] == 0.0;
    if (has_kb && !gamma_i) ensure_phase_table(h, *bases[i]);
    key.k.push_back((uintptr_t)bases[i]);
    key.k.push_back((uintptr_t)nst[i]);
    key.k.push_back((uintptr_t)psi[i]);
    key.k.push_back((uintptr_t)hpsi[i]);
    key.k.push_back((uintptr_t)bases[i]->t1[prec].p);
    key.k.push_back((uintptr_t)(opts->acc1 ? opts->acc1[i] : nullptr));
    key.k.push_back((uintptr_t)(opts->acc2 ? opts->acc2[i] : nullptr));
    key.k.push_back((uintptr_t)(opts->acc_wide ? 1 : 0));
    key.k.push_back((uintptr_t)(prec == INQX_FP64 ? bases[i]->d_ph64.p : bases[i]->d_ph32.p));
  }
  key.k.push_back((uintptr_t)(opts->write_h ? 1 : 0));
  key.k.push_back((uintptr_t)(opts->observables ? 1 : 0));
  if (opts->observables) {
    for (int i = 0; i < nblocks; ++i) {
      Basis &B = *bases[i];
      const size_t cs = prec == INQX_FP64 ? 16 : 8;
      const size_t need_p = cs * (size_t)std::max(1, h->kb.natoms) * kMaxNlm * nst[i];
      const size_t need_d = 32 * (size_t)h->kb.ntile * nst[i];
      if (B.proj[prec].bytes < need_p || B.dip.bytes < need_d) {
        INQX_CHECK(B.pinned == 0, INQX_ERR_STATE, "basis workspaces are pinned by a captured propagation graph");
        B.proj[prec].ensure(need_p);
        B.dip.ensure(need_d);
        invalidate_observable_launches(h);
      }
    }
  }
  for (int d = 0; d < 3; ++d) {
    uint64_t bits;
    std::memcpy(&bits, &h->A[d], sizeof bits);
    key.k.push_back((uintptr_t)bits);
  }
  auto ck = std::make_pair(h, key);
  auto it = g_launch_cache.find(ck);
  if (it == g_launch_cache.end()) {
    if (g_launch_cache.size() >= kLaunchCacheCap) evict_launch_cache();
    INQX_CHECK(!g_capture_keep, INQX_ERR_STATE, "launch configuration first seen during graph capture");
    auto ld = std::make_shared<LaunchData>();
    std::vector<SphereBlock<T>> blk(nblocks);
    std::vector<int> pp(nblocks + 1, 0), px(nblocks + 1, 0), pr(nblocks + 1, 0);
    int plane_p = max_plane_batch(g.n, prec);
    // the launch shape's plane batch (autotuned), at most what the LDS budget allows
    if (h->plane_batch[prec] > 0) plane_p = std::min(plane_p, h->plane_batch[prec]);
    ld->plane_batch = plane_p;
    // waves (states) per x-pass work-group within the LDS budget
    const int cs = (int)sizeof(cplx<T>);
    int nw = std::max(1, std::min(xpass_threads<T>(g.n[0]) / 64, h->xwaves[prec]));  // states (waves) per x-pass work-group
    ld->mode = mode;
    ld->observables = opts->observables;
    ld->ntile_line = (g.n[1] * g.n[2] + 31) / 32;
    if (mode == 0) {
      bool need_ph = false;  // a block with k + A != 0 reads the entry phases
      for (int i = 0; i < nblocks && has_kb; ++i) {
        const double *kc = bases[i]->kcov;
        if (kc[0] + h->A[0] != 0.0 || kc[1] + h->A[1] != 0.0 || kc[2] + h->A[2] != 0.0) need_ph = true;
      }
      auto a16 = [](long x) { return (int)((x + 15) / 16 * 16); };
      const long nrec = 64L * h->kb.maxchunk;  // entry records of the largest tile
      XLds x;
      x.v = 0;
      x.cs = a16((long)g.n[0] * kTile * (long)sizeof(T));
      x.bt = x.cs + a16(nrec * 8);
      x.ph = x.bt + a16(nrec * 4 * (long)sizeof(T));
      x.li = x.ph + (need_ph ? a16(nrec * cs) : 0);
      x.wave0 = x.li + 64 * 4;
      x.aux = a16((long)(h->kb.maxpts + 2) * cs);  // + the zero and trash slots
      x.wave_bytes = x.aux + a16((long)kMaxNlm * h->kb.maxslot * cs);
      XLds xp = x;  // the project pass has no V tile, and 64 lanes' partial sums per wave
      xp.v = 0;
      xp.cs = 0;
      xp.bt = x.bt - x.cs;
      xp.ph = x.ph - x.cs;
      xp.li = x.li - x.cs;
      xp.wave0 = x.wave0 - x.cs;
      xp.wave_bytes = xp.aux + 64 * kMaxNlm * cs;
      constexpr int kLdsMax = 160 * 1024;
      INQX_CHECK(x.wave0 + x.wave_bytes <= kLdsMax && xp.wave0 + xp.wave_bytes <= kLdsMax, INQX_ERR_UNSUPPORTED,
                 "KB tile too dense for LDS");
      while (nw > 1 && (x.wave0 + nw * x.wave_bytes > kLdsMax || xp.wave0 + nw * xp.wave_bytes > kLdsMax)) --nw;
      ld->xl_apply = x;
      ld->xl_project = xp;
      ld->xlds_apply = x.wave0 + nw * x.wave_bytes;
      ld->xlds_project = xp.wave0 + nw * xp.wave_bytes;
    }
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
      d.dps = B.dps[prec].as<cplx<T>>();
      d.slot_seg = B.d_slot_seg.as<int>();
      d.rows = B.d_rows.as<int>();
      d.pw_lds = B.d_pw_lw.as<int>();
      d.kin = mode == 2 ? nullptr : prec == INQX_FP64 ? (const T *)B.d_kin64.p : (const T *)B.d_kin32.p;
      d.rs = mode != 0 ? reinterpret_cast<cplx<T> *>(opts->rs[i]) : nullptr;
      d.rs_layout = opts->rs_layout;
      d.accumulate = opts->accumulate ? 1 : 0;
      d.acc1 = opts->acc1 ? reinterpret_cast<cplx<T> *>(opts->acc1[i]) : nullptr;
      d.acc2 = opts->acc2 ? reinterpret_cast<cplx<T> *>(opts->acc2[i]) : nullptr;
      d.acc_wide = opts->acc_wide && !std::is_same<T, double>::value ? 1 : 0;
      d.proj = opts->observables ? B.proj[prec].as<cplx<T>>() : nullptr;
      d.dip = opts->observables ? B.dip.as<double>() : nullptr;
      d.write_h = (opts->write_h || mode == 2) ? 1 : 0;
      const DevMem &vm = prec == INQX_FP64 ? h->v64 : h->v32;
      d.vloc = vm.p ? reinterpret_cast<const T *>(vm.as<char>() + (size_t)B.spin * g.N * sizeof(T)) : nullptr;
      d.nst = nst[i];
      d.npw = (int)B.npw;
      d.nslot = B.nslot;
      d.m_lo = B.m_lo;
      for (int k = 0; k < 3; ++k) d.kcov[k] = B.kcov[k] + Afull[k];
      d.gamma = (d.kcov[0] == 0.0 && d.kcov[1] == 0.0 && d.kcov[2] == 0.0) ? 1 : 0;
      d.kbph = (has_kb && !d.gamma)
                   ? (prec == INQX_FP64 ? (const cplx<T> *)B.d_ph64.p : (const cplx<T> *)B.d_ph32.p)
                   : nullptr;
      pp[i + 1] = pp[i] + d.nslot * ((nst[i] + plane_p - 1) / plane_p);
      if (mode == 0)
        px[i + 1] = px[i] + x_tiles_padded(h->kb.ntile) * ((nst[i] + nw - 1) / nw);
      else
        px[i + 1] = px[i] + ld->ntile_line * ((nst[i] + 3) / 4);
      pr[i + 1] = pr[i] + (std::max(1, h->kb.natoms) * nst[i] + 63) / 64;  // 64 (group, state) items per work-group
    }
    int nwin = 1;
    for (int i = 0; i < nblocks; ++i) nwin = std::max(nwin, bases[i]->nwin_max);
    ld->plane_pstr = (nwin + 2) * (g.n[2] + 1);
    ld->total_plane = pp[nblocks];
    ld->total_x = px[nblocks];
    ld->total_line = px[nblocks];
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
  ld.last_use = ++g_launch_clock;
  if (g_capture_keep) g_capture_keep->push_back(it->second);
  HamDev hd;
  hd.n0 = g.n[0];
  hd.n1 = g.n[1];
  hd.n2 = g.n[2];
  hd.n12 = g.n[1] * g.n[2];
  hd.N = g.N;
  hd.natoms = h->kb.natoms;
  hd.has_kb = has_kb ? 1 : 0;
  hd.ntile = h->kb.ntile;
  hd.maxpts = h->kb.maxpts;
  hd.maxslot = h->kb.maxslot;
  hd.ntg = x_tiles_padded(h->kb.ntile);
  hd.line_mask = h->kb.line_mask.as<uint32_t>();
  hd.tile_info = h->kb.tile_info.as<int>();
  hd.lane_info = h->kb.lane_info.as<int>();
  hd.slot_atom = h->kb.slot_atom.as<int>();
  hd.ent_cs = h->kb.ent_cs.as<int>();
  hd.ent_b = prec == INQX_FP64 ? h->kb.ent_b64.p : h->kb.ent_b32.p;
  hd.atom_slots = h->kb.atom_slots.as<int>();
  hd.atom_slot_list = h->kb.atom_slot_list.as<int>();
  hd.coeff = h->kb.coeff.as<double>();
  hd.group_sib = h->kb.group_sib.as<int>();
  hd.kb_scale = (g.vol / (double)g.N) / (double)g.N;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) hd.lat[3 * r + c] = g.a[r][c];
  const double cf[4] = {opts->c1r, opts->c1i, opts->c2r, opts->c2i};
  dispatch_passes<T>(g.n[0], g.n[1], g.n[2], h, ld, nblocks, hd, cf);
}

int dots_chunks(int npw) { return (npw + kDotChunk - 1) / kDotChunk; }

// largest plane batch the plane passes of this grid and precision take
int max_plane_batch(const int *n, int prec) {
  if (n[1] != n[2]) return rect_plane_batch(prec, n[1], n[2]);
  switch (n[1]) {
#define CASE(S) \
  case S:       \
    return prec == INQX_FP64 ? PlaneCfg<double, S, S>::P : PlaneCfg<float, S, S>::P;
    INQX_SIZES(CASE)
#undef CASE
    default: return 1;
  }
}

void launch_sphere_dots(Ctx &cx, int prec, const void *psi, const void *hpsi, const void *kin, int nst, int npw,
                        double *partial) {
  const int nch = dots_chunks(npw);
  ProfScope ps(cx, "sphere_dots");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_sphere_dots<double>, dim3(nst * nch), dim3(256), 0, cx.stream, (const cplx<double> *)psi,
                       (const cplx<double> *)hpsi, (const double *)kin, nst, npw, nch, partial);
  else
    hipLaunchKernelGGL(k_sphere_dots<float>, dim3(nst * nch), dim3(256), 0, cx.stream, (const cplx<float> *)psi,
                       (const cplx<float> *)hpsi, (const float *)kin, nst, npw, nch, partial);
  INQX_HIP(hipGetLastError());
}

void launch_obs_combine(Ctx &cx, int prec, const void *obs_blocks, int nblocks, int natoms, int ntile, double vol_g,
                        double dip_scale, double *ring, int ring_cap, int *counter) {
  ProfScope ps(cx, "obs_combine");
  if (prec == INQX_FP64)
    hipLaunchKernelGGL(k_obs_combine<double>, dim3(1), dim3(256), 0, cx.stream, (const ObsBlock *)obs_blocks, nblocks,
                       natoms, ntile, vol_g, dip_scale, ring, ring_cap, counter);
  else
    hipLaunchKernelGGL(k_obs_combine<float>, dim3(1), dim3(256), 0, cx.stream, (const ObsBlock *)obs_blocks, nblocks,
                       natoms, ntile, vol_g, dip_scale, ring, ring_cap, counter);
  INQX_HIP(hipGetLastError());
}

template void launch_apply<double>(inqx_ham_s *, int, Basis *const *, const int *, const void *const *, void *const *,
                                   const ApplyOpts *);
template void launch_apply<float>(inqx_ham_s *, int, Basis *const *, const int *, const void *const *, void *const *,
                                  const ApplyOpts *);

}  // namespace inqx
