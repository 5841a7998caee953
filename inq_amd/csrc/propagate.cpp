// propagate.cpp -- Taylor exponentials, the ETRS real-time step, per-step
// energies and dipole, and the hipGraph-captured propagation loop.
//
// exponential_in_place (src/operations/exponential.hpp:21-34): the Taylor
// terms (f H)^n/n! phi are accumulated into phi by the last pass of each H
// application (no separate shift/axpy kernels, src/operations/shift.hpp:33-44).
// exponential_2_for_1 (exponential.hpp:49-69): two accumulators share H^n phi.
// etrs (src/real_time/etrs.hpp:22-79) for a Hamiltonian that does not depend
// on the density: the half step exp_T4(-i dt/2 H) phi is the first output of
// the 2-for-1 call, the full-step prediction only feeds the density mixing,
// and every SCF pass of etrs.hpp:59-77 reproduces the same states, so the
// step is two in-place Taylor-4 half steps (8 H applications).
// The per-step observables (propagate.hpp:99 energy.calculate,
// energy.hpp:61-113, observables::dipole, dipole.hpp:24-59) come from a 9th H
// application whose x pass also accumulates the dipole of the real-space
// orbitals; everything is reduced in FP64 into a device ring of records.
#include "internal.hpp"

#include <memory>

#include <algorithm>
#include <complex>
#include <cstring>

using namespace inqx;

namespace inqx {

void launch_sphere_dots(Ctx &cx, int prec, const void *psi, const void *hpsi, const void *kin, int nst, int npw,
                        double *partial);
void launch_obs_combine(Ctx &cx, int prec, const void *obs_blocks, int nblocks, int natoms, int ntile, double vol_g,
                        double dip_scale, double *ring, int ring_cap, int *counter);
int dots_chunks(int npw);

struct ObsBlockHost {
  const double *dots;
  const void *proj;
  const void *dp;
  const double *dip;
  const double *occ;
  int nst, nch;
};

// grows a basis workspace unless a captured graph uses the basis
static void ensure_ws(Basis &B, DevMem &m, size_t bytes) {
  if (m.p && m.bytes >= bytes) return;
  INQX_CHECK(B.pinned == 0, INQX_ERR_STATE, "basis workspaces are pinned by a captured propagation graph");
  m.ensure(bytes);
}

template <typename T>
static void exp_in_place(inqx_ham_s *h, int nb, Basis *const *bases, const int *nst, void *const *phi,
                         std::complex<double> f, int order) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  std::vector<void *> hA(nb), hB(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    ensure_ws(B, B.hn[prec][0], cs * (size_t)nst[i] * B.npw);
    ensure_ws(B, B.hn[prec][1], cs * (size_t)nst[i] * B.npw);
    hA[i] = B.hn[prec][0].p;
    hB[i] = B.hn[prec][1].p;
  }
  std::complex<double> coeff = 1.0;
  std::vector<const void *> in(nb);
  std::vector<void *> out(nb);
  for (int it = 1; it <= order; ++it) {
    coeff *= f / (double)it;
    for (int i = 0; i < nb; ++i) {
      in[i] = (it == 1) ? phi[i] : ((it % 2 == 0) ? hA[i] : hB[i]);
      out[i] = (it % 2 == 1) ? hA[i] : hB[i];
    }
    ApplyOpts o;
    o.acc1 = phi;
    o.c1r = coeff.real();
    o.c1i = coeff.imag();
    o.write_h = it < order;
    launch_apply<T>(h, nb, bases, nst, in.data(), out.data(), &o);
  }
}

template <typename T>
static void exp_2_for_1(inqx_ham_s *h, Basis *B, int nst, void *phi, void *expphi, std::complex<double> f1,
                        std::complex<double> f2, int order) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  const size_t bytes = cs * (size_t)nst * B->npw;
  ensure_ws(*B, B->hn[prec][0], bytes);
  ensure_ws(*B, B->hn[prec][1], bytes);
  INQX_HIP(hipMemcpyAsync(expphi, phi, bytes, hipMemcpyDeviceToDevice, h->ctx->c.stream));
  std::complex<double> c1 = 1.0, c2 = 1.0;
  void *hA = B->hn[prec][0].p, *hB = B->hn[prec][1].p;
  for (int it = 1; it <= order; ++it) {
    c1 *= f1 / (double)it;
    c2 *= f2 / (double)it;
    const void *in = (it == 1) ? phi : ((it % 2 == 0) ? hA : hB);
    void *out = (it % 2 == 1) ? hA : hB;
    void *a1 = expphi, *a2 = phi;
    ApplyOpts o;
    o.acc1 = &a1;
    o.acc2 = &a2;
    o.c1r = c1.real();
    o.c1i = c1.imag();
    o.c2r = c2.real();
    o.c2i = c2.imag();
    o.write_h = it < order;
    launch_apply<T>(h, 1, &B, &nst, &in, &out, &o);
  }
}

// observables of the current states of every block into ring[counter % cap]
template <typename T>
static void energy_record(inqx_ham_s *h, int nb, Basis *const *bases, const int *nst, void *const *phi,
                          const double *const *occ, double *ring, int cap, int *counter, DevMem &obs_desc,
                          std::vector<char> &obs_host, bool capturing) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  std::vector<const void *> in(nb);
  std::vector<void *> out(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    ensure_ws(B, B.hn[prec][0], cs * (size_t)nst[i] * B.npw);
    in[i] = phi[i];
    out[i] = B.hn[prec][0].p;
  }
  ApplyOpts o;
  o.observables = true;
  launch_apply<T>(h, nb, bases, nst, in.data(), out.data(), &o);
  std::vector<ObsBlockHost> ob(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    const int nch = dots_chunks((int)B.npw);
    ensure_ws(B, B.dots, 32 * (size_t)nst[i] * nch);
    const void *kin = prec == INQX_FP64 ? B.d_kin64.p : B.d_kin32.p;
    launch_sphere_dots(h->ctx->c, prec, phi[i], out[i], kin, nst[i], (int)B.npw, B.dots.as<double>());
    ob[i].dots = B.dots.as<double>();
    ob[i].proj = h->kb.natoms > 0 ? B.proj[prec].p : nullptr;
    ob[i].dp = B.dp[prec].p;
    ob[i].dip = B.dip.as<double>();
    ob[i].occ = occ[i];
    ob[i].nst = nst[i];
    ob[i].nch = nch;
  }
  // the descriptor array is uploaded again only when a pointer changed
  const size_t ob_bytes = sizeof(ObsBlockHost) * nb;
  if (obs_host.size() != ob_bytes || std::memcmp(obs_host.data(), ob.data(), ob_bytes) != 0) {
    INQX_CHECK(!capturing, INQX_ERR_STATE, "observable buffers changed during graph capture");
    obs_desc.ensure(ob_bytes);
    INQX_HIP(hipMemcpyAsync(obs_desc.p, ob.data(), ob_bytes, hipMemcpyHostToDevice, h->ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(h->ctx->c.stream));
    obs_host.assign(reinterpret_cast<const char *>(ob.data()), reinterpret_cast<const char *>(ob.data()) + ob_bytes);
  }
  const Grid &g = h->g;
  const double vol_g = g.vol / ((double)g.N * (double)g.N);
  // dipole = dV sum r |psi|^2 with psi = u / N
  const double dip_scale = (g.vol / (double)g.N) / ((double)g.N * (double)g.N);
  launch_obs_combine(h->ctx->c, prec, obs_desc.p, nb, h->kb.natoms, h->kb.ntile, vol_g, dip_scale, ring, cap, counter);
}

}  // namespace inqx

struct inqx_prop_s {
  inqx_ham ham;
  int prec;
  double dt;
  std::vector<Basis *> bases;
  std::vector<int> nst;
  std::vector<void *> phi;
  DevMem occ_dev;
  std::vector<const double *> occ;
  DevMem ring, counter, obs_desc;
  std::vector<char> obs_host;
  bool capturing = false;
  int cap = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int steps_done = 0;
  std::vector<std::shared_ptr<void>> keep;  // launch descriptors the graph uses
};

template <typename T> static void prop_step_eager(inqx_prop_s *p) {
  const std::complex<double> half(0.0, -p->dt / 2.0);
  const int nb = (int)p->bases.size();
  exp_in_place<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), half, 4);
  exp_in_place<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), half, 4);
  energy_record<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(), p->ring.as<double>(),
                   p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, p->capturing);
}

extern "C" {

int inqx_exponential(inqx_ham h, inqx_basis bs, int nst, int precision, double f_re, double f_im, int order,
                     void *phi) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && phi && nst > 0 && order >= 1, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  Basis *b = &bs->b;
  void *pp[1] = {phi};
  if (precision == INQX_FP64) exp_in_place<double>(h, 1, &b, &nst, pp, {f_re, f_im}, order);
  else if (precision == INQX_FP32) exp_in_place<float>(h, 1, &b, &nst, pp, {f_re, f_im}, order);
  else INQX_CHECK(false, INQX_ERR_ARG, "bad precision");
  API_END(ctx)
}

int inqx_exponential_2_for_1(inqx_ham h, inqx_basis bs, int nst, int precision, double f1_re, double f1_im,
                             double f2_re, double f2_im, int order, void *phi, void *expphi) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && phi && expphi && phi != expphi && nst > 0 && order >= 1, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  if (precision == INQX_FP64)
    exp_2_for_1<double>(h, &bs->b, nst, phi, expphi, {f1_re, f1_im}, {f2_re, f2_im}, order);
  else if (precision == INQX_FP32)
    exp_2_for_1<float>(h, &bs->b, nst, phi, expphi, {f1_re, f1_im}, {f2_re, f2_im}, order);
  else INQX_CHECK(false, INQX_ERR_ARG, "bad precision");
  API_END(ctx)
}

int inqx_prop_create(inqx_ham h, int nblocks, const inqx_basis *bases, const int *nst, int precision, void *const *phi,
                     const double *const *occ, double dt, int ring_capacity, inqx_prop *out) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bases && nst && phi && occ && out && nblocks > 0 && ring_capacity > 0, INQX_ERR_ARG, "null argument");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  for (int i = 0; i < nblocks; ++i)
    INQX_CHECK(bases[i] && bases[i]->b.ham == h && nst[i] > 0 && phi[i] && occ[i], INQX_ERR_ARG, "bad block");
  std::unique_ptr<inqx_prop_s> p(new inqx_prop_s());
  p->ham = h;
  p->prec = precision;
  p->dt = dt;
  for (int i = 0; i < nblocks; ++i) {
    p->bases.push_back(&bases[i]->b);
    p->nst.push_back(nst[i]);
    p->phi.push_back(phi[i]);
  }
  std::vector<double> occ_all;
  for (int i = 0; i < nblocks; ++i) occ_all.insert(occ_all.end(), occ[i], occ[i] + nst[i]);
  upload(p->occ_dev, occ_all, ctx->c.stream);
  size_t off = 0;
  for (int i = 0; i < nblocks; ++i) {
    p->occ.push_back(p->occ_dev.as<double>() + off);
    off += nst[i];
  }
  p->cap = ring_capacity;
  p->ring.ensure(sizeof(double) * 8 * (size_t)ring_capacity);
  p->counter.ensure(sizeof(int));
  const int zero = 0;
  INQX_HIP(hipMemcpyAsync(p->counter.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  *out = p.release();
  API_END(ctx)
}

int inqx_prop_energy(inqx_prop p) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p, INQX_ERR_ARG, "null propagator");
  const int nb = (int)p->bases.size();
  if (p->prec == INQX_FP64)
    energy_record<double>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(),
                          p->ring.as<double>(), p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, false);
  else
    energy_record<float>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(),
                         p->ring.as<double>(), p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, false);
  API_END(ctx)
}

int inqx_prop_step(inqx_prop p, int nsteps, int use_graph) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && nsteps >= 0, INQX_ERR_ARG, "bad arguments");
  Ctx &cx = ctx->c;
  for (int s = 0; s < nsteps; ++s) {
    if (!use_graph || p->steps_done == 0) {
      // eager step; the first one also creates every workspace and launch descriptor
      if (p->prec == INQX_FP64) prop_step_eager<double>(p);
      else prop_step_eager<float>(p);
      ++p->steps_done;
      continue;
    }
    if (!p->exec) {
      INQX_HIP(hipStreamSynchronize(cx.stream));
      if (!p->stream) INQX_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
      hipStream_t saved = cx.stream;
      const bool prof = cx.prof;
      cx.prof = false;
      cx.stream = p->stream;
      INQX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      set_capture_keep(&p->keep);
      p->capturing = true;
      try {
        if (p->prec == INQX_FP64) prop_step_eager<double>(p);
        else prop_step_eager<float>(p);
      } catch (...) {
        p->capturing = false;
        set_capture_keep(nullptr);
        hipGraph_t g;
        (void)hipStreamEndCapture(p->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        cx.stream = saved;
        cx.prof = prof;
        p->keep.clear();
        throw;
      }
      p->capturing = false;
      set_capture_keep(nullptr);
      INQX_HIP(hipStreamEndCapture(p->stream, &p->graph));
      cx.stream = saved;
      cx.prof = prof;
      INQX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      for (Basis *b : p->bases) ++b->pinned;
      ++p->ham->pinned;
    }
    {
      ProfScope ps(cx, "etrs_step_graph");
      INQX_HIP(hipGraphLaunch(p->exec, cx.stream));
    }
    ++p->steps_done;
  }
  API_END(ctx)
}

int inqx_energy(inqx_ham h, int nblocks, const inqx_basis *bases, const int *nst, int precision, void *const *phi,
                const double *const *occ, double *out) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && nblocks > 0 && bases && nst && phi && occ && out, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  std::vector<Basis *> bb(nblocks);
  std::vector<double> occ_h;
  for (int i = 0; i < nblocks; ++i) {
    INQX_CHECK(bases[i] && bases[i]->b.ham == h, INQX_ERR_ARG, "basis belongs to another Hamiltonian");
    INQX_CHECK(nst[i] > 0 && phi[i] && occ[i], INQX_ERR_ARG, "bad block");
    bb[i] = &bases[i]->b;
    occ_h.insert(occ_h.end(), occ[i], occ[i] + nst[i]);
  }
  h->en_ring.ensure(8 * sizeof(double));
  h->en_counter.ensure(sizeof(int));
  h->en_occ.ensure(sizeof(double) * occ_h.size());
  const int zero = 0;
  INQX_HIP(hipMemcpyAsync(h->en_occ.p, occ_h.data(), sizeof(double) * occ_h.size(), hipMemcpyHostToDevice,
                          ctx->c.stream));
  INQX_HIP(hipMemcpyAsync(h->en_counter.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->c.stream));
  std::vector<const double *> occ_d(nblocks);
  size_t off = 0;
  for (int i = 0; i < nblocks; ++i) {
    occ_d[i] = h->en_occ.as<double>() + off;
    off += nst[i];
  }
  if (precision == INQX_FP64)
    energy_record<double>(h, nblocks, bb.data(), nst, phi, occ_d.data(), h->en_ring.as<double>(), 1,
                          h->en_counter.as<int>(), h->en_desc, h->en_host, false);
  else
    energy_record<float>(h, nblocks, bb.data(), nst, phi, occ_d.data(), h->en_ring.as<double>(), 1,
                         h->en_counter.as<int>(), h->en_desc, h->en_host, false);
  INQX_HIP(hipMemcpyAsync(out, h->en_ring.p, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));  // also keeps occ_h and zero alive until copied
  API_END(ctx)
}

int inqx_prop_records(inqx_prop p, int *count, double *out, int cap) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && count, INQX_ERR_ARG, "null argument");
  int n = 0;
  INQX_HIP(hipMemcpyAsync(&n, p->counter.p, sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  *count = n;
  if (out && cap > 0) {
    const int m = std::min(std::min(n, cap), p->cap);
    std::vector<double> all((size_t)8 * p->cap);
    INQX_HIP(hipMemcpyAsync(all.data(), p->ring.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(ctx->c.stream));
    // the last m records in step order
    for (int k = 0; k < m; ++k) {
      const int step = n - m + k;
      std::memcpy(out + 8 * (size_t)k, all.data() + 8 * (size_t)(step % p->cap), 8 * sizeof(double));
    }
  }
  API_END(ctx)
}

int inqx_prop_destroy(inqx_prop p) {
  if (!p) return INQX_OK;
  if (p->exec) {
    for (Basis *b : p->bases) --b->pinned;
    --p->ham->pinned;
  }
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->stream) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  delete p;
  return INQX_OK;
}

}  // extern "C"
