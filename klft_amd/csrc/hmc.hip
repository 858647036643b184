// HMC for the Wilson gauge action: momentum refresh, fused force + momentum
// update, link exponential update, kinetic / gauge action sums and the
// leapfrog trajectory with FP64 accept/reject.
//
// Semantics follow the reference's HMC stack (HMC.hpp:48-71,
// Integrator.hpp:26-53, GaugeMonomial.hpp:19-37, HamiltonianField.hpp:23-60,
// AdjointGroup.hpp), generalised to one formula for U(1), SU(2), SU(3):
//   P = sum_a p_a T_a, p_a ~ N(0,1);  H_kin = 1/2 sum p_a^2;
//   H_g = -(beta/Nc) sum_x sum_{mu<nu} Re Tr U_munu(x);
//   F_a(x,mu) = (beta/Nc) Im Tr(T_a U_mu(x) Sigma_mu(x));
//   P <- P - dtau F;  U <- exp(i dtau P) U.
// Momenta are stored structure-of-arrays: P[(mu * NGEN + a) * V + s] in the
// link precision, s the sublattice-major device index.
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>
#include <vector>

#include "internal.hpp"

using namespace kd;
namespace ki = klft_internal;

struct klft_hmc_s {
  klft_field_t f = nullptr;
  klft_hmc_params p{};
  void *P = nullptr;       // momenta (R)
  // the leapfrog's two link buffers: a trajectory reads the field's links
  // and alternates between these, then hands the final buffer to the field
  // and keeps the starting one, which a rejected trajectory hands back
  void *Ubuf[2] = {nullptr, nullptr};
  double *Fdbg = nullptr;  // force in host (AoS) order, for klft_hmc_force
  double *stage = nullptr; // host-order FP64 momenta for upload / download
};

template <typename R, int NC, int ND>
__global__ __launch_bounds__(256) void k_refresh(R *P, Geom g, uint64_t seed, uint32_t traj,
                                                  double *partials) {
  __shared__ double lds[8];
  constexpr int NG = Ngen<NC>::value;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;  // device order
  double k = 0.0;
  if (s < g.V) {
    int x[ND];
    dev_coords<ND>(g, s, x);
    const uint32_t site = (uint32_t)lex_index<ND>(g, x);
#pragma unroll
    for (int mu = 0; mu < ND; ++mu) {
      klft_rng::Stream st(seed, klft_rng::TAG_MOMENTA, site, traj, 0, 0, (uint32_t)mu);
      double z[NG + 1];
#pragma unroll
      for (int a = 0; a < NG; a += 2) {
        const double u1 = 1.0 - st.u01(), u2 = st.u01();
        const double r = sqrt(-2.0 * log(u1)), t = 2.0 * KLFT_PI * u2;
        z[a] = r * cos(t);
        z[a + 1] = r * sin(t);
      }
      double n2 = 0.0;
#pragma unroll
      for (int a = 0; a < NG; ++a) {
        const R v = (R)z[a];
        P[((long long)mu * NG + a) * g.V + s] = v;
        n2 += (double)v * (double)v;
      }
      k += 0.5 * n2;
    }
  }
  block_partial<256>(k, lds, partials);
}

// P -= dtau * F with F recomputed from the current links.  Fout (optional):
// write F in host AoS order instead of updating P.  kin (optional): also
// reduce 1/2 sum p^2 of the updated momenta.
template <typename R, int NC, bool COMP, int ND, bool LOWREG>
__global__ __launch_bounds__(256) void k_force(Links<R, NC, COMP> f, R *P, Geom g,
                                                                     R coef, R dtau, double *Fout,
                                                                     double *partials) {
  f.nd = ND;
  __shared__ double lds[8];
  constexpr int NG = Ngen<NC>::value;
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  int x[ND], mu;
  long long s;
  double k = 0.0;
  if (link_site<ND>(g, lb, x, &s, &mu)) {
    const Nbr<ND> nb = make_nbr<ND>(g, x, s);
    Mat<R, NC> S;
    if constexpr (LOWREG)
      S = staple_lowreg<R, NC, COMP, ND>(f, nb, mu);
    else
      S = staple<R, NC, COMP, ND>(f, nb, mu);
    const Mat<R, NC> M = mul(f.load(s, mu), S);
    R F[NG];
    project(M, coef, F);
    if (Fout) {
      const long long site = lex_index<ND>(g, x);
#pragma unroll
      for (int a = 0; a < NG; ++a) Fout[(site * ND + mu) * NG + a] = (double)F[a];
    } else {
      double n2 = 0.0;
#pragma unroll
      for (int a = 0; a < NG; ++a) {
        R *q = P + ((long long)mu * NG + a) * g.V + s;
        const R v = *q - dtau * F[a];
        *q = v;
        n2 += (double)v * (double)v;
      }
      k = 0.5 * n2;
    }
  }
  if (partials) block_partial<256>(k, lds, partials);
}

// One leapfrog step with the momentum and link updates fused and the links
// double-buffered: F from the links in src, P <- P - dtp F, and
// dst(x,mu) = exp(i dtu P(x,mu)) src(x,mu) (FP32 links reunitarised).  The
// staples read src only, so no link is written while another thread may
// still read it; the arithmetic is that of k_force followed by
// k_update_gauge, so the result is the same.  One thread per link as in
// k_force: a block holds one 64-site chunk, wave w taking direction w.
template <typename R, int NC, bool COMP, int ND, bool LOWREG>
__global__ __launch_bounds__(256) void k_md_step(Links<R, NC, COMP> src, Links<R, NC, COMP> dst, R *P, Geom g,
                                                  R coef, R dtp, R dtu) {
  src.nd = ND;
  dst.nd = ND;
  constexpr int NG = Ngen<NC>::value;
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  int x[ND], mu;
  long long s;
  if (!link_site<ND>(g, lb, x, &s, &mu)) return;
  const Nbr<ND> nb = make_nbr<ND>(g, x, s);
  Mat<R, NC> S;
  if constexpr (LOWREG)
    S = staple_lowreg<R, NC, COMP, ND>(src, nb, mu);
  else
    S = staple<R, NC, COMP, ND>(src, nb, mu);
  const Mat<R, NC> U = src.load(s, mu);
  R v[NG];
  project(mul(U, S), coef, v);
#pragma unroll
  for (int a = 0; a < NG; ++a) {
    R *q = P + ((long long)mu * NG + a) * g.V + s;
    const R pn = *q - dtp * v[a];
    *q = pn;
    v[a] = dtu * pn;
  }
  const Mat<R, NC> e = exp_iq(v, (Mat<R, NC> *)nullptr);
  Mat<R, NC> u = mul(e, U);
  if constexpr (sizeof(R) == 4) reunitarize(u);
  dst.store(s, mu, u);
}

// U <- exp(i dtau P) U; FP32 links are reunitarised after the update
template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(256) void k_update_gauge(Links<R, NC, COMP> f, const R *__restrict__ P,
                                                       Geom g, R dtau) {
  f.nd = ND;
  constexpr int NG = Ngen<NC>::value;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= g.V) return;
#pragma unroll 1
  for (int mu = 0; mu < ND; ++mu) {
    R v[NG];
#pragma unroll
    for (int a = 0; a < NG; ++a) v[a] = dtau * P[((long long)mu * NG + a) * g.V + s];
    const Mat<R, NC> e = exp_iq(v, (Mat<R, NC> *)nullptr);
    Mat<R, NC> u = mul(e, f.load(s, mu));
    if constexpr (sizeof(R) == 4) reunitarize(u);
    f.store(s, mu, u);
  }
}

template <typename R, int NG>
__global__ __launch_bounds__(256) void k_kinetic_soa(const R *__restrict__ P, long long V, int nd,
                                                      double *partials, unsigned int *ticket,
                                                      double *out) {
  __shared__ double lds[8];
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  double k = 0.0;
  if (s < V) {
    for (int mu = 0; mu < nd; ++mu) {
      double n2 = 0.0;
#pragma unroll
      for (int a = 0; a < NG; ++a) {
        const double v = (double)P[((long long)mu * NG + a) * V + s];
        n2 += v * v;
      }
      k += 0.5 * n2;
    }
  }
  const double bt = block_sum<256>(k, lds);
  grid_finish<256>(bt, partials, ticket, out, lds);
}

template <typename R, int NG, int ND>
__global__ __launch_bounds__(256) void k_mom_io(R *P, Geom g, double *host_img, int upload) {
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;  // lexicographic
  if (s >= g.V) return;
  int x[ND];
  lex_coords<ND>(g, s, x);
  const long long d = dev_index<ND>(g, x);
  for (int mu = 0; mu < ND; ++mu)
#pragma unroll
    for (int a = 0; a < NG; ++a) {
      R *q = P + ((long long)mu * NG + a) * g.V + d;
      double *h = host_img + ((size_t)s * ND + mu) * NG + a;
      if (upload)
        *q = (R)*h;
      else
        *h = (double)*q;
    }
}

namespace {

int ngen(int nc) { return nc == 1 ? 1 : (nc == 2 ? 3 : 8); }

int check_hmc(klft_hmc_t h) {
  if (!h || !h->P) return ki::set_error(KLFT_E_ARG, "null or destroyed HMC handle");
  return ki::check_field(h->f);
}

int enqueue_refresh(klft_hmc_t h, uint64_t traj, int slot) {
  klft_field_t f = h->f;
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    (void)COMP;
    const unsigned nb = ki::blocks_for(f->g.V, 256);
    if ((long long)nb > f->max_partials) return ki::set_error(KLFT_E_ARG, "lattice too large for reduction");
    hipLaunchKernelGGL((k_refresh<R, NC, ND>), dim3(nb), dim3(256), 0, ki::stream(), (R *)h->P, f->g,
                       h->p.seed, (uint32_t)traj, f->partials);
    KLFT_HIP(hipGetLastError());
    return ki::enqueue_sum_blocks(f->partials, nb, f->result + slot);
  });
}

// variants: 0 one thread per link, one-term-at-a-time staple; 4 the same with
// all staple loads in flight at once (full staple).  1-3 (occupancy targets,
// one thread per site) never won and were removed.
constexpr int kForceVariants = 5;

int launch_force(klft_hmc_t h, double dtau, double *Fout, int kin_slot, int variant) {
  klft_field_t f = h->f;
  if (variant != 0 && variant != 4) return ki::set_error(KLFT_E_ARG, "force variant %d is not built", variant);
  // with a kinetic slot, every block writes a partial; k_sum_blocks adds them
  double *parts = kin_slot >= 0 ? f->partials : nullptr;
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    const auto links = ki::links_of<R, NC, COMP>(f);
    const R coef = (R)(h->p.beta / (double)NC);
    const long long nb = link_blocks<ND>(f->g);
    if (nb > f->max_partials) return ki::set_error(KLFT_E_ARG, "lattice too large for reduction");
    {
      ki::ProfScope prof("force", "derivative+update_momentum");
      if (variant == 4)
        hipLaunchKernelGGL((k_force<R, NC, COMP, ND, false>), dim3((unsigned)nb), dim3(64 * ND), 0, ki::stream(),
                           links, (R *)h->P, f->g, coef, (R)dtau, Fout, parts);
      else
        hipLaunchKernelGGL((k_force<R, NC, COMP, ND, true>), dim3((unsigned)nb), dim3(64 * ND), 0, ki::stream(),
                           links, (R *)h->P, f->g, coef, (R)dtau, Fout, parts);
      KLFT_HIP(hipGetLastError());
    }
    if (parts) return ki::enqueue_sum_blocks(parts, nb, f->result + kin_slot);
    return KLFT_OK;
  });
}

// the tuning launcher uses dtau = 0, which leaves the momenta unchanged
int enqueue_force(klft_hmc_t h, double dtau, double *Fout, int kin_slot) {
  // untuned default, the fastest at 32^4: the full staple for FP64 links
  const int dflt = h->f->storage == KLFT_STORAGE_F64 ? 4 : 0;
  const int v = ki::pick_variant("force", h->f, kForceVariants, dflt,
                                 [&](int var) { return launch_force(h, 0.0, nullptr, -1, var); });
  return launch_force(h, dtau, Fout, kin_slot, v);
}

int enqueue_gauge(klft_hmc_t h, double dtau) {
  klft_field_t f = h->f;
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    ki::ProfScope prof("update_gauge", "update_gauge");
    hipLaunchKernelGGL((k_update_gauge<R, NC, COMP, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256),
                       0, ki::stream(), ki::links_of<R, NC, COMP>(f), (const R *)h->P, f->g, (R)dtau);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}

int enqueue_kinetic(klft_hmc_t h, int slot) {
  klft_field_t f = h->f;
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    (void)COMP;
    (void)ND;
    if ((long long)ki::blocks_for(f->g.V, 256) > f->max_partials)
      return ki::set_error(KLFT_E_ARG, "lattice too large for reduction");
    hipLaunchKernelGGL((k_kinetic_soa<R, Ngen<NC>::value>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256),
                       0, ki::stream(), (const R *)h->P, f->g.V, f->nd, f->partials, f->ticket,
                       f->result + slot);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}

// variants: 0 one-term-at-a-time staple, 1 full staple (all loads in flight)
constexpr int kMdVariants = 2;

int launch_md_step(klft_hmc_t h, void *src, void *dst, double dtp, double dtu, int variant) {
  klft_field_t f = h->f;
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    auto ls = ki::links_of<R, NC, COMP>(f), ld = ls;
    ls.p = reinterpret_cast<cx<R> *>(src);
    ld.p = reinterpret_cast<cx<R> *>(dst);
    const unsigned nb = (unsigned)link_blocks<ND>(f->g);
    const R coef = (R)(h->p.beta / (double)NC);
    ki::ProfScope prof("md_step", "derivative+update_momentum+update_gauge");
    if (variant == 1)
      hipLaunchKernelGGL((k_md_step<R, NC, COMP, ND, false>), dim3(nb), dim3(64 * ND), 0, ki::stream(), ls, ld,
                         (R *)h->P, f->g, coef, (R)dtp, (R)dtu);
    else
      hipLaunchKernelGGL((k_md_step<R, NC, COMP, ND, true>), dim3(nb), dim3(64 * ND), 0, ki::stream(), ls, ld,
                         (R *)h->P, f->g, coef, (R)dtp, (R)dtu);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}

// Leapfrog::integrate (Integrator.hpp:26-53): dtau = tau/n; half momentum
// step, gauge step, n x (momentum, gauge), half momentum step, restoreGauge.
// Each momentum step and the gauge step after it are one fused launch that
// writes the new links to the other buffer: field -> Ubuf[0] -> Ubuf[1] ->
// Ubuf[0] ...  The final buffer is then swapped into the field, and the
// starting links stay in the handle (slot `*start_slot`).
int enqueue_leapfrog(klft_hmc_t h, int kin_slot, int *start_slot) {
  const double dtau = h->p.tau / (double)h->p.nsteps;
  klft_field_t f = h->f;
  // untuned default: the staple form of the force kernel's default
  const int v = ki::pick_variant("md_step", f, kMdVariants, f->storage == KLFT_STORAGE_F64 ? 1 : 0, [&](int var) {
    return launch_md_step(h, f->links, h->Ubuf[0], 0.0, 0.0, var);  // dtp = 0: P unchanged
  });
  void *cur = f->links;
  int nxt = 0, rc = KLFT_OK;
  for (int i = 0; i <= h->p.nsteps && !rc; ++i) {
    rc = launch_md_step(h, cur, h->Ubuf[nxt], i == 0 ? dtau / 2.0 : dtau, dtau, v);
    cur = h->Ubuf[nxt];
    nxt ^= 1;
  }
  if (rc) return rc;
  // the field takes the final links; the buffer it held takes their slot
  const int fin = nxt ^ 1;
  h->Ubuf[fin] = f->links;
  f->links = cur;
  if (start_slot) *start_slot = fin;
  rc = enqueue_force(h, dtau / 2.0, nullptr, kin_slot);
  if (!rc) rc = ki::enqueue_reunit(h->f);
  return rc;
}

int momenta_io(klft_hmc_t h, double *host, int upload) {
  klft_field_t f = h->f;
  const size_t n = (size_t)f->g.V * f->nd * ngen(f->nc) * sizeof(double);
  if (!h->stage) KLFT_HIP(hipMalloc(&h->stage, n));  // kept with the handle
  double *stage = h->stage;
  hipError_t e = hipSuccess;
  if (upload) e = hipMemcpyAsync(stage, host, n, hipMemcpyHostToDevice, ki::stream());
  int rc = KLFT_OK;
  if (e == hipSuccess)
    rc = ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
      KLFT_UNPACK(rt, nct, ct, ndt);
      (void)COMP;
      hipLaunchKernelGGL((k_mom_io<R, Ngen<NC>::value, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256), 0,
                         ki::stream(), (R *)h->P, f->g, stage, upload);
      KLFT_HIP(hipGetLastError());
      return KLFT_OK;
    });
  if (e == hipSuccess && !rc && !upload) e = hipMemcpyAsync(host, stage, n, hipMemcpyDeviceToHost, ki::stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
  if (e != hipSuccess) return ki::hip_fail(e, "momenta transfer");
  return rc;
}

}  // namespace

extern "C" int klft_hmc_create(klft_field_t f, const klft_hmc_params *p, klft_hmc_t *out) {
  if (!out || !p) return ki::set_error(KLFT_E_ARG, "null argument");
  *out = nullptr;
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (p->nsteps < 1) return ki::set_error(KLFT_E_ARG, "n_steps must be >= 1");
  if (!(p->tau > 0.0)) return ki::set_error(KLFT_E_ARG, "tau must be > 0");
  klft_hmc_s *h = new klft_hmc_s();
  h->f = f;
  h->p = *p;
  const size_t rbytes = f->storage == KLFT_STORAGE_F64 ? 8 : 4;
  const size_t pbytes = (size_t)f->g.V * f->nd * ngen(f->nc) * rbytes;
  hipError_t e = hipMalloc(&h->P, pbytes);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipMalloc(&h->Ubuf[i], f->link_bytes);
  // cleared on the (non-blocking) library stream, where the kernels that
  // read P run, and waited for
  if (e == hipSuccess) e = hipMemsetAsync(h->P, 0, pbytes, ki::stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
  if (e != hipSuccess) {
    if (h->P) (void)hipFree(h->P);
    for (void *b : h->Ubuf)
      if (b) (void)hipFree(b);
    delete h;
    return ki::set_error(KLFT_E_NOMEM, "HMC workspace allocation failed");
  }
  *out = h;
  return KLFT_OK;
}

extern "C" int klft_hmc_destroy(klft_hmc_t h) {
  if (!h) return KLFT_OK;
  (void)ki::sync_stream();
  if (h->P) (void)hipFree(h->P);
  for (void *b : h->Ubuf)
    if (b) (void)hipFree(b);
  if (h->Fdbg) (void)hipFree(h->Fdbg);
  if (h->stage) (void)hipFree(h->stage);
  h->P = nullptr;
  delete h;
  return KLFT_OK;
}

extern "C" int klft_hmc_refresh_momenta(klft_hmc_t h, uint64_t traj) {
  int rc = check_hmc(h);
  if (!rc) rc = enqueue_refresh(h, traj, 1);
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_hmc_momenta_upload(klft_hmc_t h, const double *host) {
  int rc = check_hmc(h);
  if (rc) return rc;
  if (!host) return ki::set_error(KLFT_E_ARG, "null host buffer");
  return momenta_io(h, const_cast<double *>(host), 1);
}

extern "C" int klft_hmc_momenta_download(klft_hmc_t h, double *host) {
  int rc = check_hmc(h);
  if (rc) return rc;
  if (!host) return ki::set_error(KLFT_E_ARG, "null host buffer");
  return momenta_io(h, host, 0);
}

extern "C" int klft_hmc_force(klft_hmc_t h, double *host_force) {
  int rc = check_hmc(h);
  if (rc) return rc;
  if (!host_force) return ki::set_error(KLFT_E_ARG, "null host buffer");
  klft_field_t f = h->f;
  const size_t n = (size_t)f->g.V * f->nd * ngen(f->nc);
  if (!h->Fdbg) KLFT_HIP(hipMalloc(&h->Fdbg, n * sizeof(double)));
  rc = enqueue_force(h, 0.0, h->Fdbg, -1);
  if (rc) return rc;
  KLFT_HIP(hipMemcpyAsync(host_force, h->Fdbg, n * sizeof(double), hipMemcpyDeviceToHost, ki::stream()));
  return ki::sync_stream();
}

extern "C" int klft_hmc_update_momentum(klft_hmc_t h, double dtau) {
  int rc = check_hmc(h);
  if (!rc) rc = enqueue_force(h, dtau, nullptr, -1);
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_hmc_update_gauge(klft_hmc_t h, double dtau) {
  int rc = check_hmc(h);
  if (!rc) rc = enqueue_gauge(h, dtau);
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_hmc_kinetic(klft_hmc_t h, double *out) {
  int rc = check_hmc(h);
  if (rc) return rc;
  if (!out) return ki::set_error(KLFT_E_ARG, "null output");
  rc = enqueue_kinetic(h, 1);
  if (rc) return rc;
  return ki::read_results(h->f, 1, 1, out);
}

extern "C" int klft_gauge_action(klft_field_t f, double beta, double *out) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!out) return ki::set_error(KLFT_E_ARG, "null output");
  rc = ki::plaquette_raw(f, 2, nullptr);
  double raw = 0.0;
  if (!rc) rc = ki::read_results(f, 2, 1, &raw);
  if (rc) return rc;
  *out = -(beta / (double)f->nc) * raw;  // GaugeMonomial.hpp:19-25
  return KLFT_OK;
}

extern "C" int klft_hmc_integrate(klft_hmc_t h) {
  int rc = check_hmc(h);
  if (!rc) rc = enqueue_leapfrog(h, -1, nullptr);
  if (rc) return rc;
  return ki::sync_stream();
}

// HMC::hmc_step (HMC.hpp:48-71) with the monomials in the order of
// lib/HMC_SU2.cpp:46-47 (kinetic, then gauge).
extern "C" int klft_hmc_trajectory(klft_hmc_t h, uint64_t traj, double *dH, int *accepted) {
  int rc = check_hmc(h);
  if (rc) return rc;
  klft_field_t f = h->f;
  rc = enqueue_refresh(h, traj, 4);  // K_old
  if (rc) return rc;
  rc = ki::plaquette_raw(f, 5, nullptr);  // G_old (raw)
  int start = -1;                         // handle slot of the starting links
  if (!rc) rc = enqueue_leapfrog(h, 6, &start);  // K_new fused into the last half step
  if (!rc) rc = ki::plaquette_raw(f, 7, nullptr);
  double r[4];
  if (!rc) rc = ki::read_results(f, 4, 4, r);
  if (rc) return rc;
  const double c = -(h->p.beta / (double)f->nc);
  const double K0 = r[0], G0 = c * r[1], K1 = r[2], G1 = c * r[3];
  double d = 0.0;
  d += K1 - K0;
  d += G1 - G0;
  int acc = 1;
  if (d > 0.0) {
    klft_rng::Stream st(h->p.seed, klft_rng::TAG_ACCEPT, 0, (uint32_t)traj, 0, 0, 0);
    if (st.u01() > std::exp(-d)) acc = 0;
  }
  if (!acc) std::swap(f->links, h->Ubuf[start]);  // the starting links, untouched
  if (dH) *dH = d;
  if (accepted) *accepted = acc;
  return KLFT_OK;
}
