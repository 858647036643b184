// Checkerboarded Metropolis sweep (Metropolis.hpp:44-178).
//
// A sweep is 2^nd sequential sublattice passes; pass i updates the sites with
// parity bits b_d = (i >> (nd-1-d)) & 1 (oddeven_array, IndexHelper.hpp:84-93),
// which are exactly the contiguous device range [i*Vh, (i+1)*Vh) of the
// sublattice-major layout.  One thread owns one site and, as in
// MetropolisGaugeField::operator() (Metropolis.hpp:72-113), loops over mu
// sequentially: staple once, then nHits proposals r = randSUN(delta),
// dS = -(beta/Nc) (Re Tr(U r S) - Re Tr(U S)), accept if dS < 0 or
// u < exp(-dS) (u drawn only when dS >= 0), on accept U = restoreSUN(U r).
// The trace is evaluated as Re Tr(r M) with M = S U (Tr is cyclic), so a hit
// costs one trace instead of the reference's three matrix products; after an
// accepted hit M becomes M r, so the staple itself is dead once M is formed
// (M r and S restoreSUN(U r) differ by rounding only).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "internal.hpp"

using namespace kd;
namespace ki = klft_internal;

constexpr int kMetroVariants = 14;

namespace klft_internal {
int launch_passes_f64(klft_field_t f, void *links, int variant, int p0, int p1, const klft_metropolis_params *p,
                      uint64_t sweep, int nsweeps, uint32_t chain, unsigned long long *counter,
                      int counter_stride, int *visit);
int launch_passes_f32(klft_field_t f, void *links, int variant, int p0, int p1, const klft_metropolis_params *p,
                      uint64_t sweep, int nsweeps, uint32_t chain, unsigned long long *counter,
                      int counter_stride, int *visit);
int launch_passes_f32rows(klft_field_t f, void *links, int variant, int p0, int p1,
                          const klft_metropolis_params *p, uint64_t sweep, int nsweeps, uint32_t chain,
                          unsigned long long *counter, int counter_stride, int *visit);
}  // namespace klft_internal

__global__ void k_zero_u64(unsigned long long *p, int n) {
  for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x))
    p[i] = 0ull;
}

namespace klft_internal {

// launch passes [p0, p1) of sweeps sweep .. sweep + nsweeps - 1 on the link
// array `links`, sweep i counting its accepted hits into
// counter[i * counter_stride]; the kernels are instantiated per storage in
// metropolis_{f64,f32,f32rows}.hip (the variant table is in
// metro_kernels.hpp).  U(1) has no compressed form, so its two-row storage
// runs the FP32 kernels.
static int launch_passes(klft_field_t f, void *links, int variant, int p0, int p1,
                         const klft_metropolis_params *p, uint64_t sweep, int nsweeps, uint32_t chain,
                         unsigned long long *counter, int counter_stride, int *visit) {
  switch (f->storage) {
    case KLFT_STORAGE_F64:
      return launch_passes_f64(f, links, variant, p0, p1, p, sweep, nsweeps, chain, counter, counter_stride, visit);
    case KLFT_STORAGE_F32:
      return launch_passes_f32(f, links, variant, p0, p1, p, sweep, nsweeps, chain, counter, counter_stride, visit);
    case KLFT_STORAGE_F32_ROWS:
      if (f->nc == 1)
        return launch_passes_f32(f, links, variant, p0, p1, p, sweep, nsweeps, chain, counter, counter_stride,
                                 visit);
      return launch_passes_f32rows(f, links, variant, p0, p1, p, sweep, nsweeps, chain, counter, counter_stride,
                                   visit);
  }
  return set_error(KLFT_E_ARG, "storage=%d unsupported", f->storage);
}

// sweep variant for this field; tuning times one sweep of each variant on a
// scratch copy of the links, so the live field is never updated by it (a
// sweep, not a pass: the site-wave variants pair independent passes when
// both fit on the GPU at once, which a single pass would not show)
int metropolis_variant(klft_field_t f, const klft_metropolis_params *p, bool timed) {
  void *scratch = nullptr;
  // variants (metro_kernels.hpp): 0 one lane per site; 4 one lane with the
  // low-register staple at >= 2 waves/SIMD; 7 one lane building the next
  // direction's staple during the hits; 11 one lane with the site's own links
  // kept in registers; 12 site waves fed the next staple by staple waves; 13
  // the same with each direction's upper and lower staple terms back to back.
  // The other numbers below kMetroVariants are not built (the tuner skips
  // them).
  // Untuned defaults: site waves fed by staple waves, the fastest at 16^4,
  // 24^4 and 32^4 for every group (nHits = 10), with the interleaved term
  // order for FP64 links.  For FP32 links the interleaved order's six loads
  // in flight (196 VGPRs) leave room for one block per CU, the
  // upper-terms-first order (128 VGPRs) for two: the interleaved order wins
  // while two paired passes fill at most one block per CU, and loses once
  // they need two (32^4: two-block pairs) or blocks queue for a CU (48^4:
  // 398 against 308 us per pass).
  const bool paired_one_round = 2 * ((f->g.Vh + 255) / 256) <= (long long)cu_count();
  const int dflt = (f->storage == KLFT_STORAGE_F64 || paired_one_round) ? 13 : 12;
  // the symmetric SU(3) proposal runs in the site-wave kernels only
  if (f->proposal == KLFT_PROPOSAL_SYMMETRIC && f->nc == 3) return dflt + 2;
  // KLFT_VARIANT="metropolis=v" forces a variant for every hit count
  const int fv = forced_variant("metropolis");
  if (fv >= 0 && fv < kMetroVariants) return fv;
  // the tuning key carries nHits: the best variant depends on the hit count
  const std::string kname = "metropolis_nhits_" + std::to_string(p->nhits);
  if (!timed) return chosen_variant(kname.c_str(), f, kMetroVariants, dflt);
  const int v = pick_variant(kname.c_str(), f, kMetroVariants, dflt, [&](int var) -> int {
    if (!scratch) {
      KLFT_HIP(hipMalloc(&scratch, f->link_bytes));
      KLFT_HIP(hipMemcpyAsync(scratch, f->links, f->link_bytes, hipMemcpyDeviceToDevice, stream()));
    }
    return launch_passes(f, scratch, var, 0, 1 << f->nd, p, 0, 1, 0, f->counter + 7, 0, nullptr);
  });
  if (scratch) {
    (void)hipStreamSynchronize(stream());
    (void)hipFree(scratch);
  }
  return v;
}

// enqueue one sweep (no synchronisation); acceptance counted into *counter
int enqueue_sweep(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep, uint32_t chain,
                  unsigned long long *counter, int *visit) {
  // the schedule dump (visit != nullptr, nHits = 0) is never timed: its
  // staple-only passes are not the workload a tuned variant is chosen for
  const int v = metropolis_variant(f, p, visit == nullptr);
  return launch_passes(f, f->links, v, 0, 1 << f->nd, p, sweep, 1, chain, counter, 0, visit);
}

// enqueue sweeps sweep0 .. sweep0 + n - 1 (no synchronisation), passes of
// consecutive sweeps sharing launches where they are independent; sweep i
// counts into counter[i * counter_stride]
static int enqueue_run(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep0, int n,
                       unsigned long long *counter, int counter_stride) {
  if (n <= 0) return KLFT_OK;
  const int v = metropolis_variant(f, p, true);
  return launch_passes(f, f->links, v, 0, 1 << f->nd, p, sweep0, n, 0, counter, counter_stride, nullptr);
}

}  // namespace klft_internal

static int check_params(klft_field_t f, const klft_metropolis_params *p) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!p) return ki::set_error(KLFT_E_ARG, "null Metropolis parameters");
  if (p->nhits < 0) return ki::set_error(KLFT_E_ARG, "nHits must be >= 0");
  return KLFT_OK;
}

static double acceptance(klft_field_t f, int nhits, unsigned long long n) {
  // Metropolis.hpp:169-175
  double norm = 1.0;
  for (int d = 0; d < f->nd; ++d) norm *= (double)f->dims[d];
  norm *= (double)f->nd;
  norm *= (double)nhits;
  return (double)n / norm;
}

extern "C" int klft_metropolis_sweep(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep,
                                     double *acc_rate, int64_t *naccepted) {
  int rc = check_params(f, p);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ki::stream(), f->counter, 1);
  rc = ki::enqueue_sweep(f, p, sweep, 0, f->counter, nullptr);
  if (rc) return rc;
  unsigned long long n = 0;
  KLFT_HIP(hipMemcpyAsync(&n, f->counter, sizeof n, hipMemcpyDeviceToHost, ki::stream()));
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  if (naccepted) *naccepted = (int64_t)n;
  if (acc_rate) *acc_rate = acceptance(f, p->nhits, n);
  return KLFT_OK;
}

extern "C" int klft_metropolis_run(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep0,
                                   int n, double *acc_rates) {
  int rc = check_params(f, p);
  if (rc) return rc;
  if (n < 0) return ki::set_error(KLFT_E_ARG, "negative sweep count");
  unsigned long long *cnt = nullptr;
  if (acc_rates && n > 0) {
    // one counter per sweep, in a buffer kept with the field (grown only when
    // a longer run asks for more) and cleared by a kernel
    if (n > f->run_capacity) {
      if (f->run_counters) {
        KLFT_HIP(hipStreamSynchronize(ki::stream()));
        (void)hipFree(f->run_counters);
        f->run_counters = nullptr;
        f->run_capacity = 0;
      }
      const int cap = n > 1024 ? n : 1024;
      KLFT_HIP(hipMalloc(&f->run_counters, sizeof(unsigned long long) * cap));
      f->run_capacity = cap;
    }
    cnt = f->run_counters;
    hipLaunchKernelGGL(k_zero_u64, dim3(ki::blocks_for(n, 256)), dim3(256), 0, ki::stream(), cnt, n);
  } else {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ki::stream(), f->counter, 1);
  }
  rc = ki::enqueue_run(f, p, sweep0, n, cnt ? cnt : f->counter, cnt ? 1 : 0);
  if (!rc && cnt) {
    std::vector<unsigned long long> h(n);
    hipError_t e = hipMemcpyAsync(h.data(), cnt, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost,
                                  ki::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
    if (e != hipSuccess) rc = ki::hip_fail(e, "acceptance counters");
    for (int i = 0; i < n && !rc; ++i) acc_rates[i] = acceptance(f, p->nhits, h[i]);
  }
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_sweep_launches(int nd, int32_t *groups, int *nlaunches) {
  if (nd < 2 || nd > 4) return ki::set_error(KLFT_E_ARG, "nd=%d unsupported (2..4)", nd);
  if (!groups || !nlaunches) return ki::set_error(KLFT_E_ARG, "null argument");
  const ki::PassGroups gr = ki::pass_groups(nd);
  for (int i = 0; i < gr.n; ++i) {
    groups[2 * i] = gr.p[i][0];
    groups[2 * i + 1] = gr.p[i][1];
  }
  *nlaunches = gr.n;
  return KLFT_OK;
}

extern "C" int klft_run_launches(int nd, int nsweeps, int32_t *launches, int capacity, int *nlaunches) {
  if (nd < 2 || nd > 4) return ki::set_error(KLFT_E_ARG, "nd=%d unsupported (2..4)", nd);
  if (nsweeps < 0) return ki::set_error(KLFT_E_ARG, "negative sweep count");
  if (!nlaunches || (capacity > 0 && !launches)) return ki::set_error(KLFT_E_ARG, "null argument");
  // the same walk over the periodic schedule that the sweep launches follow
  long long n = 0;
  ki::for_each_run_launch(nd, nsweeps, [&](const ki::RunLaunch &L) {
    if (n < capacity) {
      launches[4 * n] = L.pass[0];
      launches[4 * n + 1] = L.sweep[0];
      launches[4 * n + 2] = L.pass[1];
      launches[4 * n + 3] = L.pass[1] >= 0 ? L.sweep[1] : -1;
    }
    ++n;
  });
  *nlaunches = (int)n;
  if (n > capacity) return ki::set_error(KLFT_E_ARG, "%lld launches do not fit a capacity of %d", n, capacity);
  return KLFT_OK;
}

extern "C" int klft_metropolis_schedule(klft_field_t f, int32_t *visit) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!visit) return ki::set_error(KLFT_E_ARG, "null visit buffer");
  int *dv = nullptr;  // diagnostic dump, not a hot path
  KLFT_HIP(hipMalloc(&dv, sizeof(int) * f->g.V));
  KLFT_HIP(hipMemsetAsync(dv, 0xff, sizeof(int) * f->g.V, ki::stream()));
  // nhits = 0: the real pass kernels run their thread-to-site map and record
  // the pass index without changing any link
  klft_metropolis_params p{0, 1.0, 0.1, 0};
  rc = ki::enqueue_sweep(f, &p, 0, 0, f->counter, dv);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(visit, dv, sizeof(int) * f->g.V, hipMemcpyDeviceToHost, ki::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
    if (e != hipSuccess) rc = ki::hip_fail(e, "schedule dump");
  }
  (void)hipFree(dv);
  return rc;
}

extern "C" int klft_field_set_proposal(klft_field_t f, int proposal) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (proposal != KLFT_PROPOSAL_REFERENCE && proposal != KLFT_PROPOSAL_SYMMETRIC)
    return ki::set_error(KLFT_E_ARG, "unknown proposal law %d", proposal);
  f->proposal = proposal;
  return KLFT_OK;
}
