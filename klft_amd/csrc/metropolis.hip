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
// costs one trace instead of the reference's three matrix products; M is
// rebuilt only after an accepted hit.
#include <hip/hip_runtime.h>

#include <vector>

#include "internal.hpp"

using namespace kd;
namespace ki = klft_internal;

constexpr int kMetroBlock = 256;

template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(kMetroBlock) void k_metropolis_pass(
    Links<R, NC, COMP> f, Geom g, int pass, int nhits, R beta_nc, R delta, uint64_t seed,
    uint32_t sweep, uint32_t chain, unsigned long long *nacc, int *visit) {
  __shared__ unsigned long long red[kMetroBlock / 64];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  const long long t = (long long)lb * kMetroBlock + threadIdx.x;
  unsigned long long n = 0;
  if (t < g.Vh) {
    int x[ND];
    long long q = t;
#pragma unroll
    for (int d = ND - 1; d >= 0; --d) {
      const int y = (int)(q % g.H[d]);
      q /= g.H[d];
      x[d] = 2 * y + ((pass >> (ND - 1 - d)) & 1);  // index_odd_even
    }
    const long long s = (long long)pass * g.Vh + t;
    const uint32_t site = (uint32_t)lex_index<ND>(g, x);
    if (visit) visit[site] = pass;
#pragma unroll 1
    for (int mu = 0; mu < ND; ++mu) {
      const Mat<R, NC> S = staple<R, NC, COMP, ND>(f, g, x, mu);
      Mat<R, NC> U = f.load(s, mu);
      Mat<R, NC> M = mul(S, U);
      R trM = retrace(M);
#pragma unroll 1
      for (int hit = 0; hit < nhits; ++hit) {
        klft_rng::Stream st(seed, klft_rng::TAG_METROPOLIS, site, sweep, (uint32_t)hit, chain,
                            (uint32_t)mu);
        Mat<R, NC> r;
        rand_sun(st, delta, r);
        const R dS = -beta_nc * (retr_mul(r, M) - trM);
        bool accept = dS < R(0);
        if (!accept) accept = unif<R>(st) < exp(-dS);
        if (accept) {
          U = mul(U, r);
          reunitarize(U);
          M = mul(S, U);
          trM = retrace(M);
          ++n;
        }
      }
      if (nhits > 0) f.store(s, mu, U);
    }
  }
  n = wave_sum_u64(n);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
#pragma unroll
    for (int i = 0; i < kMetroBlock / 64; ++i) b += red[i];
    if (b) atomicAdd(nacc, b);
  }
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

namespace klft_internal {

// enqueue one sweep (no synchronisation); acceptance counted into *counter
int enqueue_sweep(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep, uint32_t chain,
                  unsigned long long *counter, int *visit) {
  return dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    const unsigned nb = blocks_for(f->g.Vh, kMetroBlock);
    const R beta_nc = (R)(p->beta / (double)NC);
    for (int pass = 0; pass < (1 << ND); ++pass) {
      hipLaunchKernelGGL((k_metropolis_pass<R, NC, COMP, ND>), dim3(nb), dim3(kMetroBlock), 0, stream(),
                         links_of<R, NC, COMP>(f), f->g, pass, p->nhits, beta_nc, (R)p->delta, p->seed,
                         (uint32_t)sweep, chain, counter, visit);
    }
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
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
    KLFT_HIP(hipMalloc(&cnt, sizeof(unsigned long long) * n));
    KLFT_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * n, ki::stream()));
  } else {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ki::stream(), f->counter, 1);
  }
  for (int i = 0; i < n && !rc; ++i) rc = ki::enqueue_sweep(f, p, sweep0 + i, 0, cnt ? cnt + i : f->counter, nullptr);
  if (!rc && cnt) {
    std::vector<unsigned long long> h(n);
    hipError_t e = hipMemcpyAsync(h.data(), cnt, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost,
                                  ki::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
    if (e != hipSuccess) rc = ki::hip_fail(e, "acceptance counters");
    for (int i = 0; i < n && !rc; ++i) acc_rates[i] = acceptance(f, p->nhits, h[i]);
  }
  if (cnt) (void)hipFree(cnt);
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_metropolis_schedule(klft_field_t f, int32_t *visit) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!visit) return ki::set_error(KLFT_E_ARG, "null visit buffer");
  int *dv = nullptr;
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
