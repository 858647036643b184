// Batched Metropolis for many independent 4D chains and PTBC replica sets.
//
// One launch per sublattice pass (or per pair of independent passes, see
// pass_groups) updates every chain: the grid is a work table of 256-site
// blocks, block b belonging to chain block_chain[b], so chains of uneven
// volume and of different groups (U(1), SU(2), SU(3): the group is uniform
// per block) and different beta share one dispatch.  Each chain draws from
// its own keyed streams (chain index in the Philox counter).  A sweep of the
// whole batch -- counter reset, the 16 passes in 16 or 11 launches and, for
// replicas, the defect action sums and the swap kernel -- is captured once in
// a HIP graph and replayed; the sweep / step index lives in device memory and
// the graph's last node advances it.  A run of several sweeps of chains that
// pair their passes replays three more graphs instead (run_parts), so that
// passes of consecutive sweeps share launches too: 8 per sweep.
//
// PTBC (PTBCGaugeField.hpp:30-279): replica r's links on the defect (mu = 0,
// x0 = 0, x_k in [p_k, p_k + Ld)) are scaled by its coupling c.  Swaps of
// neighbouring couplings (c_k, c_k+1) between configurations a, b use
//   dS = -(beta/Nc) [(c_k - c_k+1)(A1_b - A1_a) + (c_k^2 - c_k+1^2)(A2_b - A2_a)],
// with A1 / A2 the FP64 sums of Re Tr U_P over plaquettes through one / two
// defect links (unscaled): only those plaquettes depend on c.  The labels
// (which configuration carries which coupling) are exchanged, never links.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "metro.hpp"
#include "metro_kernels.hpp"

using namespace kd;
namespace ki = klft_internal;

namespace {

struct ChainDev {
  void *links;
  Geom g;
  int nc;
  double beta;
  int first_block;   // pass grid (positions of one sublattice)
  int first_pblock;  // plaquette grid (all sites)
};

constexpr int kB = 256;

template <typename R, bool COMP, bool DEFECT>
__global__ __launch_bounds__(kB) void k_batch_pass(const ChainDev *__restrict__ chains,
                                                   const int *__restrict__ block_chain, int pass, int nhits,
                                                   R delta, uint64_t seed, const uint32_t *__restrict__ d_sweep,
                                                   unsigned long long *counters, Defect d,
                                                   const double *__restrict__ cval) {
  __shared__ unsigned long long red[kB / 64];
  const int c = block_chain[blockIdx.x];
  const ChainDev ch = chains[c];
  const long long t = (long long)(blockIdx.x - ch.first_block) * kB + threadIdx.x;
  const uint32_t sweep = *d_sweep;
  unsigned long long n = 0;
  if (t < ch.g.Vh) {
    const R cc = DEFECT ? (R)cval[c] : R(1);
    switch (ch.nc) {
      case 1: {
        Links<R, 1, false> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
        n = metro_site<R, 1, false, 4, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 1.0), delta, seed, sweep,
                                               (uint32_t)c, d, cc, nullptr);
        break;
      }
      case 2: {
        Links<R, 2, COMP> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
        n = metro_site<R, 2, COMP, 4, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 2.0), delta, seed, sweep,
                                              (uint32_t)c, d, cc, nullptr);
        break;
      }
      default: {
        Links<R, 3, COMP> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
        n = metro_site<R, 3, COMP, 4, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 3.0), delta, seed, sweep,
                                              (uint32_t)c, d, cc, nullptr);
        break;
      }
    }
  }
  n = wave_sum_u64(n);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
#pragma unroll
    for (int i = 0; i < kB / 64; ++i) b += red[i];
    if (b) atomicAdd(&counters[c], b);
  }
}

// The same pass with site waves fed by staple waves (klft_metro::ws_pass):
// block b of the work table holds the 256 positions of chain block_chain[b]
// that k_batch_pass gives it, in 8 waves.  The group is uniform per block, so
// all eight waves meet the same barriers.
// FP32 chains and replicas with the upper-terms-first order fit two blocks
// per CU (128 VGPRs: four waves per SIMD); the register allocator is asked
// for that (the defect kernels then spill 12 / 24 registers, FP32 two-row /
// full, and still gain: C5 FP32 917 -> 956 steps/s, two-row 961 -> 1122
// with their passes paired)
template <typename R, bool COMP, bool DEFECT, bool ILV = false>
__global__ __launch_bounds__(klft_metro::kWsBlock, (sizeof(R) == 4 && !ILV) ? 4 : 2) void k_batch_ws(
    const ChainDev *__restrict__ chains, const int *__restrict__ block_chain, int pass, int pass2, int ds, int ds2,
    int nhits, R delta, uint64_t seed, const uint32_t *__restrict__ d_sweep, unsigned long long *counters, Defect d,
    const double *__restrict__ cval) {
  __shared__ cx<R> stage[4 * 4 * 9 * 64];
  __shared__ unsigned long long red[klft_metro::kWsBlock / 64];
  // pass2 >= 0: two independent passes in one launch (pass_groups,
  // run_schedule), the second half of the grid taking pass2; the passes
  // belong to sweeps *d_sweep + ds and *d_sweep + ds2
  unsigned int bid = blockIdx.x;
  if (pass2 >= 0 && bid >= gridDim.x / 2) {
    pass = pass2;
    ds = ds2;
    bid -= gridDim.x / 2;
  }
  const int c = block_chain[bid];
  const ChainDev ch = chains[c];
  const long long t = (long long)((long long)bid - ch.first_block) * kB + (long long)((threadIdx.x >> 6) & 3) * 64 +
                      (threadIdx.x & 63);
  const uint32_t sweep = *d_sweep + (uint32_t)ds;
  const R cc = DEFECT ? (R)cval[c] : R(1);
  unsigned long long n = 0;
  switch (ch.nc) {
    case 1: {
      Links<R, 1, false> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
      n = klft_metro::ws_pass<R, 1, false, 4, ILV, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 1.0), delta,
                                                             seed, sweep, (uint32_t)c, nullptr, d, cc, stage);
      break;
    }
    case 2: {
      Links<R, 2, COMP> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
      n = klft_metro::ws_pass<R, 2, COMP, 4, ILV, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 2.0), delta,
                                                            seed, sweep, (uint32_t)c, nullptr, d, cc, stage);
      break;
    }
    default: {
      Links<R, 3, COMP> l{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4};
      n = klft_metro::ws_pass<R, 3, COMP, 4, ILV, DEFECT>(l, ch.g, pass, t, nhits, (R)(ch.beta / 3.0), delta,
                                                            seed, sweep, (uint32_t)c, nullptr, d, cc, stage);
      break;
    }
  }
  klft_metro::ws_count(n, red, &counters[c]);
}

template <typename R, bool COMP>
__global__ __launch_bounds__(kB) void k_batch_plaq(const ChainDev *__restrict__ chains,
                                                   const int *__restrict__ pblock_chain, double *partials) {
  __shared__ double lds[kB / 64];
  const int c = pblock_chain[blockIdx.x];
  const ChainDev ch = chains[c];
  const long long s = (long long)(blockIdx.x - ch.first_pblock) * kB + threadIdx.x;
  double v = 0.0;
  if (s < ch.g.V) {
    int x[4];
    dev_coords<4>(ch.g, s, x);
    switch (ch.nc) {
      case 1: v = plaq_at<R, 1, false, 4>(Links<R, 1, false>{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4}, ch.g, x, s); break;
      case 2: v = plaq_at<R, 2, COMP, 4>(Links<R, 2, COMP>{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4}, ch.g, x, s); break;
      default: v = plaq_at<R, 3, COMP, 4>(Links<R, 3, COMP>{reinterpret_cast<cx<R> *>(ch.links), ch.g.V, 4}, ch.g, x, s); break;
    }
  }
  const double bt = block_sum<kB>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = bt;
}

// per-chain sums of the plaquette partials, in block order
__global__ __launch_bounds__(kB) void k_batch_reduce(const ChainDev *__restrict__ chains, int nchains, int total,
                                                     const double *partials, double *out) {
  __shared__ double lds[kB / 64];
  const int c = blockIdx.x;
  const int b0 = chains[c].first_pblock;
  const int b1 = c + 1 < nchains ? chains[c + 1].first_pblock : total;
  double v = 0.0;
  for (int i = b0 + (int)threadIdx.x; i < b1; i += kB) v += partials[i];
  v = block_sum<kB>(v, lds);
  if (threadIdx.x == 0) out[c] = v;
}

// A1 / A2 of every replica: one wave per replica; lane l takes every 64th
// of the (0, nu) plaquettes whose mu = 0 links touch the defect (nu = 1..3,
// then the base sites in lexicographic order), and the lanes' sums are added
// by a shuffle tree, so the result does not depend on scheduling.  A single
// thread walking them one by one waited on each plaquette's loads in turn.
template <typename R, int NC, bool COMP>
__global__ __launch_bounds__(64) void k_defect_sums(const ChainDev *__restrict__ chains, int n, Defect d,
                                                    double *A) {
  const int c = blockIdx.x;
  if (c >= n) return;
  const ChainDev ch = chains[c];
  const Geom &g = ch.g;
  Links<R, NC, COMP> f{reinterpret_cast<cx<R> *>(ch.links), g.V, 4};
  // base sites y: y0 = 0, y_nu in [p_nu - 1, p_nu + Ld), the others in
  // [p_k, p_k + Ld); when Ld equals L_nu, p_nu - 1 and p_nu + Ld - 1 are
  // the same site, so the range along nu is cut to L_nu sites
  int cnt[4] = {0, 0, 0, 0};
  for (int nu = 1; nu < 4; ++nu) {
    const int enu = d.len + 1 < g.L[nu] ? d.len + 1 : g.L[nu];
    cnt[nu] = enu * d.len * d.len;
  }
  const int total = cnt[1] + cnt[2] + cnt[3];
  double a1 = 0.0, a2 = 0.0;
  for (int t = (int)threadIdx.x; t < total; t += 64) {
    int nu = 1, r = t;
    while (r >= cnt[nu]) r -= cnt[nu++];
    int ext[4] = {1, d.len, d.len, d.len};
    ext[nu] = d.len + 1 < g.L[nu] ? d.len + 1 : g.L[nu];
    const int i3 = r % ext[3], i2 = (r / ext[3]) % ext[2], i1 = r / (ext[3] * ext[2]);
    int y[4] = {0, d.pos[0] + i1, d.pos[1] + i2, d.pos[2] + i3};
    y[nu] -= 1;
    for (int k = 1; k < 4; ++k) y[k] = ((y[k] % g.L[k]) + g.L[k]) % g.L[k];
    int yn[4], y0[4];
    shifted<4>(g, y, nu, 1, -1, 0, yn);
    shifted<4>(g, y, 0, 1, -1, 0, y0);
    const int k = (int)on_defect<4>(g, d, y, 0) + (int)on_defect<4>(g, d, yn, 0);
    if (k == 0) continue;
    const Mat<R, NC> lmu = mul(f.load(dev_index<4>(g, y), 0), f.load(dev_index<4>(g, y0), nu));
    const Mat<R, NC> lnu = mul(f.load(dev_index<4>(g, y), nu), f.load(dev_index<4>(g, yn), 0));
    const double p = (double)retr_mul_ad(lmu, lnu);
    if (k == 1)
      a1 += p;
    else
      a2 += p;
  }
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (threadIdx.x == 0) {
    A[2 * c] = a1;
    A[2 * c + 1] = a2;
  }
}

// sequential nearest-neighbour swap proposals of the coupling labels
__global__ void k_swap(int n, int nc, double beta, const double *__restrict__ cvals, const double *__restrict__ A,
                       int *conf_of, double *cval, int *swaps, uint64_t seed, const uint32_t *d_step) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t step = *d_step;
  for (int k = 0; k + 1 < n; ++k) {
    const int a = conf_of[k], b = conf_of[k + 1];
    const double ck = cvals[k], ck1 = cvals[k + 1];
    const double dS = -(beta / (double)nc) *
                      ((ck - ck1) * (A[2 * b] - A[2 * a]) + (ck * ck - ck1 * ck1) * (A[2 * b + 1] - A[2 * a + 1]));
    klft_rng::Stream st(seed, klft_rng::TAG_SWAP, (uint32_t)k, step, 0, 0, 0);
    const double u = st.u01();
    const bool acc = dS <= 0.0 || u < exp(-dS);
    swaps[k] = acc ? 1 : 0;
    if (acc) {
      conf_of[k] = b;
      conf_of[k + 1] = a;
    }
  }
  for (int k = 0; k < n; ++k) cval[conf_of[k]] = cvals[k];
}

// A1 / A2 of every replica into A (one wave per replica)
template <typename R, bool COMP>
void launch_defect_sums(int nc, const ChainDev *chains, int n, const Defect &d, double *A) {
  const dim3 grid((unsigned)n), block(64);
  switch (nc) {
    case 1: hipLaunchKernelGGL((k_defect_sums<R, 1, false>), grid, block, 0, ki::stream(), chains, n, d, A); break;
    case 2: hipLaunchKernelGGL((k_defect_sums<R, 2, COMP>), grid, block, 0, ki::stream(), chains, n, d, A); break;
    default: hipLaunchKernelGGL((k_defect_sums<R, 3, COMP>), grid, block, 0, ki::stream(), chains, n, d, A); break;
  }
}

__global__ void k_zero_counters(unsigned long long *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}
__global__ void k_advance(uint32_t *p) {
  if (threadIdx.x == 0) *p += 1u;
}

}  // namespace

struct klft_batch_s {
  int n = 0, storage = 0;
  std::vector<klft_field_t> fields;
  std::vector<double> beta;
  ChainDev *d_chains = nullptr;
  int *d_block_chain = nullptr, *d_pblock_chain = nullptr;
  int nblocks = 0, npblocks = 0;
  double *d_partials = nullptr, *d_plaq = nullptr;
  unsigned long long *d_counters = nullptr;
  uint32_t *d_sweep = nullptr;
  // graph of one sweep / step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // graphs of a run of sweeps whose passes pair across sweeps: head, window,
  // tail (RunParts); run_pairs: whether the batch runs that way
  hipGraph_t rgraph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t rexec[3] = {nullptr, nullptr, nullptr};
  bool run_pairs = false;
  int g_nhits = -1;
  double g_delta = 0.0;
  uint64_t g_seed = 0;
  // PTBC
  bool ptbc = false;
  int nc = 0;
  Defect defect{};
  std::vector<double> cvals;
  double *d_cvals = nullptr, *d_cval = nullptr, *d_A = nullptr;
  int *d_conf_of = nullptr, *d_swaps = nullptr;
};

namespace {

int check_batch(klft_batch_t b) {
  if (!b || !b->d_chains) return ki::set_error(KLFT_E_ARG, "null or destroyed batch handle");
  return KLFT_OK;
}

void drop_graphs(klft_batch_t b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (int i = 0; i < 3; ++i) {
    if (b->rexec[i]) (void)hipGraphExecDestroy(b->rexec[i]);
    if (b->rgraph[i]) (void)hipGraphDestroy(b->rgraph[i]);
    b->rexec[i] = nullptr;
    b->rgraph[i] = nullptr;
  }
}

void free_batch(klft_batch_t b) {
  drop_graphs(b);
  for (auto f : b->fields) klft_field_destroy(f);
  void *ptrs[] = {b->d_chains, b->d_block_chain, b->d_pblock_chain, b->d_partials, b->d_plaq, b->d_counters,
                  b->d_sweep, b->d_cvals, b->d_cval, b->d_A, b->d_conf_of, b->d_swaps};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  b->d_chains = nullptr;
}

int build_tables(klft_batch_t b) {
  std::vector<ChainDev> ch(b->n);
  std::vector<int> bc, pbc;
  for (int c = 0; c < b->n; ++c) {
    klft_field_t f = b->fields[c];
    ch[c].links = f->links;
    ch[c].g = f->g;
    ch[c].nc = f->nc;
    ch[c].beta = b->beta[c];
    ch[c].first_block = (int)bc.size();
    ch[c].first_pblock = (int)pbc.size();
    const long long nb = (f->g.Vh + kB - 1) / kB, npb = (f->g.V + kB - 1) / kB;
    bc.insert(bc.end(), nb, c);
    pbc.insert(pbc.end(), npb, c);
  }
  b->nblocks = (int)bc.size();
  b->npblocks = (int)pbc.size();
  KLFT_HIP(hipMalloc(&b->d_chains, sizeof(ChainDev) * b->n));
  KLFT_HIP(hipMalloc(&b->d_block_chain, sizeof(int) * bc.size()));
  KLFT_HIP(hipMalloc(&b->d_pblock_chain, sizeof(int) * pbc.size()));
  KLFT_HIP(hipMalloc(&b->d_partials, sizeof(double) * pbc.size()));
  KLFT_HIP(hipMalloc(&b->d_plaq, sizeof(double) * b->n));
  KLFT_HIP(hipMalloc(&b->d_counters, sizeof(unsigned long long) * b->n));
  KLFT_HIP(hipMalloc(&b->d_sweep, sizeof(uint32_t)));
  // tables and clears go on the library stream, where the batch kernels run
  // (a non-blocking stream, not ordered after null-stream work), and are
  // waited for before the host vectors go out of scope
  hipStream_t st = ki::stream();
  KLFT_HIP(hipMemcpyAsync(b->d_chains, ch.data(), sizeof(ChainDev) * b->n, hipMemcpyHostToDevice, st));
  KLFT_HIP(hipMemcpyAsync(b->d_block_chain, bc.data(), sizeof(int) * bc.size(), hipMemcpyHostToDevice, st));
  KLFT_HIP(hipMemcpyAsync(b->d_pblock_chain, pbc.data(), sizeof(int) * pbc.size(), hipMemcpyHostToDevice, st));
  KLFT_HIP(hipMemsetAsync(b->d_counters, 0, sizeof(unsigned long long) * b->n, st));
  KLFT_HIP(hipMemsetAsync(b->d_sweep, 0, sizeof(uint32_t), st));
  KLFT_HIP(hipStreamSynchronize(st));
  return KLFT_OK;
}

// the site-wave kernel of variant v (1, 2) for chains or replicas
template <typename R, bool COMP>
static auto batch_ws_kernel(bool ptbc, int v) {
  auto k = &k_batch_ws<R, COMP, false, false>;
  if (ptbc)
    k = v == 2 ? &k_batch_ws<R, COMP, true, true> : &k_batch_ws<R, COMP, true, false>;
  else if (v == 2)
    k = &k_batch_ws<R, COMP, false, true>;
  return k;
}

// the occupancy queries behind the pairing decision, made (and cached)
// before a graph capture begins, so that the capture sees stream work only
static void prime_occupancy(klft_batch_t b) {
  auto prime = [&](auto rt, auto ct) {
    using R = typename decltype(rt)::type;
    constexpr bool COMP = decltype(ct)::value != 0;
    for (int v = 1; v <= 2; ++v)
      (void)ki::blocks_per_cu(reinterpret_cast<const void *>(batch_ws_kernel<R, COMP>(b->ptbc, v)),
                              klft_metro::kWsBlock);
  };
  switch (b->storage) {
    case KLFT_STORAGE_F64: prime(ki::tag<double>{}, ki::ic<0>{}); break;
    case KLFT_STORAGE_F32: prime(ki::tag<float>{}, ki::ic<0>{}); break;
    default: prime(ki::tag<float>{}, ki::ic<1>{}); break;
  }
}

// One launch of the site-wave kernel: pass pa of sweep *d_sweep + da, and
// pass pb of sweep *d_sweep + db (pb < 0: none).
struct BLaunch {
  int pa, da, pb, db;
};
// A run of n >= 2 sweeps of chains, with passes of consecutive sweeps
// sharing launches: the head, window and tail of run_period(4) as three
// graphs.  The head's sweep offsets are relative to the device sweep index;
// the window's too, and the window then advances that index by one; the
// tail's are relative to the last sweep, then an advance.  The run is head,
// n - 1 windows, tail.
struct RunParts {
  std::vector<BLaunch> part[3];
  bool ok = false;
};
const RunParts &run_parts() {
  static const RunParts rp = [] {
    RunParts r;
    const ki::RunPeriod per = ki::run_period(4);
    if (!per.ok) return r;
    auto conv = [](const ki::RunLaunch &L) {
      return BLaunch{L.pass[0], L.sweep[0], L.pass[1], L.pass[1] >= 0 ? L.sweep[1] : 0};
    };
    for (const ki::RunLaunch &L : per.head) r.part[0].push_back(conv(L));
    for (const ki::RunLaunch &L : per.window) r.part[1].push_back(conv(L));
    for (const ki::RunLaunch &L : per.tail) r.part[2].push_back(conv(L));
    r.ok = true;
    return r;
  }();
  return rp;
}

// part 0: enqueue one sweep (and, for replicas, the swap pass), then advance
// the sweep index; parts 1-3: the head, window and tail of a run whose
// passes pair across sweeps (run_parts; chains only, and only when the
// batch pairs passes).  No host sync.  *paired (optional) receives whether
// independent passes share launches.
int enqueue_step(klft_batch_t b, int nhits, double delta, uint64_t seed, int part = 0, bool *paired = nullptr) {
  if (part <= 1)
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(256), 0, ki::stream(), b->d_counters, b->n);
  auto passes = [&](auto rt, auto ct) -> int {
    using R = typename decltype(rt)::type;
    constexpr bool COMP = decltype(ct)::value != 0;
    // variant 0: one lane per site (k_batch_pass); 1: site waves fed by
    // staple waves (k_batch_ws), upper staple terms first; 2: the same with
    // each direction's upper and lower terms back to back.  Default as for a
    // single field: 2 for FP64 links, and for FP32 links while two paired
    // passes fill at most one 512-thread block per CU (2 takes more
    // registers, 1 leaves room for two blocks per CU)
    const int fv = ki::forced_variant("batch");
    const int v = fv >= 0 && fv <= 2 ? fv : ((sizeof(R) == 8 || 2 * b->nblocks <= ki::cu_count()) ? 2 : 1);
    const bool ws = v != 0;
    auto site_waves = [&](auto kern, int pa, int pb, int da = 0, int db = 0) {
      hipLaunchKernelGGL(kern, dim3(pb >= 0 ? 2 * b->nblocks : b->nblocks), dim3(klft_metro::kWsBlock), 0,
                         ki::stream(), b->d_chains, b->d_block_chain, pa, pb, da, db, nhits, (R)delta, seed,
                         b->d_sweep, b->d_counters, b->defect, b->d_cval);
    };
    auto ws_kernel = [&]() { return batch_ws_kernel<R, COMP>(b->ptbc, v); };
    // independent passes two to a launch (bit-identical to one by one, see
    // pass_groups) when both passes' blocks fit on the GPU at once
    // (KLFT_VARIANT=batch_pairs=0: never)
    const bool pair = ws && ki::forced_variant("batch_pairs") != 0 &&
                      2LL * b->nblocks <= (long long)ki::cu_count() *
                                              ki::blocks_per_cu(reinterpret_cast<const void *>(ws_kernel()),
                                                                klft_metro::kWsBlock);
    if (paired) *paired = pair;
    if (part > 0) {
      if (!pair || b->ptbc || !run_parts().ok) return ki::set_error(KLFT_E_STATE, "run parts need paired passes");
      for (const BLaunch &L : run_parts().part[part - 1]) site_waves(ws_kernel(), L.pa, L.pb, L.da, L.db);
    } else if (pair) {
      const ki::PassGroups gr = ki::pass_groups(4);
      for (int i = 0; i < gr.n; ++i) site_waves(ws_kernel(), gr.p[i][0], gr.p[i][1]);
    } else {
      for (int pass = 0; pass < 16; ++pass) {
        if (ws)
          site_waves(ws_kernel(), pass, -1);
        else if (b->ptbc)
          hipLaunchKernelGGL((k_batch_pass<R, COMP, true>), dim3(b->nblocks), dim3(kB), 0, ki::stream(),
                             b->d_chains, b->d_block_chain, pass, nhits, (R)delta, seed, b->d_sweep, b->d_counters,
                             b->defect, b->d_cval);
        else
          hipLaunchKernelGGL((k_batch_pass<R, COMP, false>), dim3(b->nblocks), dim3(kB), 0, ki::stream(),
                             b->d_chains, b->d_block_chain, pass, nhits, (R)delta, seed, b->d_sweep, b->d_counters,
                             b->defect, b->d_cval);
      }
    }
    if (b->ptbc && part == 0) {
      launch_defect_sums<R, COMP>(b->nc, b->d_chains, b->n, b->defect, b->d_A);
      hipLaunchKernelGGL(k_swap, dim3(1), dim3(64), 0, ki::stream(), b->n, b->nc, b->beta[0], b->d_cvals, b->d_A,
                         b->d_conf_of, b->d_cval, b->d_swaps, seed, b->d_sweep);
    }
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  };
  int rc;
  switch (b->storage) {
    case KLFT_STORAGE_F64: rc = passes(ki::tag<double>{}, ki::ic<0>{}); break;
    case KLFT_STORAGE_F32: rc = passes(ki::tag<float>{}, ki::ic<0>{}); break;
    default: rc = passes(ki::tag<float>{}, ki::ic<1>{}); break;
  }
  if (rc) return rc;
  if (part != 1) hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ki::stream(), b->d_sweep);
  KLFT_HIP(hipGetLastError());
  return KLFT_OK;
}

// capture part `part` of enqueue_step into a graph and instantiate it
int capture(klft_batch_t b, int nhits, double delta, uint64_t seed, int part, hipGraph_t *graph,
            hipGraphExec_t *exec, bool *paired) {
  KLFT_HIP(hipStreamBeginCapture(ki::stream(), hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(b, nhits, delta, seed, part, paired);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ki::stream(), &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return ki::hip_fail(e, "graph capture");
  *graph = g;
  KLFT_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return KLFT_OK;
}

// the graph of one sweep / step and, for chains whose passes pair, the
// head, window and tail graphs of a run (rebuilt when the parameters change)
int ensure_graph(klft_batch_t b, int nhits, double delta, uint64_t seed) {
  if (b->exec && b->g_nhits == nhits && b->g_delta == delta && b->g_seed == seed) return KLFT_OK;
  drop_graphs(b);
  b->run_pairs = false;
  prime_occupancy(b);
  bool paired = false;
  int rc = capture(b, nhits, delta, seed, 0, &b->graph, &b->exec, &paired);
  if (rc) return rc;
  if (paired && !b->ptbc && run_parts().ok) {
    for (int i = 0; i < 3 && !rc; ++i) rc = capture(b, nhits, delta, seed, i + 1, &b->rgraph[i], &b->rexec[i], nullptr);
    if (rc) {
      drop_graphs(b);
      return rc;
    }
    b->run_pairs = true;
  }
  b->g_nhits = nhits;
  b->g_delta = delta;
  b->g_seed = seed;
  return KLFT_OK;
}

double acceptance(klft_field_t f, int nhits, unsigned long long n) {
  double norm = 1.0;
  for (int d = 0; d < f->nd; ++d) norm *= (double)f->dims[d];
  norm *= (double)f->nd;
  norm *= (double)nhits;
  return (double)n / norm;
}

}  // namespace

extern "C" int klft_batch_create(int nchains, const klft_chain_desc *chains, int storage, klft_batch_t *out) {
  if (!out || !chains || nchains < 1) return ki::set_error(KLFT_E_ARG, "invalid batch arguments");
  if (nchains > 1024) return ki::set_error(KLFT_E_ARG, "at most 1024 chains per batch");
  *out = nullptr;
  klft_batch_s *b = new klft_batch_s();
  b->n = nchains;
  b->storage = storage;
  for (int c = 0; c < nchains; ++c) {
    klft_field_t f = nullptr;
    const int rc = klft_field_create(4, chains[c].nc, chains[c].dims, storage, &f);
    if (rc) {
      free_batch(b);
      delete b;
      return rc;
    }
    b->fields.push_back(f);
    b->beta.push_back(chains[c].beta);
  }
  const int rc = build_tables(b);
  if (rc) {
    free_batch(b);
    delete b;
    return rc;
  }
  *out = b;
  return KLFT_OK;
}

extern "C" int klft_batch_destroy(klft_batch_t b) {
  if (!b) return KLFT_OK;
  (void)ki::sync_stream();
  free_batch(b);
  delete b;
  return KLFT_OK;
}

extern "C" int klft_batch_set_identity(klft_batch_t b) {
  int rc = check_batch(b);
  for (int c = 0; c < b->n && !rc; ++c) rc = klft_field_set_identity(b->fields[c]);
  return rc;
}

extern "C" int klft_batch_sweep(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t sweep, double *acc) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (nhits < 0) return ki::set_error(KLFT_E_ARG, "nHits must be >= 0");
  const uint32_t s = (uint32_t)sweep;
  KLFT_HIP(hipMemcpyAsync(b->d_sweep, &s, sizeof s, hipMemcpyHostToDevice, ki::stream()));
  rc = enqueue_step(b, nhits, delta, seed);
  if (rc) return rc;
  std::vector<unsigned long long> n(b->n);
  KLFT_HIP(hipMemcpyAsync(n.data(), b->d_counters, sizeof(unsigned long long) * b->n, hipMemcpyDeviceToHost,
                          ki::stream()));
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  if (acc)
    for (int c = 0; c < b->n; ++c) acc[c] = acceptance(b->fields[c], nhits, n[c]);
  return KLFT_OK;
}

extern "C" int klft_batch_run(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t sweep0, int n) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (nhits < 0 || n < 0) return ki::set_error(KLFT_E_ARG, "invalid sweep count");
  rc = ensure_graph(b, nhits, delta, seed);
  if (rc) return rc;
  const uint32_t s = (uint32_t)sweep0;
  KLFT_HIP(hipMemcpyAsync(b->d_sweep, &s, sizeof s, hipMemcpyHostToDevice, ki::stream()));
  if (b->run_pairs && n >= 2) {
    // passes of consecutive sweeps share launches: head, n - 1 windows, tail
    KLFT_HIP(hipGraphLaunch(b->rexec[0], ki::stream()));
    for (int i = 0; i + 1 < n; ++i) KLFT_HIP(hipGraphLaunch(b->rexec[1], ki::stream()));
    KLFT_HIP(hipGraphLaunch(b->rexec[2], ki::stream()));
  } else {
    for (int i = 0; i < n; ++i) KLFT_HIP(hipGraphLaunch(b->exec, ki::stream()));
  }
  return ki::sync_stream();
}

extern "C" int klft_batch_plaquette(klft_batch_t b, double *out) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!out) return ki::set_error(KLFT_E_ARG, "null output");
  auto launch = [&](auto rt, auto ct) -> int {
    using R = typename decltype(rt)::type;
    constexpr bool COMP = decltype(ct)::value != 0;
    hipLaunchKernelGGL((k_batch_plaq<R, COMP>), dim3(b->npblocks), dim3(kB), 0, ki::stream(), b->d_chains,
                       b->d_pblock_chain, b->d_partials);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  };
  switch (b->storage) {
    case KLFT_STORAGE_F64: rc = launch(ki::tag<double>{}, ki::ic<0>{}); break;
    case KLFT_STORAGE_F32: rc = launch(ki::tag<float>{}, ki::ic<0>{}); break;
    default: rc = launch(ki::tag<float>{}, ki::ic<1>{}); break;
  }
  if (rc) return rc;
  hipLaunchKernelGGL(k_batch_reduce, dim3(b->n), dim3(kB), 0, ki::stream(), b->d_chains, b->n, b->npblocks,
                     b->d_partials, b->d_plaq);
  KLFT_HIP(hipGetLastError());
  std::vector<double> raw(b->n);
  KLFT_HIP(hipMemcpyAsync(raw.data(), b->d_plaq, sizeof(double) * b->n, hipMemcpyDeviceToHost, ki::stream()));
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  for (int c = 0; c < b->n; ++c) {  // GaugePlaquette.hpp:125-132
    klft_field_t f = b->fields[c];
    double norm = 1.0;
    for (int d = 0; d < f->nd; ++d) norm *= (double)f->dims[d];
    norm *= (double)((f->nd * (f->nd - 1) / 2) * f->nc);
    out[c] = raw[c] / norm;
  }
  return KLFT_OK;
}

extern "C" int klft_batch_download(klft_batch_t b, int chain, double *host) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (chain < 0 || chain >= b->n) return ki::set_error(KLFT_E_ARG, "chain index out of range");
  return klft_field_download(b->fields[chain], host);
}

extern "C" int klft_batch_upload(klft_batch_t b, int chain, const double *host) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (chain < 0 || chain >= b->n) return ki::set_error(KLFT_E_ARG, "chain index out of range");
  return klft_field_upload(b->fields[chain], host);
}

extern "C" int klft_ptbc_create(int nreplicas, int nc, const int *dims, int storage, double beta, int defect_length,
                                const int *defect_pos, const double *c, klft_batch_t *out) {
  if (!out || !dims || !defect_pos || !c || nreplicas < 1) return ki::set_error(KLFT_E_ARG, "invalid PTBC arguments");
  if (defect_length < 1) return ki::set_error(KLFT_E_ARG, "defect length must be >= 1");
  for (int k = 1; k < 4; ++k)
    if (defect_length > dims[k]) return ki::set_error(KLFT_E_ARG, "defect longer than the lattice");
  std::vector<klft_chain_desc> desc(nreplicas);
  for (int r = 0; r < nreplicas; ++r) {
    desc[r].nc = nc;
    for (int d = 0; d < 4; ++d) desc[r].dims[d] = dims[d];
    desc[r].beta = beta;
  }
  int rc = klft_batch_create(nreplicas, desc.data(), storage, out);
  if (rc) return rc;
  klft_batch_t b = *out;
  b->ptbc = true;
  b->nc = nc;
  b->defect.len = defect_length;
  for (int k = 0; k < 3; ++k) b->defect.pos[k] = ((defect_pos[k] % dims[k + 1]) + dims[k + 1]) % dims[k + 1];
  b->cvals.assign(c, c + nreplicas);
  std::vector<int> conf(nreplicas);
  for (int r = 0; r < nreplicas; ++r) conf[r] = r;
  hipError_t e = hipMalloc(&b->d_cvals, sizeof(double) * nreplicas);
  if (e == hipSuccess) e = hipMalloc(&b->d_cval, sizeof(double) * nreplicas);
  if (e == hipSuccess) e = hipMalloc(&b->d_A, sizeof(double) * 2 * nreplicas);
  if (e == hipSuccess) e = hipMalloc(&b->d_conf_of, sizeof(int) * nreplicas);
  if (e == hipSuccess) e = hipMalloc(&b->d_swaps, sizeof(int) * nreplicas);
  // on the library stream, where the replica kernels run, and waited for
  hipStream_t st = ki::stream();
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_cvals, c, sizeof(double) * nreplicas, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_cval, c, sizeof(double) * nreplicas, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b->d_conf_of, conf.data(), sizeof(int) * nreplicas, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(b->d_swaps, 0, sizeof(int) * nreplicas, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    klft_batch_destroy(b);
    *out = nullptr;
    return ki::hip_fail(e, "PTBC workspace");
  }
  return KLFT_OK;
}

extern "C" int klft_ptbc_step(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t step, int *swaps) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!b->ptbc) return ki::set_error(KLFT_E_STATE, "not a PTBC replica set");
  rc = ensure_graph(b, nhits, delta, seed);
  if (rc) return rc;
  const uint32_t s = (uint32_t)step;
  KLFT_HIP(hipMemcpyAsync(b->d_sweep, &s, sizeof s, hipMemcpyHostToDevice, ki::stream()));
  KLFT_HIP(hipGraphLaunch(b->exec, ki::stream()));
  if (swaps && b->n > 1)
    KLFT_HIP(hipMemcpyAsync(swaps, b->d_swaps, sizeof(int) * (b->n - 1), hipMemcpyDeviceToHost, ki::stream()));
  return ki::sync_stream();
}

extern "C" int klft_ptbc_labels(klft_batch_t b, int *labels) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!b->ptbc || !labels) return ki::set_error(KLFT_E_ARG, "not a PTBC replica set / null output");
  KLFT_HIP(hipMemcpyAsync(labels, b->d_conf_of, sizeof(int) * b->n, hipMemcpyDeviceToHost, ki::stream()));
  return ki::sync_stream();
}

extern "C" int klft_ptbc_action(klft_batch_t b, int replica, int coupling, double *out) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!b->ptbc || !out || replica < 0 || replica >= b->n || coupling < 0 || coupling >= b->n)
    return ki::set_error(KLFT_E_ARG, "invalid PTBC action request");
  // S_c(U) = -(beta/Nc) [P_raw(U) - A1 - A2 + c A1 + c^2 A2]
  double raw = 0.0;
  rc = klft_plaquette(b->fields[replica], 0, &raw, nullptr);
  if (rc) return rc;
  auto launch = [&](auto rt, auto ct) -> int {
    using R = typename decltype(rt)::type;
    constexpr bool COMP = decltype(ct)::value != 0;
    launch_defect_sums<R, COMP>(b->nc, b->d_chains, b->n, b->defect, b->d_A);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  };
  switch (b->storage) {
    case KLFT_STORAGE_F64: rc = launch(ki::tag<double>{}, ki::ic<0>{}); break;
    case KLFT_STORAGE_F32: rc = launch(ki::tag<float>{}, ki::ic<0>{}); break;
    default: rc = launch(ki::tag<float>{}, ki::ic<1>{}); break;
  }
  if (rc) return rc;
  double A[2];
  KLFT_HIP(hipMemcpyAsync(A, b->d_A + 2 * replica, sizeof A, hipMemcpyDeviceToHost, ki::stream()));
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  const double c = b->cvals[coupling];
  *out = -(b->beta[0] / (double)b->nc) * (raw - A[0] - A[1] + c * A[0] + c * c * A[1]);
  return KLFT_OK;
}
