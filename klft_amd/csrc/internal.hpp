// Host-side internals shared by the libklft_hip translation units: handle
// layout, error reporting, the library stream and the runtime -> template
// dispatch over (precision, Nc, storage compression, nd).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/klft_hip.h"
#include "device.hpp"

struct klft_field_s {
  int nd = 0, nc = 0, storage = 0;
  int dims[4] = {1, 1, 1, 1};
  kd::Geom g{};
  void *links = nullptr;  // blocks of 64 sites, sublattice-major (device.hpp)
  size_t link_bytes = 0;
  int ncomp = 0;          // complex components stored per link
  size_t cbytes = 0;      // bytes per complex component
  // reduction workspace: per-block partials, a ticket and result slots
  double *partials = nullptr;
  long long max_partials = 0;  // capacity of partials: the largest reduction grid
  unsigned int *ticket = nullptr;
  double *result = nullptr;  // 16 doubles
  unsigned long long *counter = nullptr;  // 8 counters
  double *host_result = nullptr;          // pinned, 16 doubles
  // per-sweep acceptance counters of klft_metropolis_run (grown on demand,
  // kept until the field is destroyed)
  unsigned long long *run_counters = nullptr;
  int run_capacity = 0;
  // measurement workspace (Wilson-loop partials, outputs, extents; grown on
  // demand, kept until the field is destroyed)
  void *meas_ws = nullptr;
  size_t meas_capacity = 0;
  int proposal = KLFT_PROPOSAL_REFERENCE;  // SU(3) Metropolis proposal law
};

namespace klft_internal {

// reduction partials per field: at least this many, and at least one per
// block of the largest grid (one thread per link, 64 sites per block)
constexpr long long kMinPartials = 1 << 16;

int set_error(int code, const char *fmt, ...);
int ensure_init();
hipStream_t stream();
int hip_fail(hipError_t e, const char *what);

#define KLFT_HIP(call)                                          \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return klft_internal::hip_fail(e_, #call); \
  } while (0)

kd::Geom make_geom(int nd, const int *dims);
// the field's measurement workspace, at least `bytes` long (grown on demand;
// growing synchronises the library stream before the old buffer is freed)
int meas_workspace(klft_field_t f, size_t bytes, void **out);
// enqueue-only helpers (no host synchronisation)
int plaquette_raw(klft_field_t f, int slot, double *dev_per_site);  // sum into f->result[slot]
int read_results(klft_field_t f, int first, int n, double *out);    // synchronises
int enqueue_reunit(klft_field_t f);
// out = sum of partials[0 .. n) in index order (one block, after the launch
// that wrote them)
int enqueue_sum_blocks(const double *partials, long long n, double *out);
int enqueue_sweep(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep, uint32_t chain,
                  unsigned long long *counter, int *visit);
int check_field(const klft_field_s *f);
int sync_stream();

template <int V>
using ic = std::integral_constant<int, V>;

// ---- kernel timing (tuner.hip): HIP events on the library stream around
// launches, enabled with klft_profile(1); read back with klft_profile_read ----
// Each launch also sits in a roctx range named with the reference's own
// kernel label (e.g. "sweep_Metropolis_GaugeField_sublat_3",
// Metropolis.hpp:161), which rocprofv3 --marker-trace records.
bool prof_enabled();
void prof_record(const char *name, hipEvent_t start, hipEvent_t stop);
struct ProfScope {
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(const char *n, const char *label) : name(n) {
    roctxRangePushA(label);
    if (prof_enabled() && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, stream());
    else
      a = b = nullptr;
  }
  ~ProfScope() {
    if (a) {
      (void)hipEventRecord(b, stream());
      prof_record(name, a, b);
    }
    roctxRangePop();
  }
};

// Passes that neither read nor write each other's links commute.  A pass
// writes only its own sites' links, and its staples read the links of the
// sublattices whose coordinate parities differ from its own in one or two
// places (x + mu, x - nu, x + mu - nu).  Passes p, q with three or more
// parities apart are therefore independent, and since every draw is keyed by
// (site, sweep, direction, hit) and not by launch order, running them in one
// launch gives the sequential sweep's result bit for bit.  The groups below
// follow the sweep order 0 .. 2^nd - 1 (Metropolis.hpp:139-162): a pass
// joins the group after the last earlier pass it depends on.  4D: 16 passes
// in 11 launches, {3,4} {5,8} {6,9} {7,10} {11,12} paired; 3D: {3,4} paired;
// 2D: none.  Each group holds at most two passes.
struct PassGroups {
  int n = 0;
  int p[16][2];
};
inline PassGroups pass_groups(int nd) {
  const int np = 1 << nd;
  int level[16];
  PassGroups gr;
  for (int q = 0; q < np; ++q) {
    int l = 0;
    for (int r = 0; r < q; ++r) {
      const int h = __builtin_popcount(q ^ r);
      if (h <= 2 && level[r] + 1 > l) l = level[r] + 1;
    }
    level[q] = l;
    if (l == gr.n) {
      gr.p[gr.n][0] = q;
      gr.p[gr.n][1] = -1;
      ++gr.n;
    } else {
      gr.p[l][1] = q;  // the level's second pass (a third never occurs for nd <= 4)
    }
  }
  return gr;
}

// The same rule across the sweeps of a run: pass p of sweep s conflicts with
// every earlier pass (of this sweep or an earlier one) whose sublattice is at
// most two parities from p's, its own sublattice included.  Passes are taken
// in program order (sweep-major, 0 .. 2^nd - 1 within a sweep) and each goes
// into the first launch after the last conflicting pass that still has room
// for it, at most two passes to a launch; two conflicting passes therefore
// keep their order, and a pass may move ahead of earlier passes it does not
// touch.  One sweep gives pass_groups' 11 launches in 4D; over many sweeps
// the first three passes of each sweep pair with the last three of the one
// before (13|0', 14|1', 15|2': three, four and three parities apart), 8
// launches per sweep, every one of them a pair.  sweep[] is relative to the
// run's first.
struct RunLaunch {
  int pass[2];
  int sweep[2];
};
inline std::vector<RunLaunch> run_schedule(int nd, int nsweeps) {
  const int np = 1 << nd;
  std::vector<RunLaunch> out;
  int last[16];  // launch of the latest scheduled pass of each sublattice
  for (int q = 0; q < np; ++q) last[q] = -1;
  std::vector<int> fill;
  for (int s = 0; s < nsweeps; ++s) {
    for (int p = 0; p < np; ++p) {
      int lo = 0;
      for (int q = 0; q < np; ++q)
        if (__builtin_popcount(p ^ q) <= 2 && last[q] + 1 > lo) lo = last[q] + 1;
      int k = lo;
      while (k < (int)out.size() && fill[k] >= 2) ++k;
      if (k == (int)out.size()) {
        out.push_back(RunLaunch{{p, -1}, {s, 0}});
        fill.push_back(1);
      } else {
        out[k].pass[fill[k]] = p;
        out[k].sweep[fill[k]] = s;
        ++fill[k];
      }
      last[p] = k;
    }
  }
  return out;
}

// run_schedule is periodic with a period of one sweep: for n >= 2 sweeps it
// is a head (the first sweep's launches before its first pair with the
// second), n - 1 windows (window i: the launches holding the last passes of
// sweep i and the first of sweep i + 1, then the rest of sweep i + 1 up to
// its pairs with sweep i + 2) and a tail (the last sweep's remaining passes).
// run_period derives the three from run_schedule(nd, 3) and checks that the
// second window is the first one sweep later; sweeps in `window` are
// relative to the window's index, in `tail` to the last sweep.  4D: head
// {0} {1} {2} {3,4} {5,8} {6,9} {7,10} {11,12}, window {13,0'} {14,1'}
// {15,2'} {3',4'} {5',8'} {6',9'} {7',10'} {11',12'}, tail {13} {14} {15}.
struct RunPeriod {
  std::vector<RunLaunch> head, window, tail;
  bool ok = false;
};
inline RunPeriod run_period(int nd) {
  RunPeriod r;
  const std::vector<RunLaunch> S = run_schedule(nd, 3);
  auto top = [](const RunLaunch &L) { return L.pass[1] >= 0 && L.sweep[1] > L.sweep[0] ? L.sweep[1] : L.sweep[0]; };
  auto shifted = [](RunLaunch L, int by) {
    L.sweep[0] -= by;
    if (L.pass[1] >= 0) L.sweep[1] -= by;
    return L;
  };
  size_t k1 = 0;
  while (k1 < S.size() && top(S[k1]) < 1) ++k1;
  size_t k2 = k1;
  while (k2 < S.size() && top(S[k2]) < 2) ++k2;
  const size_t w = k2 - k1;
  if (w == 0 || k2 + w > S.size()) return r;
  r.head.assign(S.begin(), S.begin() + (long)k1);
  r.window.assign(S.begin() + (long)k1, S.begin() + (long)k2);
  for (size_t i = 0; i < w; ++i) {
    const RunLaunch a = shifted(S[k2 + i], 1), &c = r.window[i];
    if (a.pass[0] != c.pass[0] || a.sweep[0] != c.sweep[0] || a.pass[1] != c.pass[1] ||
        (a.pass[1] >= 0 && a.sweep[1] != c.sweep[1]))
      return r;
  }
  for (size_t i = k2 + w; i < S.size(); ++i) r.tail.push_back(shifted(S[i], 2));
  r.ok = true;
  return r;
}
// the launches of a run of n sweeps, in order, without building the list:
// fn(launch) with sweeps relative to the run's first (the period above,
// cached per nd; one sweep: run_schedule(nd, 1))
template <typename Fn>
inline void for_each_run_launch(int nd, int nsweeps, Fn &&fn) {
  if (nsweeps <= 0) return;
  if (nsweeps == 1) {
    for (const RunLaunch &L : run_schedule(nd, 1)) fn(L);
    return;
  }
  static const RunPeriod periods[3] = {run_period(2), run_period(3), run_period(4)};
  const RunPeriod &rp = periods[nd - 2];
  if (!rp.ok) {
    for (const RunLaunch &L : run_schedule(nd, nsweeps)) fn(L);
    return;
  }
  auto at = [](RunLaunch L, int base) {
    L.sweep[0] += base;
    if (L.pass[1] >= 0) L.sweep[1] += base;
    return L;
  };
  for (const RunLaunch &L : rp.head) fn(L);
  for (int i = 0; i + 1 < nsweeps; ++i)
    for (const RunLaunch &L : rp.window) fn(at(L, i));
  for (const RunLaunch &L : rp.tail) fn(at(L, nsweeps - 1));
}

// compute units of the current device (256 on an MI355X), queried once
int cu_count();
// workgroups of `block` threads of kernel `kern` resident per CU at once
// (the occupancy query, cached per kernel; at least 1)
int blocks_per_cu(const void *kern, int block);

// ---- autotuner (tuner.hip) ----
bool tuning_enabled();
std::string tune_key(const char *kernel, const klft_field_s *f);
int pick_variant(const char *kernel, const klft_field_s *f, int nvar, int dflt,
                 const std::function<int(int)> &run);
// the forced, cached or default variant, never timed
int chosen_variant(const char *kernel, const klft_field_s *f, int nvar, int dflt);
// the variant KLFT_VARIANT forces for `kernel` ("name=v,..."), or -1
int forced_variant(const char *kernel);

template <typename T>
struct tag {
  using type = T;
};

// fn(tag<R>, ic<NC>, ic<COMP>, ic<ND>)
template <class F>
int dispatch(int storage, int nc, int nd, F &&fn) {
  auto by_nd = [&](auto rt, auto nct, auto ct) -> int {
    switch (nd) {
      case 2: return fn(rt, nct, ct, ic<2>{});
      case 3: return fn(rt, nct, ct, ic<3>{});
      case 4: return fn(rt, nct, ct, ic<4>{});
    }
    return set_error(KLFT_E_ARG, "nd=%d unsupported", nd);
  };
  auto by_nc = [&](auto rt, auto ct) -> int {
    switch (nc) {
      case 1: return by_nd(rt, ic<1>{}, ic<0>{});
      case 2: return by_nd(rt, ic<2>{}, ct);
      case 3: return by_nd(rt, ic<3>{}, ct);
    }
    return set_error(KLFT_E_ARG, "Nc=%d unsupported", nc);
  };
  switch (storage) {
    case KLFT_STORAGE_F64: return by_nc(tag<double>{}, ic<0>{});
    case KLFT_STORAGE_F32: return by_nc(tag<float>{}, ic<0>{});
    case KLFT_STORAGE_F32_ROWS: return by_nc(tag<float>{}, ic<1>{});
  }
  return set_error(KLFT_E_ARG, "storage=%d unsupported", storage);
}

// the same for one storage type (R, compressed rows) fixed at compile time;
// U(1) has no compressed form, so COMP = true covers Nc = 2, 3 only
template <typename RR, bool CC, class F>
int dispatch_typed(int nc, int nd, F &&fn) {
  auto by_nd = [&](auto nct, auto ct) -> int {
    switch (nd) {
      case 2: return fn(tag<RR>{}, nct, ct, ic<2>{});
      case 3: return fn(tag<RR>{}, nct, ct, ic<3>{});
      case 4: return fn(tag<RR>{}, nct, ct, ic<4>{});
    }
    return set_error(KLFT_E_ARG, "nd=%d unsupported", nd);
  };
  switch (nc) {
    case 1:
      if constexpr (!CC) return by_nd(ic<1>{}, ic<0>{});
      break;
    case 2: return by_nd(ic<2>{}, ic<CC ? 1 : 0>{});
    case 3: return by_nd(ic<3>{}, ic<CC ? 1 : 0>{});
  }
  return set_error(KLFT_E_ARG, "Nc=%d unsupported for this storage", nc);
}

#define KLFT_UNPACK(rt, nct, ct, ndt)              \
  using R = typename decltype(rt)::type;           \
  constexpr int NC = decltype(nct)::value;         \
  constexpr bool COMP = decltype(ct)::value != 0;  \
  constexpr int ND = decltype(ndt)::value;         \
  (void)0

template <typename R, int NC, bool COMP>
kd::Links<R, NC, COMP> links_of(const klft_field_s *f) {
  kd::Links<R, NC, COMP> l;
  l.p = reinterpret_cast<kd::cx<R> *>(f->links);
  l.V = f->g.V;
  l.nd = f->nd;
  return l;
}

inline unsigned int blocks_for(long long n, int block) {
  return (unsigned int)((n + block - 1) / block);
}

}  // namespace klft_internal

namespace kd {
// Workgroups are dispatched round-robin over the 8 XCDs; this gives each XCD
// a contiguous range of logical block indices, so blocks working on
// neighbouring sites share an L2.
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int nb) {
  const unsigned int per = nb / 8u, rem = nb % 8u;
  const unsigned int x = b % 8u, k = b / 8u;
  return (x < rem ? x * (per + 1u) : rem * (per + 1u) + (x - rem) * per) + k;
}
// One thread per site: logical block lb covers 256 consecutive positions r
// of sublattice lb % 2^nd, so 2^nd consecutive logical blocks -- kept on one
// XCD by xcd_block -- cover the same region of every sublattice while each
// wave reads 64 consecutive sites.
template <int ND>
__device__ __forceinline__ bool chunk_site(const Geom &g, unsigned int lb, int *x, long long *s) {
  constexpr int NSUB = 1 << ND;
  const int sub = (int)(lb % NSUB);
  const long long r = (long long)(lb / NSUB) * 256 + threadIdx.x;
  if (r >= g.Vh) return false;
  sub_coords<ND>(g, sub, r, x);
  *s = (long long)sub * g.Vh + r;
  return true;
}

template <int ND>
__host__ __device__ __forceinline__ long long chunk_threads(const Geom &g) {
  return ((g.Vh + 255) / 256) * 256 * (1LL << ND);
}

// One thread per link: a block of 64*ND threads holds one 64-site chunk of
// one sublattice, wave w taking direction mu = w, so the ND waves of a block
// share the chunk's links through L1/L2 while each wave reads coalesced.
template <int ND>
__device__ __forceinline__ bool link_site(const Geom &g, unsigned int lb, int *x, long long *s,
                                          int *mu) {
  constexpr int NSUB = 1 << ND;
  const int sub = (int)(lb % NSUB);
  const long long r = (long long)(lb / NSUB) * 64 + (threadIdx.x & 63);
  *mu = (int)(threadIdx.x >> 6);
  if (r >= g.Vh) return false;
  sub_coords<ND>(g, sub, r, x);
  *s = (long long)sub * g.Vh + r;
  return true;
}
template <int ND>
__host__ __device__ __forceinline__ long long link_blocks(const Geom &g) {
  return ((g.Vh + 63) / 64) * (1LL << ND);
}

}  // namespace kd
