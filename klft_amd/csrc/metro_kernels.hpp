// Checkerboarded Metropolis sweep kernels (Metropolis.hpp:44-178) and their
// launcher, instantiated per link storage in metropolis_f64.hip,
// metropolis_f32.hip and metropolis_f32rows.hip (the three translation units
// compile in parallel).  The algorithm is described in metropolis.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <type_traits>

#include "internal.hpp"
#include "metro.hpp"

namespace klft_metro {

using namespace kd;
using namespace klft_internal;

constexpr int kMetroBlock = 256;

// LOWREG takes the staple one term at a time; OCC asks the register allocator
// for at least OCC waves per SIMD.
template <typename R, int NC, bool COMP, int ND, bool LOWREG, int OCC>
__global__ __launch_bounds__(kMetroBlock, OCC) void k_metropolis_pass(
    Links<R, NC, COMP> f, Geom g, int pass, int nhits, R beta_nc, R delta, uint64_t seed,
    uint32_t sweep, uint32_t chain, unsigned long long *nacc, int *visit) {
  f.nd = ND;
  __shared__ unsigned long long red[kMetroBlock / 64];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  const long long t = (long long)lb * kMetroBlock + threadIdx.x;
  unsigned long long n = 0;
  if (t < g.Vh)
    n = metro_site<R, NC, COMP, ND, false, LOWREG>(f, g, pass, t, nhits, beta_nc, delta, seed, sweep,
                                                       chain, Defect{}, R(1), visit);
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

// one lane per site, the next direction's staple loaded during the hits
template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(kMetroBlock) void k_metropolis_pipe(
    Links<R, NC, COMP> f, Geom g, int pass, int nhits, R beta_nc, R delta, uint64_t seed,
    uint32_t sweep, uint32_t chain, unsigned long long *nacc, int *visit) {
  f.nd = ND;
  __shared__ unsigned long long red[kMetroBlock / 64];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  const long long t = (long long)lb * kMetroBlock + threadIdx.x;
  unsigned long long n = 0;
  if (t < g.Vh)
    n = metro_site_pipe<R, NC, COMP, ND>(f, g, pass, t, nhits, beta_nc, delta, seed, sweep, chain, visit);
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

// one lane per site, the site's own links kept in registers
template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(kMetroBlock) void k_metropolis_own(
    Links<R, NC, COMP> f, Geom g, int pass, int nhits, R beta_nc, R delta, uint64_t seed,
    uint32_t sweep, uint32_t chain, unsigned long long *nacc, int *visit) {
  f.nd = ND;
  __shared__ unsigned long long red[kMetroBlock / 64];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  const long long t = (long long)lb * kMetroBlock + threadIdx.x;
  unsigned long long n = 0;
  if (t < g.Vh)
    n = metro_site_own<R, NC, COMP, ND>(f, g, pass, t, nhits, beta_nc, delta, seed, sweep, chain, visit);
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

// ---------------------------------------------------------------------------
// Site waves and staple waves.  At 32^4 a pass has one wave of sites per
// SIMD, and the pipelined kernel's single wave cannot keep the next
// direction's staple loads in flight while it computes hits: its pass costs
// about the staple phase plus the hit phase.  Here a block of 8 waves holds
// 4 site waves and 4 staple waves, two waves per SIMD.  Site wave i runs the
// Markov chains of its 64 sites; staple wave i + 4 builds the staple of the
// next direction for the same sites and hands it over through LDS, so the
// SIMD computes hits while the other wave waits on memory.  The staple has
// the pipelined kernel's terms: every term except the upper one through
// U_mu(x), which the site wave adds once U_mu(x) is final, from
// U_mu(x+m) U_m(x+mu)^dagger left by the staple wave (with ILV each nu's
// upper and lower terms are taken back to back).  The site's own link is
// handed over the same way.  Per direction the block meets two barriers:
// after the site waves have read the handover buffers, and after the staple
// waves have refilled them and the site waves have stored the updated link
// (which the next staple reads as U_nu(x), nu < m).  Both roles meet the same
// 1 + 2 nd barriers, lanes past the sublattice included.
// ---------------------------------------------------------------------------
constexpr int kWsBlock = 512;
constexpr int kWsSites = 256;

template <typename R, int NC>
__device__ __forceinline__ void lds_put(cx<R> *buf, const Mat<R, NC> &a) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < NC * NC; ++c) buf[c * 64 + l] = a.m[c / NC][c % NC];
}
template <typename R, int NC>
__device__ __forceinline__ Mat<R, NC> lds_get(const cx<R> *buf) {
  const int l = threadIdx.x & 63;
  Mat<R, NC> a;
#pragma unroll
  for (int c = 0; c < NC * NC; ++c) a.m[c / NC][c % NC] = buf[c * 64 + l];
  return a;
}

// store the link held in an LDS buffer, one component at a time (a whole
// matrix in registers at this point would spill)
template <typename R, int NC, bool COMP>
__device__ __forceinline__ void store_from_lds(const Links<R, NC, COMP> &f, long long s, int mu,
                                               const cx<R> *buf) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < Links<R, NC, COMP>::NCOMP; ++c) f.p[f.comp_index(s, mu, c)] = buf[c * 64 + l];
}

// LinkChain with U in the site wave's LDS buffer `bu`: U is read only when
// an accepted proposal is multiplied in, and two waves per SIMD leave each
// wave half the registers of the one-wave kernels.
template <typename R, int NC>
struct LdsChain {
  Mat<R, NC> M, P;
  R trM;
  bool pending;
  cx<R> *bu;

  __device__ __forceinline__ void start(const Mat<R, NC> &S, const Mat<R, NC> &U0) {
    lds_put<R, NC>(bu, U0);
    M = mul(S, U0);
    trM = retrace(M);
    pending = false;
  }
  __device__ __forceinline__ void apply_pending() {
    if (pending) {
      Mat<R, NC> U = mul(lds_get<R, NC>(bu), P);
      reunitarize(U);
      lds_put<R, NC>(bu, U);
      pending = false;
    }
  }
  __device__ __forceinline__ void step(const Mat<R, NC> &r, R u, R beta_nc, R cl, unsigned long long &n) {
    if (metro_accept(cl * (-beta_nc * (retr_mul(r, M) - trM)), u)) {
      apply_pending();
      P = r;
      pending = true;
      M = mul(M, r);
      trM = retrace(M);
      ++n;
    }
  }
};

template <typename R, int NC>
__device__ __forceinline__ Mat<R, NC> adjoint(const Mat<R, NC> &a) {
  Mat<R, NC> c;
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) c.m[i][j] = mk(a.m[j][i].re, -a.m[j][i].im);
  return c;
}

// ILV: each nu's upper and lower staple terms back to back (else all upper
// terms first, the order of the other kernels)
// One pass of position t = wi * 64 + lane of the block's 256 sites (t past
// the sublattice: no work, every barrier met), for the calling wave's role;
// returns the site wave's accepted hits.  `stage` holds, per site wave, the
// staple (all terms but the deferred one), the deferred term's first
// product, the own link of the next direction, and the link being updated:
// 4 x 4 x Nc^2 x 64 complex.  With DEFECT every link read is scaled by c on
// the defect and the updated link's own factor multiplies dS, as in
// metro_site.  SYM draws the opt-in symmetric SU(3) proposal (propose_law).
template <typename R, int NC, bool COMP, int ND, bool ILV, bool DEFECT, bool SYM = false>
__device__ __forceinline__ unsigned long long ws_pass(const Links<R, NC, COMP> &f, const Geom &g, int pass,
                                                      long long t, int nhits, R beta_nc, R delta, uint64_t seed,
                                                      uint32_t sweep, uint32_t chain, int *visit, const Defect &d,
                                                      R c, cx<R> *stage_base) {
  constexpr int NM = NC * NC;
  const int w = (int)(threadIdx.x >> 6);
  const int wi = w & 3;
  const bool sitew = w < 4;
  const bool active = t < g.Vh;
  auto stage = [&](int k) { return stage_base + (wi * 4 + k) * NM * 64; };
  auto cl_of = [&](const int *y, int mu) -> R {
    if constexpr (DEFECT) return on_defect<ND>(g, d, y, mu) ? c : R(1);
    return R(1);
  };
  int x[ND];
  long long s = 0;
  uint32_t site = 0;
  if (active) {
    sub_coords<ND>(g, pass, t, x);
    s = (long long)pass * g.Vh + t;
    site = (uint32_t)lex_index<ND>(g, x);
  }
  cx<R> *bS = stage(0), *bA = stage(1), *bU = stage(2);
  unsigned long long n = 0;
  // Staple sg of the site without its upper term nu = defer, restricted to
  // the upper terms j with bit j of umask and the lower terms j with bit j of
  // lmask (j-th direction other than sg).  keep (if given) receives
  // U_{sg+1}(x+sg) U_sg(x+sg+1)^dagger, the first product of the upper term
  // nu = sg + 1: its adjoint is the first product of the next staple's
  // deferred term, which therefore costs no loads.
  auto build = [&](int sg, int defer, unsigned umask, unsigned lmask, Mat<R, NC> *keep) -> Mat<R, NC> {
    Mat<R, NC> acc = zero<R, NC>();
    const long long s_xp = idx_shift<ND>(g, x, sg, 1, -1, 0);
    // the link at x + s1 e_m1 + s2 e_m2 in direction dir, defect-scaled
    auto ld = [&](int m1, int s1, int m2, int s2, int dir) -> Mat<R, NC> {
      int y[ND];
      shifted<ND>(g, x, m1, s1, m2, s2, y);
      return load_at<R, NC, COMP, ND, true>(f, g, y, dir, d, c);
    };
    auto upper = [&](int j) {
      const int nu = j < sg ? j : j + 1;
      if (nu == defer || !((umask >> j) & 1u)) return;
      Mat<R, NC> ab;
      if constexpr (DEFECT)
        ab = mul_ad(ld(sg, 1, -1, 0, nu), ld(nu, 1, -1, 0, sg));
      else
        ab = mul_ad(f.load(s_xp, nu), f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), sg));
      if (keep && nu == sg + 1) *keep = ab;
      if constexpr (DEFECT)
        acc_mul_ad(acc, ab, ld(-1, 0, -1, 0, nu));
      else
        acc_mul_ad(acc, ab, f.load(s, nu));
    };
    auto lower = [&](int j) {
      if (!((lmask >> j) & 1u)) return;
      const int nu = j < sg ? j : j + 1;
      if constexpr (DEFECT) {
        acc_mul(acc, mul_dd(ld(sg, 1, nu, -1, nu), ld(nu, -1, -1, 0, sg)), ld(nu, -1, -1, 0, nu));
      } else {
        const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
        acc_mul(acc, mul_dd(f.load(idx_shift<ND>(g, x, sg, 1, nu, -1), nu), f.load(s_xmnu, sg)),
                f.load(s_xmnu, nu));
      }
    };
    // upper and lower term of one nu with their link loads issued before any
    // product: six loads in flight for FP32 links, each term's three for FP64
    // links (six FP64 matrices would not fit the two waves per SIMD)
    auto pair_terms = [&](int j) {
      const int nu = j < sg ? j : j + 1;
      const bool up = nu != defer && ((umask >> j) & 1u);
      const bool lo = ((lmask >> j) & 1u) != 0u;
      const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
      Mat<R, NC> a0, a1, a2, b0, b1, b2;
      if (up) {
        a0 = f.load(s_xp, nu);
        a1 = f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), sg);
        a2 = f.load(s, nu);
      }
      if constexpr (sizeof(R) == 4) {
        if (lo) {
          b0 = f.load(idx_shift<ND>(g, x, sg, 1, nu, -1), nu);
          b1 = f.load(s_xmnu, sg);
          b2 = f.load(s_xmnu, nu);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      if (up) {
        const Mat<R, NC> ab = mul_ad(a0, a1);
        if (keep && nu == sg + 1) *keep = ab;
        acc_mul_ad(acc, ab, a2);
      }
      if constexpr (sizeof(R) == 8) {
        if (lo) {
          b0 = f.load(idx_shift<ND>(g, x, sg, 1, nu, -1), nu);
          b1 = f.load(s_xmnu, sg);
          b2 = f.load(s_xmnu, nu);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      if (lo) acc_mul(acc, mul_dd(b0, b1), b2);
    };
    if constexpr (ILV) {
      // upper and lower term of one nu back to back: the lower term's
      // U_sg(x-nu) is the upper term's U_sg(x'+nu) of the site x' = x - 2 nu,
      // a neighbouring lane or wave, so the second read tends to hit L2
#pragma unroll 1
      for (int j = 0; j < ND - 1; ++j) {
        if constexpr (DEFECT) {
          upper(j);
          lower(j);
        } else {
          pair_terms(j);
        }
      }
    } else {
#pragma unroll 1
      for (int j = 0; j < ND - 1; ++j) upper(j);
#pragma unroll 1
      for (int j = 0; j < ND - 1; ++j) lower(j);
    }
    return acc;
  };
  // The first staple of a pass has no hits to hide behind, so the two waves
  // split it: the staple wave takes the upper term nu = 1 (which gives keep)
  // and the lower terms j >= 1, the site wave the rest, adding its part to
  // the handed-over one.
  constexpr unsigned kAll = (1u << (ND - 1)) - 1u;
  // the two roles run in separate branches, so neither role's registers are
  // held across the other's code; each executes 1 + 2 nd barriers
  if (sitew) {
    if (active && visit) visit[site] = pass;
    LdsChain<R, NC> lc;
    lc.bu = stage(3);
    // the site wave's part of the first staple waits in its own buffer
    if (active) lds_put<R, NC>(lc.bu, build(0, -1, kAll & ~1u, 1u, nullptr));
    __syncthreads();
#pragma unroll 1
    for (int mu = 0; mu < ND; ++mu) {
      if (active) {
        Mat<R, NC> S = lds_get<R, NC>(bS);
        if (mu > 0) {
          // the final U_{mu-1}(x), defect-scaled
          Mat<R, NC> Up = lds_get<R, NC>(lc.bu);
          if constexpr (DEFECT) {
            const R cp = cl_of(x, mu - 1);
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
              for (int k = 0; k < NC; ++k) Up.m[i][k] = scale(cp, Up.m[i][k]);
          }
          acc_mul_ad(S, lds_get<R, NC>(bA), Up);
        } else {
          add_to(S, lds_get<R, NC>(lc.bu));
        }
        lc.start(S, lds_get<R, NC>(bU));
      }
      __syncthreads();
      if (active) {
        const R cl = cl_of(x, mu);
        const unsigned long long n0 = n;
#pragma unroll 1
        for (int hit = 0; hit < nhits; ++hit) {
          Mat<R, NC> r;
          R u;
          propose_law<R, NC, SYM>(seed, site, sweep, (uint32_t)hit, chain, (uint32_t)mu, delta, r, u);
          lc.step(r, u, beta_nc, cl, n);
        }
        lc.apply_pending();
        // a link none of whose hits was accepted is unchanged in memory, so
        // only the lanes that accepted write theirs back (at beta = 6 about
        // one link in ten changes per pass)
        if (n != n0) store_from_lds(f, s, mu, lc.bu);
      }
      __syncthreads();
    }
  } else {
    Mat<R, NC> keep;
    if (active) {
      lds_put<R, NC>(bS, build(0, -1, 1u, kAll & ~1u, &keep));
      lds_put<R, NC>(bU, f.load(s, 0));
    }
    __syncthreads();
#pragma unroll 1
    for (int mu = 0; mu < ND; ++mu) {
      const int m = mu + 1;
      __syncthreads();
      if (active && m < ND) {
        lds_put<R, NC>(bA, adjoint(keep));
        lds_put<R, NC>(bS, build(m, mu, kAll, kAll, &keep));
        lds_put<R, NC>(bU, f.load(s, m));
      }
      __syncthreads();
    }
  }
  return n;
}

// block total of the site waves' accepted hits into *nacc
__device__ __forceinline__ void ws_count(unsigned long long n, unsigned long long *red, unsigned long long *nacc) {
  n = wave_sum_u64(n);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
#pragma unroll
    for (int i = 0; i < kWsBlock / 64; ++i) b += red[i];
    if (b) atomicAdd(nacc, b);
  }
}

// pass2 >= 0: the launch runs the independent passes `pass` of sweep
// `sweep` and `pass2` of sweep `sweep2` (pass_groups, run_schedule), the odd
// logical blocks taking pass2 and counting into nacc2, so that the blocks of
// one XCD hold the same chunk of both sublattices
template <typename R, int NC, bool COMP, int ND, bool ILV, bool SYM = false>
__global__ __launch_bounds__(kWsBlock) void k_metropolis_ws(
    Links<R, NC, COMP> f, Geom g, int pass, int pass2, int nhits, R beta_nc, R delta, uint64_t seed,
    uint32_t sweep, uint32_t sweep2, uint32_t chain, unsigned long long *nacc, unsigned long long *nacc2,
    int *visit) {
  f.nd = ND;
  __shared__ cx<R> stage[4 * 4 * NC * NC * 64];
  __shared__ unsigned long long red[kWsBlock / 64];
  unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  if (pass2 >= 0) {
    if (lb & 1u) {
      pass = pass2;
      sweep = sweep2;
      nacc = nacc2;
    }
    lb >>= 1;
  }
  const long long t = (long long)lb * kWsSites + (long long)((threadIdx.x >> 6) & 3) * 64 + (threadIdx.x & 63);
  const unsigned long long n = ws_pass<R, NC, COMP, ND, ILV, false, SYM>(f, g, pass, t, nhits, beta_nc, delta,
                                                                         seed, sweep, chain, visit, Defect{}, R(1),
                                                                         stage);
  ws_count(n, red, nacc);
}

// Sweep variants (numbers kept stable for tuning caches): 0 one lane per
// site; 4 one lane, low-register staple, >= 2 waves per SIMD; 7 one lane,
// next direction's staple loaded during the hits; 11 one lane, the site's own
// links kept in registers; 12 site waves fed by staple waves; 13 the same with
// each direction's upper and lower staple terms back to back; 14, 15: 12, 13
// with the symmetric SU(3) proposal.  1-3, 5, 6, 8-10 (four or two lanes per
// site, proposals drawn in pairs, LDS-staged staple terms) never won at
// 16^4 - 48^4 and were removed.
template <typename RR, bool CC>
int launch_passes_typed(klft_field_t f, void *links, int variant, int p0, int p1,
                        const klft_metropolis_params *p, uint64_t sweep, int nsweeps, uint32_t chain,
                        unsigned long long *counter, int counter_stride, int *visit) {
  return dispatch_typed<RR, CC>(f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    Links<R, NC, COMP> l;
    l.p = reinterpret_cast<cx<R> *>(links);
    l.V = f->g.V;
    l.nd = f->nd;
    const R beta_nc = (R)(p->beta / (double)NC);
    const unsigned nb = blocks_for(f->g.Vh, kMetroBlock);
    switch (variant) {
      case 0: case 4: case 7: case 11: case 12: case 13: case 14: case 15: break;
      default: return set_error(KLFT_E_ARG, "sweep variant %d is not built", variant);
    }
    // acceptance counter of sweep i of the run (relative to `sweep`)
    auto cnt = [&](int i) { return counter + (long long)i * counter_stride; };
    const unsigned nws = blocks_for(f->g.Vh, kWsSites);
    // launch the site-wave kernel on one pass (pb < 0), or on two
    // independent ones, each with its own sweep and counter
    auto ws = [&](auto kern, int pa, int sa, int pb, int sb) {
      hipLaunchKernelGGL(kern, dim3(pb >= 0 ? 2 * nws : nws), dim3(kWsBlock), 0, stream(), l, f->g, pa, pb,
                         p->nhits, beta_nc, (R)p->delta, p->seed, (uint32_t)(sweep + sa), (uint32_t)(sweep + sb),
                         chain, cnt(sa), cnt(sb), visit);
    };
    auto launch_ws = [&](int pa, int sa, int pb, int sb) -> int {
      switch (variant) {
        case 12: ws(k_metropolis_ws<R, NC, COMP, ND, false>, pa, sa, pb, sb); break;
        case 13: ws(k_metropolis_ws<R, NC, COMP, ND, true>, pa, sa, pb, sb); break;
        default:
          // the symmetric SU(3) proposal: site waves, terms as in 12 / 13
          if constexpr (NC == 3) {
            if (variant == 14)
              ws(k_metropolis_ws<R, NC, COMP, ND, false, true>, pa, sa, pb, sb);
            else
              ws(k_metropolis_ws<R, NC, COMP, ND, true, true>, pa, sa, pb, sb);
          } else {
            return set_error(KLFT_E_ARG, "the symmetric proposal variants are SU(3) only");
          }
      }
      return KLFT_OK;
    };
    // whole sweeps run their independent passes in pairs when both passes'
    // blocks fit on the GPU at once (KLFT_VARIANT=metropolis_pairs=0: never);
    // a run of several sweeps also pairs passes of consecutive sweeps
    // (run_schedule)
    bool pair = false;
    if (variant >= 12 && p0 == 0 && p1 == (1 << ND) && forced_variant("metropolis_pairs") != 0) {
      const void *kern = nullptr;
      switch (variant) {
        case 12: kern = reinterpret_cast<const void *>(&k_metropolis_ws<R, NC, COMP, ND, false>); break;
        case 13: kern = reinterpret_cast<const void *>(&k_metropolis_ws<R, NC, COMP, ND, true>); break;
        case 14: kern = reinterpret_cast<const void *>(&k_metropolis_ws<R, NC, COMP, ND, false, true>); break;
        default: kern = reinterpret_cast<const void *>(&k_metropolis_ws<R, NC, COMP, ND, true, true>); break;
      }
      pair = 2LL * nws <= (long long)cu_count() * blocks_per_cu(kern, kWsBlock);
    }
    if (pair) {
      int rc = KLFT_OK;
      for_each_run_launch(ND, nsweeps, [&](const RunLaunch &L) {
        if (rc) return;
        const int pa = L.pass[0], pb = L.pass[1];
        char label[56];
        if (pb >= 0)
          snprintf(label, sizeof label, "sweep_Metropolis_GaugeField_sublat_%d_%d", pa, pb);
        else
          snprintf(label, sizeof label, "sweep_Metropolis_GaugeField_sublat_%d", pa);  // Metropolis.hpp:161
        ProfScope prof(pb >= 0 ? "metropolis_pass_pair" : "metropolis_pass", label);
        rc = launch_ws(pa, L.sweep[0], pb, pb >= 0 ? L.sweep[1] : L.sweep[0]);
      });
      if (rc) return rc;
    } else {
      for (int i = 0; i < nsweeps; ++i) {
        for (int pass = p0; pass < p1; ++pass) {
          char label[48];
          snprintf(label, sizeof label, "sweep_Metropolis_GaugeField_sublat_%d", pass);  // Metropolis.hpp:161
          ProfScope prof("metropolis_pass", label);
          if (variant >= 12) {
            const int rc = launch_ws(pass, i, -1, i);
            if (rc) return rc;
          } else {
            auto one_lane = [&](auto kern) {
              hipLaunchKernelGGL(kern, dim3(nb), dim3(kMetroBlock), 0, stream(), l, f->g, pass, p->nhits, beta_nc,
                                 (R)p->delta, p->seed, (uint32_t)(sweep + i), chain, cnt(i), visit);
            };
            switch (variant) {
              case 4: one_lane(k_metropolis_pass<R, NC, COMP, ND, true, 2>); break;
              case 7: one_lane(k_metropolis_pipe<R, NC, COMP, ND>); break;
              case 11: one_lane(k_metropolis_own<R, NC, COMP, ND>); break;
              default: one_lane(k_metropolis_pass<R, NC, COMP, ND, false, 1>); break;
            }
          }
        }
      }
    }
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}


}  // namespace klft_metro
