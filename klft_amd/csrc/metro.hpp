// Per-site Metropolis update shared by the single-field sweep and the
// batched / replica sweeps (MetropolisGaugeField::operator(),
// Metropolis.hpp:72-113).
//
// With DEFECT, every link read is scaled by the replica's coupling c when it
// lies on the defect -- direction 0, x0 = 0, x_k in [p_k, p_k + Ld) (mod L_k)
// for k >= 1 (PTBCGaugeField.hpp:186-229) -- and the updated link's own
// factor multiplies dS: dS = -(beta/Nc) c_l (Re Tr(U_new S^c) - Re Tr(U_old S^c)).
#pragma once
#include "device.hpp"

namespace kd {

struct Defect {
  int len;     // Ld
  int pos[3];  // p_1 .. p_{nd-1}
};

template <int ND>
__device__ __forceinline__ bool on_defect(const Geom &g, const Defect &d, const int *x, int mu) {
  if (mu != 0 || x[0] != 0) return false;
#pragma unroll
  for (int k = 1; k < ND; ++k) {
    int r = x[k] - d.pos[k - 1];
    if (r < 0) r += g.L[k];
    if (r >= d.len) return false;
  }
  return true;
}

// coordinates of x + s1 e_m1 + s2 e_m2 (s in {-1,0,1})
template <int ND>
__device__ __forceinline__ void shifted(const Geom &g, const int *x, int m1, int s1, int m2, int s2, int *y) {
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    int w = x[d];
    if (d == m1) w += s1;
    if (d == m2) w += s2;
    y[d] = w < 0 ? w + g.L[d] : (w >= g.L[d] ? w - g.L[d] : w);
  }
}

template <typename R, int NC, bool COMP, int ND, bool DEFECT>
__device__ __forceinline__ Mat<R, NC> load_at(const Links<R, NC, COMP> &f, const Geom &g, const int *y, int mu,
                                              const Defect &d, R c) {
  Mat<R, NC> u = f.load(dev_index<ND>(g, y), mu);
  if constexpr (DEFECT) {
    if (on_defect<ND>(g, d, y, mu)) {
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) u.m[i][j] = scale(c, u.m[i][j]);
    }
  }
  return u;
}

// staple with defect-scaled links (same term order as staple()); LOWREG
// selects the one-term-in-flight staple for a defect-free field
template <typename R, int NC, bool COMP, int ND, bool DEFECT, bool LOWREG = false>
__device__ __forceinline__ Mat<R, NC> staple_c(const Links<R, NC, COMP> &f, const Geom &g, const int *x, int mu,
                                               const Defect &d, R c) {
  if constexpr (!DEFECT) {
    if constexpr (LOWREG)
      return staple_lowreg<R, NC, COMP, ND>(f, g, x, mu);
    else
      return staple<R, NC, COMP, ND>(f, g, x, mu);
  } else {
    Mat<R, NC> t = zero<R, NC>();
    int xp[ND], y[ND];
    shifted<ND>(g, x, mu, 1, -1, 0, xp);
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      shifted<ND>(g, x, nu, 1, -1, 0, y);
      const Mat<R, NC> a = load_at<R, NC, COMP, ND, true>(f, g, xp, nu, d, c);
      const Mat<R, NC> b = load_at<R, NC, COMP, ND, true>(f, g, y, mu, d, c);
      const Mat<R, NC> e = load_at<R, NC, COMP, ND, true>(f, g, x, nu, d, c);
      acc_mul_ad(t, mul_ad(a, b), e);
    }
#pragma unroll
    for (int nu = 0; nu < ND; ++nu) {
      if (nu == mu) continue;
      int ya[ND], yb[ND];
      shifted<ND>(g, x, mu, 1, nu, -1, ya);
      shifted<ND>(g, x, nu, -1, -1, 0, yb);
      const Mat<R, NC> a = load_at<R, NC, COMP, ND, true>(f, g, ya, nu, d, c);
      const Mat<R, NC> b = load_at<R, NC, COMP, ND, true>(f, g, yb, mu, d, c);
      const Mat<R, NC> e = load_at<R, NC, COMP, ND, true>(f, g, yb, nu, d, c);
      acc_mul(t, mul_dd(a, b), e);
    }
    return t;
  }
}

// The Markov chain of one link through its hits.  dS = -(beta/Nc) c_l
// (Re Tr(r M) - Re Tr M) with M = S U (Tr is cyclic), so a hit costs one
// trace; after an accept M becomes M r, so the staple itself is dead once M is
// formed (M r and S restoreSUN(U r) differ by rounding only).  An accepted
// proposal is kept in P and multiplied into U (with restoreSUN) only when the
// next accept comes or the hits end: U goes through exactly the reference's
// sequence restoreSUN(U r1), restoreSUN(U r2), ... but a wave pays for that
// work once per link instead of at every hit where any of its 64 lanes
// accepts (about half the hits at 1 % acceptance).
// accept iff dS < 0 or u < e^{-dS} (Metropolis.hpp:96-105)
template <typename R>
__device__ __forceinline__ bool metro_accept(R dS, R u) {
  bool accept = dS < R(0);
  if (!accept) {
    // e^x > 1 + x + x^2/2 + x^3/6 for x > 0, so u * that cubic >= 1 means
    // u > e^{-dS}: a rejection found without the exponential.  dS > 1/2
    // keeps the margin (> 0.1 %) far above rounding, so the decision is
    // always the one u < e^{-dS} gives.
    const R p3 = R(1) + dS * (R(1) + dS * (R(0.5) + dS * (R(1) / R(6))));
    if (!(dS > R(0.5) && u * p3 >= R(1))) accept = u < exp(-dS);
  }
  return accept;
}

template <typename R, int NC>
struct LinkChain {
  Mat<R, NC> U, M, P;
  R trM;
  bool pending;

  __device__ __forceinline__ void start(const Mat<R, NC> &S, const Mat<R, NC> &U0) {
    U = U0;
    M = mul(S, U);
    trM = retrace(M);
    pending = false;
  }
  __device__ __forceinline__ void apply_pending() {
    if (pending) {
      U = mul(U, P);
      reunitarize(U);
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

// update all directions of the site at position t of sublattice `pass`;
// returns the number of accepted hits
template <typename R, int NC, bool COMP, int ND, bool DEFECT, bool LOWREG = false>
__device__ __forceinline__ unsigned long long metro_site(const Links<R, NC, COMP> &f, const Geom &g, int pass,
                                                         long long t, int nhits, R beta_nc, R delta,
                                                         uint64_t seed, uint32_t sweep, uint32_t chain,
                                                         const Defect &d, R c, int *visit) {
  unsigned long long n = 0;
  int x[ND];
  sub_coords<ND>(g, pass, t, x);  // index_odd_even
  const long long s = (long long)pass * g.Vh + t;
  const uint32_t site = (uint32_t)lex_index<ND>(g, x);
  if (visit) visit[site] = pass;
#pragma unroll 1
  for (int mu = 0; mu < ND; ++mu) {
    const Mat<R, NC> S = staple_c<R, NC, COMP, ND, DEFECT, LOWREG>(f, g, x, mu, d, c);
    R cl = R(1);
    if constexpr (DEFECT) cl = on_defect<ND>(g, d, x, mu) ? c : R(1);
    LinkChain<R, NC> lc;
    lc.start(S, f.load(s, mu));
    const unsigned long long n0 = n;
#pragma unroll 1
    for (int hit = 0; hit < nhits; ++hit) {
      Mat<R, NC> r;
      R u;
      propose<R, NC>(seed, site, sweep, (uint32_t)hit, chain, (uint32_t)mu, delta, r, u);
      lc.step(r, u, beta_nc, cl, n);
    }
    lc.apply_pending();
    // only a link with an accepted hit has changed
    if (n != n0) f.store(s, mu, lc.U);
  }
  return n;
}

// ---------------------------------------------------------------------------
// Site update keeping the site's own links in registers (a defect-free
// field).  The staple of U_mu(x) reads U_nu(x) for every nu != mu, so the
// one-lane kernel loads each own link four times per pass (once to update
// it, three times in other directions' staples), 16 of its 76 link loads
// per site; here they are loaded once.  The direction loop is unrolled so
// the links stay in named registers.
// ---------------------------------------------------------------------------
template <typename R, int NC, bool COMP, int ND, int MU>
__device__ __forceinline__ void metro_dir_own(const Links<R, NC, COMP> &f, const Geom &g, const int *x,
                                              long long s, uint32_t site, int nhits, R beta_nc, R delta,
                                              uint64_t seed, uint32_t sweep, uint32_t chain, Mat<R, NC> *Uown,
                                              unsigned long long &n) {
  Mat<R, NC> S = zero<R, NC>();
  const long long s_xpmu = idx_shift<ND>(g, x, MU, 1, -1, 0);
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == MU) continue;
    const Mat<R, NC> ab = mul_ad(f.load(s_xpmu, nu), f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), MU));
    acc_mul_ad(S, ab, Uown[nu]);
  }
#pragma unroll 1
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == MU) continue;
    const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
    const Mat<R, NC> ab = mul_dd(f.load(idx_shift<ND>(g, x, MU, 1, nu, -1), nu), f.load(s_xmnu, MU));
    acc_mul(S, ab, f.load(s_xmnu, nu));
  }
  LinkChain<R, NC> lc;
  lc.start(S, Uown[MU]);
#pragma unroll 1
  for (int hit = 0; hit < nhits; ++hit) {
    Mat<R, NC> r;
    R u;
    propose<R, NC>(seed, site, sweep, (uint32_t)hit, chain, (uint32_t)MU, delta, r, u);
    lc.step(r, u, beta_nc, R(1), n);
  }
  lc.apply_pending();
  if (nhits > 0) {
    Uown[MU] = lc.U;
    f.store(s, MU, lc.U);
  }
}

template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ unsigned long long metro_site_own(const Links<R, NC, COMP> &f, const Geom &g,
                                                             int pass, long long t, int nhits, R beta_nc,
                                                             R delta, uint64_t seed, uint32_t sweep,
                                                             uint32_t chain, int *visit) {
  unsigned long long n = 0;
  int x[ND];
  sub_coords<ND>(g, pass, t, x);
  const long long s = (long long)pass * g.Vh + t;
  const uint32_t site = (uint32_t)lex_index<ND>(g, x);
  if (visit) visit[site] = pass;
  Mat<R, NC> Uown[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) Uown[d] = f.load(s, d);
  metro_dir_own<R, NC, COMP, ND, 0>(f, g, x, s, site, nhits, beta_nc, delta, seed, sweep, chain, Uown, n);
  metro_dir_own<R, NC, COMP, ND, 1>(f, g, x, s, site, nhits, beta_nc, delta, seed, sweep, chain, Uown, n);
  if constexpr (ND > 2)
    metro_dir_own<R, NC, COMP, ND, 2>(f, g, x, s, site, nhits, beta_nc, delta, seed, sweep, chain, Uown, n);
  if constexpr (ND > 3)
    metro_dir_own<R, NC, COMP, ND, 3>(f, g, x, s, site, nhits, beta_nc, delta, seed, sweep, chain, Uown, n);
  return n;
}

// ---------------------------------------------------------------------------
// Site update with the next direction's staple built during the current
// direction's hits (a defect-free field).  At 32^4 a pass has one wave per
// SIMD, so the staple's link loads, issued in bursts between hit phases,
// leave the memory system idle while the SIMDs compute hits and vice versa.
// Here hit h of direction mu first issues the three link loads of staple
// term h of direction m = mu + 1, then runs the hit, then folds the term in,
// so the loads are in flight while the hit computes.
//
// Which links of m's staple can change while mu is updated?  Only U_mu(x)
// (every other link it reads belongs to another sublattice or was finished
// earlier in this site's direction loop).  The upper term with nu = mu is
// therefore formed after the hits, from the final U_mu(x).  The staple's terms
// are summed in a different order from staple() (that one term last), so M
// differs from the one-lane kernel's by rounding only.
// ---------------------------------------------------------------------------
template <typename R, int NC, bool COMP, int ND>
struct TermLoads {
  Mat<R, NC> a, b, c;
};

// links of staple term t of link (x, m): t < ND-1 upper (nu = the t-th
// direction other than m), else lower; the upper term with nu == defer is
// skipped here
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ void term_load(const Links<R, NC, COMP> &f, const Geom &g, const int *x, long long s,
                                          int m, int t, int defer, TermLoads<R, NC, COMP, ND> &L) {
  const bool upper = t < ND - 1;
  const int j = upper ? t : t - (ND - 1);
  const int nu = j < m ? j : j + 1;
  if (upper) {
    if (nu == defer) return;
    L.a = f.load(idx_shift<ND>(g, x, m, 1, -1, 0), nu);
    L.b = f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), m);
    L.c = f.load(s, nu);
  } else {
    const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
    L.a = f.load(idx_shift<ND>(g, x, m, 1, nu, -1), nu);
    L.b = f.load(s_xmnu, m);
    L.c = f.load(s_xmnu, nu);
  }
}

template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ void term_fold(int m, int t, int defer, const TermLoads<R, NC, COMP, ND> &L,
                                          Mat<R, NC> &acc) {
  const bool upper = t < ND - 1;
  const int j = upper ? t : t - (ND - 1);
  const int nu = j < m ? j : j + 1;
  if (upper) {
    if (nu != defer) acc_mul_ad(acc, mul_ad(L.a, L.b), L.c);
  } else {
    acc_mul(acc, mul_dd(L.a, L.b), L.c);
  }
}

template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ unsigned long long metro_site_pipe(const Links<R, NC, COMP> &f, const Geom &g,
                                                              int pass, long long t, int nhits, R beta_nc,
                                                              R delta, uint64_t seed, uint32_t sweep,
                                                              uint32_t chain, int *visit) {
  constexpr int NT = 2 * (ND - 1);
  unsigned long long n = 0;
  int x[ND];
  sub_coords<ND>(g, pass, t, x);
  const long long s = (long long)pass * g.Vh + t;
  const uint32_t site = (uint32_t)lex_index<ND>(g, x);
  if (visit) visit[site] = pass;
  Mat<R, NC> S = staple_lowreg<R, NC, COMP, ND>(f, g, x, 0);
#pragma unroll 1
  for (int mu = 0; mu < ND; ++mu) {
    LinkChain<R, NC> lc;
    lc.start(S, f.load(s, mu));
    const int m = mu + 1;
    const bool next = m < ND;
    Mat<R, NC> acc = zero<R, NC>();
    int k = 0;
    const unsigned long long n0 = n;
#pragma unroll 1
    for (int hit = 0; hit < nhits; ++hit) {
      const bool fold = next && k < NT;
      Mat<R, NC> r;
      R u;
      TermLoads<R, NC, COMP, ND> L;
      if (fold) term_load<R, NC, COMP, ND>(f, g, x, s, m, k, mu, L);
      propose<R, NC>(seed, site, sweep, (uint32_t)hit, chain, (uint32_t)mu, delta, r, u);
      lc.step(r, u, beta_nc, R(1), n);
      if (fold) term_fold<R, NC, COMP, ND>(m, k++, mu, L, acc);
    }
    if (next) {
#pragma unroll 1
      for (; k < NT; ++k) {
        TermLoads<R, NC, COMP, ND> L;
        term_load<R, NC, COMP, ND>(f, g, x, s, m, k, mu, L);
        term_fold<R, NC, COMP, ND>(m, k, mu, L, acc);
      }
    }
    lc.apply_pending();
    if (n != n0) f.store(s, mu, lc.U);
    if (next) {
      // the upper term through U_mu(x), now final
      const Mat<R, NC> ab = mul_ad(f.load(idx_shift<ND>(g, x, m, 1, -1, 0), mu),
                                   f.load(idx_shift<ND>(g, x, mu, 1, -1, 0), m));
      acc_mul_ad(acc, ab, lc.U);
      S = acc;
    }
  }
  return n;
}

}  // namespace kd
