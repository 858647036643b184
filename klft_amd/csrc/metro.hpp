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
      add_to(t, mul_ad(mul_ad(a, b), e));
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
      add_to(t, mul(mul_dd(a, b), e));
    }
    return t;
  }
}

// one Markov step of a link: dS = -(beta/Nc) c_l (Re Tr(r M) - Re Tr M), M = S U
template <typename R, int NC>
__device__ __forceinline__ void metro_step(const Mat<R, NC> &r, R u, R beta_nc, R cl, Mat<R, NC> &U,
                                           Mat<R, NC> &M, R &trM, unsigned long long &n) {
  const R dS = cl * (-beta_nc * (retr_mul(r, M) - trM));
  bool accept = dS < R(0);
  if (!accept) accept = u < exp(-dS);
  if (accept) {
    U = mul(U, r);
    reunitarize(U);
    M = mul(M, r);  // S U r: the staple is not kept through the hits
    trM = retrace(M);
    ++n;
  }
}

// update all directions of the site at position t of sublattice `pass`;
// returns the number of accepted hits.  HB proposals are drawn before their
// Markov steps (they do not depend on the field), so their generator rounds
// overlap; a proposal past nHits is drawn and discarded.
template <typename R, int NC, bool COMP, int ND, bool DEFECT, bool LOWREG = false, int HB = 1>
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
    Mat<R, NC> U = f.load(s, mu);
    Mat<R, NC> M = mul(S, U);  // Re Tr(U r S) = Re Tr(r M)
    R trM = retrace(M);
#pragma unroll 1
    for (int hit = 0; hit < nhits; hit += HB) {
      Mat<R, NC> r[HB];
      R u[HB];
#pragma unroll
      for (int k = 0; k < HB; ++k)
        propose<R, NC>(seed, site, sweep, (uint32_t)(hit + k), chain, (uint32_t)mu, delta, r[k], u[k]);
#pragma unroll
      for (int k = 0; k < HB; ++k)
        if (k == 0 || hit + k < nhits) metro_step(r[k], u[k], beta_nc, cl, U, M, trM, n);
    }
    if (nhits > 0) f.store(s, mu, U);
  }
  return n;
}

}  // namespace kd
