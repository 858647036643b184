/*
 * klft_oracle.c -- FP64 CPU restatement of the klft gauge-update and
 * measurement path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * Each function cites the reference file:line whose algorithm it restates
 * (paths relative to the aniketsen/klft source tree).  Arithmetic follows
 * the reference operation by operation: complex products in the naive
 * (re*re - im*im, re*im + im*re) form of Kokkos::complex, matrix products
 * accumulated k = 0..Nc-1 in order (SUN.hpp:28-44), products associated
 * left to right.  Build with -ffp-contract=off so no FMA changes rounding.
 *
 * Random numbers (the reference uses Kokkos' stateful XorShift pools, whose
 * streams depend on backend and thread count; they are replaced here and in
 * the product by a counter-based generator so both draw identical numbers):
 *   Philox4x32-10 (Salmon et al., SC'11), key = (seed lo32, seed hi32),
 *   counter of Philox block j of stream (tag, c0, c1, c2, chain, mu) =
 *     { c0, c1, c2, tag<<28 | chain<<18 | mu<<16 | j }.
 *   A double in [0,1) is built from two consecutive 32-bit words (a, b) as
 *   ((a>>5) * 2^26 + (b>>6)) * 2^-53.
 * Streams:
 *   tag 1  Metropolis   c0 = lexicographic site, c1 = sweep, c2 = hit
 *   tag 2  momenta      c0 = site, c1 = trajectory, c2 = 0
 *   tag 3  random-algebra start  c0 = site, c1 = c2 = 0
 *   tag 7  HMC accept   c0 = 0, c1 = trajectory, c2 = 0, mu = 0
 */
#include "klft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846

typedef struct {
  double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
  cplx c = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return c;
}
static inline cplx cadd(cplx a, cplx b) {
  cplx c = {a.re + b.re, a.im + b.im};
  return c;
}
static inline cplx csub(cplx a, cplx b) {
  cplx c = {a.re - b.re, a.im - b.im};
  return c;
}
static inline cplx cconj(cplx a) {
  cplx c = {a.re, -a.im};
  return c;
}
static inline cplx cdivr(cplx a, double r) {
  cplx c = {a.re / r, a.im / r};
  return c;
}
static inline cplx crmul(double r, cplx a) {
  cplx c = {r * a.re, r * a.im};
  return c;
}

#define MAXN 3
typedef struct {
  cplx m[MAXN][MAXN];
} sun;

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                            */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

typedef struct {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t hi; /* high bits of ctr[3] (tag, chain, mu) */
  uint32_t block;
  uint32_t buf[4];
  int pos;
  int narrow; /* uniforms of 24 bits from one word (the FP32 draws) */
} stream;

static void stream_init(stream *s, uint64_t seed, uint32_t tag, uint32_t c0,
                        uint32_t c1, uint32_t c2, uint32_t chain, uint32_t mu) {
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  s->ctr[0] = c0;
  s->ctr[1] = c1;
  s->ctr[2] = c2;
  s->hi = (tag << 28) | (chain << 18) | (mu << 16);
  s->block = 0;
  s->pos = 4;
  s->narrow = 0;
}

static uint32_t next_word(stream *s) {
  if (s->pos == 4) {
    s->ctr[3] = s->hi | (s->block & 0xffffu);
    orc_philox4x32_10(s->ctr, s->key, s->buf);
    s->block++;
    s->pos = 0;
  }
  return s->buf[s->pos++];
}

static double next_u01(stream *s) {
  /* FP32 links draw one word per uniform, 24 bits: exact in float and double */
  if (s->narrow) return (double)(next_word(s) >> 8) * (1.0 / 16777216.0);
  uint32_t a = next_word(s) >> 5;
  uint32_t b = next_word(s) >> 6;
  return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

/* Kokkos drand(start, end) semantics: start + (end - start) * u */
static double drand(stream *s, double lo, double hi) {
  return lo + (hi - lo) * next_u01(s);
}

/* Box-Muller pair; the first argument is taken as 1 - u so log is finite */
static void normal_pair(stream *s, double *z0, double *z1) {
  double u1 = 1.0 - next_u01(s);
  double u2 = next_u01(s);
  double r = sqrt(-2.0 * log(u1));
  double t = 2.0 * PI * u2;
  *z0 = r * cos(t);
  *z1 = r * sin(t);
}

static void normals(stream *s, int n, double *out) {
  for (int i = 0; i < n; i += 2) {
    double z0, z1;
    normal_pair(s, &z0, &z1);
    out[i] = z0;
    if (i + 1 < n) out[i + 1] = z1;
  }
}

void orc_stream_doubles(uint64_t seed, uint32_t tag, uint32_t c0, uint32_t c1,
                        uint32_t c2, uint32_t chain, uint32_t mu, int n, double *out) {
  stream s;
  stream_init(&s, seed, tag, c0, c1, c2, chain, mu);
  for (int i = 0; i < n; ++i) out[i] = next_u01(&s);
}

/* ------------------------------------------------------------------------ */
/* SU(N) algebra (SUN.hpp)                                                  */
/* ------------------------------------------------------------------------ */
static inline sun zero_sun(void) {
  sun z;
  memset(&z, 0, sizeof z);
  return z;
}
static inline sun identity_sun(int nc) {
  sun z = zero_sun();
  for (int i = 0; i < nc; ++i) z.m[i][i].re = 1.0;
  return z;
}

/* SUN.hpp:28-44 */
static inline sun mm(int nc, const sun *a, const sun *b) {
  sun c;
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      c.m[i][j] = cmul(a->m[i][0], b->m[0][j]);
      for (int k = 1; k < nc; ++k) c.m[i][j] = cadd(c.m[i][j], cmul(a->m[i][k], b->m[k][j]));
    }
  return c;
}
/* SUN.hpp:134-146 (klft's conj is the Hermitian conjugate) */
static inline sun dag(int nc, const sun *a) {
  sun c;
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) c.m[i][j] = cconj(a->m[j][i]);
  return c;
}
/* SUN.hpp:67-78 */
static inline void addto(int nc, sun *a, const sun *b) {
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) a->m[i][j] = cadd(a->m[i][j], b->m[i][j]);
}
/* SUN.hpp:148-157 */
static inline cplx trace(int nc, const sun *a) {
  cplx c = {0.0, 0.0};
  for (int i = 0; i < nc; ++i) c = cadd(c, a->m[i][i]);
  return c;
}

/* SUN.hpp:256-340, restoreSUN<1,2,3> (in place) */
static void restore_sun(int nc, sun *a) {
  if (nc == 1) {
    double n = sqrt(a->m[0][0].re * a->m[0][0].re + a->m[0][0].im * a->m[0][0].im);
    a->m[0][0] = cdivr(a->m[0][0], n);
  } else if (nc == 2) {
    double n = sqrt(a->m[0][0].re * a->m[0][0].re + a->m[0][0].im * a->m[0][0].im +
                    a->m[0][1].re * a->m[0][1].re + a->m[0][1].im * a->m[0][1].im);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) a->m[i][j] = cdivr(a->m[i][j], n);
  } else {
    cplx s0 = cadd(cadd(cmul(cconj(a->m[0][0]), a->m[0][0]), cmul(cconj(a->m[0][1]), a->m[0][1])),
                   cmul(cconj(a->m[0][2]), a->m[0][2]));
    cplx s1 = cadd(cadd(cmul(cconj(a->m[1][0]), a->m[1][0]), cmul(cconj(a->m[1][1]), a->m[1][1])),
                   cmul(cconj(a->m[1][2]), a->m[1][2]));
    double n0 = sqrt(s0.re), n1 = sqrt(s1.re);
    for (int j = 0; j < 3; ++j) {
      a->m[0][j] = cdivr(a->m[0][j], n0);
      a->m[1][j] = cdivr(a->m[1][j], n1);
    }
    sun *c = a;
    c->m[2][0] = cconj(csub(cmul(c->m[0][1], c->m[1][2]), cmul(c->m[0][2], c->m[1][1])));
    c->m[2][1] = cconj(csub(cmul(c->m[0][2], c->m[1][0]), cmul(c->m[0][0], c->m[1][2])));
    c->m[2][2] = cconj(csub(cmul(c->m[0][0], c->m[1][1]), cmul(c->m[0][1], c->m[1][0])));
    c->m[1][0] = cconj(csub(cmul(c->m[2][1], c->m[0][2]), cmul(c->m[2][2], c->m[0][1])));
    c->m[1][1] = cconj(csub(cmul(c->m[2][2], c->m[0][0]), cmul(c->m[2][0], c->m[0][2])));
    c->m[1][2] = cconj(csub(cmul(c->m[2][0], c->m[0][1]), cmul(c->m[2][1], c->m[0][0])));
  }
}

/* SUN.hpp:163-251, randSUN<1,2,3>; draws come from the stream in the
 * reference's order, including the rejection loops of randSUN<3>. */
static void rand_sun(int nc, stream *s, double delta, sun *r) {
  *r = zero_sun();
  if (nc == 1) {
    double th = drand(s, -delta * PI, delta * PI);
    r->m[0][0].re = cos(th);
    r->m[0][0].im = sin(th);
  } else if (nc == 2) {
    double alpha = drand(s, 0.0, delta * 2 * PI);
    double u = drand(s, -1.0, 1.0);
    double theta = drand(s, 0.0, 2.0 * PI);
    double sa = sin(alpha);
    double n1 = sqrt(1.0 - u * u) * sin(theta);
    double n2 = sqrt(1.0 - u * u) * cos(theta);
    r->m[0][0].re = cos(alpha);
    r->m[0][0].im = u * sa;
    r->m[0][1].re = n1 * sa;
    r->m[0][1].im = n2 * sa;
    r->m[1][0].re = -r->m[0][1].re;
    r->m[1][0].im = r->m[0][1].im;
    r->m[1][1].re = r->m[0][0].re;
    r->m[1][1].im = -r->m[0][0].im;
  } else {
    double r1[6], r2[6], norm, fact;
    cplx z1[3], z2[3], z;
    for (;;) {
      for (int i = 0; i < 6; ++i) r1[i] = drand(s, 0.0, delta);
      norm = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2] + r1[3] * r1[3] + r1[4] * r1[4] +
                  r1[5] * r1[5]);
      if (1.0 != (1.0 + norm)) break;
    }
    fact = 1.0 / norm;
    for (int i = 0; i < 3; ++i) {
      cplx t = {r1[2 * i], r1[2 * i + 1]};
      z1[i] = crmul(fact, t);
    }
    for (;;) {
      for (;;) {
        for (int i = 0; i < 6; ++i) r2[i] = drand(s, 0.0, delta);
        norm = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2] + r2[3] * r2[3] +
                    r2[4] * r2[4] + r2[5] * r2[5]);
        if (1.0 != (1.0 + norm)) break;
      }
      fact = 1.0 / norm;
      for (int i = 0; i < 3; ++i) {
        cplx t = {r2[2 * i], r2[2 * i + 1]};
        z2[i] = crmul(fact, t);
      }
      z = cadd(cadd(cmul(cconj(z1[0]), z2[0]), cmul(cconj(z1[1]), z2[1])), cmul(cconj(z1[2]), z2[2]));
      for (int i = 0; i < 3; ++i) z2[i] = csub(z2[i], cmul(z, z1[i]));
      norm = sqrt(z2[0].re * z2[0].re + z2[0].im * z2[0].im + z2[1].re * z2[1].re +
                  z2[1].im * z2[1].im + z2[2].re * z2[2].re + z2[2].im * z2[2].im);
      if (1.0 != (1.0 + norm)) break;
    }
    fact = 1.0 / norm;
    for (int i = 0; i < 3; ++i) {
      cplx t = z2[i];
      z2[i] = cmul(t, (cplx){fact, 0.0});
    }
    for (int i = 0; i < 3; ++i) {
      r->m[0][i] = z1[i];
      r->m[1][i] = z2[i];
    }
    r->m[2][0] = cconj(csub(cmul(z1[1], z2[2]), cmul(z1[2], z2[1])));
    r->m[2][1] = cconj(csub(cmul(z1[2], z2[0]), cmul(z1[0], z2[2])));
    r->m[2][2] = cconj(csub(cmul(z1[0], z2[1]), cmul(z1[1], z2[0])));
  }
}

/* ------------------------------------------------------------------------ */
/* exp(i Q), Q = sum_a v_a T_a                                              */
/* ------------------------------------------------------------------------ */
static sun hermitian_from_components(int nc, const double *v) {
  sun q = zero_sun();
  if (nc == 1) {
    q.m[0][0].re = v[0];
  } else if (nc == 2) { /* T = (sigma_x, sigma_y, sigma_z) */
    q.m[0][0].re = v[2];
    q.m[1][1].re = -v[2];
    q.m[0][1] = (cplx){v[0], -v[1]};
    q.m[1][0] = (cplx){v[0], v[1]};
  } else { /* T = Gell-Mann lambda_1..lambda_8, Tr(T_a T_b) = 2 delta_ab */
    const double r3 = 1.0 / sqrt(3.0);
    q.m[0][0].re = v[2] + v[7] * r3;
    q.m[1][1].re = -v[2] + v[7] * r3;
    q.m[2][2].re = -2.0 * v[7] * r3;
    q.m[0][1] = (cplx){v[0], -v[1]};
    q.m[1][0] = (cplx){v[0], v[1]};
    q.m[0][2] = (cplx){v[3], -v[4]};
    q.m[2][0] = (cplx){v[3], v[4]};
    q.m[1][2] = (cplx){v[5], -v[6]};
    q.m[2][1] = (cplx){v[5], v[6]};
  }
  return q;
}

/* AdjointGroup.hpp:170-173 (U(1)), :98-105 (SU(2), with the alpha = 0 case
 * returning the identity instead of dividing by zero).  SU(3) has no
 * reference counterpart; it uses a Taylor series with scaling and squaring
 * (an independent method from the product's Cayley-Hamilton form). */
static sun exp_iq(int nc, const double *v) {
  sun e = zero_sun();
  if (nc == 1) {
    e.m[0][0].re = cos(v[0]);
    e.m[0][0].im = sin(v[0]);
    return e;
  }
  if (nc == 2) {
    double alpha = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (alpha == 0.0) return identity_sun(2);
    double n[3] = {v[0] / alpha, v[1] / alpha, v[2] / alpha};
    double sa = sin(alpha);
    e.m[0][0] = (cplx){cos(alpha), n[2] * sa};
    e.m[0][1] = (cplx){n[1] * sa, n[0] * sa};
    e.m[1][0] = (cplx){-n[1] * sa, n[0] * sa};
    e.m[1][1] = (cplx){cos(alpha), -n[2] * sa};
    return e;
  }
  sun q = hermitian_from_components(3, v);
  /* A = i Q */
  sun a;
  double norm = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a.m[i][j] = (cplx){-q.m[i][j].im, q.m[i][j].re};
      norm += a.m[i][j].re * a.m[i][j].re + a.m[i][j].im * a.m[i][j].im;
    }
  norm = sqrt(norm);
  int sq = 0;
  while (norm > 0.125) {
    norm *= 0.5;
    sq++;
  }
  double scale = ldexp(1.0, -sq);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a.m[i][j] = crmul(scale, a.m[i][j]);
  /* Horner: E = I + A/K (I + A/(K-1) (...)) */
  e = identity_sun(3);
  for (int k = 20; k >= 1; --k) {
    sun t = mm(3, &a, &e);
    double f = 1.0 / (double)k;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) t.m[i][j] = crmul(f, t.m[i][j]);
    for (int i = 0; i < 3; ++i) t.m[i][i].re += 1.0;
    e = t;
  }
  for (int s = 0; s < sq; ++s) e = mm(3, &e, &e);
  return e;
}

/* ------------------------------------------------------------------------ */
/* single-matrix entry points (interleaved re/im, row major)                */
/* ------------------------------------------------------------------------ */
static sun load_m(int nc, const double *p) {
  sun a = zero_sun();
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      a.m[i][j].re = p[2 * (i * nc + j)];
      a.m[i][j].im = p[2 * (i * nc + j) + 1];
    }
  return a;
}
static void store_m(int nc, const sun *a, double *p) {
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      p[2 * (i * nc + j)] = a->m[i][j].re;
      p[2 * (i * nc + j) + 1] = a->m[i][j].im;
    }
}
void orc_matmul(int nc, const double *a, const double *b, double *c) {
  sun A = load_m(nc, a), B = load_m(nc, b);
  sun C = mm(nc, &A, &B);
  store_m(nc, &C, c);
}
void orc_dagger(int nc, const double *a, double *c) {
  sun A = load_m(nc, a);
  sun C = dag(nc, &A);
  store_m(nc, &C, c);
}
void orc_restore_sun(int nc, double *a) {
  sun A = load_m(nc, a);
  restore_sun(nc, &A);
  store_m(nc, &A, a);
}
int orc_rand_sun(int nc, double delta, uint64_t seed, uint32_t sweep, uint32_t site,
                 uint32_t mu, uint32_t hit, double *r) {
  stream s;
  stream_init(&s, seed, 1, site, sweep, hit, 0, mu);
  sun R;
  rand_sun(nc, &s, delta, &R);
  store_m(nc, &R, r);
  return (int)(s.block * 4 - (4 - s.pos));
}
void orc_exp_iq(int nc, const double *v, double *out) {
  sun e = exp_iq(nc, v);
  store_m(nc, &e, out);
}

/* ------------------------------------------------------------------------ */
/* lattice helpers (IndexHelper.hpp:29-93)                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nd, nc;
  int L[4];
  long V;
} lat;

static lat mklat(int nd, int nc, const int *dims) {
  lat l;
  l.nd = nd;
  l.nc = nc;
  l.V = 1;
  for (int d = 0; d < 4; ++d) l.L[d] = d < nd ? dims[d] : 1;
  for (int d = 0; d < nd; ++d) l.V *= dims[d];
  return l;
}
static inline void coords(const lat *l, long site, int *x) {
  for (int d = l->nd - 1; d >= 0; --d) {
    x[d] = (int)(site % l->L[d]);
    site /= l->L[d];
  }
}
static inline long lex(const lat *l, const int *x) {
  long s = 0;
  for (int d = 0; d < l->nd; ++d) s = s * l->L[d] + x[d];
  return s;
}
/* shift_index_plus / shift_index_minus */
static inline long shift(const lat *l, long site, int d, int by) {
  int x[4];
  coords(l, site, x);
  x[d] = ((x[d] + by) % l->L[d] + l->L[d]) % l->L[d];
  return lex(l, x);
}

static inline sun link(const lat *l, const double *U, long site, int mu) {
  return load_m(l->nc, U + 2 * ((size_t)site * l->nd + mu) * l->nc * l->nc);
}
static inline void set_link(const lat *l, double *U, long site, int mu, const sun *a) {
  store_m(l->nc, a, U + 2 * ((size_t)site * l->nd + mu) * l->nc * l->nc);
}
/* link scaled by the PTBC defect factor (PTBCGaugeField.hpp:209-229) */
static inline sun link_c(const lat *l, const double *U, const double *C, long site, int mu) {
  sun a = link(l, U, site, mu);
  if (C) {
    double c = C[(size_t)site * l->nd + mu];
    for (int i = 0; i < l->nc; ++i)
      for (int j = 0; j < l->nc; ++j) {
        a.m[i][j].re *= c;
        a.m[i][j].im *= c;
      }
  }
  return a;
}

/* deterministic pairwise sum */
static double psum(const double *a, long n) {
  if (n <= 16) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i];
    return s;
  }
  long h = n / 2;
  return psum(a, h) + psum(a + h, n - h);
}

/* ------------------------------------------------------------------------ */
/* staple (GaugeField.hpp:187-237; 3D :389-433; 2D :579-618)                */
/* ------------------------------------------------------------------------ */
static sun staple_c(const lat *l, const double *U, const double *C, long site, int mu) {
  int nc = l->nc;
  sun t = zero_sun();
  long xpmu = shift(l, site, mu, 1);
  for (int nu = 0; nu < l->nd; ++nu) {
    if (nu == mu) continue;
    long xpnu = shift(l, site, nu, 1);
    sun a = link_c(l, U, C, xpmu, nu);
    sun b = link_c(l, U, C, xpnu, mu);
    sun c = link_c(l, U, C, site, nu);
    sun bd = dag(nc, &b), cd = dag(nc, &c);
    sun p = mm(nc, &a, &bd);
    sun q = mm(nc, &p, &cd);
    addto(nc, &t, &q);
  }
  for (int nu = 0; nu < l->nd; ++nu) {
    if (nu == mu) continue;
    long xpmumnu = shift(l, xpmu, nu, -1);
    long xmnu = shift(l, site, nu, -1);
    sun a = link_c(l, U, C, xpmumnu, nu);
    sun b = link_c(l, U, C, xmnu, mu);
    sun c = link_c(l, U, C, xmnu, nu);
    sun ad = dag(nc, &a), bd = dag(nc, &b);
    sun p = mm(nc, &ad, &bd);
    sun q = mm(nc, &p, &c);
    addto(nc, &t, &q);
  }
  return t;
}

void orc_staple(int nd, int nc, const int *dims, const double *U, long site, int mu, double *S) {
  lat l = mklat(nd, nc, dims);
  sun s = staple_c(&l, U, NULL, site, mu);
  store_m(nc, &s, S);
}

/* Gauge_Util.hpp:27-97 */
void orc_staple_field(int nd, int nc, const int *dims, const double *U, double *S) {
  lat l = mklat(nd, nc, dims);
#pragma omp parallel for schedule(static)
  for (long site = 0; site < l.V; ++site)
    for (int mu = 0; mu < nd; ++mu) {
      sun s = staple_c(&l, U, NULL, site, mu);
      set_link(&l, S, site, mu, &s);
    }
}

/* ------------------------------------------------------------------------ */
/* plaquette (GaugePlaquette.hpp:52-135)                                    */
/* ------------------------------------------------------------------------ */
static double plaquette_impl(const lat *l, const double *U, const double *C, int normalize,
                             double *per_site) {
  int nc = l->nc, nd = l->nd;
  double *re = (double *)malloc(sizeof(double) * l->V);
#pragma omp parallel for schedule(static)
  for (long site = 0; site < l->V; ++site) {
    cplx t = {0.0, 0.0};
    for (int mu = 0; mu < nd; ++mu)
      for (int nu = 0; nu < nd; ++nu) {
        if (nu <= mu) continue;
        sun a = link_c(l, U, C, site, mu);
        sun b = link_c(l, U, C, shift(l, site, mu, 1), nu);
        sun c = link_c(l, U, C, site, nu);
        sun d = link_c(l, U, C, shift(l, site, nu, 1), mu);
        sun lmu = mm(nc, &a, &b);
        sun lnu = mm(nc, &c, &d);
        for (int c1 = 0; c1 < nc; ++c1)
          for (int c2 = 0; c2 < nc; ++c2) t = cadd(t, cmul(lmu.m[c1][c2], cconj(lnu.m[c1][c2])));
      }
    re[site] = t.re;
  }
  double sum = psum(re, l->V);
  if (per_site) memcpy(per_site, re, sizeof(double) * l->V);
  free(re);
  if (normalize) {
    double norm = 1.0;
    for (int d = 0; d < nd; ++d) norm *= (double)l->L[d];
    norm *= (double)((nd * (nd - 1) / 2) * nc);
    sum /= norm;
  }
  return sum;
}

double orc_plaquette(int nd, int nc, const int *dims, const double *U, int normalize,
                     double *per_site) {
  lat l = mklat(nd, nc, dims);
  return plaquette_impl(&l, U, NULL, normalize, per_site);
}

/* ------------------------------------------------------------------------ */
/* Wilson loops (WilsonLoop.hpp:61-304)                                     */
/* ------------------------------------------------------------------------ */
void orc_wilson_loop_mu_nu(int nd, int nc, const int *dims, const double *U, int mu, int nu,
                           const int *pairs, int npairs, int normalize, double *out,
                           double *per_site_last) {
  lat l = mklat(nd, nc, dims);
  sun *wlmu = (sun *)malloc(sizeof(sun) * l.V);
  sun *wlnu = (sun *)malloc(sizeof(sun) * l.V);
  double *w = (double *)malloc(sizeof(double) * l.V);
  for (long s = 0; s < l.V; ++s) wlmu[s] = wlnu[s] = identity_sun(nc);
  int cur_mu = 0, cur_nu = 0;
  for (int p = 0; p < npairs; ++p) {
    int Lmu = pairs[2 * p], Lnu = pairs[2 * p + 1];
    if (Lmu < cur_mu || Lnu < cur_nu) { /* reset_Lmu_Lnu :86-96 */
      for (long s = 0; s < l.V; ++s) wlmu[s] = wlnu[s] = identity_sun(nc);
      cur_mu = cur_nu = 0;
    }
    int old_mu = cur_mu, old_nu = cur_nu;
    cur_mu = Lmu;
    cur_nu = Lnu;
    /* WLmunu::operator() :61-77 */
#pragma omp parallel for schedule(static)
    for (long s = 0; s < l.V; ++s) {
      sun a = wlmu[s], b = wlnu[s];
      for (int i = old_mu; i < Lmu; ++i) {
        sun u = link(&l, U, shift(&l, s, mu, i), mu);
        a = mm(nc, &a, &u);
      }
      for (int i = old_nu; i < Lnu; ++i) {
        sun u = link(&l, U, shift(&l, s, nu, i), nu);
        b = mm(nc, &b, &u);
      }
      wlmu[s] = a;
      wlnu[s] = b;
    }
    /* WLoop_munu::operator() :129-160 */
#pragma omp parallel for schedule(static)
    for (long s = 0; s < l.V; ++s) {
      sun lmu = mm(nc, &wlmu[s], &wlnu[shift(&l, s, mu, Lmu)]);
      sun lnu = mm(nc, &wlnu[s], &wlmu[shift(&l, s, nu, Lnu)]);
      cplx t = {0.0, 0.0};
      for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) t = cadd(t, cmul(lmu.m[c1][c2], cconj(lnu.m[c1][c2])));
      w[s] = t.re;
    }
    double W = psum(w, l.V);
    if (normalize) { /* :233-239, sequential divisions */
      for (int d = 0; d < nd; ++d) W /= (double)l.L[d];
      W /= (double)nc;
    }
    out[p] = W;
    if (per_site_last && p == npairs - 1) memcpy(per_site_last, w, sizeof(double) * l.V);
  }
  free(wlmu);
  free(wlnu);
  free(w);
}

/* WilsonLoop.hpp:250-304 */
int orc_wilson_loop_temporal(int nd, int nc, const int *dims, const double *U, const int *pairs,
                             int npairs, int normalize, double *out) {
  double *tmp = (double *)malloc(sizeof(double) * (npairs > 0 ? npairs : 1));
  orc_wilson_loop_mu_nu(nd, nc, dims, U, 0, nd - 1, pairs, npairs, normalize, out, NULL);
  for (int mu = 1; mu < nd - 1; ++mu) {
    orc_wilson_loop_mu_nu(nd, nc, dims, U, mu, nd - 1, pairs, npairs, normalize, tmp, NULL);
    for (int p = 0; p < npairs; ++p) out[p] += tmp[p];
  }
  for (int p = 0; p < npairs; ++p) out[p] /= (double)(nd - 1);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* field initialisation                                                     */
/* ------------------------------------------------------------------------ */
void orc_field_identity(int nd, int nc, const int *dims, double *U) {
  lat l = mklat(nd, nc, dims);
  sun id = identity_sun(nc);
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) set_link(&l, U, s, mu, &id);
}

int orc_ngen(int nc) { return nc == 1 ? 1 : (nc == 2 ? 3 : 8); }

void orc_field_random_algebra(int nd, int nc, const int *dims, double eps, uint64_t seed,
                              double *U) {
  lat l = mklat(nd, nc, dims);
  int ng = orc_ngen(nc);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      stream st;
      stream_init(&st, seed, 3, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      double h[8], v[8];
      normals(&st, ng, h);
      for (int a = 0; a < ng; ++a) v[a] = eps * h[a] * 0.5;
      sun e = exp_iq(nc, v);
      set_link(&l, U, s, mu, &e);
    }
}

/* GaugeField.hpp:55-60, :96-112: every link a randSUN(delta) draw, from the
 * stream (tag 5, c0 = lexicographic site, mu) */
void orc_field_random_sun(int nd, int nc, const int *dims, double delta, uint64_t seed,
                          double *U) {
  lat l = mklat(nd, nc, dims);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      stream st;
      stream_init(&st, seed, 5, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      sun r;
      rand_sun(nc, &st, delta, &r);
      set_link(&l, U, s, mu, &r);
    }
}

/* GaugeField.hpp:58-63, :118-141: every entry complex(drand(-1,1),
 * drand(-1,1)), real part drawn first, c1 outer, c2 inner; stream (tag 6,
 * c0 = lexicographic site, mu) */
void orc_field_random_complex(int nd, int nc, const int *dims, uint64_t seed, double *U) {
  lat l = mklat(nd, nc, dims);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      stream st;
      stream_init(&st, seed, 6, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      sun a = zero_sun();
      for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) {
          double re = drand(&st, -1.0, 1.0);
          double im = drand(&st, -1.0, 1.0);
          a.m[c1][c2].re = re;
          a.m[c1][c2].im = im;
        }
      set_link(&l, U, s, mu, &a);
    }
}

/* ------------------------------------------------------------------------ */
/* Metropolis (Metropolis.hpp:72-178)                                       */
/* ------------------------------------------------------------------------ */
/* The opt-in symmetric SU(3) proposal (no counterpart in the reference,
 * whose randSUN<3> is asymmetric, SUN.hpp:192-251): r = exp(i delta sum_a
 * h_a lambda_a) with h_a = drand(-1, 1), a = 0..7, drawn from the stream
 * (tag 9); the accept uniform follows.  h -> -h maps r to r^dagger with the
 * same density, so the proposal is symmetric. */
static void rand_sym3(stream *s, double delta, sun *r) {
  double v[8];
  for (int a = 0; a < 8; ++a) v[a] = delta * drand(s, -1.0, 1.0);
  *r = exp_iq(3, v);
}

static double metropolis_impl(const lat *l, double *U, const double *C, int nhits, double beta,
                              double delta, uint64_t seed, uint32_t sweep, uint32_t chain,
                              long *naccepted, int *visit, int sym, int narrow) {
  int nd = l->nd, nc = l->nc;
  long H[4] = {1, 1, 1, 1}, Vh = 1;
  for (int d = 0; d < nd; ++d) {
    H[d] = l->L[d] / 2;
    Vh *= H[d];
  }
  long ntot = 0;
  for (int pass = 0; pass < (1 << nd); ++pass) {
    int b[4];
    for (int d = 0; d < nd; ++d) b[d] = (pass >> (nd - 1 - d)) & 1; /* oddeven_array */
    long nacc_pass = 0;
#pragma omp parallel for schedule(static) reduction(+ : nacc_pass)
    for (long y = 0; y < Vh; ++y) {
      int x[4];
      long r = y;
      for (int d = nd - 1; d >= 0; --d) {
        x[d] = 2 * (int)(r % H[d]) + b[d]; /* index_odd_even */
        r /= H[d];
      }
      long site = lex(l, x);
      if (visit) visit[site] = pass;
      long nacc = 0;
      for (int mu = 0; mu < nd; ++mu) {
        sun S = staple_c(l, U, C, site, mu);
        double cl = C ? C[(size_t)site * nd + mu] : 1.0;
        for (int hit = 0; hit < nhits; ++hit) {
          stream st;
          sun rr;
          if (sym && nc == 3) {
            stream_init(&st, seed, 9, (uint32_t)site, sweep, (uint32_t)hit, chain, (uint32_t)mu);
            st.narrow = narrow;
            rand_sym3(&st, delta, &rr);
          } else {
            stream_init(&st, seed, 1, (uint32_t)site, sweep, (uint32_t)hit, chain, (uint32_t)mu);
            st.narrow = narrow;
            rand_sun(nc, &st, delta, &rr);
          }
          sun Uold = link(l, U, site, mu);
          sun Unew = mm(nc, &Uold, &rr);
          double dS;
          if (C) {
            sun a = Unew, o = Uold;
            for (int i = 0; i < nc; ++i)
              for (int j = 0; j < nc; ++j) {
                a.m[i][j].re *= cl;
                a.m[i][j].im *= cl;
                o.m[i][j].re *= cl;
                o.m[i][j].im *= cl;
              }
            sun t1 = mm(nc, &a, &S), t2 = mm(nc, &o, &S);
            dS = -(beta / (double)nc) * (trace(nc, &t1).re - trace(nc, &t2).re);
          } else {
            sun t1 = mm(nc, &Unew, &S), t2 = mm(nc, &Uold, &S);
            dS = -(beta / (double)nc) * (trace(nc, &t1).re - trace(nc, &t2).re);
          }
          int accept = dS < 0.0;
          if (!accept) accept = next_u01(&st) < exp(-dS);
          if (accept) {
            restore_sun(nc, &Unew);
            set_link(l, U, site, mu, &Unew);
            nacc++;
          }
        }
      }
      nacc_pass += nacc;
    }
    ntot += nacc_pass;
  }
  if (naccepted) *naccepted = ntot;
  double acc = (double)ntot;
  double norm = 1.0;
  for (int d = 0; d < nd; ++d) norm *= (double)l->L[d];
  norm *= (double)nd;
  norm *= (double)nhits;
  return acc / norm;
}

double orc_metropolis_sweep(int nd, int nc, const int *dims, double *U, int nhits, double beta,
                            double delta, uint64_t seed, uint32_t sweep, long *naccepted,
                            int *visit) {
  lat l = mklat(nd, nc, dims);
  return metropolis_impl(&l, U, NULL, nhits, beta, delta, seed, sweep, 0, naccepted, visit, 0, 0);
}

double orc_metropolis_sweep_sym(int nd, int nc, const int *dims, double *U, int nhits, double beta,
                                double delta, uint64_t seed, uint32_t sweep, long *naccepted) {
  lat l = mklat(nd, nc, dims);
  return metropolis_impl(&l, U, NULL, nhits, beta, delta, seed, sweep, 0, naccepted, NULL, 1, 0);
}

/* The same sweep with the FP32 links' draws: every uniform is 24 bits of one
 * stream word (the library's FP32 draw), so the proposals and accept draws
 * are those of the FP32 kernels, with FP64 arithmetic throughout.  chain is
 * the stream's chain index (a batched chain's position in its batch). */
double orc_metropolis_sweep_narrow(int nd, int nc, const int *dims, double *U, int nhits, double beta,
                                   double delta, uint64_t seed, uint32_t sweep, uint32_t chain, int sym,
                                   long *naccepted) {
  lat l = mklat(nd, nc, dims);
  return metropolis_impl(&l, U, NULL, nhits, beta, delta, seed, sweep, chain, naccepted, NULL, sym, 1);
}

int orc_rand_sym3(double delta, uint64_t seed, uint32_t sweep, uint32_t site, uint32_t hit,
                  uint32_t mu, double *out) {
  stream s;
  stream_init(&s, seed, 9, site, sweep, hit, 0, mu);
  sun r;
  rand_sym3(&s, delta, &r);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      out[2 * (i * 3 + j)] = r.m[i][j].re;
      out[2 * (i * 3 + j) + 1] = r.m[i][j].im;
    }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* HMC (HamiltonianField.hpp:23-60, GaugeMonomial.hpp:19-37,                 */
/*      Integrator.hpp:26-53, HMC.hpp:48-71, AdjointGroup.hpp)               */
/* ------------------------------------------------------------------------ */
void orc_refresh_momenta(int nd, int nc, const int *dims, uint64_t seed, uint32_t traj,
                         double *P) {
  lat l = mklat(nd, nc, dims);
  int ng = orc_ngen(nc);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      stream st;
      stream_init(&st, seed, 2, (uint32_t)s, traj, 0, 0, (uint32_t)mu);
      normals(&st, ng, P + ((size_t)s * nd + mu) * ng);
    }
}

/* F_a = (beta/Nc) Im Tr(T_a M), M = U_mu(x) Sigma_mu(x).  For SU(2) this is
 * AdjointSU2(const SU2&) (AdjointGroup.hpp:13-19) written for a general 2x2
 * M; for U(1) AdjointU1 (:114-117). */
static void project(int nc, const sun *M, double c, double *F) {
  if (nc == 1) {
    F[0] = c * M->m[0][0].im;
  } else if (nc == 2) {
    F[0] = c * (M->m[0][1].im + M->m[1][0].im);
    F[1] = c * (M->m[0][1].re - M->m[1][0].re);
    F[2] = c * (M->m[0][0].im - M->m[1][1].im);
  } else {
    const double r3 = 1.0 / sqrt(3.0);
    F[0] = c * (M->m[0][1].im + M->m[1][0].im);
    F[1] = c * (M->m[0][1].re - M->m[1][0].re);
    F[2] = c * (M->m[0][0].im - M->m[1][1].im);
    F[3] = c * (M->m[0][2].im + M->m[2][0].im);
    F[4] = c * (M->m[0][2].re - M->m[2][0].re);
    F[5] = c * (M->m[1][2].im + M->m[2][1].im);
    F[6] = c * (M->m[1][2].re - M->m[2][1].re);
    F[7] = c * ((M->m[0][0].im + M->m[1][1].im - 2.0 * M->m[2][2].im) * r3);
  }
}

void orc_force(int nd, int nc, const int *dims, const double *U, double beta, double *F) {
  lat l = mklat(nd, nc, dims);
  int ng = orc_ngen(nc);
  double c = beta / (double)nc;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      sun S = staple_c(&l, U, NULL, s, mu);
      sun u = link(&l, U, s, mu);
      sun M = mm(nc, &u, &S);
      project(nc, &M, c, F + ((size_t)s * nd + mu) * ng);
    }
}

void orc_update_momentum(int nd, int nc, const int *dims, double *P, const double *F,
                         double dtau) {
  lat l = mklat(nd, nc, dims);
  long n = l.V * nd * orc_ngen(nc);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) P[i] -= dtau * F[i];
}

void orc_update_gauge(int nd, int nc, const int *dims, double *U, const double *P, double dtau) {
  lat l = mklat(nd, nc, dims);
  int ng = orc_ngen(nc);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      double v[8];
      const double *p = P + ((size_t)s * nd + mu) * ng;
      for (int a = 0; a < ng; ++a) v[a] = dtau * p[a];
      sun e = exp_iq(nc, v);
      sun u = link(&l, U, s, mu);
      sun r = mm(nc, &e, &u);
      set_link(&l, U, s, mu, &r);
    }
}

void orc_restore_field(int nd, int nc, const int *dims, double *U) {
  lat l = mklat(nd, nc, dims);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < l.V; ++s)
    for (int mu = 0; mu < nd; ++mu) {
      sun u = link(&l, U, s, mu);
      restore_sun(nc, &u);
      set_link(&l, U, s, mu, &u);
    }
}

double orc_kinetic(int nd, int nc, const int *dims, const double *P) {
  lat l = mklat(nd, nc, dims);
  int ng = orc_ngen(nc);
  long nl = l.V * nd;
  double *k = (double *)malloc(sizeof(double) * nl);
  for (long i = 0; i < nl; ++i) {
    double n2 = 0.0;
    for (int a = 0; a < ng; ++a) n2 += P[i * ng + a] * P[i * ng + a];
    k[i] = 0.5 * n2;
  }
  double r = psum(k, nl);
  free(k);
  return r;
}

double orc_gauge_action(int nd, int nc, const int *dims, const double *U, double beta) {
  lat l = mklat(nd, nc, dims);
  return -(beta / (double)nc) * plaquette_impl(&l, U, NULL, 0, NULL);
}

void orc_leapfrog(int nd, int nc, const int *dims, double *U, double *P, double beta, double tau,
                  int nsteps) {
  lat l = mklat(nd, nc, dims);
  long n = l.V * nd * orc_ngen(nc);
  double *F = (double *)malloc(sizeof(double) * n);
  double dtau = tau / (double)nsteps;
  orc_force(nd, nc, dims, U, beta, F);
  orc_update_momentum(nd, nc, dims, P, F, dtau / 2.0);
  orc_update_gauge(nd, nc, dims, U, P, dtau);
  for (int i = 0; i < nsteps; ++i) {
    orc_force(nd, nc, dims, U, beta, F);
    orc_update_momentum(nd, nc, dims, P, F, dtau);
    orc_update_gauge(nd, nc, dims, U, P, dtau);
  }
  orc_force(nd, nc, dims, U, beta, F);
  orc_update_momentum(nd, nc, dims, P, F, dtau / 2.0);
  orc_restore_field(nd, nc, dims, U);
  free(F);
}

int orc_hmc_trajectory(int nd, int nc, const int *dims, double *U, double beta, double tau,
                       int nsteps, uint64_t seed, uint32_t traj, double *dH, double *plaq) {
  lat l = mklat(nd, nc, dims);
  size_t nU = (size_t)l.V * nd * nc * nc * 2;
  long nP = l.V * nd * orc_ngen(nc);
  double *P = (double *)malloc(sizeof(double) * nP);
  double *Uold = (double *)malloc(sizeof(double) * nU);
  orc_refresh_momenta(nd, nc, dims, seed, traj, P);
  memcpy(Uold, U, sizeof(double) * nU);
  double K0 = orc_kinetic(nd, nc, dims, P);
  double G0 = orc_gauge_action(nd, nc, dims, U, beta);
  orc_leapfrog(nd, nc, dims, U, P, beta, tau, nsteps);
  double K1 = orc_kinetic(nd, nc, dims, P);
  double G1 = orc_gauge_action(nd, nc, dims, U, beta);
  double d = 0.0;
  d += K1 - K0;
  d += G1 - G0;
  int accept = 1;
  if (d > 0.0) {
    stream st;
    stream_init(&st, seed, 7, 0, traj, 0, 0, 0);
    if (next_u01(&st) > exp(-d)) accept = 0;
  }
  if (!accept) memcpy(U, Uold, sizeof(double) * nU);
  if (dH) *dH = d;
  if (plaq) *plaq = orc_plaquette(nd, nc, dims, U, 1, NULL);
  free(P);
  free(Uold);
  return accept;
}

/* ------------------------------------------------------------------------ */
/* PTBC defect (PTBCGaugeField.hpp:30-36, 186-229)                           */
/* ------------------------------------------------------------------------ */
void orc_defect_field(int nd, const int *dims, int defect_length, const int *defect_pos,
                      double cr, double *C) {
  lat l = mklat(nd, 1, dims);
  for (long i = 0; i < l.V * nd; ++i) C[i] = 1.0;
  int n = 1;
  for (int k = 1; k < nd; ++k) n *= defect_length;
  for (int t = 0; t < n; ++t) {
    int x[4] = {0, 0, 0, 0};
    int r = t;
    for (int k = nd - 1; k >= 1; --k) {
      x[k] = (r % defect_length + defect_pos[k - 1]) % l.L[k];
      r /= defect_length;
    }
    C[(size_t)lex(&l, x) * nd + 0] = cr;
  }
}

double orc_plaquette_defect(int nd, int nc, const int *dims, const double *U, const double *C,
                            int normalize) {
  lat l = mklat(nd, nc, dims);
  return plaquette_impl(&l, U, C, normalize, NULL);
}

double orc_metropolis_sweep_defect(int nd, int nc, const int *dims, double *U, const double *C,
                                   int nhits, double beta, double delta, uint64_t seed,
                                   uint32_t sweep, uint32_t chain, long *naccepted) {
  lat l = mklat(nd, nc, dims);
  return metropolis_impl(&l, U, C, nhits, beta, delta, seed, sweep, chain, naccepted, NULL, 0, 0);
}

/* the defect sweep with the FP32 links' draws (orc_metropolis_sweep_narrow) */
double orc_metropolis_sweep_defect_narrow(int nd, int nc, const int *dims, double *U, const double *C,
                                          int nhits, double beta, double delta, uint64_t seed,
                                          uint32_t sweep, uint32_t chain, long *naccepted) {
  lat l = mklat(nd, nc, dims);
  return metropolis_impl(&l, U, C, nhits, beta, delta, seed, sweep, chain, naccepted, NULL, 0, 1);
}
