// Device-side building blocks of the klft MI355X path: complex and SU(N)
// algebra, the sublattice-major structure-of-arrays link layout, site
// navigation, Metropolis proposals, reunitarisation, the Lie-algebra
// exponential and block / grid reductions.
//
// Link layout in HBM (one allocation per field), blocks of 64 sites:
//   element (s, mu, c) at ((s / 64) * nd + mu) * NCOMP * 64 + c * 64 + s % 64
//   (FP32: components paired per lane, see Links below)
//   s   = sub * Vh + r  (device index)
//   sub = sum_d (x_d & 1) << (nd - 1 - d)   -- the sublattice (pass) index of
//         the reference's oddeven_array (IndexHelper.hpp:84-93)
//   r   = position of y_d = x_d >> 1 inside the sublattice (Geom below)
//   c   = row-major colour component; NCOMP = Nc*Nc, or (Nc-1)*Nc for the
//         compressed two-row (SU(3)) / one-row (SU(2)) storage.
// Each Metropolis pass therefore reads and writes whole contiguous
// (sublattice, direction) groups, and a wave's load of one link is one
// contiguous run of NCOMP * 64 complex numbers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"

namespace kd {

// ---------------------------------------------------------------------------
// complex numbers (operation order of Kokkos::complex)
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> mk(R a, R b) {
  cx<R> c;
  c.re = a;
  c.im = b;
  return c;
}
template <typename R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
  return mk(a.re - b.re, a.im - b.im);
}
template <typename R>
__device__ __forceinline__ cx<R> conj(cx<R> a) {
  return mk(a.re, -a.im);
}
template <typename R>
__device__ __forceinline__ cx<R> scale(R s, cx<R> a) {
  return mk(s * a.re, s * a.im);
}
// a * conj(b), the conjugate folded into the product
template <typename R>
__device__ __forceinline__ cx<R> mulc(cx<R> a, cx<R> b) {
  return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// conj(a) * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> mulcc(cx<R> a, cx<R> b) {
  return mk(a.re * b.re - a.im * b.im, -(a.re * b.im + a.im * b.re));
}
// conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cmul_(cx<R> a, cx<R> b) {
  return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}

// ---------------------------------------------------------------------------
// Nc x Nc matrices, row major (GLOBAL.hpp:51-52)
// ---------------------------------------------------------------------------
template <typename R, int N>
struct Mat {
  cx<R> m[N][N];
};

template <typename R, int N>
__device__ __forceinline__ Mat<R, N> zero() {
  Mat<R, N> z;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) z.m[i][j] = mk<R>(0, 0);
  return z;
}
template <typename R, int N>
__device__ __forceinline__ Mat<R, N> identity() {
  Mat<R, N> z = zero<R, N>();
#pragma unroll
  for (int i = 0; i < N; ++i) z.m[i][i].re = R(1);
  return z;
}

// Products of Nc x Nc matrices, k accumulated in order (SUN.hpp:28-44).  Each
// complex multiply-add is four FMAs: the first product of an entry is formed
// as mul + fma per component and every further term is fused into the
// accumulator, which rounds differently from the reference's separately
// rounded product and sum only at the last bit.
template <typename R>
__device__ __forceinline__ R fmar(R a, R b, R c) {
  return fma(a, b, c);
}
// t + a * b
template <typename R>
__device__ __forceinline__ cx<R> cfma(cx<R> a, cx<R> b, cx<R> t) {
  return mk(fmar(-a.im, b.im, fmar(a.re, b.re, t.re)), fmar(a.im, b.re, fmar(a.re, b.im, t.im)));
}
// t + a * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cfma_c(cx<R> a, cx<R> b, cx<R> t) {
  return mk(fmar(a.im, b.im, fmar(a.re, b.re, t.re)), fmar(-a.re, b.im, fmar(a.im, b.re, t.im)));
}
// t + conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cfma_l(cx<R> a, cx<R> b, cx<R> t) {
  return mk(fmar(a.im, b.im, fmar(a.re, b.re, t.re)), fmar(-a.im, b.re, fmar(a.re, b.im, t.im)));
}
// t + conj(a) * conj(b)
template <typename R>
__device__ __forceinline__ cx<R> cfma_cc(cx<R> a, cx<R> b, cx<R> t) {
  return mk(fmar(-a.im, b.im, fmar(a.re, b.re, t.re)), fmar(-a.im, b.re, fmar(-a.re, b.im, t.im)));
}
// a * b alone (mul + fma per component)
template <typename R>
__device__ __forceinline__ cx<R> cmul1(cx<R> a, cx<R> b) {
  return mk(fmar(a.re, b.re, -(a.im * b.im)), fmar(a.re, b.im, a.im * b.re));
}

template <typename R, int N>
__device__ __forceinline__ Mat<R, N> mul(const Mat<R, N>& a, const Mat<R, N>& b) {
  Mat<R, N> c;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> t = cmul1(a.m[i][0], b.m[0][j]);
#pragma unroll
      for (int k = 1; k < N; ++k) t = cfma(a.m[i][k], b.m[k][j], t);
      c.m[i][j] = t;
    }
  return c;
}
// a * b^dagger
template <typename R, int N>
__device__ __forceinline__ Mat<R, N> mul_ad(const Mat<R, N>& a, const Mat<R, N>& b) {
  Mat<R, N> c;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> t = cmul1(a.m[i][0], conj(b.m[j][0]));
#pragma unroll
      for (int k = 1; k < N; ++k) t = cfma_c(a.m[i][k], b.m[j][k], t);
      c.m[i][j] = t;
    }
  return c;
}
// a^dagger * b
template <typename R, int N>
__device__ __forceinline__ Mat<R, N> mul_da(const Mat<R, N>& a, const Mat<R, N>& b) {
  Mat<R, N> c;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> t = cmul1(conj(a.m[0][i]), b.m[0][j]);
#pragma unroll
      for (int k = 1; k < N; ++k) t = cfma_l(a.m[k][i], b.m[k][j], t);
      c.m[i][j] = t;
    }
  return c;
}
// a^dagger * b^dagger
template <typename R, int N>
__device__ __forceinline__ Mat<R, N> mul_dd(const Mat<R, N>& a, const Mat<R, N>& b) {
  Mat<R, N> c;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> t = cmul1(conj(a.m[0][i]), conj(b.m[j][0]));
#pragma unroll
      for (int k = 1; k < N; ++k) t = cfma_cc(a.m[k][i], b.m[j][k], t);
      c.m[i][j] = t;
    }
  return c;
}
template <typename R, int N>
__device__ __forceinline__ void add_to(Mat<R, N>& a, const Mat<R, N>& b) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) a.m[i][j] = a.m[i][j] + b.m[i][j];
}
// t += a * b and t += a * b^dagger, each product term fused into t
template <typename R, int N>
__device__ __forceinline__ void acc_mul(Mat<R, N>& t, const Mat<R, N>& a, const Mat<R, N>& b) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> v = t.m[i][j];
#pragma unroll
      for (int k = 0; k < N; ++k) v = cfma(a.m[i][k], b.m[k][j], v);
      t.m[i][j] = v;
    }
}
template <typename R, int N>
__device__ __forceinline__ void acc_mul_ad(Mat<R, N>& t, const Mat<R, N>& a, const Mat<R, N>& b) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      cx<R> v = t.m[i][j];
#pragma unroll
      for (int k = 0; k < N; ++k) v = cfma_c(a.m[i][k], b.m[j][k], v);
      t.m[i][j] = v;
    }
}
// Re Tr(a * b), two FMAs per term, one accumulator per row (independent FMA
// chains instead of one chain of 2 N^2 dependent FMAs), rows summed in order
template <typename R, int N>
__device__ __forceinline__ R retr_mul(const Mat<R, N>& a, const Mat<R, N>& b) {
  R r[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    R t = a.m[i][0].re * b.m[0][i].re;
    t = fmar(-a.m[i][0].im, b.m[0][i].im, t);
#pragma unroll
    for (int k = 1; k < N; ++k) t = fmar(-a.m[i][k].im, b.m[k][i].im, fmar(a.m[i][k].re, b.m[k][i].re, t));
    r[i] = t;
  }
  R t = r[0];
#pragma unroll
  for (int i = 1; i < N; ++i) t += r[i];
  return t;
}
// Re Tr(a * b^dagger) = sum_ij Re(a_ij conj(b_ij)), rows (c1) summed in order
template <typename R, int N>
__device__ __forceinline__ R retr_mul_ad(const Mat<R, N>& a, const Mat<R, N>& b) {
  R r[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    R t = a.m[i][0].re * b.m[i][0].re;
    t = fmar(a.m[i][0].im, b.m[i][0].im, t);
#pragma unroll
    for (int j = 1; j < N; ++j) t = fmar(a.m[i][j].im, b.m[i][j].im, fmar(a.m[i][j].re, b.m[i][j].re, t));
    r[i] = t;
  }
  R t = r[0];
#pragma unroll
  for (int i = 1; i < N; ++i) t += r[i];
  return t;
}
template <typename R, int N>
__device__ __forceinline__ R retrace(const Mat<R, N>& a) {
  R t = R(0);
#pragma unroll
  for (int i = 0; i < N; ++i) t += a.m[i][i].re;
  return t;
}

template <typename RO, typename RI, int N>
__device__ __forceinline__ Mat<RO, N> convert(const Mat<RI, N>& a) {
  Mat<RO, N> c;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) c.m[i][j] = mk<RO>((RO)a.m[i][j].re, (RO)a.m[i][j].im);
  return c;
}

// third row of an SU(3) matrix from the first two: conj(r0 x r1), the
// reference's own completion (SUN.hpp:334-336)
template <typename R>
__device__ __forceinline__ void complete_row2(Mat<R, 3>& a) {
  a.m[2][0] = conj(a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]);
  a.m[2][1] = conj(a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2]);
  a.m[2][2] = conj(a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]);
}
template <typename R>
__device__ __forceinline__ void complete_row1(Mat<R, 2>& a) {
  a.m[1][0] = mk(-a.m[0][1].re, a.m[0][1].im);
  a.m[1][1] = mk(a.m[0][0].re, -a.m[0][0].im);
}

// ---------------------------------------------------------------------------
// reunitarisation
// ---------------------------------------------------------------------------
// restoreSUN<1,2,3> exactly as the reference (SUN.hpp:256-340): the plain FP64
// method uses this one.
template <typename R>
__device__ __forceinline__ void restore_sun(Mat<R, 1>& a) {
  const R n = sqrt(a.m[0][0].re * a.m[0][0].re + a.m[0][0].im * a.m[0][0].im);
  a.m[0][0] = mk(a.m[0][0].re / n, a.m[0][0].im / n);
}
template <typename R>
__device__ __forceinline__ void restore_sun(Mat<R, 2>& a) {
  const R n = sqrt(a.m[0][0].re * a.m[0][0].re + a.m[0][0].im * a.m[0][0].im +
                   a.m[0][1].re * a.m[0][1].re + a.m[0][1].im * a.m[0][1].im);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) a.m[i][j] = mk(a.m[i][j].re / n, a.m[i][j].im / n);
}
template <typename R>
__device__ __forceinline__ void restore_sun(Mat<R, 3>& a) {
  const cx<R> s0 = cmul_(a.m[0][0], a.m[0][0]) + cmul_(a.m[0][1], a.m[0][1]) +
                   cmul_(a.m[0][2], a.m[0][2]);
  const cx<R> s1 = cmul_(a.m[1][0], a.m[1][0]) + cmul_(a.m[1][1], a.m[1][1]) +
                   cmul_(a.m[1][2], a.m[1][2]);
  const R n0 = sqrt(s0.re), n1 = sqrt(s1.re);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    a.m[0][j] = mk(a.m[0][j].re / n0, a.m[0][j].im / n0);
    a.m[1][j] = mk(a.m[1][j].re / n1, a.m[1][j].im / n1);
  }
  complete_row2(a);
  a.m[1][0] = conj(a.m[2][1] * a.m[0][2] - a.m[2][2] * a.m[0][1]);
  a.m[1][1] = conj(a.m[2][2] * a.m[0][0] - a.m[2][0] * a.m[0][2]);
  a.m[1][2] = conj(a.m[2][0] * a.m[0][1] - a.m[2][1] * a.m[0][0]);
}

// Gram-Schmidt reunitarisation for FP32 links: products in FP32, every norm
// accumulated and square-rooted in FP64.
__device__ __forceinline__ float inv_norm64(double n2) { return (float)(1.0 / sqrt(n2)); }

__device__ __forceinline__ void reunit_gs(Mat<float, 1>& a) {
  const double n2 = (double)a.m[0][0].re * a.m[0][0].re + (double)a.m[0][0].im * a.m[0][0].im;
  const float f = inv_norm64(n2);
  a.m[0][0] = scale(f, a.m[0][0]);
}
__device__ __forceinline__ void reunit_gs(Mat<float, 2>& a) {
  double n2 = 0.0;
#pragma unroll
  for (int j = 0; j < 2; ++j)
    n2 += (double)a.m[0][j].re * a.m[0][j].re + (double)a.m[0][j].im * a.m[0][j].im;
  const float f = inv_norm64(n2);
  a.m[0][0] = scale(f, a.m[0][0]);
  a.m[0][1] = scale(f, a.m[0][1]);
  complete_row1(a);
}
__device__ __forceinline__ void reunit_gs(Mat<float, 3>& a) {
  double n2 = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    n2 += (double)a.m[0][j].re * a.m[0][j].re + (double)a.m[0][j].im * a.m[0][j].im;
  float f = inv_norm64(n2);
#pragma unroll
  for (int j = 0; j < 3; ++j) a.m[0][j] = scale(f, a.m[0][j]);
  // r1 -= (r0^dagger r1) r0
  cx<float> p = cmul_(a.m[0][0], a.m[1][0]) + cmul_(a.m[0][1], a.m[1][1]) +
                cmul_(a.m[0][2], a.m[1][2]);
#pragma unroll
  for (int j = 0; j < 3; ++j) a.m[1][j] = a.m[1][j] - p * a.m[0][j];
  n2 = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    n2 += (double)a.m[1][j].re * a.m[1][j].re + (double)a.m[1][j].im * a.m[1][j].im;
  f = inv_norm64(n2);
#pragma unroll
  for (int j = 0; j < 3; ++j) a.m[1][j] = scale(f, a.m[1][j]);
  complete_row2(a);
}

// policy: FP64 links are restored with the reference's restoreSUN; FP32 links
// with the FP64-norm Gram-Schmidt above.
template <int N>
__device__ __forceinline__ void reunitarize(Mat<double, N>& a) {
  restore_sun(a);
}
template <int N>
__device__ __forceinline__ void reunitarize(Mat<float, N>& a) {
  reunit_gs(a);
}

// ---------------------------------------------------------------------------
// lattice geometry
// ---------------------------------------------------------------------------
// Within each sublattice the half-extent coordinates y_d = x_d >> 1 are
// ordered in 4D tiles of power-of-two extent T_d = 2^tb_d:
//   r = sum_d (y_d >> tb_d) ts_d + sum_d (y_d & (T_d - 1)) << lsh_d,
// tiles in lexicographic order (last dimension fastest), sites inside a tile
// likewise.  The library uses T_d = 1 (plain lexicographic order) unless
// KLFT_TILES=1 asks for T_d = 4 where H_d = L_d/2 allows it.
struct Geom {
  int nd;
  int L[4];
  int H[4];          // L / 2
  int tb[4];         // log2 of the tile extent
  int lsh[4];        // log2 of the in-tile stride
  int NT[4];         // tiles per dimension
  int TV;            // tile volume (power of two)
  int tvb;           // log2(TV)
  long long V, Vh;   // volume, volume / 2^nd
  long long ts[4];   // tile strides (multiples of TV)
  long long ls[4];   // strides of the full lexicographic index
};

template <int ND>
__device__ __forceinline__ long long encode_r(const Geom& g, const int* y) {
  long long r = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d)
    r += (long long)(y[d] >> g.tb[d]) * g.ts[d] + ((y[d] & ((1 << g.tb[d]) - 1)) << g.lsh[d]);
  return r;
}
template <int ND>
__device__ __forceinline__ void decode_r(const Geom& g, long long r, int* y) {
  long long tile = r >> g.tvb;
  const int loc = (int)(r & (g.TV - 1));
#pragma unroll
  for (int d = ND - 1; d >= 0; --d) {
    const int t = (int)(tile % g.NT[d]);
    tile /= g.NT[d];
    y[d] = (t << g.tb[d]) + ((loc >> g.lsh[d]) & ((1 << g.tb[d]) - 1));
  }
}
// coordinates of position r of sublattice sub (index_odd_even)
template <int ND>
__device__ __forceinline__ void sub_coords(const Geom& g, int sub, long long r, int* x) {
  int y[ND];
  decode_r<ND>(g, r, y);
#pragma unroll
  for (int d = 0; d < ND; ++d) x[d] = 2 * y[d] + ((sub >> (ND - 1 - d)) & 1);
}

template <int ND>
__device__ __forceinline__ long long dev_index(const Geom& g, const int* x) {
  int sub = 0, y[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    sub |= (x[d] & 1) << (ND - 1 - d);
    y[d] = x[d] >> 1;
  }
  return (long long)sub * g.Vh + encode_r<ND>(g, y);
}
template <int ND>
__device__ __forceinline__ long long lex_index(const Geom& g, const int* x) {
  long long r = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d) r += (long long)x[d] * g.ls[d];
  return r;
}
// device index -> coordinates
template <int ND>
__device__ __forceinline__ void dev_coords(const Geom& g, long long s, int* x) {
  const int sub = (int)(s / g.Vh);
  sub_coords<ND>(g, sub, s - (long long)sub * g.Vh, x);
}
// lexicographic index -> coordinates
template <int ND>
__device__ __forceinline__ void lex_coords(const Geom& g, long long s, int* x) {
#pragma unroll
  for (int d = ND - 1; d >= 0; --d) {
    x[d] = (int)(s % g.L[d]);
    s /= g.L[d];
  }
}
template <int ND>
__device__ __forceinline__ void copy_x(const int* a, int* b) {
#pragma unroll
  for (int d = 0; d < ND; ++d) b[d] = a[d];
}
// periodic shift of one coordinate by +-1 (shift_index_plus/minus)
__device__ __forceinline__ int up(int v, int L) { return v + 1 == L ? 0 : v + 1; }
__device__ __forceinline__ int dn(int v, int L) { return v == 0 ? L - 1 : v - 1; }
__device__ __forceinline__ int shiftk(int v, int k, int L) {
  int r = (v + k) % L;
  return r < 0 ? r + L : r;
}

// device index of x + s1 e_m1 + s2 e_m2 for s1, s2 in {-1, 0, +1} (m2 = -1:
// none).  The coordinate loop is unrolled and the runtime directions are
// compared, not used as array indices, so nothing spills to scratch.
template <int ND>
__device__ __forceinline__ long long idx_shift(const Geom& g, const int* x, int m1, int s1, int m2,
                                               int s2) {
  int v[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    int w = x[d];
    const int Ld = g.L[d];
    if (d == m1) w += s1;
    if (d == m2) w += s2;
    v[d] = w < 0 ? w + Ld : (w >= Ld ? w - Ld : w);
  }
  return dev_index<ND>(g, v);
}
// the same for arbitrary non-negative shifts k1, k2
// Device-index offsets of a site's +-1 neighbours.  x +- e_d lies in the
// sublattice with bit d flipped (+-Vh << (nd-1-d)) and at half-coordinate
// y_d or y_d +- 1, so every neighbour of x -- and x + e_mu - e_nu -- is s
// plus one or two of these offsets: additions in place of dev_index's
// coordinate arithmetic and multiplications on every neighbour load.
__device__ __forceinline__ long long enc_dim(const Geom& g, int d, int y) {
  return (long long)(y >> g.tb[d]) * g.ts[d] + ((y & ((1 << g.tb[d]) - 1)) << g.lsh[d]);
}
template <int ND>
struct Nbr {
  long long s;
  long long dp[ND], dm[ND];
  // runtime direction: compare-select chain, no private-array indexing
  __device__ __forceinline__ long long p(int d) const {
    long long v = dp[0];
#pragma unroll
    for (int k = 1; k < ND; ++k) v = d == k ? dp[k] : v;
    return v;
  }
  __device__ __forceinline__ long long m(int d) const {
    long long v = dm[0];
#pragma unroll
    for (int k = 1; k < ND; ++k) v = d == k ? dm[k] : v;
    return v;
  }
  __device__ __forceinline__ long long up(int mu) const { return s + p(mu); }              // x + e_mu
  __device__ __forceinline__ long long dn(int nu) const { return s + m(nu); }              // x - e_nu
  __device__ __forceinline__ long long updn(int mu, int nu) const { return s + p(mu) + m(nu); }
};
template <int ND>
__device__ __forceinline__ Nbr<ND> make_nbr(const Geom& g, const int* x, long long s) {
  Nbr<ND> n;
  n.s = s;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const long long D = (long long)g.Vh << (ND - 1 - d);
    const int y = x[d] >> 1;
    const bool odd = (x[d] & 1) != 0;
    const long long e0 = enc_dim(g, d, y);
    const int yp = odd ? (y + 1 == g.H[d] ? 0 : y + 1) : y;
    const int ym = odd ? y : (y == 0 ? g.H[d] - 1 : y - 1);
    n.dp[d] = (odd ? -D : D) + (enc_dim(g, d, yp) - e0);
    n.dm[d] = (odd ? -D : D) + (enc_dim(g, d, ym) - e0);
  }
  return n;
}

// ---------------------------------------------------------------------------
// link field access
// ---------------------------------------------------------------------------
// Array-of-structs-of-arrays: sites are grouped in blocks of 64 consecutive
// device indices, and a block stores all directions and colour components of
// its 64 sites contiguously, one run of NCOMP * 64 complex numbers per
// (block, mu) at ((s / 64) * nd + mu) * NCOMP * 64, so a wave's load of one
// link for 64 sites reads one contiguous run (9 KiB for FP64 SU(3)) instead of
// NCOMP runs in arrays a field-size apart.  Inside a run:
//   FP64: component c of lane l at c * 64 + l (one 16-byte complex per lane
//         and load instruction);
//   FP32: components 2k and 2k + 1 of lane l side by side at k * 128 + 2 l
//         (+1), and for odd NCOMP the last one at (NCOMP - 1) * 64 + l, so a
//         lane loads 16 bytes per instruction there too: an FP32 link takes 5
//         load instructions instead of 9 (two-row SU(3): 3 instead of 6).
//         The vector memory unit's work scales with the number of
//         instructions more than with their width.
template <typename R, int NC, bool COMP>
struct Links {
  static constexpr bool C = COMP && NC > 1;
  static constexpr int NCOMP = C ? (NC - 1) * NC : NC * NC;
  static constexpr bool PAIRED = sizeof(R) == 4;
  static constexpr int NPAIR = PAIRED ? NCOMP / 2 : 0;
  cx<R>* __restrict__ p;
  long long V;
  int nd;  // kernels templated on the direction count set nd = ND on entry,
           // which makes the index arithmetic below compile-time

  // start of the (block, mu) run holding site s
  __device__ __forceinline__ long long run(long long s, int mu) const {
    return (((s >> 6) * nd + mu) * NCOMP) * 64;
  }
  // element offset of component c of link (s, mu)
  __device__ __forceinline__ long long comp_index(long long s, int mu, int c) const {
    const long long l = s & 63;
    if constexpr (PAIRED) {
      if (c < 2 * NPAIR) return run(s, mu) + (c >> 1) * 128 + 2 * l + (c & 1);
      return run(s, mu) + (NCOMP - 1) * 64 + l;
    } else {
      return run(s, mu) + c * 64 + l;
    }
  }
  __device__ __forceinline__ Mat<R, NC> load(long long s, int mu) const {
    Mat<R, NC> u;
    const cx<R>* q = p + run(s, mu);
    const int l = (int)(s & 63);
    if constexpr (PAIRED) {
#pragma unroll
      for (int k = 0; k < NPAIR; ++k) {
        const float4 v = reinterpret_cast<const float4*>(q + k * 128)[l];
        u.m[(2 * k) / NC][(2 * k) % NC] = mk<R>(v.x, v.y);
        u.m[(2 * k + 1) / NC][(2 * k + 1) % NC] = mk<R>(v.z, v.w);
      }
      if constexpr (NCOMP & 1) u.m[(NCOMP - 1) / NC][(NCOMP - 1) % NC] = q[(NCOMP - 1) * 64 + l];
    } else {
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) u.m[c / NC][c % NC] = q[c * 64 + l];
    }
    if constexpr (C) {
      if constexpr (NC == 3) complete_row2(u);
      if constexpr (NC == 2) complete_row1(u);
    }
    return u;
  }
  __device__ __forceinline__ void store(long long s, int mu, const Mat<R, NC>& u) const {
    cx<R>* q = p + run(s, mu);
    const int l = (int)(s & 63);
    if constexpr (PAIRED) {
#pragma unroll
      for (int k = 0; k < NPAIR; ++k) {
        const cx<R> a = u.m[(2 * k) / NC][(2 * k) % NC], b = u.m[(2 * k + 1) / NC][(2 * k + 1) % NC];
        reinterpret_cast<float4*>(q + k * 128)[l] = make_float4(a.re, a.im, b.re, b.im);
      }
      if constexpr (NCOMP & 1) q[(NCOMP - 1) * 64 + l] = u.m[(NCOMP - 1) / NC][(NCOMP - 1) % NC];
    } else {
#pragma unroll
      for (int c = 0; c < NCOMP; ++c) q[c * 64 + l] = u.m[c / NC][c % NC];
    }
  }
};

// number of link-array elements (complex) for V sites
__host__ __device__ inline long long links_elems(long long V, int nd, int ncomp) {
  return ((V + 63) / 64) * 64 * nd * ncomp;
}

// ---------------------------------------------------------------------------
// staple, GaugeField.hpp:187-237 (3D :389-433, 2D :579-618):
//   S = sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//     + sum_{nu != mu} U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
// products associated left to right, the upper staples first.
// ---------------------------------------------------------------------------
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ Mat<R, NC> staple(const Links<R, NC, COMP>& f, const Geom& g,
                                             const int* x, int mu) {
  Mat<R, NC> t = zero<R, NC>();
  const long long s_x = dev_index<ND>(g, x);
  const long long s_xpmu = idx_shift<ND>(g, x, mu, 1, -1, 0);
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const Mat<R, NC> a = f.load(s_xpmu, nu);
    const Mat<R, NC> b = f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), mu);
    const Mat<R, NC> c = f.load(s_x, nu);
    acc_mul_ad(t, mul_ad(a, b), c);
  }
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
    const Mat<R, NC> a = f.load(idx_shift<ND>(g, x, mu, 1, nu, -1), nu);
    const Mat<R, NC> b = f.load(s_xmnu, mu);
    const Mat<R, NC> c = f.load(s_xmnu, nu);
    acc_mul(t, mul_dd(a, b), c);
  }
  return t;
}
// The same with the neighbours' device indices from precomputed offsets (the
// one-thread-per-link kernels: less integer work per link; the Metropolis
// kernels measured faster recomputing indices from the coordinates).
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ Mat<R, NC> staple(const Links<R, NC, COMP>& f, const Nbr<ND>& n, int mu) {
  Mat<R, NC> t = zero<R, NC>();
  const long long s_xpmu = n.up(mu);
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const Mat<R, NC> a = f.load(s_xpmu, nu);
    const Mat<R, NC> b = f.load(n.up(nu), mu);
    const Mat<R, NC> c = f.load(n.s, nu);
    acc_mul_ad(t, mul_ad(a, b), c);
  }
#pragma unroll
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const long long s_xmnu = n.dn(nu);
    const Mat<R, NC> a = f.load(n.updn(mu, nu), nu);
    const Mat<R, NC> b = f.load(s_xmnu, mu);
    const Mat<R, NC> c = f.load(s_xmnu, nu);
    acc_mul(t, mul_dd(a, b), c);
  }
  return t;
}

// The same staple with one term in flight at a time: the compiler cannot
// hoist all 2(nd-1)*3 link loads, which keeps register use low enough for
// several waves per SIMD (used by the one-thread-per-link kernels).
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ Mat<R, NC> staple_lowreg(const Links<R, NC, COMP>& f, const Geom& g,
                                                    const int* x, int mu) {
  Mat<R, NC> t = zero<R, NC>();
  const long long s_x = dev_index<ND>(g, x);
  const long long s_xpmu = idx_shift<ND>(g, x, mu, 1, -1, 0);
#pragma unroll 1
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const Mat<R, NC> ab = mul_ad(f.load(s_xpmu, nu), f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), mu));
    acc_mul_ad(t, ab, f.load(s_x, nu));
  }
#pragma unroll 1
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const long long s_xmnu = idx_shift<ND>(g, x, nu, -1, -1, 0);
    const Mat<R, NC> ab = mul_dd(f.load(idx_shift<ND>(g, x, mu, 1, nu, -1), nu), f.load(s_xmnu, mu));
    acc_mul(t, ab, f.load(s_xmnu, nu));
  }
  return t;
}
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ Mat<R, NC> staple_lowreg(const Links<R, NC, COMP>& f, const Nbr<ND>& n, int mu) {
  Mat<R, NC> t = zero<R, NC>();
  const long long s_xpmu = n.up(mu);
#pragma unroll 1
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const Mat<R, NC> ab = mul_ad(f.load(s_xpmu, nu), f.load(n.up(nu), mu));
    acc_mul_ad(t, ab, f.load(n.s, nu));
  }
#pragma unroll 1
  for (int nu = 0; nu < ND; ++nu) {
    if (nu == mu) continue;
    const long long s_xmnu = n.dn(nu);
    const Mat<R, NC> ab = mul_dd(f.load(n.updn(mu, nu), nu), f.load(s_xmnu, mu));
    acc_mul(t, ab, f.load(s_xmnu, nu));
  }
  return t;
}

// plaquette at one site (GaugePlaquette.hpp:52-85), real part only:
// p(x) = sum_{mu<nu} Re Tr[(U_mu(x) U_nu(x+mu)) (U_nu(x) U_mu(x+nu))^+]
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ double plaq_at(const Links<R, NC, COMP>& f, const Geom& g, const int* x,
                                          long long s) {
  double p = 0.0;
#pragma unroll 1
  for (int mu = 0; mu < ND - 1; ++mu) {
    const long long sm = idx_shift<ND>(g, x, mu, 1, -1, 0);
#pragma unroll 1
    for (int nu = mu + 1; nu < ND; ++nu) {
      // the plane's four loads are issued together before any product (the
      // scheduler otherwise interleaves them with the FMA chains and waits
      // on each in turn)
      const Mat<R, NC> a = f.load(s, mu), b = f.load(sm, nu);
      const Mat<R, NC> c = f.load(s, nu), d = f.load(idx_shift<ND>(g, x, nu, 1, -1, 0), mu);
      __builtin_amdgcn_sched_barrier(0);
      p += (double)retr_mul_ad(mul(a, b), mul(c, d));
    }
  }
  return p;
}

// the same sum with the site's own links loaded once and every plane's
// loads in flight together (same term order, so the same result)
template <typename R, int NC, bool COMP, int ND>
__device__ __forceinline__ double plaq_at_full(const Links<R, NC, COMP>& f, const Nbr<ND>& n) {
  Mat<R, NC> U[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) U[d] = f.load(n.s, d);
  double p = 0.0;
  // the planes of one mu form a group whose loads are issued together
#pragma unroll
  for (int mu = 0; mu < ND - 1; ++mu) {
    Mat<R, NC> a[ND], b[ND];
#pragma unroll
    for (int nu = mu + 1; nu < ND; ++nu) {
      a[nu] = f.load(n.s + n.dp[mu], nu);
      b[nu] = f.load(n.s + n.dp[nu], mu);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int nu = mu + 1; nu < ND; ++nu) p += (double)retr_mul_ad(mul(U[mu], a[nu]), mul(U[nu], b[nu]));
  }
  return p;
}

// waves per SIMD requested from the register allocator
template <typename R, int NC>
struct Occ {
  static constexpr int value = (sizeof(R) == 8 && NC == 3) ? 3 : 4;
};

// ---------------------------------------------------------------------------
// Metropolis proposals, randSUN<1,2,3> (SUN.hpp:163-251), drawn from the
// keyed stream in the reference's order.  FP64 draws take 53-bit doubles;
// FP32 draws take 24-bit floats (half the generator work).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double draw(klft_rng::Stream& s, double lo, double hi) {
  return lo + (hi - lo) * s.u01();
}
__device__ __forceinline__ float draw(klft_rng::Stream& s, float lo, float hi) {
  return lo + (hi - lo) * s.u01f();
}
template <typename R>
__device__ __forceinline__ R unif(klft_rng::Stream& s);
template <>
__device__ __forceinline__ double unif<double>(klft_rng::Stream& s) {
  return s.u01();
}
template <>
__device__ __forceinline__ float unif<float>(klft_rng::Stream& s) {
  return s.u01f();
}

#define KLFT_PI 3.14159265358979323846

template <typename R>
__device__ __forceinline__ void rand_sun(klft_rng::Stream& s, R delta, Mat<R, 1>& r) {
  const R pi = (R)KLFT_PI;
  const R th = draw(s, -delta * pi, delta * pi);
  r.m[0][0] = mk(cos(th), sin(th));
}
template <typename R>
__device__ __forceinline__ void rand_sun(klft_rng::Stream& s, R delta, Mat<R, 2>& r) {
  const R pi = (R)KLFT_PI;
  const R alpha = draw(s, R(0), delta * 2 * pi);
  const R u = draw(s, R(-1), R(1));
  const R theta = draw(s, R(0), R(2) * pi);
  const R sa = sin(alpha);
  const R n1 = sqrt(R(1) - u * u) * sin(theta);
  const R n2 = sqrt(R(1) - u * u) * cos(theta);
  r.m[0][0] = mk(cos(alpha), u * sa);
  r.m[0][1] = mk(n1 * sa, n2 * sa);
  r.m[1][0] = mk(-r.m[0][1].re, r.m[0][1].im);
  r.m[1][1] = mk(r.m[0][0].re, -r.m[0][0].im);
}
template <typename R>
__device__ __forceinline__ void rand_sun(klft_rng::Stream& s, R delta, Mat<R, 3>& r) {
  R v[6], norm, fact;
  cx<R> z1[3], z2[3];
  for (;;) {
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] = draw(s, R(0), delta);
    norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    if (R(1) != (R(1) + norm)) break;
  }
  fact = R(1) / norm;
#pragma unroll
  for (int i = 0; i < 3; ++i) z1[i] = mk(fact * v[2 * i], fact * v[2 * i + 1]);
  for (;;) {
    for (;;) {
#pragma unroll
      for (int i = 0; i < 6; ++i) v[i] = draw(s, R(0), delta);
      norm =
          sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
      if (R(1) != (R(1) + norm)) break;
    }
    fact = R(1) / norm;
#pragma unroll
    for (int i = 0; i < 3; ++i) z2[i] = mk(fact * v[2 * i], fact * v[2 * i + 1]);
    const cx<R> z = cmul_(z1[0], z2[0]) + cmul_(z1[1], z2[1]) + cmul_(z1[2], z2[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) z2[i] = z2[i] - z * z1[i];
    norm = sqrt(z2[0].re * z2[0].re + z2[0].im * z2[0].im + z2[1].re * z2[1].re +
                z2[1].im * z2[1].im + z2[2].re * z2[2].re + z2[2].im * z2[2].im);
    if (R(1) != (R(1) + norm)) break;
  }
  fact = R(1) / norm;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    z2[i] = scale(fact, z2[i]);
    r.m[0][i] = z1[i];
    r.m[1][i] = z2[i];
  }
  complete_row2(r);
}

// ---------------------------------------------------------------------------
// One Metropolis hit's draws -- the proposal and the accept uniform that
// follows it in the hit's stream -- with every Philox block of the hit
// computed up front.  The blocks are independent, so their rounds interleave
// instead of forming one long dependent chain (a pass has one wave per SIMD
// at 32^4, so nothing else hides that latency).  The SU(3) proposal's
// rejection loops (SUN.hpp:198-233) repeat only when a norm is below 2^-53;
// then the hit is redrawn by the looping rand_sun above from the start of the
// same stream, which gives the same values, so the outcome is identical.
// ---------------------------------------------------------------------------
template <typename R>
struct DrawWords {  // stream words per uniform
  static constexpr int value = sizeof(R) == 8 ? 2 : 1;
};
template <int NC>
struct HitDraws;  // uniforms per hit, the accept draw included
template <>
struct HitDraws<1> {
  static constexpr int value = 2;
};
template <>
struct HitDraws<2> {
  static constexpr int value = 4;
};
template <>
struct HitDraws<3> {
  static constexpr int value = 13;
};

// uniform number i of a hit from its words (Stream::u01 / u01f arithmetic)
template <typename R>
__device__ __forceinline__ R u01_at(const uint32_t* w, int i);
template <>
__device__ __forceinline__ double u01_at<double>(const uint32_t* w, int i) {
  const uint32_t a = w[2 * i] >> 5;
  const uint32_t b = w[2 * i + 1] >> 6;
  return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}
template <>
__device__ __forceinline__ float u01_at<float>(const uint32_t* w, int i) {
  return (float)(w[i] >> 8) * (1.0f / 16777216.0f);
}
template <typename R>
__device__ __forceinline__ R draw_at(const uint32_t* w, int i, R lo, R hi) {
  return lo + (hi - lo) * u01_at<R>(w, i);
}

// first-attempt proposals; false when a rejection loop would repeat
template <typename R>
__device__ __forceinline__ bool rand_sun_w(const uint32_t* w, R delta, Mat<R, 1>& r) {
  const R pi = (R)KLFT_PI;
  const R th = draw_at<R>(w, 0, -delta * pi, delta * pi);
  r.m[0][0] = mk(cos(th), sin(th));
  return true;
}
template <typename R>
__device__ __forceinline__ bool rand_sun_w(const uint32_t* w, R delta, Mat<R, 2>& r) {
  const R pi = (R)KLFT_PI;
  const R alpha = draw_at<R>(w, 0, R(0), delta * 2 * pi);
  const R u = draw_at<R>(w, 1, R(-1), R(1));
  const R theta = draw_at<R>(w, 2, R(0), R(2) * pi);
  const R sa = sin(alpha);
  const R n1 = sqrt(R(1) - u * u) * sin(theta);
  const R n2 = sqrt(R(1) - u * u) * cos(theta);
  r.m[0][0] = mk(cos(alpha), u * sa);
  r.m[0][1] = mk(n1 * sa, n2 * sa);
  r.m[1][0] = mk(-r.m[0][1].re, r.m[0][1].im);
  r.m[1][1] = mk(r.m[0][0].re, -r.m[0][0].im);
  return true;
}
template <typename R>
__device__ __forceinline__ bool rand_sun_w(const uint32_t* w, R delta, Mat<R, 3>& r) {
  R v[6];
  cx<R> z1[3], z2[3];
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = draw_at<R>(w, i, R(0), delta);
  R norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
  if (!(R(1) != (R(1) + norm))) return false;
  R fact = R(1) / norm;
#pragma unroll
  for (int i = 0; i < 3; ++i) z1[i] = mk(fact * v[2 * i], fact * v[2 * i + 1]);
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = draw_at<R>(w, 6 + i, R(0), delta);
  norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
  if (!(R(1) != (R(1) + norm))) return false;
  fact = R(1) / norm;
#pragma unroll
  for (int i = 0; i < 3; ++i) z2[i] = mk(fact * v[2 * i], fact * v[2 * i + 1]);
  const cx<R> z = cmul_(z1[0], z2[0]) + cmul_(z1[1], z2[1]) + cmul_(z1[2], z2[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) z2[i] = z2[i] - z * z1[i];
  norm = sqrt(z2[0].re * z2[0].re + z2[0].im * z2[0].im + z2[1].re * z2[1].re + z2[1].im * z2[1].im +
              z2[2].re * z2[2].re + z2[2].im * z2[2].im);
  if (!(R(1) != (R(1) + norm))) return false;
  fact = R(1) / norm;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    z2[i] = scale(fact, z2[i]);
    r.m[0][i] = z1[i];
    r.m[1][i] = z2[i];
  }
  complete_row2(r);
  return true;
}

// proposal r and accept uniform u of Metropolis hit `hit` of link (site, mu)
template <typename R, int NC>
__device__ __forceinline__ void propose(uint64_t seed, uint32_t site, uint32_t sweep, uint32_t hit,
                                        uint32_t chain, uint32_t mu, R delta, Mat<R, NC>& r, R& u) {
  constexpr int NB = (HitDraws<NC>::value * DrawWords<R>::value + 3) / 4;
  const klft_rng::Stream st(seed, klft_rng::TAG_METROPOLIS, site, sweep, hit, chain, mu);
  uint32_t w[4 * NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) st.block_words((uint32_t)j, w + 4 * j);
  if (rand_sun_w(w, delta, r)) {
    u = u01_at<R>(w, HitDraws<NC>::value - 1);
    return;
  }
  klft_rng::Stream s2 = st;
  rand_sun(s2, delta, r);
  u = unif<R>(s2);
}

// ---------------------------------------------------------------------------
// Lie algebra: Q = sum_a v_a T_a with T = 1 (U(1)), Pauli sigma (SU(2)),
// Gell-Mann lambda (SU(3), Tr T_a T_b = 2 delta_ab).
// ---------------------------------------------------------------------------
template <int NC>
struct Ngen;
template <>
struct Ngen<1> {
  static constexpr int value = 1;
};
template <>
struct Ngen<2> {
  static constexpr int value = 3;
};
template <>
struct Ngen<3> {
  static constexpr int value = 8;
};

// F_a = c * Im Tr(T_a M)  (AdjointGroup.hpp:13-19 for SU(2), :114-117 for U(1))
template <typename R>
__device__ __forceinline__ void project(const Mat<R, 1>& M, R c, R* F) {
  F[0] = c * M.m[0][0].im;
}
template <typename R>
__device__ __forceinline__ void project(const Mat<R, 2>& M, R c, R* F) {
  F[0] = c * (M.m[0][1].im + M.m[1][0].im);
  F[1] = c * (M.m[0][1].re - M.m[1][0].re);
  F[2] = c * (M.m[0][0].im - M.m[1][1].im);
}
template <typename R>
__device__ __forceinline__ void project(const Mat<R, 3>& M, R c, R* F) {
  const R r3 = (R)0.57735026918962576451;
  F[0] = c * (M.m[0][1].im + M.m[1][0].im);
  F[1] = c * (M.m[0][1].re - M.m[1][0].re);
  F[2] = c * (M.m[0][0].im - M.m[1][1].im);
  F[3] = c * (M.m[0][2].im + M.m[2][0].im);
  F[4] = c * (M.m[0][2].re - M.m[2][0].re);
  F[5] = c * (M.m[1][2].im + M.m[2][1].im);
  F[6] = c * (M.m[1][2].re - M.m[2][1].re);
  F[7] = c * ((M.m[0][0].im + M.m[1][1].im - R(2) * M.m[2][2].im) * r3);
}

// exp(i v) for U(1) (AdjointGroup.hpp:170-173)
template <typename R>
__device__ __forceinline__ Mat<R, 1> exp_iq(const R* v, Mat<R, 1>*) {
  Mat<R, 1> e;
  e.m[0][0] = mk(cos(v[0]), sin(v[0]));
  return e;
}
// exp(i v.sigma) for SU(2) (AdjointGroup.hpp:98-105); alpha = 0 gives the
// identity instead of the reference's division by zero.
template <typename R>
__device__ __forceinline__ Mat<R, 2> exp_iq(const R* v, Mat<R, 2>*) {
  const R alpha = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (alpha == R(0)) return identity<R, 2>();
  const R n0 = v[0] / alpha, n1 = v[1] / alpha, n2 = v[2] / alpha;
  const R sa = sin(alpha), ca = cos(alpha);
  Mat<R, 2> e;
  e.m[0][0] = mk(ca, n2 * sa);
  e.m[0][1] = mk(n1 * sa, n0 * sa);
  e.m[1][0] = mk(-n1 * sa, n0 * sa);
  e.m[1][1] = mk(ca, -n2 * sa);
  return e;
}
// exp(i Q) for SU(3) by the Cayley-Hamilton reduction of the exponential
// series (Morningstar & Peardon, Phys. Rev. D 69, 054501 (2004), sec. III):
// Q^3 = c1 Q + c0 I with c0 = det Q, c1 = Tr Q^2 / 2, so
// exp(iQ) = f0 I + f1 Q + f2 Q^2, and the coefficients of Q^n in the basis
// {I, Q, Q^2} follow a_{n+1} = c0 c_n, b_{n+1} = a_n + c1 c_n, c_{n+1} = b_n.
// The series is summed until the next term is below the working precision.
template <typename R>
__device__ __forceinline__ Mat<R, 3> exp_iq(const R* v, Mat<R, 3>*) {
  const R r3 = (R)0.57735026918962576451;
  Mat<R, 3> q;
  q.m[0][0] = mk(v[2] + v[7] * r3, R(0));
  q.m[1][1] = mk(-v[2] + v[7] * r3, R(0));
  q.m[2][2] = mk(R(-2) * v[7] * r3, R(0));
  q.m[0][1] = mk(v[0], -v[1]);
  q.m[1][0] = mk(v[0], v[1]);
  q.m[0][2] = mk(v[3], -v[4]);
  q.m[2][0] = mk(v[3], v[4]);
  q.m[1][2] = mk(v[5], -v[6]);
  q.m[2][1] = mk(v[5], v[6]);
  const Mat<R, 3> q2 = mul(q, q);
  const R c1 = R(0.5) * retrace(q2);
  const R c0 = retr_mul(q2, q) / R(3);  // Tr Q^3 / 3 = det Q
  // a, b, c: coefficients of Q^n / n! ; f accumulates i^n Q^n / n!
  R a = 1, b = 0, c = 0;
  cx<R> f0 = mk<R>(1, 0), f1 = mk<R>(0, 0), f2 = mk<R>(0, 0);
  const R eps = sizeof(R) == 8 ? (R)1e-18 : (R)1e-9;
  // spectral radius bound: sum of squared eigenvalues is 2 c1
  const R rho = sqrt(R(2) * c1);
  for (int n = 1; n < 64; ++n) {
    const R inv = R(1) / (R)n;
    const R an = c0 * c * inv, bn = (a + c1 * c) * inv, cn = b * inv;
    a = an;
    b = bn;
    c = cn;
    // i^n
    switch (n & 3) {
      case 0: f0.re += a; f1.re += b; f2.re += c; break;
      case 1: f0.im += a; f1.im += b; f2.im += c; break;
      case 2: f0.re -= a; f1.re -= b; f2.re -= c; break;
      default: f0.im -= a; f1.im -= b; f2.im -= c; break;
    }
    if (n > 3 && fabs(a) + (fabs(b) + fabs(c) * rho) * rho < eps) break;
  }
  Mat<R, 3> e;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cx<R> t = f1 * q.m[i][j] + f2 * q2.m[i][j];
      if (i == j) t = t + f0;
      e.m[i][j] = t;
    }
  return e;
}

// The opt-in symmetric SU(3) proposal (no counterpart in the reference, whose
// randSUN<3> is asymmetric, SUN.hpp:192-251): r = exp(i delta sum_a h_a
// lambda_a), h_a ~ U(-1, 1) for a = 0..7 in stream order (tag 9), then the
// accept uniform.  h -> -h maps r to r^dagger with the same density, so the
// proposal is symmetric and the Metropolis chain satisfies detailed balance.
// U(1) and SU(2) proposals (SUN.hpp:163-191) are symmetric already.
template <typename R, int NC, bool SYM>
__device__ __forceinline__ void propose_law(uint64_t seed, uint32_t site, uint32_t sweep, uint32_t hit,
                                            uint32_t chain, uint32_t mu, R delta, Mat<R, NC>& r, R& u) {
  if constexpr (SYM && NC == 3) {
    constexpr int NB = (9 * DrawWords<R>::value + 3) / 4;
    const klft_rng::Stream st(seed, klft_rng::TAG_METROPOLIS_SYM, site, sweep, hit, chain, mu);
    uint32_t w[4 * NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) st.block_words((uint32_t)j, w + 4 * j);
    R v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) v[a] = delta * draw_at<R>(w, a, R(-1), R(1));
    r = exp_iq(v, (Mat<R, 3>*)nullptr);
    u = u01_at<R>(w, 8);
  } else {
    propose<R, NC>(seed, site, sweep, hit, chain, mu, delta, r, u);
  }
}

// ---------------------------------------------------------------------------
// reductions (wave64 shuffles -> LDS -> per-block partial)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over the block in a fixed order; the result is valid in thread 0
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (int)((blockDim.x + 63) / 64);
#pragma unroll
    for (int i = 0; i < BLOCK / 64; ++i)
      if (i < nw) t += lds[i];
  }
  __syncthreads();
  return t;
}

// Block partial for kernels that write large outputs: a grid_finish release
// there would write back the XCD's dirty L2 lines once per block, so the
// partials are summed by a second launch (k_sum_blocks) in block order.
template <int BLOCK>
__device__ __forceinline__ void block_partial(double v, double* lds, double* partials) {
  const double bt = block_sum<BLOCK>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = bt;
}

// Single-pass grid reduction: every block writes its partial, the last block
// to arrive (agent-scope ticket) sums the partials in block order and resets
// the ticket, so the result is independent of scheduling.
template <int BLOCK>
__device__ __forceinline__ void grid_finish(double block_total, double* partials,
                                            unsigned int* ticket, double* out, double* lds) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = block_total;
    const unsigned int t =
        __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  double v = 0.0;
  for (unsigned int i = threadIdx.x; i < gridDim.x; i += blockDim.x)
    v += __hip_atomic_load(&partials[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  v = block_sum<BLOCK>(v, lds);
  if (threadIdx.x == 0) {
    *out = v;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace kd
