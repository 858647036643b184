// libklft_hip core: device context, gauge-field handles, host interchange,
// field initialisation, plaquette, staple field and Wilson loops.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace kd;
namespace ki = klft_internal;

// ============================================================================
// context and errors
// ============================================================================
namespace klft_internal {

static thread_local std::string g_err;
static std::mutex g_init_mu;
static bool g_inited = false;
static int g_device = 0;
static hipStream_t g_stream = nullptr;

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return set_error(KLFT_E_HIP, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
}

static int init_device(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return set_error(KLFT_E_HIP, "no HIP device available");
  if (device < 0 || device >= n) return set_error(KLFT_E_ARG, "device %d out of range", device);
  KLFT_HIP(hipSetDevice(device));
  KLFT_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  g_inited = true;
  return KLFT_OK;
}

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_inited) {
    KLFT_HIP(hipSetDevice(g_device));
    return KLFT_OK;
  }
  return init_device(0);
}

hipStream_t stream() { return g_stream; }

int sync_stream() {
  KLFT_HIP(hipStreamSynchronize(g_stream));
  return KLFT_OK;
}

Geom make_geom(int nd, const int *dims) {
  Geom g{};
  g.nd = nd;
  g.V = 1;
  g.Vh = 1;
  for (int d = 0; d < 4; ++d) {
    g.L[d] = d < nd ? dims[d] : 1;
    g.H[d] = d < nd ? dims[d] / 2 : 1;
  }
  for (int d = 0; d < nd; ++d) {
    g.V *= g.L[d];
    g.Vh *= g.H[d];
  }
  // Sites of a sublattice are in lexicographic order (last index fastest), so
  // at 32^4 a 64-site block is 4 rows of 16 and every neighbour load of a
  // wave touches whole 128-byte lines.  KLFT_TILES=1 in the environment
  // switches to 4D tiles of extent 4 (else 2, else 1) inside each sublattice
  // (measured no faster: the L2 cannot hold a pass's working set either way).
  static const bool tiles = [] {
    const char *e = std::getenv("KLFT_TILES");
    return e && e[0] == '1';
  }();
  g.TV = 1;
  g.tvb = 0;
  for (int d = 0; d < 4; ++d) {
    g.tb[d] = 0;
    if (d < nd && tiles) g.tb[d] = (g.H[d] % 4 == 0) ? 2 : ((g.H[d] % 2 == 0) ? 1 : 0);
    g.NT[d] = g.H[d] >> g.tb[d];
    g.TV <<= g.tb[d];
    g.tvb += g.tb[d];
    g.ts[d] = 0;
    g.ls[d] = 0;
    g.lsh[d] = 0;
  }
  g.ls[nd - 1] = 1;
  g.ts[nd - 1] = g.TV;
  g.lsh[nd - 1] = 0;
  for (int d = nd - 2; d >= 0; --d) {
    g.ls[d] = g.ls[d + 1] * g.L[d + 1];
    g.ts[d] = g.ts[d + 1] * g.NT[d + 1];
    g.lsh[d] = g.lsh[d + 1] + g.tb[d + 1];
  }
  return g;
}

int check_field(const klft_field_s *f) {
  if (!f || !f->links) return set_error(KLFT_E_ARG, "null or destroyed field handle");
  return KLFT_OK;
}

}  // namespace klft_internal

extern "C" int klft_init(int device) {
  std::lock_guard<std::mutex> lk(ki::g_init_mu);
  if (ki::g_inited) {
    if (device == ki::g_device) return KLFT_OK;
    return ki::set_error(KLFT_E_STATE, "library already initialised on device %d", ki::g_device);
  }
  return ki::init_device(device);
}

extern "C" const char *klft_last_error(void) { return ki::g_err.c_str(); }

extern "C" const char *klft_version(void) { return "klft_hip 1.0 gfx950"; }

// ============================================================================
// kernels: interchange and initialisation
// ============================================================================
template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(256) void k_upload(Links<R, NC, COMP> f, Geom g,
                                                 const double *__restrict__ host_img) {
  f.nd = ND;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;  // lexicographic
  if (s >= g.V) return;
  int x[ND];
  lex_coords<ND>(g, s, x);
  const long long d = dev_index<ND>(g, x);
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    Mat<double, NC> u;
    const double *p = host_img + 2 * ((size_t)s * ND + mu) * NC * NC;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) u.m[i][j] = mk(p[2 * (i * NC + j)], p[2 * (i * NC + j) + 1]);
    f.store(d, mu, convert<R>(u));
  }
}

template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(256) void k_download(Links<R, NC, COMP> f, Geom g,
                                                   double *__restrict__ host_img) {
  f.nd = ND;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= g.V) return;
  int x[ND];
  lex_coords<ND>(g, s, x);
  const long long d = dev_index<ND>(g, x);
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    const Mat<R, NC> u = f.load(d, mu);
    double *p = host_img + 2 * ((size_t)s * ND + mu) * NC * NC;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        p[2 * (i * NC + j)] = (double)u.m[i][j].re;
        p[2 * (i * NC + j) + 1] = (double)u.m[i][j].im;
      }
  }
}

// kind: 0 identity, 1 randSUN(delta), 2 exp(i eps Q), 3 Haar, 4 every entry
// an independent complex (drand(-1,1), drand(-1,1)), 5 every entry the
// constant (param, param2)
template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(256) void k_init(Links<R, NC, COMP> f, Geom g, int kind,
                                               double param, double param2, uint64_t seed) {
  f.nd = ND;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;  // lexicographic
  if (s >= g.V) return;
  int x[ND];
  lex_coords<ND>(g, s, x);
  const long long d = dev_index<ND>(g, x);
  constexpr int NG = Ngen<NC>::value;
  for (int mu = 0; mu < ND; ++mu) {
    Mat<double, NC> u;
    if (kind == 0) {
      u = identity<double, NC>();
    } else if (kind == 1) {
      klft_rng::Stream st(seed, klft_rng::TAG_RANDSUN, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      rand_sun(st, param, u);
    } else if (kind == 4) {
      // GaugeField.hpp:118-141: real part, then imaginary part, c1 outer, c2 inner
      klft_rng::Stream st(seed, klft_rng::TAG_COMPLEX, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const double re = draw(st, -1.0, 1.0);
          const double im = draw(st, -1.0, 1.0);
          u.m[i][j] = mk(re, im);
        }
    } else if (kind == 5) {
      // GaugeField.hpp:65-84
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) u.m[i][j] = mk(param, param2);
    } else if (kind == 2) {
      klft_rng::Stream st(seed, klft_rng::TAG_ALGEBRA, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      double h[NG + 1];
#pragma unroll
      for (int a = 0; a < NG; a += 2) {
        const double u1 = 1.0 - st.u01(), u2 = st.u01();
        const double r = sqrt(-2.0 * log(u1)), t = 2.0 * KLFT_PI * u2;
        h[a] = r * cos(t);
        h[a + 1] = r * sin(t);
      }
      double v[NG];
#pragma unroll
      for (int a = 0; a < NG; ++a) v[a] = param * h[a] * 0.5;
      u = exp_iq(v, (Mat<double, NC> *)nullptr);
    } else {
      klft_rng::Stream st(seed, klft_rng::TAG_HAAR, (uint32_t)s, 0, 0, 0, (uint32_t)mu);
      double z[2 * NC * NC + 2];
#pragma unroll
      for (int a = 0; a < 2 * NC * NC; a += 2) {
        const double u1 = 1.0 - st.u01(), u2 = st.u01();
        const double r = sqrt(-2.0 * log(u1)), t = 2.0 * KLFT_PI * u2;
        z[a] = r * cos(t);
        z[a + 1] = r * sin(t);
      }
      if constexpr (NC == 1) {
        u.m[0][0] = mk(z[0], z[1]);
        restore_sun(u);
      } else if constexpr (NC == 2) {
        u.m[0][0] = mk(z[0], z[3]);
        u.m[0][1] = mk(z[2], z[1]);
        complete_row1(u);
        restore_sun(u);
      } else {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) u.m[i][j] = mk(z[2 * (3 * i + j)], z[2 * (3 * i + j) + 1]);
        // Gram-Schmidt in FP64, third row conj(r0 x r1)
        double n0 = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) n0 += u.m[0][j].re * u.m[0][j].re + u.m[0][j].im * u.m[0][j].im;
        n0 = 1.0 / sqrt(n0);
#pragma unroll
        for (int j = 0; j < 3; ++j) u.m[0][j] = scale(n0, u.m[0][j]);
        cx<double> p = cmul_(u.m[0][0], u.m[1][0]) + cmul_(u.m[0][1], u.m[1][1]) +
                       cmul_(u.m[0][2], u.m[1][2]);
#pragma unroll
        for (int j = 0; j < 3; ++j) u.m[1][j] = u.m[1][j] - p * u.m[0][j];
        double n1 = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) n1 += u.m[1][j].re * u.m[1][j].re + u.m[1][j].im * u.m[1][j].im;
        n1 = 1.0 / sqrt(n1);
#pragma unroll
        for (int j = 0; j < 3; ++j) u.m[1][j] = scale(n1, u.m[1][j]);
        complete_row2(u);
      }
    }
    f.store(d, mu, convert<R>(u));
  }
}

template <typename R, int NC, bool COMP, int ND>
__global__ __launch_bounds__(256) void k_reunit(Links<R, NC, COMP> f, Geom g) {
  f.nd = ND;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= g.V) return;
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) {
    Mat<R, NC> u = f.load(s, mu);
    reunitarize(u);
    f.store(s, mu, u);
  }
}

template <typename RO, int NC, bool CO, typename RI, bool CI, int ND>
__global__ __launch_bounds__(256) void k_convert(Links<RO, NC, CO> o, Links<RI, NC, CI> in, Geom g) {
  o.nd = ND;
  in.nd = ND;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= g.V) return;
#pragma unroll
  for (int mu = 0; mu < ND; ++mu) o.store(s, mu, convert<RO>(in.load(s, mu)));
}

// ============================================================================
// field lifecycle
// ============================================================================
extern "C" int klft_field_create(int nd, int nc, const int *dims, int storage, klft_field_t *out) {
  if (!out || !dims) return ki::set_error(KLFT_E_ARG, "null argument");
  *out = nullptr;
  if (nd < 2 || nd > 4) return ki::set_error(KLFT_E_ARG, "nd=%d unsupported (2..4)", nd);
  if (nc < 1 || nc > 3) return ki::set_error(KLFT_E_ARG, "Nc=%d unsupported (1..3)", nc);
  if (storage < KLFT_STORAGE_F64 || storage > KLFT_STORAGE_F32_ROWS)
    return ki::set_error(KLFT_E_ARG, "storage=%d unsupported", storage);
  long long V = 1;
  for (int d = 0; d < nd; ++d) {
    if (dims[d] < 2 || (dims[d] & 1))
      return ki::set_error(KLFT_E_ARG, "extent L%d=%d must be even and >= 2", d, dims[d]);
    V *= dims[d];
  }
  if (V >= (1LL << 32)) return ki::set_error(KLFT_E_ARG, "volume %lld exceeds 2^32 sites", V);
  int rc = ki::ensure_init();
  if (rc) return rc;
  klft_field_s *f = new klft_field_s();
  f->nd = nd;
  f->nc = nc;
  f->storage = storage;
  for (int d = 0; d < nd; ++d) f->dims[d] = dims[d];
  f->g = ki::make_geom(nd, dims);
  const bool comp = storage == KLFT_STORAGE_F32_ROWS && nc > 1;
  f->ncomp = comp ? (nc - 1) * nc : nc * nc;
  f->cbytes = storage == KLFT_STORAGE_F64 ? 16 : 8;
  f->link_bytes = (size_t)kd::links_elems(V, nd, f->ncomp) * f->cbytes;
  hipError_t e = hipMalloc(&f->links, f->link_bytes);
  if (e != hipSuccess) {
    delete f;
    return ki::set_error(KLFT_E_NOMEM, "hipMalloc of %zu link bytes failed", (size_t)V * nd * 16);
  }
  f->max_partials = std::max<long long>(ki::kMinPartials, kd::link_blocks<4>(f->g) + 64);
  const size_t pbytes = ((sizeof(double) * (size_t)f->max_partials + 255) / 256) * 256;
  size_t ws = pbytes + 256 + 16 * sizeof(double) + 8 * sizeof(unsigned long long);
  char *w = nullptr;
  e = hipMalloc((void **)&w, ws);
  if (e != hipSuccess) {
    (void)hipFree(f->links);
    delete f;
    return ki::set_error(KLFT_E_NOMEM, "workspace allocation failed");
  }
  f->partials = (double *)w;
  f->ticket = (unsigned int *)(w + pbytes);
  f->result = (double *)(w + pbytes + 256);
  f->counter = (unsigned long long *)(w + pbytes + 256 + 16 * sizeof(double));
  // Both clears go on the library stream and are waited for: that stream is
  // non-blocking, so a clear on the null stream would not be ordered before
  // the initialiser kernel that follows, and could land on links it had
  // already written.
  e = hipMemsetAsync(w, 0, ws, ki::stream());
  // the links start zeroed, not as whatever the allocation held: the
  // reference has no field without an initialiser (GaugeField.hpp:33-62),
  // and here one of the set_* calls or an upload gives the start
  if (e == hipSuccess) e = hipMemsetAsync(f->links, 0, f->link_bytes, ki::stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
  if (e == hipSuccess) e = hipHostMalloc((void **)&f->host_result, 16 * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipFree(w);
    (void)hipFree(f->links);
    delete f;
    return ki::hip_fail(e, "workspace init");
  }
  *out = f;
  return KLFT_OK;
}

extern "C" int klft_field_destroy(klft_field_t f) {
  if (!f) return KLFT_OK;
  (void)ki::sync_stream();
  if (f->links) (void)hipFree(f->links);
  if (f->partials) (void)hipFree(f->partials);
  if (f->host_result) (void)hipHostFree(f->host_result);
  if (f->run_counters) (void)hipFree(f->run_counters);
  if (f->meas_ws) (void)hipFree(f->meas_ws);
  f->links = nullptr;
  delete f;
  return KLFT_OK;
}

int klft_internal::meas_workspace(klft_field_t f, size_t bytes, void **out) {
  if (bytes > f->meas_capacity) {
    if (f->meas_ws) {
      KLFT_HIP(hipStreamSynchronize(ki::stream()));
      (void)hipFree(f->meas_ws);
      f->meas_ws = nullptr;
      f->meas_capacity = 0;
    }
    KLFT_HIP(hipMalloc(&f->meas_ws, bytes));
    f->meas_capacity = bytes;
  }
  *out = f->meas_ws;
  return KLFT_OK;
}

extern "C" int klft_field_info(klft_field_t f, int *nd, int *nc, int *dims, int *storage) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (nd) *nd = f->nd;
  if (nc) *nc = f->nc;
  if (dims)
    for (int d = 0; d < f->nd; ++d) dims[d] = f->dims[d];
  if (storage) *storage = f->storage;
  return KLFT_OK;
}

static int launch_init(klft_field_t f, int kind, double param, uint64_t seed, double param2 = 0.0) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  rc = ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    hipLaunchKernelGGL((k_init<R, NC, COMP, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256), 0,
                       ki::stream(), ki::links_of<R, NC, COMP>(f), f->g, kind, param, param2, seed);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
  if (rc) return rc;
  return ki::sync_stream();
}

extern "C" int klft_field_set_identity(klft_field_t f) { return launch_init(f, 0, 0.0, 0); }
extern "C" int klft_field_set_random_sun(klft_field_t f, double delta, uint64_t seed) {
  return launch_init(f, 1, delta, seed);
}
extern "C" int klft_field_set_random_algebra(klft_field_t f, double eps, uint64_t seed) {
  return launch_init(f, 2, eps, seed);
}
extern "C" int klft_field_set_haar(klft_field_t f, uint64_t seed) { return launch_init(f, 3, 0.0, seed); }

// general complex matrices: the two-row storage keeps SU(N) links only (the
// last row is rebuilt from the others), so it cannot hold them
static int check_general_matrices(klft_field_t f) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (f->storage == KLFT_STORAGE_F32_ROWS && f->nc > 1)
    return ki::set_error(KLFT_E_ARG, "two-row storage holds SU(N) links only; use F64 or F32 storage");
  return KLFT_OK;
}
extern "C" int klft_field_set_random_complex(klft_field_t f, uint64_t seed) {
  int rc = check_general_matrices(f);
  return rc ? rc : launch_init(f, 4, 0.0, seed);
}
extern "C" int klft_field_set_constant(klft_field_t f, double re, double im) {
  int rc = check_general_matrices(f);
  return rc ? rc : launch_init(f, 5, re, 0, im);
}

extern "C" int klft_field_upload(klft_field_t f, const double *host) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!host) return ki::set_error(KLFT_E_ARG, "null host buffer");
  const size_t n = (size_t)f->g.V * f->nd * f->nc * f->nc * 2 * sizeof(double);
  double *stage = nullptr;
  KLFT_HIP(hipMalloc(&stage, n));
  hipError_t e = hipMemcpyAsync(stage, host, n, hipMemcpyHostToDevice, ki::stream());
  if (e == hipSuccess) {
    rc = ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
      KLFT_UNPACK(rt, nct, ct, ndt);
      hipLaunchKernelGGL((k_upload<R, NC, COMP, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256), 0,
                         ki::stream(), ki::links_of<R, NC, COMP>(f), f->g, stage);
      KLFT_HIP(hipGetLastError());
      return KLFT_OK;
    });
    e = hipStreamSynchronize(ki::stream());
  }
  (void)hipFree(stage);
  if (e != hipSuccess) return ki::hip_fail(e, "upload");
  return rc;
}

extern "C" int klft_field_download(klft_field_t f, double *host) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!host) return ki::set_error(KLFT_E_ARG, "null host buffer");
  const size_t n = (size_t)f->g.V * f->nd * f->nc * f->nc * 2 * sizeof(double);
  double *stage = nullptr;
  KLFT_HIP(hipMalloc(&stage, n));
  rc = ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    hipLaunchKernelGGL((k_download<R, NC, COMP, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256), 0,
                       ki::stream(), ki::links_of<R, NC, COMP>(f), f->g, stage);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(host, stage, n, hipMemcpyDeviceToHost, ki::stream());
  if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
  (void)hipFree(stage);
  if (e != hipSuccess) return ki::hip_fail(e, "download");
  return rc;
}

extern "C" int klft_field_copy(klft_field_t dst, klft_field_t src) {
  int rc = ki::check_field(dst);
  if (!rc) rc = ki::check_field(src);
  if (rc) return rc;
  if (dst->nd != src->nd || dst->nc != src->nc)
    return ki::set_error(KLFT_E_ARG, "copy between fields of different shape");
  for (int d = 0; d < src->nd; ++d)
    if (dst->dims[d] != src->dims[d]) return ki::set_error(KLFT_E_ARG, "copy between different extents");
  if (dst->storage == src->storage) {
    KLFT_HIP(hipMemcpyAsync(dst->links, src->links, src->link_bytes, hipMemcpyDeviceToDevice, ki::stream()));
    return ki::sync_stream();
  }
  // conversion: dispatch on the source, then on the destination storage
  rc = ki::dispatch(src->storage, src->nc, src->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    using RI = typename decltype(rt)::type;
    constexpr int NC = decltype(nct)::value;
    constexpr bool CI = decltype(ct)::value != 0;
    constexpr int ND = decltype(ndt)::value;
    auto in = ki::links_of<RI, NC, CI>(src);
    const unsigned nb = ki::blocks_for(src->g.V, 256);
    switch (dst->storage) {
      case KLFT_STORAGE_F64:
        hipLaunchKernelGGL((k_convert<double, NC, false, RI, CI, ND>), dim3(nb), dim3(256), 0, ki::stream(),
                           ki::links_of<double, NC, false>(dst), in, src->g);
        break;
      case KLFT_STORAGE_F32:
        hipLaunchKernelGGL((k_convert<float, NC, false, RI, CI, ND>), dim3(nb), dim3(256), 0, ki::stream(),
                           ki::links_of<float, NC, false>(dst), in, src->g);
        break;
      default:
        hipLaunchKernelGGL((k_convert<float, NC, true, RI, CI, ND>), dim3(nb), dim3(256), 0, ki::stream(),
                           ki::links_of<float, NC, true>(dst), in, src->g);
        break;
    }
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
  if (rc) return rc;
  return ki::sync_stream();
}

int klft_internal::enqueue_reunit(klft_field_t f) {
  return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    hipLaunchKernelGGL((k_reunit<R, NC, COMP, ND>), dim3(ki::blocks_for(f->g.V, 256)), dim3(256), 0,
                       ki::stream(), ki::links_of<R, NC, COMP>(f), f->g);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}

extern "C" int klft_field_reunitarize(klft_field_t f) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  rc = ki::enqueue_reunit(f);
  if (rc) return rc;
  return ki::sync_stream();
}

// ============================================================================
// plaquette (GaugePlaquette.hpp:36-135)
// ============================================================================
// variants 0: plane by plane; 3: own links loaded once, all planes' loads in
// flight together (4, 5: k_plaquette_walk, below, all loads in flight / plane
// by plane); 1, 2 (occupancy targets of 0) never won and were removed
template <typename R, int NC, bool COMP, int ND, int OCC, bool FULL = false>
__global__ __launch_bounds__(256, OCC) void k_plaquette(Links<R, NC, COMP> f, Geom g, double *partials,
                                                    unsigned int *ticket, double *out,
                                                    double *per_site) {
  f.nd = ND;
  __shared__ double lds[8];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  int x[ND];
  long long s;
  double v = 0.0;
  if (chunk_site<ND>(g, lb, x, &s)) {
    if constexpr (FULL)
      v = plaq_at_full<R, NC, COMP, ND>(f, make_nbr<ND>(g, x, s));
    else
      v = plaq_at<R, NC, COMP, ND>(f, g, x, s);
    if (per_site) per_site[lex_index<ND>(g, x)] = v;
  }
  const double bt = block_sum<256>(v, lds);
  grid_finish<256>(bt, partials, ticket, out, lds);
}

// variants 4, 5: each thread walks S consecutive sites along dimension 0 at a
// fixed position in the other dimensions.  A block holds 256 consecutive
// such positions (last dimension fastest) of one segment of dimension 0, and
// all segments run at once, so the neighbours a site reads in dimensions
// 1..nd-1 are read by the same or adjacent blocks in the same slice at about
// the same time, and its neighbour in dimension 0 is the site the same thread
// reads next: each link is re-read while it is still in L2.
// FULL: all planes' loads in flight together (else plane by plane, fewer
// registers and more waves per SIMD)
template <typename R, int NC, bool COMP, int ND, bool FULL>
__global__ __launch_bounds__(256) void k_plaquette_walk(Links<R, NC, COMP> f, Geom g, int S, long long Vs,
                                                        double *partials, unsigned int *ticket, double *out,
                                                        double *per_site) {
  f.nd = ND;
  __shared__ double lds[8];
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned int nreg = (unsigned int)((Vs + 255) / 256);
  const int seg = (int)(lb / nreg);
  const long long sp = (long long)(lb % nreg) * 256 + threadIdx.x;
  double v = 0.0;
  if (sp < Vs) {
    int x[ND];
    long long r = sp;
#pragma unroll
    for (int d = ND - 1; d >= 1; --d) {
      x[d] = (int)(r % g.L[d]);
      r /= g.L[d];
    }
    const int t0 = seg * S;
#pragma unroll 1
    for (int j = 0; j < S && t0 + j < g.L[0]; ++j) {
      x[0] = t0 + j;
      const long long s = dev_index<ND>(g, x);
      double p;
      if constexpr (FULL)
        p = plaq_at_full<R, NC, COMP, ND>(f, make_nbr<ND>(g, x, s));
      else
        p = plaq_at<R, NC, COMP, ND>(f, g, x, s);
      if (per_site) per_site[lex_index<ND>(g, x)] = p;
      v += p;
    }
  }
  const double bt = block_sum<256>(v, lds);
  grid_finish<256>(bt, partials, ticket, out, lds);
}

namespace klft_internal {
// raw (unnormalised) plaquette sum into f->result[slot]; optional per-site
// reduction workspace of one launch: per-block partials (capacity cap), the
// grid ticket and the result slot
struct RedBuf {
  double *partials;
  long long cap;
  unsigned int *ticket;
  double *out;
};
static int launch_plaquette(klft_field_t f, const RedBuf &rb, double *dev_per_site, int variant) {
  return dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    if (variant == 4 || variant == 5) {
      // segments of dimension 0 for about 2^16 (FP64 links) or 2^17 (FP32)
      // walking threads: longer walks re-read more links from L2, fewer
      // threads hide less latency (measured at 24^4, 32^4 and 48^4)
      const long long Vs = f->g.V / f->g.L[0];
      const long long target = sizeof(R) == 8 ? (1LL << 16) : (1LL << 17);
      const int nseg = (int)std::min<long long>(f->g.L[0], std::max<long long>(1, (target + Vs - 1) / Vs));
      const int S = (f->g.L[0] + nseg - 1) / nseg;
      const int nsegs = (f->g.L[0] + S - 1) / S;
      const unsigned nb = (unsigned)(nsegs * ((Vs + 255) / 256));
      if ((long long)nb > rb.cap) return set_error(KLFT_E_ARG, "lattice too large for reduction");
      ProfScope prof("plaquette", "GaugePlaquette_GaugeField");
      if (variant == 4)
        hipLaunchKernelGGL((k_plaquette_walk<R, NC, COMP, ND, true>), dim3(nb), dim3(256), 0, stream(),
                           links_of<R, NC, COMP>(f), f->g, S, Vs, rb.partials, rb.ticket, rb.out,
                           dev_per_site);
      else
        hipLaunchKernelGGL((k_plaquette_walk<R, NC, COMP, ND, false>), dim3(nb), dim3(256), 0, stream(),
                           links_of<R, NC, COMP>(f), f->g, S, Vs, rb.partials, rb.ticket, rb.out,
                           dev_per_site);
      KLFT_HIP(hipGetLastError());
      return KLFT_OK;
    }
    const unsigned nb = blocks_for(chunk_threads<ND>(f->g), 256);
    if ((long long)nb > rb.cap) return set_error(KLFT_E_ARG, "lattice too large for reduction");
    if (variant == 3) {
      ProfScope prof("plaquette", "GaugePlaquette_GaugeField");
      hipLaunchKernelGGL((k_plaquette<R, NC, COMP, ND, 1, true>), dim3(nb), dim3(256), 0, stream(),
                         links_of<R, NC, COMP>(f), f->g, rb.partials, rb.ticket, rb.out, dev_per_site);
      KLFT_HIP(hipGetLastError());
      return KLFT_OK;
    }
    if (variant != 0) return set_error(KLFT_E_ARG, "plaquette variant %d is not built", variant);
    ProfScope prof("plaquette", "GaugePlaquette_GaugeField");
    hipLaunchKernelGGL((k_plaquette<R, NC, COMP, ND, 1>), dim3(nb), dim3(256), 0, stream(), links_of<R, NC, COMP>(f),
                       f->g, rb.partials, rb.ticket, rb.out, dev_per_site);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
}
// untuned defaults, the fastest at 32^4 and 48^4: threads walking dimension
// 0, plane by plane for FP64 links, all loads in flight for FP32
static int plaquette_default(const klft_field_s *f) { return f->storage == KLFT_STORAGE_F64 ? 5 : 4; }
constexpr int kPlaquetteVariants = 6;
int plaquette_raw(klft_field_t f, int slot, double *dev_per_site) {
  const RedBuf rb{f->partials, f->max_partials, f->ticket, f->result + slot};
  const int v = pick_variant("plaquette", f, kPlaquetteVariants, plaquette_default(f),
                             [&](int var) { return launch_plaquette(f, rb, nullptr, var); });
  return launch_plaquette(f, rb, dev_per_site, v);
}
int read_results(klft_field_t f, int first, int n, double *out) {
  KLFT_HIP(hipMemcpyAsync(f->host_result, f->result + first, n * sizeof(double), hipMemcpyDeviceToHost,
                          stream()));
  KLFT_HIP(hipStreamSynchronize(stream()));
  for (int i = 0; i < n; ++i) out[i] = f->host_result[i];
  return KLFT_OK;
}
}  // namespace klft_internal

extern "C" int klft_plaquette(klft_field_t f, int normalize, double *out, double *per_site) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!out) return ki::set_error(KLFT_E_ARG, "null output");
  double *dps = nullptr;
  if (per_site) KLFT_HIP(hipMalloc(&dps, sizeof(double) * f->g.V));
  rc = ki::plaquette_raw(f, 0, dps);
  double sum = 0.0;
  if (!rc) rc = ki::read_results(f, 0, 1, &sum);
  if (!rc && per_site) {
    hipError_t e = hipMemcpyAsync(per_site, dps, sizeof(double) * f->g.V, hipMemcpyDeviceToHost, ki::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
    if (e != hipSuccess) rc = ki::hip_fail(e, "per-site copy");
  }
  if (dps) (void)hipFree(dps);
  if (rc) return rc;
  if (normalize) {  // GaugePlaquette.hpp:125-132
    double norm = 1.0;
    for (int d = 0; d < f->nd; ++d) norm *= (double)f->dims[d];
    norm *= (double)((f->nd * (f->nd - 1) / 2) * f->nc);
    sum /= norm;
  }
  *out = sum;
  return KLFT_OK;
}

// ============================================================================
// plaquette without allocation, for callers that cannot allocate device
// memory.  Its reduction partials, grid ticket, result and per-site staging
// live in `scratch`, one static array per GPU (held until the process
// ends), in place of the field's reduction workspace and the per-call
// hipMalloc of klft_plaquette.  Calls that overlap on one GPU share the array
// and race.
// ============================================================================
namespace {
constexpr long long kNaPartials = 1 << 16;  // reduction blocks
constexpr long long kNaSlots = 32;          // result (slot 0) and ticket (slot 16)
constexpr long long kNaSites = 1 << 20;     // per-site staging (32^4)
static __device__ __align__(256) double scratch[kNaPartials + kNaSlots + kNaSites];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
double *g_na_ptr[kNaMaxDevices] = {};

// the device address of this GPU's scratch (the array's name is a host-side
// stand-in and must never reach a kernel), resolved once per device with the
// device current
int scratch_of_current(double **out) {
  int dev = -1;
  KLFT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) return ki::set_error(KLFT_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    KLFT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_na_ptr[dev];
  return KLFT_OK;
}

// zero the result slot and the grid ticket (a kernel, not a memset node, so a
// captured call replays correctly)
__global__ void k_na_clear(double *out, unsigned int *ticket) {
  if (threadIdx.x == 0) {
    *out = 0.0;
    *ticket = 0u;
  }
}
}  // namespace

extern "C" int klft_plaquette_noalloc(klft_field_t f, int normalize, double *out, double *per_site) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (!out) return ki::set_error(KLFT_E_ARG, "null output");
  if (per_site && f->g.V > kNaSites)
    return ki::set_error(KLFT_E_ARG, "per-site output of %lld sites exceeds the no-allocation limit of %lld",
                         f->g.V, kNaSites);
  double *base = nullptr;
  rc = scratch_of_current(&base);
  if (rc) return rc;
  double *res = base + kNaPartials;
  unsigned int *ticket = reinterpret_cast<unsigned int *>(base + kNaPartials + 16);
  double *stage = per_site ? base + kNaPartials + kNaSlots : nullptr;
  hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, ki::stream(), res, ticket);
  KLFT_HIP(hipGetLastError());
  const ki::RedBuf rb{base, kNaPartials, ticket, res};
  rc = ki::launch_plaquette(f, rb, stage, ki::plaquette_default(f));
  if (rc) return rc;
  KLFT_HIP(hipMemcpyAsync(f->host_result, res, sizeof(double), hipMemcpyDeviceToHost, ki::stream()));
  if (per_site)
    KLFT_HIP(hipMemcpyAsync(per_site, stage, sizeof(double) * f->g.V, hipMemcpyDeviceToHost, ki::stream()));
  KLFT_HIP(hipStreamSynchronize(ki::stream()));
  double sum = f->host_result[0];
  if (normalize) {  // GaugePlaquette.hpp:125-132
    double norm = 1.0;
    for (int d = 0; d < f->nd; ++d) norm *= (double)f->dims[d];
    norm *= (double)((f->nd * (f->nd - 1) / 2) * f->nc);
    sum /= norm;
  }
  *out = sum;
  return KLFT_OK;
}

// ============================================================================
// staple field (Gauge_Util.hpp:27-97)
// ============================================================================
// variants 0: one-term-at-a-time staple; 3: full staple (all its loads in
// flight at once)
template <typename R, int NC, bool COMP, int ND, int OCC, bool LOWREG = true>
__global__ __launch_bounds__(256, OCC) void k_staple_field(Links<R, NC, COMP> f,
                                                                            Links<R, NC, false> o, Geom g) {
  f.nd = ND;
  o.nd = ND;
  const unsigned int lb = xcd_block(blockIdx.x, gridDim.x);
  int x[ND], mu;
  long long s;
  if (!link_site<ND>(g, lb, x, &s, &mu)) return;
  const Nbr<ND> nb = make_nbr<ND>(g, x, s);
  if constexpr (LOWREG)
    o.store(s, mu, staple_lowreg<R, NC, COMP, ND>(f, nb, mu));
  else
    o.store(s, mu, staple<R, NC, COMP, ND>(f, nb, mu));
}

extern "C" int klft_staple_field(klft_field_t f, klft_field_t out) {
  int rc = ki::check_field(f);
  if (!rc) rc = ki::check_field(out);
  if (rc) return rc;
  const int want = f->storage == KLFT_STORAGE_F64 ? KLFT_STORAGE_F64 : KLFT_STORAGE_F32;
  if (out->storage != want || out->nd != f->nd || out->nc != f->nc)
    return ki::set_error(KLFT_E_ARG, "staple output must be a full-storage field of the same precision and shape");
  for (int d = 0; d < f->nd; ++d)
    if (out->dims[d] != f->dims[d]) return ki::set_error(KLFT_E_ARG, "staple output extents differ");
  auto launch = [&](int variant) {
    return ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
      KLFT_UNPACK(rt, nct, ct, ndt);
      const unsigned nb = (unsigned)link_blocks<ND>(f->g);
      if (variant == 3) {
        ki::ProfScope prof("staple_field", "stapleField_GaugeField");
        hipLaunchKernelGGL((k_staple_field<R, NC, COMP, ND, 1, false>), dim3(nb), dim3(64 * ND), 0, ki::stream(),
                           ki::links_of<R, NC, COMP>(f), ki::links_of<R, NC, false>(out), f->g);
        KLFT_HIP(hipGetLastError());
        return KLFT_OK;
      }
      if (variant != 0) return ki::set_error(KLFT_E_ARG, "staple-field variant %d is not built", variant);
      ki::ProfScope prof("staple_field", "stapleField_GaugeField");
      hipLaunchKernelGGL((k_staple_field<R, NC, COMP, ND, 1>), dim3(nb), dim3(64 * ND), 0, ki::stream(),
                         ki::links_of<R, NC, COMP>(f), ki::links_of<R, NC, false>(out), f->g);
      KLFT_HIP(hipGetLastError());
      return KLFT_OK;
    });
  };
  // variants: 0 one-term-at-a-time staple, 3 full staple (1, 2: occupancy
  // targets that never won, removed)
  rc = launch(ki::pick_variant("staple_field", f, 4, f->storage == KLFT_STORAGE_F64 ? 3 : 0, launch));
  if (rc) return rc;
  return ki::sync_stream();
}

// ============================================================================
// Wilson loops (WilsonLoop.hpp:34-304)
// ============================================================================
// One block holds C full columns along nu (periodic), one thread per site.
// For each requested L (ascending) a thread extends its mu-line
// F_L(x) = U_mu(x) ... U_mu(x+(L-1)mu) and publishes it in LDS; then for
// T = 0, 1, ... it carries the three-sided staple
//   D(L,T) = l_nu(x;T)^+ F_L(x) l_nu(x+L mu;T),
//   D(L,T) = U_nu(x+(T-1)nu)^+ D(L,T-1) U_nu(x+L mu+(T-1)nu),
// and closes the loop with the top line read from LDS:
//   w(x;L,T) = Re Tr[D(L,T) F_L(x+T nu)^+],
// which equals the reference's Re Tr[A B^+] with A = l_mu(x;L) l_nu(x+L mu;T),
// B = l_nu(x;T) l_mu(x+T nu;L) (WilsonLoop.hpp:129-160).  No line field is
// materialised in HBM.  Wave partial sums go to
// partials[(((plane * nL + li) * (Tmax+1) + T) * gridDim.x + blockIdx.x)
//          * waves per block + wave].
template <typename R, int NC, bool COMP, int ND, bool LNU = true>
__global__ __launch_bounds__(512) void k_wilson(Links<R, NC, COMP> f, Geom g, int mu0, int nu,
                                                 int ncolblk, int nL, const int *__restrict__ Ls,
                                                 int Tmax, double *partials, double *per_site,
                                                 int ps_li, int ps_T) {
  f.nd = ND;
  extern __shared__ __align__(16) unsigned char smem[];
  // lds: the threads' current mu-lines; lnu (LNU): the columns' nu-links
  // U_nu(x), read by every step T of the staple recurrence (else loaded)
  Mat<R, NC> *lds = reinterpret_cast<Mat<R, NC> *>(smem);
  Mat<R, NC> *lnu = lds + blockDim.x;
  const int plane = blockIdx.y;
  const int mu = mu0 < 0 ? plane : mu0;
  const int Ln = g.L[nu];
  const int tid = threadIdx.x;
  const int cl = tid / Ln, k = tid - cl * Ln;
  const long long ncol = g.V / Ln;
  const long long col = (long long)blockIdx.x * ncolblk + cl;
  const bool valid = cl < ncolblk && col < ncol;
  int x[ND];
  {
    long long r = valid ? col : 0;
#pragma unroll
    for (int d = ND - 1; d >= 0; --d) {
      if (d == nu) continue;
      x[d] = (int)(r % g.L[d]);
      r /= g.L[d];
    }
    x[nu] = k;
  }
  // A device index is a sum over dimensions of the coordinate's sublattice
  // bit (times Vh) and its encoded half-coordinate, so the shifted sites'
  // indices below replace only the shifted coordinates' terms of x's.
  auto dterm = [&](int d, int v) -> long long {
    return (long long)((v & 1) << (ND - 1 - d)) * g.Vh + enc_dim(g, d, v >> 1);
  };
  auto wrap = [](int v, int n) { return v < n ? v : v % n; };  // v >= 0
  const int Lmu = g.L[mu];
  long long base = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d) base += dterm(d, x[d]);
  const long long base_nomu = base - dterm(mu, x[mu]);  // x without its mu term
  const long long term_nu = dterm(nu, x[nu]);
  if constexpr (LNU) {
    if (valid) lnu[tid] = f.load(base, nu);
  }
  Mat<R, NC> F = identity<R, NC>();
  int Lcur = 0;
  for (int li = 0; li < nL; ++li) {
    const int L = Ls[li];
    if (valid) {
      while (Lcur < L) {
        F = mul(F, f.load(base_nomu + dterm(mu, wrap(x[mu] + Lcur, Lmu)), mu));  // U_mu(x + Lcur mu)
        ++Lcur;
      }
    }
    __syncthreads();
    lds[tid] = F;
    __syncthreads();
    Mat<R, NC> D = F;
    // index of x + L mu without its nu term
    const long long base_L = base_nomu + dterm(mu, wrap(x[mu] + L, Lmu)) - term_nu;
    for (int T = 0; T <= Tmax; ++T) {
      double w = 0.0;
      if (valid) {
        if (T > 0) {
          const int kp = wrap(k + T - 1, Ln);
          const long long tnu = dterm(nu, kp);
          const Mat<R, NC> b = f.load(base_L + tnu, nu);   // U_nu(x + L mu + (T-1) nu)
          Mat<R, NC> a;                                     // U_nu(x + (T-1) nu), same column
          if constexpr (LNU)
            a = lnu[cl * Ln + kp];
          else
            a = f.load(base - term_nu + tnu, nu);
          D = mul(mul_da(a, D), b);
        }
        const int kt = wrap(k + T, Ln);
        w = (double)retr_mul_ad(D, lds[cl * Ln + kt]);
        if (per_site && li == ps_li && T == ps_T && plane == 0) per_site[lex_index<ND>(g, x)] = w;
      }
      // one partial per wave (no block barrier per (L,T)); the partials of
      // an output are summed in a fixed order afterwards
      const double ws = wave_sum(w);
      if ((tid & 63) == 0)
        partials[(((long long)(plane * nL + li) * (Tmax + 1) + T) * gridDim.x + blockIdx.x) * (blockDim.x >> 6) +
                 (tid >> 6)] = ws;
    }
  }
}

// fixed-order sum of n block partials into one output
__global__ __launch_bounds__(256) void k_sum_blocks(const double *partials, long long n, double *out) {
  __shared__ double red[8];
  double v = 0.0;
  for (long long i = threadIdx.x; i < n; i += 256) v += partials[i];
  v = block_sum<256>(v, red);
  if (threadIdx.x == 0) *out = v;
}

int klft_internal::enqueue_sum_blocks(const double *partials, long long n, double *out) {
  hipLaunchKernelGGL(k_sum_blocks, dim3(1), dim3(256), 0, ki::stream(), partials, n, out);
  KLFT_HIP(hipGetLastError());
  return KLFT_OK;
}

// fixed-order sum of nblk partials for each output
__global__ __launch_bounds__(256) void k_sum_partials(const double *partials, int nblk, double *out) {
  __shared__ double red[8];
  const double *p = partials + (long long)blockIdx.x * nblk;
  double v = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) v += p[i];
  v = block_sum<256>(v, red);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

namespace {
// sums over sites of w for the planes (mu_first.., nu); res[plane][pair]
int wilson_sums(klft_field_t f, int mu0, int nplanes, int nu, const int *pairs, int npairs,
                std::vector<double> &res, double *per_site_host) {
  if (npairs <= 0) {
    res.clear();
    return KLFT_OK;
  }
  std::vector<int> Ls;
  int Tmax = 0;
  for (int p = 0; p < npairs; ++p) {
    if (pairs[2 * p] < 0 || pairs[2 * p + 1] < 0)
      return ki::set_error(KLFT_E_ARG, "negative Wilson loop extent");
    Ls.push_back(pairs[2 * p]);
    Tmax = std::max(Tmax, pairs[2 * p + 1]);
  }
  std::sort(Ls.begin(), Ls.end());
  Ls.erase(std::unique(Ls.begin(), Ls.end()), Ls.end());
  const int nL = (int)Ls.size();
  const int Ln = f->dims[nu];
  const size_t mbytes = (size_t)f->nc * f->nc * f->cbytes;
  int C = std::max(1, 256 / Ln);
  // whole waves: the trailing threads (cl >= C) hold no column
  int block = ((C * Ln + 63) / 64) * 64;
  // the block's mu-lines and nu-links in LDS, 2 * block matrices: fewer
  // columns per block where that does not fit; a column too long for both
  // keeps only the mu-lines (the nu-links are loaded at every step)
  while (C > 1 && 2 * block * mbytes > 64 * 1024) {
    C /= 2;
    block = ((C * Ln + 63) / 64) * 64;
  }
  const bool lnu = 2 * block * mbytes <= 64 * 1024;
  if (block > 512 || block * mbytes > 64 * 1024)
    return ki::set_error(KLFT_E_ARG, "extent %d along nu too large for the Wilson-loop kernel", Ln);
  const long long ncol = f->g.V / Ln;
  const unsigned nbx = (unsigned)((ncol + C - 1) / C);
  const long long nout = (long long)nplanes * nL * (Tmax + 1);
  double *dpart = nullptr, *dout = nullptr, *dps = nullptr;
  int *dLs = nullptr;
  const int nw = block / 64;  // one partial per wave
  {
    // partials, outputs and extents in the field's measurement workspace
    // (256-byte aligned pieces); the per-site output, a V-sized debug
    // array, is allocated per call
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bpart = up(sizeof(double) * nout * nbx * nw), bout = up(sizeof(double) * nout);
    void *ws = nullptr;
    const int wrc = ki::meas_workspace(f, bpart + bout + up(sizeof(int) * nL), &ws);
    if (wrc) return wrc;
    dpart = static_cast<double *>(ws);
    dout = reinterpret_cast<double *>(static_cast<char *>(ws) + bpart);
    dLs = reinterpret_cast<int *>(static_cast<char *>(ws) + bpart + bout);
  }
  if (per_site_host) KLFT_HIP(hipMalloc(&dps, sizeof(double) * f->g.V));
  KLFT_HIP(hipMemcpyAsync(dLs, Ls.data(), sizeof(int) * nL, hipMemcpyHostToDevice, ki::stream()));
  int ps_li = 0, ps_T = 0;
  if (per_site_host) {
    const int lastL = pairs[2 * (npairs - 1)];
    ps_li = (int)(std::lower_bound(Ls.begin(), Ls.end(), lastL) - Ls.begin());
    ps_T = pairs[2 * (npairs - 1) + 1];
  }
  int rc = ki::dispatch(f->storage, f->nc, f->nd, [&](auto rt, auto nct, auto ct, auto ndt) -> int {
    KLFT_UNPACK(rt, nct, ct, ndt);
    ki::ProfScope prof("wilson", "WilsonLoop_GaugeField_WLoop_munu");
    if (lnu)
      hipLaunchKernelGGL((k_wilson<R, NC, COMP, ND, true>), dim3(nbx, nplanes), dim3(block),
                         2 * block * sizeof(Mat<R, NC>), ki::stream(), ki::links_of<R, NC, COMP>(f), f->g,
                         nplanes > 1 ? -1 : mu0, nu, C, nL, dLs, Tmax, dpart, dps, ps_li, ps_T);
    else
      hipLaunchKernelGGL((k_wilson<R, NC, COMP, ND, false>), dim3(nbx, nplanes), dim3(block),
                         block * sizeof(Mat<R, NC>), ki::stream(), ki::links_of<R, NC, COMP>(f), f->g,
                         nplanes > 1 ? -1 : mu0, nu, C, nL, dLs, Tmax, dpart, dps, ps_li, ps_T);
    KLFT_HIP(hipGetLastError());
    return KLFT_OK;
  });
  if (!rc) {
    hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)nout), dim3(256), 0, ki::stream(), dpart, (int)nbx * nw,
                       dout);
    hipError_t e = hipGetLastError();
    std::vector<double> all(nout);
    if (e == hipSuccess) e = hipMemcpyAsync(all.data(), dout, sizeof(double) * nout, hipMemcpyDeviceToHost, ki::stream());
    if (e == hipSuccess && per_site_host)
      e = hipMemcpyAsync(per_site_host, dps, sizeof(double) * f->g.V, hipMemcpyDeviceToHost, ki::stream());
    if (e == hipSuccess) e = hipStreamSynchronize(ki::stream());
    if (e != hipSuccess) rc = ki::hip_fail(e, "wilson loop");
    if (!rc) {
      res.assign((size_t)nplanes * npairs, 0.0);
      for (int pl = 0; pl < nplanes; ++pl)
        for (int p = 0; p < npairs; ++p) {
          const int li = (int)(std::lower_bound(Ls.begin(), Ls.end(), pairs[2 * p]) - Ls.begin());
          res[(size_t)pl * npairs + p] = all[((size_t)pl * nL + li) * (Tmax + 1) + pairs[2 * p + 1]];
        }
    }
  }
  if (dps) (void)hipFree(dps);
  return rc;
}

// WilsonLoop.hpp:233-239: sequential divisions by each extent, then by Nc
double wilson_normalize(klft_field_t f, double w, int normalize) {
  if (normalize) {
    for (int d = 0; d < f->nd; ++d) w /= (double)f->dims[d];
    w /= (double)f->nc;
  }
  return w;
}
}  // namespace

extern "C" int klft_wilson_loop_mu_nu(klft_field_t f, int mu, int nu, const int *pairs, int npairs,
                                      int normalize, double *out, double *per_site) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (mu < 0 || mu >= f->nd || nu < 0 || nu >= f->nd || mu == nu)
    return ki::set_error(KLFT_E_ARG, "invalid plane (%d, %d)", mu, nu);
  if (npairs > 0 && (!pairs || !out)) return ki::set_error(KLFT_E_ARG, "null argument");
  std::vector<double> res;
  rc = wilson_sums(f, mu, 1, nu, pairs, npairs, res, per_site);
  if (rc) return rc;
  for (int p = 0; p < npairs; ++p) out[p] = wilson_normalize(f, res[p], normalize);
  return KLFT_OK;
}

extern "C" int klft_wilson_loop_temporal(klft_field_t f, const int *pairs, int npairs, int normalize,
                                         double *out) {
  int rc = ki::check_field(f);
  if (rc) return rc;
  if (npairs > 0 && (!pairs || !out)) return ki::set_error(KLFT_E_ARG, "null argument");
  const int nu = f->nd - 1;
  std::vector<double> res;
  rc = wilson_sums(f, 0, f->nd - 1, nu, pairs, npairs, res, nullptr);
  if (rc) return rc;
  // WilsonLoop.hpp:257-303: per-plane normalised values summed in mu order,
  // then divided by (Nd - 1)
  for (int p = 0; p < npairs; ++p) {
    double w = wilson_normalize(f, res[p], normalize);
    for (int pl = 1; pl < f->nd - 1; ++pl) w += wilson_normalize(f, res[(size_t)pl * npairs + p], normalize);
    out[p] = w / (double)(f->nd - 1);
  }
  return KLFT_OK;
}
