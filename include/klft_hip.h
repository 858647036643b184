/*
 * klft_hip.h -- C ABI of the MI355X-native klft gauge-update and measurement
 * path (libklft_hip.so).
 *
 * Every function returns KLFT_OK (0) or a negative KLFT_E* code;
 * klft_last_error() returns a thread-local message for the last failure.
 * Calls are synchronous, as every klft launch is followed by Kokkos::fence()
 * (e.g. Metropolis.hpp:163, GaugePlaquette.hpp:118-122).  A handle is used
 * from one host thread at a time; distinct handles are independent.
 *
 * Host gauge-field buffers use the reference's OpenMP (LayoutRight)
 * array-of-structs image of View<SUN<Nc>****[Nd]> (GLOBAL.hpp:61-68):
 *   complex offset = ((site * nd + mu) * nc + c1) * nc + c2,
 *   site = ((i0 * L1 + i1) * L2 + i2) * L3 + i3   (i0 slowest),
 * as interleaved (re, im) doubles.  Momentum / force buffers: double offset
 * (site * nd + mu) * ngen + a, with ngen = 1, 3, 8 for Nc = 1, 2, 3 and
 * generators T = 1, Pauli sigma, Gell-Mann lambda.
 *
 * Citations are to the aniketsen/klft sources.
 */
#ifndef KLFT_HIP_H
#define KLFT_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KLFT_OK 0
#define KLFT_E_ARG -1      /* invalid argument (shape, group, storage, odd extent ...) */
#define KLFT_E_HIP -2      /* HIP runtime error or no device */
#define KLFT_E_NOMEM -3    /* device allocation failed */
#define KLFT_E_STATE -4    /* handle in the wrong state */
#define KLFT_E_PARSE -5    /* input file could not be parsed */
#define KLFT_E_IO -6       /* file could not be read or written */

/* link storage (precision / compression) */
#define KLFT_STORAGE_F64 0        /* FP64, Nc x Nc complex per link (the plain method) */
#define KLFT_STORAGE_F32 1        /* FP32, Nc x Nc complex per link */
#define KLFT_STORAGE_F32_ROWS 2   /* FP32, first Nc-1 rows (SU(3): 12 reals, SU(2): 4 reals) */

typedef struct klft_field_s *klft_field_t;
typedef struct klft_hmc_s *klft_hmc_t;
typedef struct klft_batch_s *klft_batch_t;

/* ---- library ------------------------------------------------------------- */
/* Select and initialise the HIP device (the first call does this implicitly
 * with device 0). */
int klft_init(int device);
const char *klft_last_error(void);
/* Build string: "klft_hip <version> gfx950" */
const char *klft_version(void);

/* ---- gauge fields (replaces deviceGaugeField<Nd,Nc>, GaugeField.hpp:30-239,
 *      3D :241-435, 2D :437-620) ---------------------------------------------- */
/* nd in {2,3,4}; nc in {1,2,3}; every extent even and >= 2 (the checkerboard
 * of Metropolis.hpp:127-130 silently skips a slice for odd extents).  The
 * links start zeroed: a set_* call or an upload gives the start. */
int klft_field_create(int nd, int nc, const int *dims, int storage, klft_field_t *out);
int klft_field_destroy(klft_field_t f);
int klft_field_info(klft_field_t f, int *nd, int *nc, int *dims, int *storage);
/* deviceGaugeField(L.., identitySUN<Nc>()) cold start (GaugeField.hpp:47-52, :80-93) */
int klft_field_set_identity(klft_field_t f);
/* deviceGaugeField(L.., rng, delta): every link a randSUN(delta) proposal
 * (GaugeField.hpp:55-60, :96-112), drawn from the keyed stream (tag 5). */
int klft_field_set_random_sun(klft_field_t f, double delta, uint64_t seed);
/* U = exp(i eps Q), Q = sum_a h_a T_a / 2, h_a ~ N(0,1) keyed (seed, site, mu);
 * computed in FP64 and rounded to the field's storage. */
int klft_field_set_random_algebra(klft_field_t f, double eps, uint64_t seed);
/* Haar-random links (SU(2): normalised Gaussian quaternion; U(1): uniform
 * phase; SU(3): Gram-Schmidt of a Gaussian matrix, third row conj(r0 x r1)). */
int klft_field_set_haar(klft_field_t f, uint64_t seed);
/* deviceGaugeField(L.., rng) (GaugeField.hpp:58-63, :118-141): every entry
 * of every link an independent complex(drand(-1,1), drand(-1,1)), real part
 * first, c1 outer, c2 inner, from the keyed stream (tag 6).  The links are
 * general complex matrices, so two-row storage (Nc > 1) is rejected. */
int klft_field_set_random_complex(klft_field_t f, uint64_t seed);
/* deviceGaugeField(L.., complex_t init) (GaugeField.hpp:36-40, :65-84):
 * every entry of every link = re + i im; two-row storage (Nc > 1) rejected. */
int klft_field_set_constant(klft_field_t f, double re, double im);
/* host AoS FP64 <-> device (rounded to the storage precision on upload) */
int klft_field_upload(klft_field_t f, const double *host);
int klft_field_download(klft_field_t f, double *host);
/* dst <- src with precision / compression conversion (same nd, nc, dims) */
int klft_field_copy(klft_field_t dst, klft_field_t src);
/* restoreSUN on every link for FP64 storage (SUN.hpp:256-340); FP64-norm
 * Gram-Schmidt for FP32 storage */
int klft_field_reunitarize(klft_field_t f);

/* ---- measurement ------------------------------------------------------------ */
/* GaugePlaquette<rank,Nc>(g, normalize) (GaugePlaquette.hpp:89-135).
 * per_site: optional host array of V doubles receiving Re p(x) in
 * lexicographic order. */
int klft_plaquette(klft_field_t f, int normalize, double *out, double *per_site);
/* klft_plaquette without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The reduction
 * workspace and the per-site staging come from one static device array per
 * GPU (8.5 MiB, held until the process ends) instead of the field's workspace
 * and a per-call hipMalloc; per_site requires V <= 2^20 sites (KLFT_E_ARG
 * beyond).  Calls that overlap on one GPU share that array and race. */
int klft_plaquette_noalloc(klft_field_t f, int normalize, double *out, double *per_site);
/* stapleField (Gauge_Util.hpp:27-97): out must be a full-storage field of the
 * same shape and precision; out(x,mu) = staple(x,mu) (GaugeField.hpp:187-237). */
int klft_staple_field(klft_field_t f, klft_field_t out);
/* WilsonLoop_mu_nu (WilsonLoop.hpp:170-245): pairs = npairs (Lmu, Lnu) pairs;
 * out[p] = W for pair p (list order; results do not depend on the order).
 * per_site (optional): V doubles, Re w(x) for the last pair. */
int klft_wilson_loop_mu_nu(klft_field_t f, int mu, int nu, const int *pairs, int npairs,
                           int normalize, double *out, double *per_site);
/* WilsonLoop_temporal (WilsonLoop.hpp:250-304): planes (mu, nd-1) for
 * mu = 0..nd-2 in one launch; out[p] = average over the planes. */
int klft_wilson_loop_temporal(klft_field_t f, const int *pairs, int npairs, int normalize,
                              double *out);

/* ---- Metropolis (Metropolis.hpp:44-178) --------------------------------------- */
typedef struct {
  int nhits;    /* MetropolisParams::nHits */
  double beta;  /* MetropolisParams::beta */
  double delta; /* MetropolisParams::delta */
  uint64_t seed;
} klft_metropolis_params;

/* sweep_Metropolis<rank,Nc>: 2^nd sequential sublattice passes.  sweep keys
 * the random stream; acc_rate (optional) = accepted / (V * nd * nhits);
 * naccepted (optional) = number of accepted hits. */
int klft_metropolis_sweep(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep,
                          double *acc_rate, int64_t *naccepted);
/* SU(3) Metropolis proposal law of a field.  KLFT_PROPOSAL_REFERENCE (the
 * default) is randSUN<3>(delta) (SUN.hpp:192-251), whose rows are normalised
 * uniform(0, delta) vectors: delta cancels and the law is asymmetric, so the
 * chain does not satisfy detailed balance.  KLFT_PROPOSAL_SYMMETRIC has no
 * counterpart in the reference: r = exp(i delta sum_a h_a lambda_a),
 * h_a ~ U(-1, 1), a symmetric law (r and r^dagger equally likely) whose step
 * size delta sets.  U(1) and SU(2) proposals are symmetric either way. */
#define KLFT_PROPOSAL_REFERENCE 0
#define KLFT_PROPOSAL_SYMMETRIC 1
int klft_field_set_proposal(klft_field_t f, int proposal);
/* n sweeps sweep0 .. sweep0+n-1 with no host synchronisation in between;
 * acc_rates (optional) receives n values. */
int klft_metropolis_run(klft_field_t f, const klft_metropolis_params *p, uint64_t sweep0,
                        int n, double *acc_rates);
/* schedule dump: visit[site] = index of the sublattice pass that updates site
 * (lexicographic order, V int32) -- written by the sweep kernel's own
 * thread-to-site map. */
int klft_metropolis_schedule(klft_field_t f, int32_t *visit);
/* the launch order of a sweep whose independent passes (three or more
 * coordinate parities apart) run two to a launch: launch i runs passes
 * groups[2i] and groups[2i+1] (-1: none), *nlaunches launches in all;
 * groups holds 2 * 2^nd ints.  No counterpart in the reference, which
 * launches every pass on its own (Metropolis.hpp:139-162); for callers and
 * tests that check the schedule.  Host only: no GPU call. */
int klft_sweep_launches(int nd, int32_t *groups, int *nlaunches);
/* the launch order of klft_metropolis_run over nsweeps sweeps: launch i
 * runs pass launches[4i] of sweep launches[4i+1] (relative to the run's
 * first) and pass launches[4i+2] of sweep launches[4i+3] (-1, -1: none).
 * Independent passes of consecutive sweeps share launches too (13|0', 14|1',
 * 15|2' in 4D: 8 launches per sweep over a long run).  *nlaunches receives
 * the count; launches holds 4 * capacity ints (an error when too small).
 * No counterpart in the reference (Metropolis.hpp:139-162 launches pass by
 * pass, sweep by sweep); for callers and tests that check the schedule.
 * Host only: no GPU call. */
int klft_run_launches(int nd, int nsweeps, int32_t *launches, int capacity, int *nlaunches);

/* ---- HMC (HMC.hpp:48-71, Integrator.hpp:26-53, GaugeMonomial.hpp:19-37,
 *      HamiltonianField.hpp:23-60) ------------------------------------------------ */
typedef struct {
  double beta;  /* GaugeMonomial::beta */
  double tau;   /* HMC_Params::tau */
  int nsteps;   /* HMC_Params::n_steps (leapfrog with nsteps+1 gauge updates) */
  uint64_t seed;
} klft_hmc_params;

int klft_hmc_create(klft_field_t f, const klft_hmc_params *p, klft_hmc_t *out);
int klft_hmc_destroy(klft_hmc_t h);
/* HamiltonianField::randomize_momentum: p_a ~ N(0,1), keyed (seed, traj, site, mu) */
int klft_hmc_refresh_momenta(klft_hmc_t h, uint64_t traj);
int klft_hmc_momenta_upload(klft_hmc_t h, const double *host);
int klft_hmc_momenta_download(klft_hmc_t h, double *host);
/* GaugeMonomial::derivative: F_a(x,mu) = (beta/Nc) Im Tr(T_a U_mu(x) staple) */
int klft_hmc_force(klft_hmc_t h, double *host_force);
/* HamiltonianField::update_momentum (P -= dtau F, F recomputed from U) */
int klft_hmc_update_momentum(klft_hmc_t h, double dtau);
/* HamiltonianField::update_gauge: U <- exp(i dtau P) U */
int klft_hmc_update_gauge(klft_hmc_t h, double dtau);
/* KineticMonomial: 1/2 sum p^2 (FP64) */
int klft_hmc_kinetic(klft_hmc_t h, double *out);
/* GaugeMonomial: -(beta/Nc) sum_x sum_{mu<nu} Re Tr U_munu(x) (FP64) */
int klft_gauge_action(klft_field_t f, double beta, double *out);
/* Leapfrog::integrate with the reference's structure (n+1 gauge updates,
 * restoreGauge at the end) */
int klft_hmc_integrate(klft_hmc_t h);
/* HMC::hmc_step: refresh, integrate, FP64 dH, accept/reject keyed (seed, traj) */
int klft_hmc_trajectory(klft_hmc_t h, uint64_t traj, double *dH, int *accepted);

/* ---- batched chains and PTBC replicas ---------------------------------------------- */
/* A batch owns n independent 4D chains of mixed extents, groups U(1)/SU(2)
 * (FP32 links) or SU(3), each with its own beta, updated by one launch per
 * sublattice pass (or per pair of independent passes) over all chains
 * (captured in a HIP graph). */
typedef struct {
  int nc;        /* 1, 2 or 3 */
  int dims[4];
  double beta;
} klft_chain_desc;

int klft_batch_create(int nchains, const klft_chain_desc *chains, int storage, klft_batch_t *out);
int klft_batch_destroy(klft_batch_t b);
int klft_batch_set_identity(klft_batch_t b);
/* one sweep of every chain; acc (optional) n acceptance rates */
int klft_batch_sweep(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t sweep,
                     double *acc);
/* n sweeps replayed from one captured HIP graph */
int klft_batch_run(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t sweep0,
                   int n);
/* FP64 plaquette of every chain */
int klft_batch_plaquette(klft_batch_t b, double *out);
/* chain i <-> host AoS FP64 */
int klft_batch_download(klft_batch_t b, int chain, double *host);
int klft_batch_upload(klft_batch_t b, int chain, const double *host);

/* PTBC: n replicas of one SU(N) lattice; replica r's links on the defect
 * (mu = 0, x0 = 0, x_k in [p_k, p_k + Ld) mod L_k, PTBCGaugeField.hpp:186-198)
 * are scaled by its coupling c (PTBCGaugeField.hpp:209-229). */
int klft_ptbc_create(int nreplicas, int nc, const int *dims, int storage, double beta,
                     int defect_length, const int *defect_pos, const double *c,
                     klft_batch_t *out);
/* one step: Metropolis sweep of every replica, then nearest-neighbour swap
 * proposals of the coupling labels with FP64 action differences.
 * swaps (optional, n-1 ints): 1 where pair (r, r+1) swapped. */
int klft_ptbc_step(klft_batch_t b, int nhits, double delta, uint64_t seed, uint64_t step,
                   int *swaps);
/* label[r] = replica (configuration) currently holding coupling c[r] */
int klft_ptbc_labels(klft_batch_t b, int *labels);
/* FP64 action S_c(U) of configuration `replica` evaluated with coupling
 * index `coupling` (sum over all plaquettes, defect-scaled links) */
int klft_ptbc_action(klft_batch_t b, int replica, int coupling, double *out);

/* ---- autotuner (tune_and_launch_for + read/writeTuneCache, Tuner.hpp:73-358) ---------- */
/* Enable / disable variant tuning (default: the KLFT_TUNING environment
 * variable).  Tuning times every launch variant of a kernel on the first call
 * for a field shape, without changing the live field. */
int klft_tuning(int on);
/* Cache file lines: "<rank> <key> <variant> <ns> 0 .." (the reference's
 * "<rank> <key> t0 .. t_{rank-1}" shape).  A missing file is reported, not fatal. */
int klft_tune_cache_read(const char *path);
int klft_tune_cache_write(const char *path);
int klft_tune_lookup(const char *key, int *variant);

/* ---- kernel timing: HIP events on the library stream around each launch of
 * the named kernel ("metropolis_pass", "force", "update_gauge", "plaquette",
 * "staple_field", "wilson"); the reference prints Kokkos::Timer values ---- */
int klft_profile(int on); /* enable/disable; clears recorded events */
/* sum of durations and number of launches recorded since the last read */
int klft_profile_read(const char *name, double *total_ms, int64_t *count);

/* ---- drop-in driver (klft::Metropolis, lib/Metropolis.cpp:36-166) ------------------ */
/* Reads the YAML input, runs the sweeps and writes the observable files
 * exactly as the reference's metropolis binary.  Returns 0 or -1. */
int klft_metropolis_main(const char *input_file);

#ifdef __cplusplus
}
#endif
#endif
