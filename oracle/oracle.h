/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference's particle-history hot path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (minimc_amd/) never links, loads or calls it.
 *
 * It consumes the same flattened World descriptor as the product
 * (include/minimc_amd.h) and follows the reference's functions one by one;
 * each function in oracle.cpp cites the reference file:line it restates.
 * Parity-pinning status of every part is listed at the top of oracle.cpp.
 */
#ifndef MMC_ORACLE_H
#define MMC_ORACLE_H

#include "../include/minimc_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

/* RNG known answers (libstdc++ minstd_rand semantics). */
int orc_rng_draws(uint64_t seed, uint64_t n, uint64_t* out);
int orc_rng_canonicals(uint64_t seed, uint64_t n, double* out);
int orc_rng_exponentials(uint64_t seed, double lambda, uint64_t n,
                         double* out);
int orc_rng_directions(uint64_t seed, uint64_t n, double* out3);

int orc_create(const mmc_world_desc* world, const mmc_tally_desc* tallies,
               const mmc_options* options, orc_world** out);
void orc_destroy(orc_world* w);
const char* orc_last_error(const orc_world* w);

int orc_sample_source(orc_world* w, const mmc_source_desc* src,
                      uint64_t seed0, uint64_t n, mmc_particle_soa* out);
int orc_eval_event(orc_world* w, int kind, mmc_particle_soa* p, uint64_t n);
int orc_run_fixed_source(orc_world* w, const mmc_source_desc* src,
                         uint64_t seed0, uint64_t first, uint64_t n,
                         int threads);
int orc_keig_init(orc_world* w, const mmc_source_desc* src, uint64_t n);
int orc_run_cycle(orc_world* w, int active, int threads,
                  mmc_cycle_stats* out);
int orc_get_source(orc_world* w, mmc_site* out, uint64_t cap, uint64_t* n);
int orc_set_source(orc_world* w, const mmc_site* sites, uint64_t n);
int orc_tally_size(orc_world* w, int estimator, uint64_t* n);
int orc_get_tallies(orc_world* w, int estimator, double* S, double* S2,
                    uint64_t n);
int orc_reset_tallies(orc_world* w);
int orc_get_counters(orc_world* w, mmc_counters* out);
int orc_tsl_sample(orc_world* w, int tsl, uint64_t n, const double* E,
                   const double* T, uint32_t* rng, double* alpha,
                   double* beta, int threads);
/* POD reconstruction of one partition entry (BetaPartition/AlphaPartition::
 * Evaluate): which = 0 beta partitions, 1 alpha partitions. */
int orc_tsl_evaluate(orc_world* w, int tsl, int which, int part, uint64_t cdf_i,
                     uint64_t x_i, double T, double* out);
/* ThermalScattering::GetTotal at (E, T). */
int orc_tsl_total(orc_world* w, int tsl, double E, double T, double* out);
int orc_set_entropy_mesh(orc_world* w, const int32_t n[3], const double lo[3],
                         const double hi[3]);

/* Output text, written with the C++ streams the reference writes with
 * (std::scientific, default precision).  Each returns the length of the
 * text (without the terminating zero) and writes at most cap bytes
 * including it. */
/* One number as `os << std::scientific << v` prints it. */
uint64_t orc_format_scientific(double v, char* buf, uint64_t cap);
/* Scorable::GetScoreAsString (Scorable.cpp:51-70). */
uint64_t orc_format_scores(const double* S, const double* S2, uint64_t n,
                           double total_weight, char* buf, uint64_t cap);
/* Bins::to_string of one bins spec (Bins.cpp:51, :84-90, :125-131,
 * :164-170); kind as mmc_bins.kind, bounds for MMC_BINS_BOUNDARIES. */
uint64_t orc_format_bins(int kind, int n, double lo, double hi, double base,
                         const double* bounds, char* buf, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
