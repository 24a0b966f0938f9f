/*
 * refcpu: CPU oracle for MoDLE's per-cell loop-extrusion epoch loop and banded contact matrix.
 *
 * TEST INFRASTRUCTURE. This is the checker, never the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. Nothing under
 * modlefork_amd/ includes, links or calls it.
 *
 * It is a serial restatement of the reference (MoDLE 1.0.0-rc.7) hot path; every function in
 * refcpu.c cites the reference file:line it follows. Third-party pieces the reference pulls in
 * (Boost.Random 1.79 distributions, xoshiro-cpp 1.1, cpp-sort 1.13 sorters, xxHash 0.8.1) are
 * restated from their published algorithms; their pins are listed in refcpu.c.
 */
#ifndef MODLE_REFCPU_H
#define MODLE_REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Post-transform simulation parameters (src/common/include/modle/common/simulation_config.hpp:24-135
 * after src/modle/cli.cpp:1192-1215). */
typedef struct rc_config {
  uint64_t bin_size;
  uint64_t diagonal_width;
  uint64_t rev_speed, fwd_speed;
  uint64_t rev_speed_burnin, fwd_speed_burnin;
  double rev_speed_std, fwd_speed_std;
  double prob_release, prob_release_burnin;
  double p_bypass;
  double pblock_major, pblock_minor;
  double hard_stall_mult, soft_stall_mult;
  double tad_to_loop_ratio;
  double gev_mu, gev_sigma, gev_xi;
  int32_t noisify;
  int32_t skip_burnin;
  uint64_t contact_sampling_interval;
  uint64_t burnin_history_length;
  uint64_t burnin_smoothing_window;
  uint64_t min_burnin_epochs;
  uint64_t max_burnin_epochs;
  uint64_t burnin_target_epochs_for_lef_activation;
  double target_contact_density; /* < 0 => stop on epochs */
  uint64_t target_simulation_epochs;
  uint64_t seed;
} rc_config;

typedef struct rc_chrom {
  const char* name;
  uint64_t name_len;
  uint64_t size;  /* full chrom.sizes length (hashed into the per-cell seed) */
  uint64_t start; /* simulated [start, end) */
  uint64_t end;
} rc_chrom;

/* Barriers sorted by position. dir = blocking direction: 1 = REV, 2 = FWD (dna.hpp:75-78). */
typedef struct rc_barriers {
  uint64_t n;
  const uint64_t* pos;
  const uint8_t* dir;
  const double* stp_active;   /* pbb */
  const double* stp_inactive; /* puu */
} rc_barriers;

typedef struct rc_cell_stats {
  uint64_t epochs;
  uint64_t burnin_epochs;
  uint64_t contacts;
  uint64_t draws; /* 64-bit words consumed from the cell's engine */
} rc_cell_stats;

/* ---- RNG stack ---- */
typedef struct rc_rng {
  uint64_t s[4];
  uint64_t draws;
} rc_rng;

uint64_t rc_xxh3_64_withseed(const void* data, size_t len, uint64_t seed);
uint64_t rc_cell_seed(const char* name, uint64_t name_len, uint64_t chrom_size, uint64_t cell_id,
                      uint64_t seed);
void rc_rng_seed(rc_rng* g, uint64_t seed);
uint64_t rc_rng_next(rc_rng* g);
uint64_t rc_uniform_u64(rc_rng* g, uint64_t lo, uint64_t hi);
int rc_bernoulli(rc_rng* g, double p);
double rc_canonical(rc_rng* g);
double rc_uniform01(rc_rng* g);
double rc_normal(rc_rng* g, double mean, double sigma);
double rc_exponential(rc_rng* g, double lambda);
uint64_t rc_poisson(rc_rng* g, double mean);
int64_t rc_binomial(rc_rng* g, int64_t t, double p);
double rc_gev(rc_rng* g, double mu, double sigma, double xi);
/* Fill out[0..n) with draws of one kind: 0 raw, 1 normal(a,b), 2 bernoulli(a), 3 canonical,
 * 4 uniform_u64[a,b], 5 poisson(a), 6 binomial(int a, b), 7 gev(mu=a, sigma=b, xi=c),
 * 8 exponential(a). Doubles are returned through their bit patterns. */
void rc_draw_sequence(uint64_t seed, int kind, double a, double b, double c, uint64_t n,
                      uint64_t* out);

/* ---- sorters (index arrays keyed by u64) ---- */
void rc_pdq_sort_idx(size_t* idx, size_t n, const uint64_t* key);
void rc_split_sort_idx(size_t* idx, size_t n, const uint64_t* key);

/* ---- step-level entry points (the reference's ENABLE_TESTING hooks, simulation.hpp:499-651) ----
 * LEF arrays: rev[i], fwd[i], epoch[i] (binding epoch, UINT64_MAX when unbound).
 * Collisions use the reference encoding: index | (event << 56). */
void rc_rank_lefs(const uint64_t* rev, const uint64_t* fwd, const uint64_t* epoch, size_t n,
                  size_t* rev_ranks, size_t* fwd_ranks, int partially_sorted, int init_buffers);
void rc_adjust_and_clamp_moves(const rc_chrom* chrom, const uint64_t* rev, const uint64_t* fwd,
                               const uint64_t* epoch, size_t n, const size_t* rev_ranks,
                               const size_t* fwd_ranks, uint64_t* rev_moves, uint64_t* fwd_moves,
                               int adjust);
/* mode 0: full process_collisions (incl. fix_secondary, simulation.cpp:925-955)
 * mode 1: test_process_collisions (no fix_secondary, simulation.hpp:587-618)
 * mode 2: detect_units_at_chrom_boundaries + detect_lef_bar_collisions only
 * mode 3: detect_primary_lef_lef + correct_moves_for_primary + process_secondary (test hook :631)
 * mode 4: fix_secondary_lef_lef_collisions only, with n5 = n3 = 0 (test hook :620)
 * mode 5: detect_lef_bar_collisions only, with n5 = n3 = 0
 * mode 6: detect_primary_lef_lef_collisions only, with n5 = n3 = 0
 * mode 7: correct_moves_for_lef_bar_collisions only
 * mode 8: detect_units_at_chrom_boundaries only */
void rc_process_collisions(const rc_config* cfg, const rc_chrom* chrom, uint64_t* rev,
                           uint64_t* fwd, const uint64_t* epoch, size_t n, size_t* rev_ranks,
                           size_t* fwd_ranks, uint64_t* rev_moves, uint64_t* fwd_moves,
                           const rc_barriers* barriers, const uint8_t* barrier_state,
                           uint64_t* rev_coll, uint64_t* fwd_coll, rc_rng* g, int mode,
                           uint64_t* n5_out, uint64_t* n3_out);
void rc_generate_moves(const rc_config* cfg, const rc_chrom* chrom, const uint64_t* rev,
                       const uint64_t* fwd, const uint64_t* epoch, size_t n,
                       const size_t* rev_ranks, const size_t* fwd_ranks, uint64_t* rev_moves,
                       uint64_t* fwd_moves, int burnin_completed, rc_rng* g, int adjust);
void rc_bind_lefs(uint64_t start, uint64_t end, uint64_t* rev, uint64_t* fwd, uint64_t* epoch,
                  size_t n, size_t* rev_ranks, size_t* fwd_ranks, const uint8_t* mask,
                  rc_rng* g, uint64_t current_epoch);
double rc_occupancy_from_stp(double stp_active, double stp_inactive);
/* collision words (reference 64-bit layout): encode, decode and the occurred / avoided tests
 * (ev 0: any event) */
uint64_t rc_coll_encode(uint64_t idx, unsigned ev);
uint64_t rc_coll_index(uint64_t c);
unsigned rc_coll_event(uint64_t c);
int rc_coll_occurred(uint64_t c, unsigned ev);
int rc_coll_avoided(uint64_t c, unsigned ev);
double rc_stp_active_from_occupancy(double stp_inactive, double occupancy);

/* ---- contact matrix ---- */
typedef struct rc_band {
  uint64_t nrows, ncols;
  uint32_t* data; /* nrows*ncols + 1, column-major, data[col*nrows + (col-row)] */
  uint64_t missed;
} rc_band;
void rc_band_increment(rc_band* m, uint64_t row, uint64_t col);
uint32_t rc_band_get(const rc_band* m, uint64_t row, uint64_t col);
uint32_t rc_band_get_block(const rc_band* m, uint64_t row, uint64_t col, uint64_t block_size);
uint64_t rc_band_nrows(uint64_t diagonal_width, uint64_t bin_size, uint64_t length);
uint64_t rc_band_ncols(uint64_t bin_size, uint64_t length);

/* ---- whole-cell / whole-chromosome simulation ---- */
uint64_t rc_num_lefs(double lefs_per_mbp, uint64_t simulated_size);
uint64_t rc_contacts_per_epoch(const rc_config* cfg, uint64_t nlefs);
/* Per-cell targets (scheduler_simulate.cpp:131-153); returns tot target epochs. */
uint64_t rc_plan_targets(const rc_config* cfg, uint64_t ncells, uint64_t nlefs, uint64_t npixels,
                         uint64_t* target_contacts);
/* epoch_hash (optional, may be NULL): per-epoch hash of LEF positions, up to max_trace entries. */
int rc_simulate_cell(const rc_config* cfg, const rc_chrom* chrom, const rc_barriers* barriers,
                     uint64_t nlefs, uint64_t cell_id, uint64_t target_contacts,
                     uint64_t target_epochs, rc_band* band, rc_cell_stats* stats,
                     uint64_t* epoch_hash, uint64_t max_trace);
/* Simulates cells [cell_begin, cell_end) on nthreads threads into band (atomic increments). */
int rc_simulate_cells(const rc_config* cfg, const rc_chrom* chrom, const rc_barriers* barriers,
                      uint64_t nlefs, uint64_t cell_begin, uint64_t cell_end,
                      const uint64_t* target_contacts, uint64_t target_epochs, rc_band* band,
                      int nthreads, rc_cell_stats* stats);

/* ---- optional per-cell outputs ---- */
/* One line of the model internal-state log (Simulation::dump_stats, simulation.cpp:1166-1224),
 * taken after extrude and before release_lefs (simulation.cpp:1133-1145). */
typedef struct rc_state_record {
  uint64_t cell_id;
  uint64_t epoch;
  int32_t burnin; /* 1 while burn-in is running */
  uint32_t num_active_lefs;
  uint64_t num_barriers, active_barriers;
  uint64_t stalls_rev, stalls_fwd, stalls_both;
  uint64_t lef_bar, primary, secondary;
  double avg_loop_size;
} rc_state_record;

/* occupancy: u64[ceil(chrom size / bin)] summed over cells (register_1d_lef_occupancy,
 * register_contacts.cpp:200-232), NULL when --track-1d-lef-position is off. log: records
 * appended in any cell order (slot = atomic counter); log_n counts every record produced, so
 * log_n > log_cap means the buffer was too small. NULL when the state log is off. */
typedef struct rc_outputs {
  uint64_t* occupancy;
  rc_state_record* log;
  uint64_t log_cap;
  uint64_t log_n;
} rc_outputs;

/* dump_stats' counts over one epoch's state: n active LEFs, their collisions (reference
 * encoding) and nbar barrier states (1 = active). cell_id, epoch and burnin are left to the
 * caller. */
void rc_count_state(const uint64_t* rev, const uint64_t* fwd, size_t n, const uint64_t* rev_coll,
                    const uint64_t* fwd_coll, const uint8_t* barrier_state, size_t nbar,
                    rc_state_record* out);
int rc_simulate_cell_out(const rc_config* cfg, const rc_chrom* chrom, const rc_barriers* barriers,
                         uint64_t nlefs, uint64_t cell_id, uint64_t target_contacts,
                         uint64_t target_epochs, rc_band* band, rc_cell_stats* stats,
                         uint64_t* epoch_hash, uint64_t max_trace, rc_outputs* out);
int rc_simulate_cells_out(const rc_config* cfg, const rc_chrom* chrom,
                          const rc_barriers* barriers, uint64_t nlefs, uint64_t cell_begin,
                          uint64_t cell_end, const uint64_t* target_contacts,
                          uint64_t target_epochs, rc_band* band, int nthreads,
                          rc_cell_stats* stats, rc_outputs* out);

#ifdef __cplusplus
}
#endif
#endif
