/*
 * modle_gpu.h: C-ABI of the MI355X engine for MoDLE's simulate hot path.
 *
 * The library replaces the per-cell epoch loop of libmodle's Simulation (the
 * simulate_worker -> simulate_one_cell chain, src/libmodle/cpu/scheduler_simulate.cpp:186-284,
 * src/libmodle/cpu/simulation.cpp:1056-1152) and the ContactMatrixDense<u32> it fills
 * (src/contact_matrix/contact_matrix_dense_impl.hpp). The caller keeps everything above that
 * line: CLI and Config parsing, genome/BED import and the Cooler writer.
 *
 * Conventions: plain C types, no exceptions across the boundary. Every function returns 0 on
 * success and a negative MGPU_E* code on failure; mgpu_last_error() returns the message of the
 * last failure on a context (or of the last context-free call on this thread when ctx is NULL).
 * A context is not thread-safe: use one per driver thread.
 */
#ifndef MODLE_GPU_H
#define MODLE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPU_OK 0
#define MGPU_EINVAL -1   /* invalid argument */
#define MGPU_EDEVICE -2  /* HIP runtime error */
#define MGPU_ENOMEM -3   /* device allocation failed */
#define MGPU_ERANGE -4   /* input outside the supported range */

#define MGPU_MODE_DETERMINISTIC 0 /* reference draw order, bit-exact bands */
#define MGPU_MODE_FAST 1          /* counter-based RNG, statistically equivalent */

/* Stopping criterion (simulation_config.hpp:31). */
#define MGPU_STOP_CONTACT_DENSITY 0
#define MGPU_STOP_EPOCHS 1

/* User-facing parameters with the names, units and defaults of modle::Config
 * (src/common/include/modle/common/simulation_config.hpp:24-135), limited to what the simulate
 * hot path reads. Fill with mgpu_config_defaults(), set what the user passed, then call
 * mgpu_config_transform() with the MGPU_SET_* bits of the options given explicitly. */
typedef struct mgpu_config {
  uint64_t bin_size;
  uint64_t diagonal_width;
  int32_t sample_loop_contacts; /* ContactSamplingStrategy::loop */
  int32_t sample_tad_contacts;  /* ContactSamplingStrategy::tad */
  int32_t noisify_contacts;     /* ContactSamplingStrategy::noisify */
  int32_t stopping_criterion;   /* MGPU_STOP_* */
  double tad_to_loop_contact_ratio;
  double genextreme_mu, genextreme_sigma, genextreme_xi;
  uint64_t rev_extrusion_speed, fwd_extrusion_speed;
  double rev_extrusion_speed_std, fwd_extrusion_speed_std;
  double number_of_lefs_per_mbp;
  uint64_t avg_lef_processivity;
  uint64_t contact_sampling_interval;
  double extrusion_barrier_occupancy;
  double barrier_occupied_stp, barrier_not_occupied_stp;
  double hard_stall_lef_stability_multiplier, soft_stall_lef_stability_multiplier;
  double probability_of_extrusion_unit_bypass;
  double lef_bar_major_collision_pblock, lef_bar_minor_collision_pblock;
  uint64_t target_simulation_epochs;
  double target_contact_density;
  uint64_t num_cells;
  uint64_t seed;
  uint64_t probability_normalization_factor;
  int32_t normalize_probabilities;
  int32_t skip_burnin;
  uint64_t burnin_history_length, burnin_smoothing_window_size;
  uint64_t min_burnin_epochs, max_burnin_epochs;
  double burnin_speed_coefficient;
  int32_t track_1d_lef_position;    /* --track-1d-lef-position */
  int32_t log_model_internal_state; /* --log-model-internal-state */
} mgpu_config;

/* Options given explicitly on the command line (the CLI's "option parsed" tests). */
#define MGPU_SET_REV_SPEED (1u << 0)
#define MGPU_SET_FWD_SPEED (1u << 1)
#define MGPU_SET_BARRIER_OCCUPANCY (1u << 2)

/* Derived parameters, i.e. modle::Config after Cli::transform_args (src/modle/cli.cpp:1192-1215).
 * This is the block every simulate entry point consumes. */
typedef struct mgpu_params {
  uint64_t bin_size;
  uint64_t diagonal_width;
  uint64_t rev_speed, fwd_speed;               /* bp per epoch */
  uint64_t rev_speed_burnin, fwd_speed_burnin; /* bp per epoch during burn-in */
  double rev_speed_std, fwd_speed_std;         /* absolute, bp */
  double prob_release, prob_release_burnin;
  double p_bypass;
  double pblock_major, pblock_minor;
  double hard_stall_mult, soft_stall_mult;
  double tad_to_loop_ratio; /* 0: loop only, +inf: TAD only */
  double gev_mu, gev_sigma, gev_xi;
  int32_t noisify;
  int32_t skip_burnin;
  uint64_t contact_sampling_interval;
  uint64_t burnin_history_length;
  uint64_t burnin_smoothing_window;
  uint64_t min_burnin_epochs;
  uint64_t max_burnin_epochs;
  uint64_t burnin_target_epochs_for_lef_activation;
  double target_contact_density; /* < 0 when the stopping criterion is epochs */
  uint64_t target_simulation_epochs;
  uint64_t seed;
  double lefs_per_mbp;
  double barrier_occupied_stp, barrier_not_occupied_stp; /* defaults for barriers w/o score */
  int32_t override_barrier_occupancy; /* --extrusion-barrier-occupancy given */
  int32_t track_1d_lef_position;      /* a job keeps each chromosome's 1D LEF occupancy */
  int32_t log_model_internal_state;   /* a job logs every cell's per-epoch state */
  int32_t reserved;
} mgpu_params;

/* One chromosome (Chromosome, src/libmodle/internal/include/modle/genome.hpp:32-120). */
typedef struct mgpu_chrom {
  const char* name;
  uint64_t name_len;
  uint64_t size;  /* chrom.sizes length; hashed into the per-cell seed */
  uint64_t start; /* simulated range [start, end) */
  uint64_t end;
} mgpu_chrom;

/* Barriers of one chromosome, sorted by position. dir is the blocking direction
 * (dna::Direction: 1 = REV, 2 = FWD). */
typedef struct mgpu_barriers {
  uint64_t n;
  const uint64_t* pos;
  const uint8_t* dir;
  const double* stp_active;   /* self-transition probability of the occupied state */
  const double* stp_inactive; /* self-transition probability of the free state */
} mgpu_barriers;

/* A BED record as the genome importer sees it (bed::BED fields used at genome.cpp:194-210). */
typedef struct mgpu_bed_record {
  uint64_t chrom_start, chrom_end;
  double score;
  char strand; /* '+', '-' or '.' */
} mgpu_bed_record;

typedef struct mgpu_stats {
  uint64_t cells;
  uint64_t epochs;        /* sum over cells */
  uint64_t burnin_epochs; /* sum over cells */
  uint64_t contacts;      /* sum over cells (registered, incl. missed) */
  uint64_t missed;        /* out-of-band updates */
  double kernel_ms;       /* device time of the simulation launches (HIP events) */
  double total_ms;        /* wall time of the call */
  /* Fast mode only: contacts (and 1D occupancy samples) abandoned after 65 536 rejected draws
   * found no LEF inside the chromosome; sum(band) + missed + dropped == contacts. Always 0 in
   * deterministic mode, which retries without bound as the reference does. */
  uint64_t dropped;
  uint64_t dropped_occupancy;
} mgpu_stats;

typedef struct mgpu_ctx mgpu_ctx;
typedef struct mgpu_job mgpu_job;

/* ---- context ---- */
int mgpu_init(int device, mgpu_ctx** out);
void mgpu_destroy(mgpu_ctx* ctx);
const char* mgpu_last_error(const mgpu_ctx* ctx);
const char* mgpu_version(void);

/* ---- parameters (host only) ---- */
/* modle::Config defaults, simulation_config.hpp:24-135. */
int mgpu_config_defaults(mgpu_config* c);
/* Cli::transform_args, cli.cpp:1085-1215; set_mask = MGPU_SET_* of explicit options. */
int mgpu_config_transform(const mgpu_config* c, uint32_t set_mask, mgpu_params* out);
/* ExtrusionBarrier::compute_stp_active_from_occupancy / compute_occupancy_from_stp,
 * extrusion_barriers_impl.hpp:106-128. */
double mgpu_stp_active_from_occupancy(double stp_inactive, double occupancy);
double mgpu_occupancy_from_stp(double stp_active, double stp_inactive);
/* A chromosome's barriers as Genome::import_barriers leaves them (genome.cpp:433-480, with
 * Chromosome::add_extrusion_barrier, genome.cpp:181-210, and the occupancy override of
 * Simulation::Simulation, simulation.cpp:68-78): recs are the chromosome's BED records in file
 * order; those whose midpoint (start + end + 1) / 2 lies in [chrom_start, chrom_end) are kept, in
 * the order of the chromosome's interval tree (the records ordered by start twice by pdq_sort,
 * once by the parser's tree and once by the chromosome's, interval_tree_impl.hpp:58-68).
 * bed_start / bed_end (may be NULL) receive each kept record's interval, which perturbate's
 * window queries search (map_barriers_to_window, scheduler_common.cpp:113-127). name_puu as in
 * mgpu_prepare_barriers_name_puu (NULL: the default not-bound stp). Outputs hold n entries;
 * *nout receives the number kept. */
int mgpu_import_barriers(const mgpu_params* p, const mgpu_bed_record* recs,
                         const double* name_puu, uint64_t n, uint64_t chrom_start,
                         uint64_t chrom_end, uint64_t* bed_start, uint64_t* bed_end,
                         uint64_t* pos, uint8_t* dir, double* stp_active, double* stp_inactive,
                         uint64_t* nout);
/* The order cpp-sort's pdq_sort (1.13, the published pdqsort) gives indices [0, n) sorted by
 * keys[i] (unstable): the order of the interval trees the reference builds of BED files
 * (IITree::make_BST, interval_tree_impl.hpp:58-68), e.g. perturbate's deletions and feature
 * tracks. Host only. */
int mgpu_pdq_order(const uint64_t* keys, uint64_t n, uint32_t* order);
/* The position sort of the ExtrusionBarriers constructor (extrusion_barriers.cpp:234-259, a
 * pdq_sort of the indices by position), in place: the order the simulation reads. */
int mgpu_sort_barriers(uint64_t n, uint64_t* pos, uint8_t* dir, double* stp_active,
                       double* stp_inactive);
/* mgpu_import_barriers then mgpu_sort_barriers: a chromosome's barrier tables from its BED
 * records. Outputs hold n entries; *nout receives the number kept. */
int mgpu_prepare_barriers(const mgpu_params* p, const mgpu_bed_record* recs, uint64_t n,
                          uint64_t chrom_start, uint64_t chrom_end, uint64_t* pos, uint8_t* dir,
                          double* stp_active, double* stp_inactive, uint64_t* nout);
/* mgpu_prepare_barriers for a BED file whose name field holds each barrier's not-bound
 * self-transition probability (Genome::import_barriers with interpret_name_field_as_puu,
 * genome.cpp:459-470; the simulate CLI's --interpret-extrusion-barrier-name-as-not-bound-stp,
 * Config::interpret_bed_name_field_as_barrier_not_occupied_stp, simulation_config.hpp:93).
 * name_puu[i] is record i's name field parsed as a number (the caller parses it, as
 * utils::parse_numeric_or_throw does); it must lie in [0, 1] (MGPU_EINVAL otherwise, also for
 * records outside [chrom_start, chrom_end), which the reference checks before dropping them) and
 * takes the place of p->barrier_not_occupied_stp for that record. name_puu NULL is
 * mgpu_prepare_barriers. */
int mgpu_prepare_barriers_name_puu(const mgpu_params* p, const mgpu_bed_record* recs,
                                   const double* name_puu, uint64_t n, uint64_t chrom_start,
                                   uint64_t chrom_end, uint64_t* pos, uint8_t* dir,
                                   double* stp_active, double* stp_inactive, uint64_t* nout);
/* Per-chromosome plan of run_simulate (scheduler_simulate.cpp:128-153): number of LEFs, band shape
 * and per-cell contact targets (target_contacts has ncells entries) and target epochs. */
int mgpu_plan_chrom(const mgpu_params* p, const mgpu_chrom* chrom, uint64_t ncells,
                    uint64_t* nlefs, uint64_t* nrows, uint64_t* ncols, uint64_t* target_contacts,
                    uint64_t* target_epochs);
/* Per-cell seed, Chromosome::hash (genome.cpp:317-336): XXH3-64 of name || size || cell_id. */
uint64_t mgpu_cell_seed(const mgpu_chrom* chrom, uint64_t cell_id, uint64_t seed);

/* ---- simulation ---- */
/* Simulates cells [cell_begin, cell_begin + ncells) of one chromosome and adds their contacts to
 * out_band (nrows*ncols + 1 u32, column-major band, data[col*nrows + (col-row)]; the layout of
 * ContactMatrixDense<u32>::_contacts). out_band is a host pointer; *out_missed receives the
 * out-of-band count. Replaces simulate_one_cell for every task of the chromosome. */
int mgpu_simulate_chrom(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                        const mgpu_barriers* barriers, uint64_t nlefs, uint64_t cell_begin,
                        uint64_t ncells, const uint64_t* target_contacts, uint64_t target_epochs,
                        int mode, uint32_t* out_band, uint64_t* out_missed, mgpu_stats* stats);

/* mgpu_simulate_chrom in fast mode without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Inputs are host arrays as for
 * mgpu_simulate_chrom; d_band is the caller's device memory for the chromosome's band
 * (nrows*ncols + 1 u32), to which the cells' contacts are added; *out_missed receives the
 * out-of-band count. Counters, barrier tables, work items and per-cell statistics live in a
 * statically allocated 4 MiB scratch area per GPU (set up by mgpu_init), so the call takes at
 * most 65 536 barriers and 16 384 cells, cells whose working set fits in LDS (about 5 000 LEFs
 * with its barriers), no 1D occupancy or state log; larger inputs are MGPU_ERANGE / MGPU_EINVAL.
 * Every call on one GPU shares that scratch area: calls that overlap on one GPU race. */
int mgpu_simulate_chrom_noalloc(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                                const mgpu_barriers* barriers, uint64_t nlefs,
                                uint64_t cell_begin, uint64_t ncells,
                                const uint64_t* target_contacts, uint64_t target_epochs,
                                uint32_t* d_band, uint64_t* out_missed, mgpu_stats* stats);

/* Device-resident batch over many chromosomes: inputs are uploaded once, bands live in HBM until
 * read back. per_chrom arrays have nchroms entries; target_contacts[c] has ncells[c] entries. */
int mgpu_job_create(mgpu_ctx* ctx, const mgpu_params* p, uint64_t nchroms,
                    const mgpu_chrom* chroms, const mgpu_barriers* barriers, const uint64_t* nlefs,
                    const uint64_t* ncells, const uint64_t* const* target_contacts,
                    const uint64_t* target_epochs, int mode, mgpu_job** out);
/* The same with one parameter block per chromosome (params has nchroms entries), for parameter
 * sweeps and training windows whose chromosomes differ in LEF density and processivity. The
 * blocks may differ only in lefs_per_mbp, prob_release, prob_release_burnin and
 * burnin_target_epochs_for_lef_activation; in deterministic mode they must be identical. No counterpart in the reference, which runs one Config per
 * process (each block is what Cli::transform_args, cli.cpp:1192-1215, makes of one Config). */
int mgpu_job_create_multi(mgpu_ctx* ctx, const mgpu_params* params, uint64_t nchroms,
                          const mgpu_chrom* chroms, const mgpu_barriers* barriers,
                          const uint64_t* nlefs, const uint64_t* ncells,
                          const uint64_t* const* target_contacts, const uint64_t* target_epochs,
                          int mode, mgpu_job** out);
/* ---- perturbate / replay windows ----
 * One window task of modle perturbate and modle replay (Simulation::TaskPW,
 * simulation.hpp:78-95, simulated by Simulation::simulate_window through simulate_one_cell,
 * scheduler_perturbate.cpp:438-474 and 372-436, scheduler_replay.cpp:103-154): chrom.name and
 * chrom.size are the whole chromosome's (the cell's seed is Chromosome::hash(seed, cell_id),
 * simulation.cpp:1065), chrom.start and chrom.end the outer window [window_start, window_end);
 * barriers are the window's (TaskPW::barriers, map_barriers_to_window,
 * scheduler_common.cpp:113-127), sorted, with the task's deletion applied (mgpu_apply_deletion);
 * nlefs, cell_id, target_contacts and target_epochs are the task's num_lefs, cell_id,
 * num_target_contacts and num_target_epochs (scheduler_perturbate.cpp:286-313). */
typedef struct mgpu_window_task {
  mgpu_chrom chrom;
  mgpu_barriers barriers;
  uint64_t nlefs;
  uint64_t cell_id;
  uint64_t target_contacts;
  uint64_t target_epochs;
} mgpu_window_task;
/* A job of window tasks: task t is the job's chromosome t, with one cell (id tasks[t].cell_id)
 * and its own band over the window (ContactMatrixDense resized to the window,
 * scheduler_perturbate.cpp:418-419); mgpu_job_run simulates every task in one persistent
 * launch, mgpu_job_read_band(job, t, ...) and mgpu_job_block_sums read the results. */
int mgpu_job_create_windows(mgpu_ctx* ctx, const mgpu_params* p, uint64_t ntasks,
                            const mgpu_window_task* tasks, int mode, mgpu_job** out);
/* The barrier configuration of simulate_window (scheduler_perturbate.cpp:463-472): barriers
 * whose position lies in [deletion_begin, deletion_begin + deletion_size) get stp_active 0 and
 * stp_inactive 1 (occupancy 0: inactive from the start and for good). *num_active receives the
 * number of barriers left (the BEDPE's num_active_barriers). Host only. */
int mgpu_apply_deletion(const uint64_t* pos, uint64_t n, uint64_t deletion_begin,
                        uint64_t deletion_size, double* stp_active, double* stp_inactive,
                        uint64_t* num_active);
/* ContactMatrixDense<u32>::unsafe_get_block (contact_matrix_dense_unsafe_impl.hpp:89-113) on the
 * device for n queries: out[i] is the sum, in u32, of the block_size x block_size pixels of
 * chromosome (task) chrom[i]'s band centred on (row[i], col[i]), rows and columns clamped to
 * the band's bins (edge extension), pixels beyond the band counting 0; block_size 1 reads the
 * pixel (unsafe_get). block_size must be odd; the reference also expects it below nrows.
 * Bins are relative to the chromosome's (window's) start. */
int mgpu_job_block_sums(mgpu_job* job, uint64_t n, const uint64_t* chrom, const uint64_t* row,
                        const uint64_t* col, uint64_t block_size, uint32_t* out);
/* The same on a band in host memory (nrows*ncols + 1 u32, the band layout above), such as the
 * reference matrix a perturbate run reads from a .cool file (read_reference_contacts,
 * scheduler_perturbate.cpp:81-91), whose blocks simulate_window compares with the simulated
 * ones (scheduler_perturbate.cpp:503-506). */
int mgpu_band_block_sums(mgpu_ctx* ctx, const uint32_t* band, uint64_t nrows, uint64_t ncols,
                         uint64_t n, const uint64_t* row, const uint64_t* col,
                         uint64_t block_size, uint32_t* out);

/* Zeroes the bands and simulates every cell of the job. */
int mgpu_job_run(mgpu_job* job, mgpu_stats* stats);
/* Records one run (the clears of bands and counters, then the persistent simulation launch, in
 * which every epoch of every cell runs) into a HIP graph; later mgpu_job_run calls replay it.
 * A tuning change or mgpu_job_release_graph drops it. No counterpart in the reference. */
int mgpu_job_capture(mgpu_job* job);
int mgpu_job_release_graph(mgpu_job* job);
/* Copies chromosome c's band to host memory (nrows*ncols + 1 u32) and its missed count. */
int mgpu_job_read_band(mgpu_job* job, uint64_t c, uint32_t* out_band, uint64_t* out_missed);
/* Device pointer of chromosome c's band (valid until mgpu_job_destroy). */
int mgpu_job_band_device_ptr(mgpu_job* job, uint64_t c, void** dptr, uint64_t* nelems);
void mgpu_job_destroy(mgpu_job* job);

/* Per-launch figures of the last run that was not replayed from a graph (no counterpart in the
 * reference: they time the device launches). Fast mode launches once per working-set size tier,
 * each tier on a stream of its own, so the launches of a run overlap (a tier's workgroups start
 * on the CUs the tier before it leaves); deterministic mode reports one launch (when its cells'
 * arrays need more device-memory slots than the slab holds, the launches over the same slots are
 * timed together). ms runs from the moment a launch's stream reaches it to its end. *n receives the number of launches; up to max
 * are written. */
typedef struct mgpu_launch_stats {
  uint32_t threads, grid; /* workgroup size and workgroups */
  uint64_t cells;
  uint64_t epochs;   /* summed over the launch's cells */
  uint64_t contacts; /* registered by the launch's cells (incl. missed) */
  uint64_t barriers; /* barrier-table entries the launch's cells read (per cell, summed) */
  double ms;         /* device time between HIP events around the launch */
} mgpu_launch_stats;
int mgpu_job_launch_stats(mgpu_job* job, mgpu_launch_stats* out, uint32_t max, uint32_t* n);
/* Name of the simulation kernel the job's runs launch (as rocprof lists it, without template
 * arguments): "fast_cells_kernel", "det_wave_kernel" (deterministic mode, one wavefront per
 * cell, its working set in LDS), "det_wave_slab_kernel" (the same with the cells' arrays in
 * device memory, for cells too large for the LDS) or "det_cells_kernel" (deterministic mode on
 * one lane per cell, after mgpu_debug_det_kernel). No counterpart in the reference. */
int mgpu_job_kernel(mgpu_job* job, char* name, uint64_t cap);
/* Per-cell figures of the last run, four per work item in job order (chromosomes in the order
 * given, their cells in order): epochs, burn-in epochs, contacts registered, and in
 * deterministic mode the engine words the cell consumed (Xoshiro256++ outputs, as
 * Simulation::State's generator advanced; 0 in fast mode). max: entries out holds (4 per
 * cell); *n: the job's cells. The reference keeps these per task in its state and logs
 * (simulation.cpp:1056-1152); no entry point of its own. */
int mgpu_job_cell_stats(mgpu_job* job, uint64_t* out, uint64_t max, uint64_t* n);

/* ---- optional outputs ----
 * 1D LEF occupancy (Chromosome::lef_1d_occupancy, genome.cpp:236-245, filled by
 * Simulation::register_1d_lef_occupancy, register_contacts.cpp:200-232): jobs created with
 * track_1d_lef_position keep one u64 histogram of ceil(size / bin_size) bins per chromosome,
 * cleared by every run. Each contact sampling event adds the bins of both units of one more
 * sampled LEF. Copies chromosome c's histogram to out (nbins entries, nbins may be NULL). */
int mgpu_job_read_occupancy(mgpu_job* job, uint64_t c, uint64_t* out, uint64_t* nbins);

/* One line of the model internal-state log (Simulation::dump_stats, simulation.cpp:1166-1224),
 * taken after extrusion and before LEF release each epoch. */
typedef struct mgpu_state_record {
  uint64_t chrom; /* index of the chromosome in the job */
  uint64_t cell_id;
  uint64_t epoch;
  int32_t burnin; /* 1 while burn-in runs */
  uint32_t num_active_lefs;
  uint64_t num_barriers, active_barriers;
  uint64_t stalls_rev, stalls_fwd, stalls_both;
  uint64_t lef_bar, primary, secondary; /* collisions that occurred, both sides summed */
  double avg_loop_size;
} mgpu_state_record;

/* Replaces the per-thread compressed_io::Writer of simulate_worker
 * (scheduler_simulate.cpp:205-226): device memory for `capacity` records, filled by every later
 * run (a job created with log_model_internal_state starts with room for about twice the epochs
 * its cells are expected to need); 0 turns logging off. */
int mgpu_job_set_state_log(mgpu_job* job, uint64_t capacity);
/* Copies the last run's records, sorted by (chromosome, cell, epoch), into out (room for cap).
 * *n_records receives the number the run produced; MGPU_ERANGE when that exceeds the job's
 * capacity or cap (nothing is copied then: enlarge and run again). */
int mgpu_job_read_state_log(mgpu_job* job, mgpu_state_record* out, uint64_t cap,
                            uint64_t* n_records);

/* ---- launch tuning (no counterpart in the reference: it sizes the device launch) ----
 * Fast mode runs its cells in persistent launches. By default (threads_per_cell 0) it makes one
 * launch per size tier: cells whose working set lets 1, 2 or 4 of them share one CU's LDS run at
 * 1024, 512 or 256 lanes per cell (16 waves per CU in every tier); cells too large for LDS whose
 * positions, ranks and barrier states fit there run hybrid at 1024 lanes (the per-epoch
 * temporaries in device memory); larger cells run wholly from device memory at 512 lanes;
 * tier_threads[k] (set, when nonzero) overrides tier k's size.
 * threads_per_cell 64, 128, 256, 512 or 1024 instead runs every cell in one launch of that size.
 * global_working_set: 1 keeps every cell's working set in device memory, in one launch (0 or the
 * chosen threads_per_cell: 512 lanes by default). Bands, missed counts and statistics do not
 * depend on the tuning. Get only: launches (the launches the next run makes, at most 5), their
 * sizes in tier_threads and grids in tier_grid, and grid (the largest of them).
 * Deterministic mode (get only): 64 threads per cell, grid = cells per launch, launches > 1
 * when the cells' arrays sit in device memory (global_working_set 1) and more cells than the
 * slab's slots run, one launch after another over the same slots. */
typedef struct mgpu_tuning {
  uint32_t threads_per_cell;
  uint32_t grid;
  uint32_t global_working_set;
  uint32_t launches;
  uint32_t tier_threads[5];
  uint32_t tier_grid[5];
} mgpu_tuning;
int mgpu_job_set_tuning(mgpu_job* job, const mgpu_tuning* t);
int mgpu_job_get_tuning(mgpu_job* job, mgpu_tuning* t);
/* Times each fast-mode workgroup size (64 to 1024 lanes) on sample_cells cells spread over the
 * job (0: all; each tier gets its share, but at least 4 cells per CU or all of its cells, so
 * every size runs at the occupancy of the full run), per size tier (or for the single launch a
 * global_working_set job makes), keeps
 * the fastest for later runs and reports the tuning in *chosen (may be null); ms (5 entries,
 * may be null) receives the times summed over the tiers. Deterministic mode: no-op. */
int mgpu_job_autotune(mgpu_job* job, uint64_t sample_cells, mgpu_tuning* chosen, float* ms);

/* ---- step-level entry points (the reference's ENABLE_TESTING hooks, simulation.hpp:499-651),
 * run by the device step code on one LEF state ----
 * rev/fwd/epoch: LEF arrays (UINT64_MAX when unbound); ranks and moves as in the reference;
 * collisions use the reference's 64-bit encoding (index | event << 56).
 * op: 0 adjust+clamp moves; 1 process_collisions without fix_secondary (test_process_collisions);
 *     2 full process_collisions; 3 detect_primary + correct_primary + process_secondary with no
 *     boundary units; 4 fix_secondary with no boundary units; 5 detect_lef_bar with no boundary
 *     units; 6 detect_primary with no boundary units; 7 correct_moves_for_lef_bar;
 *     8 detect_units_at_chrom_boundaries; 9 rank_lefs (pdq, init ranks); 10 rank_lefs (split);
 *     11 generate_moves (burn-in speeds off). rng_state is the 4-word engine state (in/out). */
int mgpu_debug_step(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom, int op,
                    uint64_t n, uint64_t* rev, uint64_t* fwd, uint64_t* epoch,
                    uint64_t* rev_ranks, uint64_t* fwd_ranks, uint64_t* rev_moves,
                    uint64_t* fwd_moves, uint64_t* rev_coll, uint64_t* fwd_coll,
                    const mgpu_barriers* barriers, const uint8_t* barrier_state,
                    uint64_t* rng_state);
/* n draws of one sampler from a seeded engine (kind as the oracle: 0 raw, 1 normal(a,b),
 * 2 bernoulli(a), 3 canonical, 4 uniform[a,b], 5 poisson(a), 6 binomial(a,b), 7 gev(a,b,c),
 * 8 exponential(a)); doubles are returned as bit patterns. Runs on the device. */
int mgpu_debug_draws(mgpu_ctx* ctx, uint64_t seed, int kind, double a, double b, double c,
                     uint64_t n, uint64_t* out);
/* Fast mode's counter-based generator, Philox4x32-10, on the device: in holds n tuples of
 * (counter[4], key[2]) words, out receives n x 4 output words. No counterpart in the reference
 * (fast mode replaces its engine); exposed for known-answer tests. */
int mgpu_debug_philox(mgpu_ctx* ctx, const uint32_t* in, uint64_t n, uint32_t* out);
/* ContactMatrixDense<u32>::increment (contact_matrix_dense_safe_impl.hpp:53-87) as the device
 * applies it in deterministic (mode 0, 64-bit coordinates) or fast mode (mode 1, 32-bit): the n
 * increments at (rows[i], cols[i]) are added in order to band (nrows*ncols + 1 u32, the band
 * layout of mgpu_simulate_chrom; in/out) and *missed (in/out) counts those whose diagonal
 * offset |row - col| is nrows or more. Coordinates beyond the matrix are MGPU_ERANGE, as
 * the reference's debug builds throw. Exposed for the reference's matrix known-answer tests. */
int mgpu_debug_band_increments(mgpu_ctx* ctx, int mode, uint64_t nrows, uint64_t ncols,
                               uint64_t n, const uint64_t* rows, const uint64_t* cols,
                               uint32_t* band, uint64_t* missed);
/* Deterministic mode's collision words (Collision<u32f>, collision_encoding_impl.hpp:73-240) on
 * the device for n (index, event) pairs: the encoded word, its decoded index and event, and
 * flags: bit 0 the COLLISION bit, bits 1-4 collision occurred with CHROM_BOUNDARY, LEF_BAR,
 * LEF_LEF_PRIMARY, LEF_LEF_SECONDARY, bits 5-8 the same events avoided. */
int mgpu_debug_collisions(mgpu_ctx* ctx, uint64_t n, const uint64_t* idx, const uint8_t* ev,
                          uint64_t* words, uint64_t* index, uint32_t* event, uint32_t* flags);
/* Fast mode's double-precision elementary functions on the device (no counterpart in the
 * reference, whose draws call libm): fn 0 ln(x), 1 expm1(x), 2 sin(2 pi x) and cos(2 pi x) (out
 * holds 2n values, the sines first). Exposed for accuracy tests. */
int mgpu_debug_fast_math(mgpu_ctx* ctx, int fn, const double* x, uint64_t n, double* out);
/* Fast mode's epoch-phase timers (no counterpart in the reference): enable != 0 makes later
 * job runs on this context accumulate, per phase, the 100 MHz wall-clock ticks of every cell;
 * out (n entries, may be null) receives the totals, which are then cleared. Phases: 0 stop test
 * and burn-in, 1 bind, 2 rank (both sides; the rev side alone in epochs that bind more than
 * 1024 LEFs), 3 rank fwd in those epochs, 4 contacts, 5 moves, 6 adjust, 7 barrier states
 * and chromosome ends, 8 LEF-BAR, 9 primary LEF-LEF, 10 move corrections, 11 secondary scans
 * of both sides, 13 bypass fix-ups of both sides, 14 extrude and release (12 and 15 unused).
 * Entries 16-27 (when n > 16) are deterministic mode's wave-parallel kernel: 16 burn-in, 17 bind,
 * 18 rank, 19 contacts, 20 moves and adjust, 21 barrier states, 22 chromosome ends, 23 LEF-BAR,
 * 24 primary LEF-LEF and move corrections, 25 secondary, 26 bypass fix-ups, 27 extrude and
 * release. */
int mgpu_debug_phase_times(mgpu_ctx* ctx, int enable, uint64_t* out, uint64_t n);
/* Fast mode's bypass fix-ups (fix_secondary_lef_lef_collisions,
 * simulation_detect_collisions.cpp:520-647) run in parallel over runs of consecutive ranks;
 * serial != 0 makes later runs of the job walk each side's list on one lane in the reference
 * loop's order instead. Results must not change: exposed to check exactly that. */
int mgpu_debug_fast_serial_fixups(mgpu_job* job, int serial);
/* Deterministic mode runs each cell on one wavefront (all 64 lanes, the reference's draws in
 * the reference's order); serial != 0 makes later jobs and traces of this context run every
 * cell on the one-lane kernel instead. Both give the reference's bands and optional outputs:
 * exposed to compare them. */
int mgpu_debug_det_kernel(mgpu_ctx* ctx, int serial);
/* Caps the device-memory cell slots of later deterministic jobs whose cells exceed the LDS
 * (0: as many as the 8 GiB slab holds). A job with more such cells than slots runs as several
 * launches over the same slots; the cap lets a test make that happen with a few cells. */
int mgpu_debug_det_slab_slots(mgpu_ctx* ctx, uint64_t slots);
/* Per-epoch state hash of one cell in deterministic mode (for locating a first divergence). */
int mgpu_debug_cell_trace(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                          const mgpu_barriers* barriers, uint64_t nlefs, uint64_t cell_id,
                          uint64_t target_contacts, uint64_t target_epochs, uint64_t max_trace,
                          uint64_t* trace, uint64_t* out_epochs);

/* HIP stream (hipStream_t) and device of a context, for code that enqueues its own work in
 * order with the library's. */
int mgpu_context_stream(mgpu_ctx* ctx, void** stream, int* device);

/* ---- learned surrogate (no counterpart in the reference) ----
 * A network that stands in for the simulation in parameter sweeps: it maps the barrier track of
 * a window (plus LEF density and processivity, broadcast as input channels) to the log1p contact
 * band of the window's central output bins. BF16 weights and activations on MFMA with FP32
 * accumulation; fp32 master weights updated by AdamW. Input x: [n][in_len][in_ch] fp32; output
 * and targets: [n][ndiag][out_len] fp32, entry [d][i] = pixel (i, i + d) of the output window
 * (i + d >= out_len is outside the window: predicted but not trained). Parameters travel in a
 * canonical fp32 order: per layer weights [cout][taps][cin] then biases, for the encoder layers,
 * the two 1 x 1 projections (with a bias on the first), the [ndiag][head_ch] diagonal embedding,
 * the head layers and the 1-channel output layer. */
typedef struct mgpu_sg_config {
  uint32_t in_len;      /* input bins: output window plus context on both sides */
  uint32_t out_offset;  /* first output bin within the input */
  uint32_t out_len;     /* output bins */
  uint32_t ndiag;       /* predicted diagonals (band width in bins) */
  uint32_t in_ch;       /* input channels per bin, at most 32 */
  uint32_t enc_ch;      /* encoder channels: 64 or 128 */
  uint32_t enc_layers;  /* 1D conv layers in the encoder */
  uint32_t enc_k;       /* encoder kernel width (odd, at most 9) */
  uint32_t head_ch;     /* pairwise / head channels: 64 */
  uint32_t head_layers; /* 3 x 3 conv layers of the head, the 1-channel output layer included */
  uint32_t max_batch;   /* samples per device batch */
  uint32_t reserved;
  uint64_t seed;        /* weight initialisation */
  uint32_t enc_dilation[16]; /* dilation of encoder layer l (0 means 1) */
} mgpu_sg_config;
typedef struct mgpu_sg mgpu_sg;

/* 600 input bins (2 Mbp window of 400 5-kb bins with 500 kb of context each side), 100
 * diagonals, 6 input channels, 8 x 128-channel encoder (k = 5, dilations 1 1 2 4 8 16 32 1: each
 * output bin sees 130 bins to either side), 4-layer 64-channel head. */
int mgpu_sg_config_default(mgpu_sg_config* c);
int mgpu_sg_create(mgpu_ctx* ctx, const mgpu_sg_config* c, mgpu_sg** out);
void mgpu_sg_destroy(mgpu_sg* s);
const char* mgpu_sg_last_error(const mgpu_sg* s);
uint64_t mgpu_sg_num_params(const mgpu_sg* s);
int mgpu_sg_get_params(mgpu_sg* s, float* out);
/* Replaces the weights and resets the optimiser state. */
int mgpu_sg_set_params(mgpu_sg* s, const float* in);
/* Predictions for n samples (any n; batched by max_batch). */
int mgpu_sg_forward(mgpu_sg* s, const float* x, uint64_t n, float* y);
/* Loss (mean squared error over valid pixels) and its gradient for n <= max_batch samples;
 * grads (may be null) receives num_params values in the canonical order. */
int mgpu_sg_gradients(mgpu_sg* s, const float* x, const float* y, uint64_t n, float* loss,
                      float* grads);
/* One AdamW step (beta 0.9 / 0.999, eps 1e-8, decoupled weight decay) with the given gradients. */
int mgpu_sg_adamw_step(mgpu_sg* s, const float* grads, float lr, float weight_decay);
/* Uploads a training set (n samples) to device memory. */
int mgpu_sg_set_data(mgpu_sg* s, const float* x, const float* y, uint64_t n);
/* One training step on samples idx[0, n) of the training set (n <= max_batch): forward,
 * backward and AdamW on the device. *loss receives the batch loss before the update, *ms (may be
 * null) the device time of the step. */
int mgpu_sg_train_step(mgpu_sg* s, const uint32_t* idx, uint64_t n, float lr, float weight_decay,
                       float* loss, float* ms);
/* Device time of `repeats` forward passes over the last uploaded batch of n samples. */
int mgpu_sg_predict_timed(mgpu_sg* s, uint64_t n, uint64_t repeats, float* ms);
/* Whole-chromosome predictions kept in device memory (no counterpart in the reference: the
 * surrogate's sweep path). reserve: room for npoints bands of nbins bins (ndiag x nbins floats
 * each), replacing an earlier reservation. predict: chromosomes first_point .. + npoints, each
 * tiled by nwin windows whose inputs x holds as for mgpu_sg_forward (npoints * nwin samples) and
 * whose outputs start at bins starts[0] = 0 < ... ; pixel (i, i + d) comes from the last window
 * starting at or before i and is stored as the relative count expm1(y) / 10. read: copies bands
 * to the host ([npoints][ndiag][nbins]). */
int mgpu_sg_reserve_bands(mgpu_sg* s, uint64_t npoints, uint64_t nbins);
int mgpu_sg_predict_chromosomes(mgpu_sg* s, const float* x, uint64_t first_point,
                                uint64_t npoints, uint64_t nwin, const uint32_t* starts);
int mgpu_sg_read_bands(mgpu_sg* s, uint64_t first_point, uint64_t npoints, float* out);
/* One v_mfma_f32_32x32x16_bf16 of a (32 x 16) and b (16 x 32), bf16 row-major, into d (32 x 32
 * fp32 row-major) through the lane layout the surrogate kernels assume. */
int mgpu_debug_mfma(mgpu_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d);

#ifdef __cplusplus
}
#endif
#endif
