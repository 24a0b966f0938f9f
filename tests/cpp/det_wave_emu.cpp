// Host build of the wave-parallel deterministic kernel (modlefork_amd/csrc/det_wave.h) for the
// CPU tests: TEST INFRASTRUCTURE, not the product. It runs det_wave_cell with 64 emulated lanes
// (one coroutine per lane on its own stack, switched round-robin at every wave synchronisation,
// which is the barrier semantics the code is written to; stacks are set up with ucontext and
// switched with _setjmp/_longjmp, which save no signal mask), so the parallel restatement can be compared
// with the oracle draw for draw without a GPU. Inputs are the oracle's parameter structs; they
// are turned into the kernel's DetParams / DetChrom the way libmodle_gpu.so's job set-up does.
#define MGPU_EMU 1
#define __device__
#define __host__
#define __constant__
#define __global__

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <setjmp.h>
#include <ucontext.h>

#include <vector>

namespace mgpu {
inline unsigned atomicAdd(unsigned* p, unsigned v) {
  const unsigned o = *p;
  *p = o + v;
  return o;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  const unsigned long long o = *p;
  *p = o + v;
  return o;
}
inline unsigned atomicMax(unsigned* p, unsigned v) {
  const unsigned o = *p;
  if (v > o) *p = v;
  return o;
}
inline unsigned atomicMin(unsigned* p, unsigned v) {
  const unsigned o = *p;
  if (v < o) *p = v;
  return o;
}
}  // namespace mgpu

#include "../../modlefork_amd/csrc/det_wave.h"
#include "refcpu.h"

namespace {

constexpr int kLanes = 64;
constexpr size_t kStack = 512 * 1024;

struct Sched {
  ucontext_t init_ctx;
  ucontext_t ctx[kLanes];
  jmp_buf main_buf;
  jmp_buf buf[kLanes];
  std::vector<char> stack[kLanes];
  bool done[kLanes];
  uint64_t syncs[kLanes];
  uint64_t slot[kLanes];
  bool broken = false;

  bool down = false;  // lanes run 63 -> 0 between synchronisations instead of 0 -> 63
  int next_live(int l) const {
    for (int k = 1; k <= kLanes; ++k) {
      const int m = down ? (l - k + 2 * kLanes) % kLanes : (l + k) % kLanes;
      if (!done[m]) return m;
    }
    return -1;
  }
  void yield(int l) {
    ++syncs[l];
    const int m = next_live(l);
    if (down ? m >= l : m <= l) {
      // a full round: every live lane must have met the same number of synchronisations
      for (int k = 0; k < kLanes; ++k)
        if (!done[k] && syncs[k] != syncs[l]) broken = true;
    }
    if (m != l && _setjmp(buf[l]) == 0) _longjmp(buf[m], 1);
  }
};

Sched* g_sched = nullptr;
int g_lane_order = 0;
int g_separate_cell = 0;
// optional outputs of the next emulated cells (null: off)
unsigned long long* g_occupancy = nullptr;
mgpu::StateLog g_log{nullptr, nullptr, 0};
uint64_t g_last_syncs = 0;

struct EmuWave {
  uint32_t ln;
  uint32_t lane() const { return ln; }
  void sync() const { g_sched->yield((int)ln); }
  uint64_t ballot(bool p) const {
    g_sched->slot[ln] = p ? 1 : 0;
    sync();
    uint64_t m = 0;
    for (int l = 0; l < kLanes; ++l)
      if (g_sched->slot[l]) m |= 1ull << l;
    sync();
    return m;
  }
  uint64_t bcast(uint64_t v, uint32_t src) const {
    g_sched->slot[ln] = v;
    sync();
    const uint64_t r = g_sched->slot[src];
    sync();
    return r;
  }
  uint64_t shfl(uint64_t v, uint32_t src) const { return bcast(v, src); }
};

struct Job {
  mgpu::DetParams P;
  mgpu::DetChrom ch;
  uint8_t* lds;
  uint8_t* cell;  // the cell's arrays outside the LDS image, or null
  uint64_t seed, target;
  uint64_t* trace;
  uint64_t max_trace;
  uint64_t stats[4];
  uint64_t* dbg;
};
Job* g_job = nullptr;

void lane_entry(int ln) {
  if (_setjmp(g_sched->buf[ln]) == 0) swapcontext(&g_sched->ctx[ln], &g_sched->init_ctx);
  const EmuWave w{(uint32_t)ln};
  mgpu::det_wave_cell(w, g_job->P, g_job->ch, g_job->lds, g_job->cell, g_job->P.hist_len,
                      g_job->seed,
                      g_job->target, g_job->trace, g_job->max_trace,
                      ln == 0 ? g_job->stats : nullptr, 0, g_job->dbg);
  Sched& s = *g_sched;
  s.done[ln] = true;
  for (int k = 0; k < kLanes; ++k)
    if (!s.done[k] && s.syncs[k] != s.syncs[ln]) s.broken = true;
  const int m = s.next_live(ln);
  if (m < 0) _longjmp(s.main_buf, 1);
  _longjmp(s.buf[m], 1);
}

}  // namespace

extern "C" {

// Order in which the emulated lanes run between two synchronisations: 0 ascending, 1
// descending. Results must not depend on it (a missing synchronisation would show).
void emu_set_lane_order(int order) { g_lane_order = order; }

// 1: the cell's arrays in an allocation of their own (the layout of the device-memory
// placement), 0: after the fixed part of the LDS image.
void emu_set_separate_cell(int on) { g_separate_cell = on; }

// Optional outputs of the cells emulated next: 1D LEF occupancy bins (u64, added to) and the
// state log (records of mgpu::StateRecord, `count` advanced per record, records at or past
// `cap` not written). Null pointers turn an output off.
void emu_set_outputs(unsigned long long* occupancy, mgpu::StateRecord* rec,
                     unsigned long long* count, uint64_t cap) {
  g_occupancy = occupancy;
  g_log.rec = rec;
  g_log.count = count;
  g_log.cap = cap;
}

// Wave synchronisations lane 0 met in the last emulated cell (ballots count two).
uint64_t emu_last_syncs() { return g_last_syncs; }

// Bytes of LDS one cell takes in the wave-parallel kernel.
uint64_t emu_det_wave_bytes(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  return mgpu::det_wave_bytes(nlefs, nbar, hist);
}

// Simulates cell `cell_id` of `chrom` as det_wave_kernel does, adding its contacts to band
// (u32[nrows * ncols + 1], ContactMatrixDense layout) and *missed. stats: epochs, burn-in epochs,
// contacts, engine words (or ~0 on an internal error). trace: per-epoch hashes as the oracle's.
// dbg (may be null): the kernel's path counters (mgpu::WDbg), added to.
// Returns 0, or -1 when the lanes fell out of step (a synchronisation reached by some lanes
// only) or the kernel reported an error.
int emu_det_wave_cell(const rc_config* c, const rc_chrom* chrom, const rc_barriers* b,
                      uint64_t nlefs, uint64_t cell_id, uint64_t target, uint64_t target_epochs,
                      uint32_t* band, uint64_t nrows, unsigned long long* missed,
                      uint64_t* trace, uint64_t max_trace, uint64_t* stats, uint64_t* dbg) {
  Job job{};
  job.dbg = dbg;
  mgpu::DetParams& P = job.P;
  P.bin_size = c->bin_size;
  P.rev_speed = c->rev_speed;
  P.fwd_speed = c->fwd_speed;
  P.rev_speed_burnin = c->rev_speed_burnin;
  P.fwd_speed_burnin = c->fwd_speed_burnin;
  P.rev_std = c->rev_speed_std;
  P.fwd_std = c->fwd_speed_std;
  P.prob_release = c->prob_release;
  P.prob_release_burnin = c->prob_release_burnin;
  P.p_bypass = c->p_bypass;
  P.pblock_major = c->pblock_major;
  P.pblock_minor = c->pblock_minor;
  P.inv_soft = 1.0 / c->soft_stall_mult;
  P.inv_hard = 1.0 / c->hard_stall_mult;
  P.tad_to_loop_ratio = c->tad_to_loop_ratio;
  P.gev_mu = c->gev_mu;
  P.gev_sigma = c->gev_sigma;
  P.gev_xi = c->gev_xi;
  P.noisify = c->noisify;
  P.skip_burnin = c->skip_burnin;
  P.hist_len = c->burnin_history_length;
  P.window = c->burnin_smoothing_window;
  P.min_burnin_epochs = c->min_burnin_epochs;
  P.max_burnin_epochs = c->max_burnin_epochs;
  P.target_contact_density = c->target_contact_density;
  P.log = g_log;

  const uint64_t cpe = rc_contacts_per_epoch(c, nlefs);
  std::vector<double> qn(cpe + 1);
  const double p_loop = 1.0 / (c->tad_to_loop_ratio + 1.0);
  const double pr = p_loop > 0.5 ? 1 - p_loop : p_loop;
  for (uint64_t t = 0; t <= cpe; ++t) qn[t] = pow(1 - pr, (double)t);
  unsigned long long dropped[2] = {0, 0};
  mgpu::DetChrom& ch = job.ch;
  ch.start = chrom->start;
  ch.end = chrom->end;
  ch.nlefs = nlefs;
  ch.nbar = b->n;
  ch.bar_pos = b->pos;
  ch.bar_dir = b->dir;
  ch.bar_stp_active = b->stp_active;
  ch.bar_stp_inactive = b->stp_inactive;
  ch.band = band;
  ch.missed = missed;
  ch.dropped = dropped;
  ch.nrows = nrows;
  ch.cpe = cpe;
  ch.target_epochs = target_epochs;
  ch.bind_rate = (double)nlefs / (double)c->burnin_target_epochs_for_lef_activation;
  ch.qn = qn.data();
  ch.occupancy = g_occupancy;

  // the cell's arrays in a separate allocation when the caller asks (as the device-memory
  // placement of large cells does), else after the fixed part
  std::vector<uint64_t> lds(mgpu::det_wave_bytes(nlefs, b->n, P.hist_len) / 8 + 1, 0);
  std::vector<uint64_t> cell;
  job.lds = reinterpret_cast<uint8_t*>(lds.data());
  job.cell = nullptr;
  if (g_separate_cell) {
    lds.assign(mgpu::det_wave_fixed_bytes() / 8 + 1, 0);
    cell.assign(mgpu::det_wave_cell_bytes(nlefs, b->n, P.hist_len) / 8 + 1, 0);
    job.lds = reinterpret_cast<uint8_t*>(lds.data());
    job.cell = reinterpret_cast<uint8_t*>(cell.data());
  }
  job.seed = rc_cell_seed(chrom->name, chrom->name_len, chrom->size, cell_id, c->seed);
  job.target = target;
  job.trace = trace;
  job.max_trace = max_trace;

  Sched* s = new Sched();
  for (int l = 0; l < kLanes; ++l) {
    s->done[l] = false;
    s->syncs[l] = 0;
    s->stack[l].resize(kStack);
    getcontext(&s->ctx[l]);
    s->ctx[l].uc_stack.ss_sp = s->stack[l].data();
    s->ctx[l].uc_stack.ss_size = kStack;
    s->ctx[l].uc_link = nullptr;
    makecontext(&s->ctx[l], (void (*)())lane_entry, 1, l);
  }
  g_sched = s;
  g_job = &job;
  // each lane runs up to its first _setjmp and returns here; then lane 0 starts the round robin
  for (int l = 0; l < kLanes; ++l) swapcontext(&s->init_ctx, &s->ctx[l]);
  s->down = g_lane_order != 0;
  if (_setjmp(s->main_buf) == 0) _longjmp(s->buf[s->down ? kLanes - 1 : 0], 1);
  const bool broken = s->broken;
  g_last_syncs = s->syncs[0];
  delete s;
  g_sched = nullptr;
  g_job = nullptr;
  for (int k = 0; k < 4; ++k) stats[k] = job.stats[k];
  if (broken) {
    fprintf(stderr, "det_wave_emu: lanes out of step\n");
    return -1;
  }
  return job.stats[3] == ~0ull ? -1 : 0;
}

}  // extern "C"
