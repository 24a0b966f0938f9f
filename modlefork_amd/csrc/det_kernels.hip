// Deterministic-mode kernels: the whole-cell epoch loop behind a device work queue, and the
// step-level kernels behind the debug entry points. Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "det_sim.h"
#include "det_wave.h"
#include "kernels.h"

namespace mgpu {

// Each block (one wavefront) pulls (chromosome, cell) items from a device counter until the
// queue is empty, so uneven chromosomes and cells with different epoch counts balance across CUs.
// Only lane 0 runs the serial per-cell chain; the block's working set is carved from dynamic
// LDS when use_lds is set, else from its slab of global scratch.
__global__ void __launch_bounds__(64) det_cells_kernel(DetParams P, const DetChrom* chroms,
                                                       const WorkItem* items, uint64_t nitems,
                                                       unsigned long long* next_item,
                                                       uint8_t* scratch, uint64_t stride,
                                                       int use_lds, uint64_t* cell_stats) {
  extern __shared__ __align__(16) uint8_t lds[];
  if (threadIdx.x != 0) return;
  uint8_t* base = use_lds ? lds : scratch + (uint64_t)blockIdx.x * stride;
  for (;;) {
    const uint64_t it = atomicAdd(next_item, 1ULL);
    if (it >= nitems) return;
    const WorkItem w = items[it];
    const DetChrom ch = chroms[w.chrom];
    DetCell s;
    det_carve(s, base, ch.nlefs, ch.nbar, P.hist_len);
    det_simulate_cell(P, ch, s, w.seed, w.target, nullptr, 0, (uint32_t)it);
    if (cell_stats) {
      cell_stats[4 * it + 0] = s.epoch;
      cell_stats[4 * it + 1] = s.burnin_epochs;
      cell_stats[4 * it + 2] = s.contacts;
      cell_stats[4 * it + 3] = s.g.draws;
    }
  }
}

// Wave-parallel deterministic mode (det_wave.h): block b is one wavefront that simulates work
// item order[b] with all 64 lanes; the working set and the engine-word ring live in dynamic LDS,
// and the hardware dispatcher starts the next block wherever one finishes (the host orders the
// items by expected cost, largest first, so the longest cells do not start last). There is deliberately no
// persistent loop over a work queue: with one, the optimizer merged the lane-0 region that
// stores a finished cell's statistics with the next iteration's lane-0 queue fetch, and the
// compiled loop sent lanes 1-63 back into the cell without lane 0 (the item index they then
// broadcast from lane 0 was stale). A cell that fails an internal bound (an engine draw past
// the ring's reach, a contact loop that never completes) counts in *err.
__global__ void __launch_bounds__(64) det_wave_kernel(DetParams P, const DetChrom* chroms,
                                                      const WorkItem* items, const uint32_t* order,
                                                      uint64_t nitems, uint64_t* cell_stats,
                                                      unsigned long long* err) {
  extern __shared__ __align__(16) uint8_t lds[];
  const DevWave w;
  if (blockIdx.x >= nitems) return;
  const uint64_t it = order[blockIdx.x];
  const WorkItem wi = items[it];
  const DetChrom ch = chroms[wi.chrom];
  uint64_t st[4] = {0, 0, 0, 0};
  det_wave_cell<DevWave, kPlaceLds>(w, P, ch, lds, nullptr, P.hist_len, wi.seed, wi.target,
                                    nullptr, 0, st, (uint32_t)it);
  if (w.lane() == 0) {
    if (cell_stats)
      for (int k = 0; k < 4; ++k) cell_stats[4 * it + k] = st[k];
    if (st[3] == ~0ull) atomicAdd(err, 1ULL);
  }
}

// The same for cells whose arrays do not fit a CU's LDS: block b keeps its engine-word ring,
// wave scratch and ziggurat tables in LDS and the arrays of work item order[b] in
// slab + b * stride (device memory, one slot per block of the launch).
__global__ void __launch_bounds__(64) det_wave_slab_kernel(DetParams P, const DetChrom* chroms,
                                                           const WorkItem* items,
                                                           const uint32_t* order, uint64_t nitems,
                                                           uint8_t* slab, uint64_t stride,
                                                           uint64_t* cell_stats,
                                                           unsigned long long* err) {
  extern __shared__ __align__(16) uint8_t lds[];
  const DevWave w;
  if (blockIdx.x >= nitems) return;
  const uint64_t it = order[blockIdx.x];
  const WorkItem wi = items[it];
  const DetChrom ch = chroms[wi.chrom];
  uint64_t st[4] = {0, 0, 0, 0};
  det_wave_cell<DevWave, kPlaceMem>(w, P, ch, lds, slab + (uint64_t)blockIdx.x * stride, P.hist_len, wi.seed,
                wi.target, nullptr, 0, st, (uint32_t)it);
  if (w.lane() == 0) {
    if (cell_stats)
      for (int k = 0; k < 4; ++k) cell_stats[4 * it + k] = st[k];
    if (st[3] == ~0ull) atomicAdd(err, 1ULL);
  }
}

// One cell with its per-epoch state hashes; cell_mem: the cell's arrays in device memory, or null
// for all of it in LDS.
__global__ void __launch_bounds__(64) det_wave_trace_kernel(DetParams P, DetChrom ch,
                                                            uint64_t seed, uint64_t target,
                                                            uint64_t* trace, uint64_t max_trace,
                                                            uint64_t* out, uint8_t* cell_mem) {
  extern __shared__ __align__(16) uint8_t lds[];
  const DevWave w;
  uint64_t st[4] = {0, 0, 0, 0};
  det_wave_cell(w, P, ch, lds, cell_mem, P.hist_len, seed, target, trace, max_trace, st, 0);
  if (w.lane() == 0)
    for (int k = 0; k < 4; ++k) out[k] = st[k];
}

__global__ void det_trace_kernel(DetParams P, DetChrom ch, uint64_t seed, uint64_t target,
                                 uint8_t* scratch, uint64_t* trace, uint64_t max_trace,
                                 uint64_t* out_epochs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DetCell s;
  det_carve(s, scratch, ch.nlefs, ch.nbar, P.hist_len);
  det_simulate_cell(P, ch, s, seed, target, trace, max_trace);
  out_epochs[0] = s.epoch;
  out_epochs[1] = s.burnin_epochs;
  out_epochs[2] = s.contacts;
  out_epochs[3] = s.g.draws;
}

// One step of the epoch loop on a state held in global memory (debug / known-answer tests).
__global__ void det_step_kernel(DetParams P, DetChrom ch, int op, uint64_t n, uint64_t* rev,
                                uint64_t* fwd, uint64_t* ep, uint32_t* rr, uint32_t* fr,
                                uint32_t* tmp, uint64_t* rm, uint64_t* fm, uint64_t* rc,
                                uint64_t* fc, uint8_t* bst, uint64_t* rng) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DetCell s;
  s.rev = rev;
  s.fwd = fwd;
  s.ep = ep;
  s.rr = rr;
  s.fr = fr;
  s.tmp = tmp;
  s.rm = rm;
  s.fm = fm;
  s.rc = rc;
  s.fc = fc;
  s.bst = bst;
  s.n = s.nact = n;
  s.burnin_done = true;
  s.g.s0 = rng[0];
  s.g.s1 = rng[1];
  s.g.s2 = rng[2];
  s.g.s3 = rng[3];
  s.g.draws = 0;
  uint64_t n5 = 0, n3 = 0;
  switch (op) {
    case 0: det_adjust_clamp(ch, s, n, true); break;
    case 1: det_process_collisions(P, ch, s, n, false); break;
    case 2: det_process_collisions(P, ch, s, n, true); break;
    case 3:
      det_primary(P, ch, s, n, 0, 0);
      det_correct_primary(s, n);
      det_secondary(P, s, n, 0, 0);
      break;
    case 4: det_fix_secondary(ch, s, n, 0, 0); break;
    case 5: det_lef_bar(P, ch, s, n, 0, 0); break;
    case 6: det_primary(P, ch, s, n, 0, 0); break;
    case 7: det_correct_lef_bar(ch, s, n); break;
    case 8: det_boundaries(ch, s, n, n5, n3); break;
    case 9:
      for (uint64_t i = 0; i < n; ++i) rr[i] = fr[i] = (uint32_t)i;
      det_rank(s, n, false);
      break;
    case 10: det_rank(s, n, true); break;
    case 11: det_generate_moves(P, ch, s, n); break;
    default: break;
  }
  rng[0] = s.g.s0;
  rng[1] = s.g.s1;
  rng[2] = s.g.s2;
  rng[3] = s.g.s3;
}

__global__ void det_draws_kernel(uint64_t seed, int kind, double a, double b, double c, uint64_t n,
                                 uint64_t* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  DetRng g;
  det_seed(g, seed);
  for (uint64_t i = 0; i < n; ++i) {
    double d = 0.0;
    uint64_t v = 0;
    switch (kind) {
      case 0: v = det_next(g); break;
      case 1: d = det_normal(g, a, b); v = __double_as_longlong(d); break;
      case 2: v = det_bernoulli(g, a) ? 1 : 0; break;
      case 3: d = det_canonical(g); v = __double_as_longlong(d); break;
      case 4: v = det_uniform(g, (uint64_t)a, (uint64_t)b); break;
      case 5: v = det_poisson(g, a); break;
      case 6: v = (uint64_t)det_binomial(g, (int64_t)a, b, -1.0); break;
      case 7: d = det_gev(g, a, b, c); v = __double_as_longlong(d); break;
      case 8: d = det_unit_exp(g) / a; v = __double_as_longlong(d); break;
      default: break;
    }
    out[i] = v;
  }
}

// ---- host-side launchers ----
hipError_t launch_det_cells(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                            const WorkItem* d_items, uint64_t nitems, unsigned long long* d_next,
                            uint8_t* d_scratch, uint64_t stride, bool use_lds, uint32_t grid,
                            uint64_t* d_cell_stats) {
  const size_t shmem = use_lds ? stride : 0;
  if (use_lds && shmem > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(det_cells_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(det_cells_kernel, dim3(grid), dim3(64), shmem, stream, P, d_chroms, d_items,
                     nitems, d_next, d_scratch, stride, use_lds ? 1 : 0, d_cell_stats);
  return hipGetLastError();
}

hipError_t launch_det_wave(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                           const WorkItem* d_items, const uint32_t* d_order, uint64_t nitems,
                           uint64_t lds_bytes, uint64_t* d_cell_stats, unsigned long long* d_err) {
  if (nitems == 0) return hipSuccess;
  if (nitems > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(det_wave_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(det_wave_kernel, dim3((uint32_t)nitems), dim3(64), lds_bytes, stream, P,
                     d_chroms, d_items, d_order, nitems, d_cell_stats, d_err);
  return hipGetLastError();
}

hipError_t launch_det_wave_slab(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                                const WorkItem* d_items, const uint32_t* d_order, uint64_t nitems,
                                uint8_t* d_slab, uint64_t stride, uint64_t* d_cell_stats,
                                unsigned long long* d_err) {
  if (nitems == 0) return hipSuccess;
  if (nitems > 0x7FFFFFFFull) return hipErrorInvalidValue;
  const uint64_t lds_bytes = det_wave_fixed_bytes();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(det_wave_slab_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(det_wave_slab_kernel, dim3((uint32_t)nitems), dim3(64), lds_bytes, stream,
                     P, d_chroms, d_items, d_order, nitems, d_slab, stride, d_cell_stats, d_err);
  return hipGetLastError();
}

hipError_t launch_det_wave_trace(hipStream_t stream, const DetParams& P, const DetChrom& ch,
                                 uint64_t seed, uint64_t target, uint64_t lds_bytes,
                                 uint8_t* d_cell, uint64_t* d_trace, uint64_t max_trace,
                                 uint64_t* d_out) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(det_wave_trace_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(det_wave_trace_kernel, dim3(1), dim3(64), lds_bytes, stream, P, ch, seed,
                     target, d_trace, max_trace, d_out, d_cell);
  return hipGetLastError();
}

uint64_t det_wave_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  return det_wave_bytes(nlefs, nbar, hist);
}
uint64_t det_wave_fixed_bytes_host() { return det_wave_fixed_bytes(); }
uint64_t det_wave_cell_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  return det_wave_cell_bytes(nlefs, nbar, hist);
}

__device__ unsigned long long g_det_wave_phase[kDetWavePhases];

unsigned long long* det_wave_phase_ptr() {
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_det_wave_phase)) != hipSuccess) return nullptr;
  return static_cast<unsigned long long*>(p);
}

hipError_t det_wave_phase_times(uint64_t* out, int n) {
  unsigned long long h[kDetWavePhases] = {};
  hipError_t e = hipMemcpyFromSymbol(h, HIP_SYMBOL(g_det_wave_phase), sizeof(h), 0,
                                     hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (int k = 0; k < n && k < kDetWavePhases; ++k) out[k] = h[k];
  const unsigned long long z[kDetWavePhases] = {};
  return hipMemcpyToSymbol(HIP_SYMBOL(g_det_wave_phase), z, sizeof(z), 0, hipMemcpyHostToDevice);
}

hipError_t launch_det_trace(hipStream_t stream, const DetParams& P, const DetChrom& ch,
                            uint64_t seed, uint64_t target, uint8_t* d_scratch, uint64_t* d_trace,
                            uint64_t max_trace, uint64_t* d_out) {
  hipLaunchKernelGGL(det_trace_kernel, dim3(1), dim3(64), 0, stream, P, ch, seed, target,
                     d_scratch, d_trace, max_trace, d_out);
  return hipGetLastError();
}

hipError_t launch_det_step(hipStream_t stream, const DetParams& P, const DetChrom& ch, int op,
                           uint64_t n, uint64_t* rev, uint64_t* fwd, uint64_t* ep, uint32_t* rr,
                           uint32_t* fr, uint32_t* tmp, uint64_t* rm, uint64_t* fm, uint64_t* rc,
                           uint64_t* fc, uint8_t* bst, uint64_t* rng) {
  hipLaunchKernelGGL(det_step_kernel, dim3(1), dim3(64), 0, stream, P, ch, op, n, rev, fwd, ep, rr,
                     fr, tmp, rm, fm, rc, fc, bst, rng);
  return hipGetLastError();
}

hipError_t launch_det_draws(hipStream_t stream, uint64_t seed, int kind, double a, double b,
                            double c, uint64_t n, uint64_t* d_out) {
  hipLaunchKernelGGL(det_draws_kernel, dim3(1), dim3(64), 0, stream, seed, kind, a, b, c, n, d_out);
  return hipGetLastError();
}

// ContactMatrixDense<u32>::increment as deterministic mode applies it, for (rows[i], cols[i]) in
// order, on one lane.
__global__ void det_band_kernel(DetChrom ch, const uint64_t* rows, const uint64_t* cols,
                                uint64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint64_t i = 0; i < n; ++i) band_add(ch, rows[i], cols[i]);
}

// Collision words of deterministic mode (the reference's 64-bit layout): encoded word, decoded
// index and event, and flags: bit 0 COLLISION set, bits 1-4 collision occurred with CHROM,
// LEF_BAR, PRIMARY, SECONDARY, bits 5-8 the same events avoided.
__global__ void det_collision_kernel(const uint64_t* idx, const uint8_t* ev, uint64_t n,
                                     uint64_t* words, uint64_t* index, uint32_t* event,
                                     uint32_t* flags) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t w = cenc(idx[i], ev[i]);
  const unsigned kinds[4] = {kEvChrom, kEvLefBar, kEvPrimary, kEvSecondary};
  uint32_t f = c_any(w) ? 1u : 0u;
  for (int k = 0; k < 4; ++k) {
    f |= (c_is(w, kinds[k]) ? 1u : 0u) << (1 + k);
    f |= (c_avoided(w, kinds[k]) ? 1u : 0u) << (5 + k);
  }
  words[i] = w;
  index[i] = cidx(w);
  event[i] = cev(w);
  flags[i] = f;
}

// ContactMatrixDense<u32>::unsafe_get_block (contact_matrix_dense_unsafe_impl.hpp:89-113) for
// query i on band q.chrom[i]: the block_size x block_size pixels centred on (row, col), each
// coordinate clamped to [0, ncols - 1], each pixel read as unsafe_get does (transposed to the
// upper triangle, 0 beyond the band), summed in u32. One lane per query.
__global__ void block_sums_kernel(const BlockBand* bands, const uint64_t* chrom,
                                  const uint64_t* row, const uint64_t* col, uint64_t n,
                                  uint32_t block_size, uint32_t* out) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const BlockBand b = bands[chrom[q]];
  if (block_size == 1) {  // unsafe_get(row, col) itself
    const uint64_t r = row[q], c = col[q];
    const uint64_t i = r > c ? r - c : c - r, j = r > c ? r : c;
    out[q] = i < b.nrows ? b.band[j * b.nrows + i] : 0u;
    return;
  }
  const int64_t half = (int64_t)(block_size / 2);
  const int64_t r0 = (int64_t)row[q] - half, c0 = (int64_t)col[q] - half;
  const int64_t last = (int64_t)b.ncols - 1;
  uint32_t sum = 0;
  for (int64_t r = r0; r < r0 + (int64_t)block_size; ++r) {
    const uint64_t rr = (uint64_t)(r < 0 ? 0 : (r > last ? last : r));
    for (int64_t c = c0; c < c0 + (int64_t)block_size; ++c) {
      const uint64_t cc = (uint64_t)(c < 0 ? 0 : (c > last ? last : c));
      const uint64_t i = rr > cc ? rr - cc : cc - rr, j = rr > cc ? rr : cc;
      if (i < b.nrows) sum += b.band[j * b.nrows + i];
    }
  }
  out[q] = sum;
}

hipError_t launch_block_sums(hipStream_t stream, const BlockBand* d_bands, const uint64_t* d_chrom,
                             const uint64_t* d_row, const uint64_t* d_col, uint64_t n,
                             uint32_t block_size, uint32_t* d_out) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(block_sums_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                     d_bands, d_chrom, d_row, d_col, n, block_size, d_out);
  return hipGetLastError();
}

hipError_t launch_det_band(hipStream_t stream, const DetChrom& ch, const uint64_t* d_rows,
                           const uint64_t* d_cols, uint64_t n) {
  hipLaunchKernelGGL(det_band_kernel, dim3(1), dim3(64), 0, stream, ch, d_rows, d_cols, n);
  return hipGetLastError();
}

hipError_t launch_det_collisions(hipStream_t stream, const uint64_t* d_idx, const uint8_t* d_ev,
                                 uint64_t n, uint64_t* d_words, uint64_t* d_index,
                                 uint32_t* d_event, uint32_t* d_flags) {
  hipLaunchKernelGGL(det_collision_kernel, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, stream,
                     d_idx, d_ev, n, d_words, d_index, d_event, d_flags);
  return hipGetLastError();
}

uint64_t det_cell_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  uint64_t b = nlefs * (7 * 8 + 3 * 4) + nbar + 2 * (hist + 1) * 8;
  return (b + 255) & ~255ULL;
}

}  // namespace mgpu
