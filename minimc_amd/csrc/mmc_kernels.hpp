// Event kernels of the tracker: source, advance (cross-section lookup +
// distance to collision + CSG distance to surface + stream), surface
// crossing, collision, thermal scattering; plus tally and bank kernels.
//
// Particle state lives in structure-of-arrays banks in HBM and stays in its
// slot; each event kernel walks a queue of slot indices.  Queues are filled
// with wave-wide ballots: every wave counts its pushes with one ballot and
// reserves its range with one atomic.  Loops are block-uniform so all 64
// lanes of a wave reach every ballot.
#pragma once
#include "mmc_device.hpp"
#include "../../include/minimc_amd.h"

namespace mmc {

constexpr int kRing = 16;        // iteration counter ring
constexpr int kRingStride = 8;   // counters per iteration
// kQ counts the live queue Q, kQ2 the collision survivors' live queue Q2
// (the next iteration's live particles are Q then Q2), kC the collision
// queue and kT the thermal-scattering queue.
enum RingSlot : int { kQ = 0, kC = 1, kT = 2, kQ2 = 3 };
constexpr int kMiscFallback = 5;
// event counters (u64)
enum Ev : int {
  kEvAdvance = 0, kEvCollide, kEvCross, kEvVirtual, kEvTsl, kEvFallback,
  kEvSites, kEvTail,
  // the tail kernel's own events (also in the totals above), and the thermal
  // collisions the network sampled
  kEvTailAdv, kEvTailCol, kEvTailTsl, kEvTailCross, kEvMlp, kEvN
};
// The event counters are held kEvSlots times, kEvStride apart (128 B), and a
// block adds to copy blockIdx.x % kEvSlots: the waves of a launch flush their
// counts at the same moment, and atomics on one address serialise.  The host
// sums the copies.
constexpr int kEvSlots = 64, kEvStride = 16, kEvWords = kEvSlots * kEvStride;
static_assert(kEvN <= kEvStride, "event counters per copy");

struct DevBank {
  double *x, *y, *z, *u, *v, *w, *E;
  uint32_t* rng;
  int32_t* cell;
  uint32_t* hist;
  uint32_t* steps;
  double *k_col, *k_trk, *k_abs;  // k-eigenvalue sums, indexed by history
  double* tslT;   // temperature of a deferred thermal-scattering collision
  int32_t* tslI;  // its thermal-scattering table
  int32_t* xs_grid;  // last energy-grid search of the particle: grid (-1: none)
  int32_t* xs_gi;    // and upper-bound index (a hint, verified before use)
  // indirect effect of each perturbation, [perturbation][slot]
  // (Particle::indirect_effects, Particle.hpp:105-106); null without
  // perturbations
  double* ie;
};

struct RunState {
  DevBank B;
  uint32_t* Q[2];
  uint32_t* Q2[2];  // live queue of the collision survivors (after Q)
  uint32_t *QC, *QT;
  uint32_t* ring;      // kRing * kRingStride
  uint32_t* misc;      // [0] err bits, [1] records, [2] sites, [3] slots,
                       // [kMiscFallback] parked surrogate fallbacks
  unsigned long long* ev;
  uint32_t capacity;   // slots
  int keig;
  uint32_t max_steps;
  uint32_t chunk_first;  // first history of the chunk (fixed source)
  // tallies: one record per score, key = bin << 32 | history (relative to
  // chunk_first), reduced in a fixed order when the batch ends
  uint64_t* rec_key;
  double* rec_val;
  uint32_t rec_cap;
  // fission sites (k-eig)
  mmc_site* sites;
  uint32_t site_cap;
  uint32_t* fis_cnt;   // children per history (k-eig)
  int tsl_surrogate;
  float* mlp_in;       // [cap * 4] surrogate inputs of the TSL queue
  float* mlp_out;      // [cap * 2]
};

__device__ __forceinline__ unsigned long long* Ev(const RunState& R, int k) {
  return R.ev + (blockIdx.x % kEvSlots) * kEvStride + k;
}
__device__ __forceinline__ uint32_t WavePush(uint32_t* counter, bool pred,
                                             uint32_t& total_out) {
  const uint64_t mask = __ballot(pred);
  const uint32_t total = __popcll(mask);
  total_out = total;
  if (total == 0) return 0;
  const uint32_t lane = __lane_id();
  const uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(counter, total);
  base = __shfl(base, 0);
  return base + prefix;
}
// A push of the whole block with one atomic: the waves' counts are summed in
// LDS (at most 16 waves).  Every thread of the block must call it, once.
__device__ __forceinline__ uint32_t BlockPush(uint32_t* counter, bool pred) {
  __shared__ uint32_t wave_base[17];
  const uint64_t mask = __ballot(pred);
  const uint32_t lane = __lane_id();
  const int wave = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
  if (lane == 0) wave_base[wave] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (int w = 0; w < nw; w++) {
      const uint32_t c = wave_base[w];
      wave_base[w] = total;
      total += c;
    }
    wave_base[16] = total ? atomicAdd(counter, total) : 0u;
  }
  __syncthreads();
  return wave_base[16] + wave_base[wave] + __popcll(mask & ((1ull << lane) - 1ull));
}
// Two pushes of one wave with one atomic: counter pair c[0] (pred0) and
// c[1] (pred1), adjacent u32s at an 8-byte aligned address, updated as one
// u64 (low word c[0]; no count comes near 2^32, so no carry crosses).
__device__ __forceinline__ void WavePush2(uint32_t* c, bool pred0, bool pred1, uint32_t& i0,
                                          uint32_t& i1) {
  const uint64_t m0 = __ballot(pred0), m1 = __ballot(pred1);
  const uint32_t lane = __lane_id();
  const uint64_t below = (1ull << lane) - 1ull;
  uint64_t base = 0;
  if (lane == 0 && (m0 | m1))
    base = atomicAdd(reinterpret_cast<unsigned long long*>(c),
                     (uint64_t(__popcll(m1)) << 32) | uint64_t(__popcll(m0)));
  const uint32_t b0 = __shfl(uint32_t(base), 0), b1 = __shfl(uint32_t(base >> 32), 0);
  i0 = b0 + __popcll(m0 & below);
  i1 = b1 + __popcll(m1 & below);
}
static_assert(kQ % 2 == 0 && kC == kQ + 1 && kRingStride % 2 == 0, "kQ, kC: one u64");
// WavePush2 for the whole block: one atomic per block and call.  Every
// thread of the block must call it (in a grid-stride loop, the pass count is
// the same for the whole block).  Between two calls, a wave writes only its
// own count, and thread 0 rewrites the shared entries only after every wave
// has passed the first barrier, i.e. has read the previous call's bases.
__device__ __forceinline__ void BlockPush2(uint32_t* c, bool pred0, bool pred1, uint32_t& i0,
                                           uint32_t& i1) {
  __shared__ uint32_t n0[17], n1[17];  // per-wave counts, then offsets; [16]: block base
  const uint64_t m0 = __ballot(pred0), m1 = __ballot(pred1);
  const uint32_t lane = __lane_id();
  const int wave = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
  if (lane == 0) {
    n0[wave] = __popcll(m0);
    n1[wave] = __popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t0 = 0, t1 = 0;
    for (int w = 0; w < nw; w++) {
      const uint32_t a = n0[w], b = n1[w];
      n0[w] = t0;
      n1[w] = t1;
      t0 += a;
      t1 += b;
    }
    uint64_t base = 0;
    if (t0 | t1)
      base = atomicAdd(reinterpret_cast<unsigned long long*>(c), (uint64_t(t1) << 32) | t0);
    n0[16] = uint32_t(base);
    n1[16] = uint32_t(base >> 32);
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  i0 = n0[16] + n0[wave] + __popcll(m0 & below);
  i1 = n1[16] + n1[wave] + __popcll(m1 & below);
}
static_assert(kT % 2 == 0 && kQ2 == kT + 1, "kT, kQ2: one u64");
// A wave's event count over all its grid-stride passes, added to the device
// counter once when the wave is done: one atomic per wave and counter, not
// one per pass (the counters are shared by every wave of the launch).
struct WaveTally {
  unsigned long long n = 0;
  __device__ __forceinline__ void add(bool pred) { n += __popcll(__ballot(pred)); }
  __device__ __forceinline__ void flush(unsigned long long* c) const {
    if (__lane_id() == 0 && n) atomicAdd(c, n);
  }
};
__host__ __device__ __forceinline__ uint32_t* Ring(const RunState& R, int iter, int slot) {
  return R.ring + ((iter + kRing) % kRing) * kRingStride + slot;
}
// Entry q of the live queue made by iteration `iter` (Q[qin] of length
// n1 = kQ, then Q2[qin]), and its length.
__device__ __forceinline__ uint32_t LiveLength(const RunState& R, int iter, uint32_t& n1) {
  n1 = *Ring(R, iter, kQ);
  return n1 + *Ring(R, iter, kQ2);
}
__device__ __forceinline__ uint32_t& LiveAt(const RunState& R, int qin, uint32_t n1,
                                            uint32_t q) {
  return q < n1 ? R.Q[qin][q] : R.Q2[qin][q - n1];
}

__device__ __forceinline__ V3 LoadPos(const DevBank& B, uint32_t p) {
  return {B.x[p], B.y[p], B.z[p]};
}
__device__ __forceinline__ V3 LoadDir(const DevBank& B, uint32_t p) {
  return {B.u[p], B.v[p], B.w[p]};
}
// A new particle's indirect effects start at zero: the reference calls
// Particle::SetPerturbations (Particle.cpp:71-73) on every particle it takes
// from the bank, source particles and secondaries alike (FixedSource.cpp:
// 63-65), and PerturbationSet::GetIndirectEffects gives each perturbation
// the value 0 (Perturbation.cpp:122-128).
__device__ __forceinline__ void ClearIndirect(const DevWorld& W, const RunState& R,
                                              uint32_t slot) {
  for (int j = 0; j < W.n_pert; j++) R.B.ie[size_t(j) * R.capacity + slot] = 0;
}

// Score one value into a global tally bin: a record (bin, history, value).
// The batch's records are reduced in a fixed order when it ends (per-history
// sums, then per-bin sums: TallyReduce), so the tallies do not depend on the
// order in which the GPU produced the scores.  misc[1] counts the records
// asked for, also past the capacity, so that the caller can size a rerun.
__device__ __forceinline__ void ScoreBin(const RunState& R, uint32_t hist, int bin,
                                         double val) {
  const uint32_t k = atomicAdd(R.misc + 1, 1u);
  if (k >= R.rec_cap) {
    atomicOr(R.misc, kErrRecords);
    return;
  }
  R.rec_key[k] = (uint64_t(uint32_t(bin)) << 32) | uint64_t(hist - R.chunk_first);
  R.rec_val[k] = val;
}

// CurrentEstimator::GetScore + ParticleBins::GetIndex
// (Estimator.cpp:142-151, Bins.cpp:196-204)
// A sensitivity estimator (sens = j + 1) scores the particle's indirect
// effect of perturbation j times the current's score of 1
// (CurrentTotalCrossSectionSensitivity::GetScore, Sensitivity.cpp:55-58).
__device__ __forceinline__ void ScoreCurrent(const DevWorld& W, const RunState& R, uint32_t p,
                                             uint32_t hist, int surf, V3 d, double E) {
  for (int e = 0; e < W.n_est; e++) {
    const DEst& est = W.ests[e];
    if (est.kind != 0 || est.surface != surf) continue;
    const int ci = est.has_dir ? BinIndex(W, est.cosine, Dot(V3{est.dx, est.dy, est.dz}, d)) : 0;
    const int ei = BinIndex(W, est.energy, E);
    const double v = est.sens ? R.B.ie[size_t(est.sens - 1) * R.capacity + p] * 1.0 : 1.0;
    ScoreBin(R, hist, est.bin_off + est.energy.n_bins * ci + ei, v);
  }
}
__device__ __forceinline__ void ScoreFlux(const DevWorld& W, const RunState& R,
                                          uint32_t hist, V3 p, double sigma_macro) {
  for (int e = 0; e < W.n_est; e++) {
    const DEst& est = W.ests[e];
    if (est.kind != 1) continue;
    const double q[3] = {p.x, p.y, p.z};
    int idx = 0;
    bool in = true;
    for (int a = 0; a < 3; a++) {
      const double f = (q[a] - est.mesh_lo[a]) / (est.mesh_hi[a] - est.mesh_lo[a]);
      if (!(f >= 0 && f < 1)) {
        in = false;
        break;
      }
      int i = int(f * est.mesh_n[a]);
      if (i >= est.mesh_n[a]) i = est.mesh_n[a] - 1;
      idx = idx * est.mesh_n[a] + i;
    }
    if (in) ScoreBin(R, hist, est.bin_off + idx, 1.0 / sigma_macro);
  }
}
__device__ __forceinline__ void ScoreAlphaBeta(const DevWorld& W, const RunState& R,
                                               uint32_t hist, double a, double b) {
  for (int e = 0; e < W.n_est; e++) {
    const DEst& est = W.ests[e];
    if (est.kind == 2) ScoreBin(R, hist, est.bin_off + BinIndex(W, est.energy, a), 1.0);
    if (est.kind == 3) ScoreBin(R, hist, est.bin_off + BinIndex(W, est.energy, b), 1.0);
  }
}

#ifdef MMC_TRACKER_KERNELS
// ---- source ---------------------------------------------------------------
// Source::Sample (Source.cpp:143-154) for history first + i, seed seed0 + hist.
__global__ void KSourceFixed(DevWorld W, RunState R, mmc_source_desc s, uint64_t seed0,
                             uint32_t first, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) R.misc[3] = n;
  const bool valid = i < n;
  bool ok = false;
  if (valid) {
  const uint32_t hist = first + i;
  uint32_t g = LcgSeed(seed0 + hist);
  double q[3] = {s.x, s.y, s.z};
  for (int a = 0; a < 3; a++)
    if (s.box_hi[a] > s.box_lo[a]) q[a] = Uniform(g, s.box_lo[a], s.box_hi[a]);
  V3 d;
  if (s.dir_kind == 1) {
    d = IsotropicDirection(g);
  } else if (s.dir_kind == 2) {
    const double mu = sqrt(Canonical(g));
    const double phi = Uniform(g, 0., 2 * kPi);
    d = Rotate(Normalized(V3{s.u, s.v, s.w}), mu, phi);
  } else {
    d = Normalized(V3{s.u, s.v, s.w});
  }
  const double E = s.energy_kind == 1 ? Uniform(g, s.E, s.E_hi) : s.E;
  const uint32_t seed = LcgSeed(LcgNext(g));
  const V3 p{q[0], q[1], q[2]};
  const int cell = FindCell(W, p);
  if (cell < 0) atomicOr(R.misc, kErrLost);
  R.B.x[i] = p.x; R.B.y[i] = p.y; R.B.z[i] = p.z;
  R.B.u[i] = d.x; R.B.v[i] = d.y; R.B.w[i] = d.z;
  R.B.E[i] = E;
  R.B.rng[i] = seed;
  R.B.cell[i] = cell;
  R.B.hist[i] = hist;
  R.B.steps[i] = 0;
  R.B.xs_grid[i] = -1;
  ClearIndirect(W, R, i);
  ok = cell >= 0 && W.cells[cell].material >= 0;  // born in void: leaks at once
  }
  const uint32_t iq = BlockPush(Ring(R, -1, kQ), ok);
  if (ok) R.Q[0][iq] = i;
}

__device__ __forceinline__ DevWorld DescInLds(const DevWorld& W, unsigned char* lds);

// k-eigenvalue: source site j -> particle slot j.  kLds: the cell search
// reads the World's descriptors from an LDS copy.
template <bool kLds>
__global__ void KSourceSites(DevWorld Wg, RunState R, const mmc_site* src, uint32_t n) {
  extern __shared__ uint4 lds_desc[];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) R.misc[3] = n;
  if (blockIdx.x * blockDim.x >= n) return;  // whole blocks only: the staging syncs
  const DevWorld W = kLds ? DescInLds(Wg, reinterpret_cast<unsigned char*>(lds_desc)) : Wg;
  const bool valid = i < n;
  bool ok = false;
  if (valid) {
  const mmc_site s = src[i];
  const V3 p{s.x, s.y, s.z};
  const int cell = FindCell(W, p);
  if (cell < 0) atomicOr(R.misc, kErrLost);
  R.B.x[i] = s.x; R.B.y[i] = s.y; R.B.z[i] = s.z;
  R.B.u[i] = s.u; R.B.v[i] = s.v; R.B.w[i] = s.w;
  R.B.E[i] = s.E;
  R.B.rng[i] = LcgSeed(s.seed);
  R.B.cell[i] = cell;
  R.B.hist[i] = i;
  R.B.steps[i] = 0;
  R.B.xs_grid[i] = -1;
  ClearIndirect(W, R, i);
  R.B.k_col[i] = 0;
  R.B.k_trk[i] = 0;
  R.B.k_abs[i] = 0;
  R.fis_cnt[i] = 0;
  ok = cell >= 0 && W.cells[cell].material >= 0;
  }
  const uint32_t iq = BlockPush(Ring(R, -1, kQ), ok);
  if (ok) R.Q[0][iq] = i;
}

#endif  // MMC_TRACKER_KERNELS

// LDS layout of the kernels that stage tables: the first grid's bucket hash
// (n_hash + 1 ints, padded to an even count) when there is one, then the
// thermal-scattering totals (tot_n doubles) and/or the sampler's tables
// (hot_n doubles), as the kernel needs them.
__device__ __forceinline__ int HashLdsInts(const DevWorld& W) {
  return (W.energy_mode == 1 && W.grids[0].n_hash > 0) ? ((W.grids[0].n_hash + 2) & ~1) : 0;
}

// Copies the World's descriptor block into LDS (block-wide; every thread of
// the block must call it) and returns the World with every descriptor array
// read from there.  The geometry, material and table descriptors sit on every
// event's chain of dependent loads (cell -> faces -> surfaces, material ->
// components -> nuclide -> grid, table, reaction); in LDS each link costs an
// LDS read instead of a cache round trip.
__device__ __forceinline__ DevWorld DescInLds(const DevWorld& W, unsigned char* lds) {
  // 8-byte words: lds is 8-byte aligned (it follows ints padded to an even
  // count and doubles)
  const uint2* src = reinterpret_cast<const uint2*>(W.desc);
  uint2* dst = reinterpret_cast<uint2*>(lds);
  for (int i = threadIdx.x; i < W.desc_bytes / 8; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  DevWorld L = W;
  L.surfaces = reinterpret_cast<const DSurface*>(lds + W.desc_off[kDescSurfaces]);
  L.cells = reinterpret_cast<const DCell*>(lds + W.desc_off[kDescCells]);
  L.faces = reinterpret_cast<const DFace*>(lds + W.desc_off[kDescFaces]);
  L.fields = reinterpret_cast<const DField*>(lds + W.desc_off[kDescFields]);
  L.mats = reinterpret_cast<const DMaterial*>(lds + W.desc_off[kDescMats]);
  L.comps = reinterpret_cast<const DComp*>(lds + W.desc_off[kDescComps]);
  L.ce = reinterpret_cast<const DCeNuclide*>(lds + W.desc_off[kDescCe]);
  L.rx = reinterpret_cast<const DCeReaction*>(lds + W.desc_off[kDescRx]);
  L.tables = reinterpret_cast<const DTable*>(lds + W.desc_off[kDescTables]);
  L.grids = reinterpret_cast<const DGrid*>(lds + W.desc_off[kDescGrids]);
  L.tsl = reinterpret_cast<const DTsl*>(lds + W.desc_off[kDescTsl]);
  L.parts = reinterpret_cast<const DTslPart*>(lds + W.desc_off[kDescParts]);
  L.ests = reinterpret_cast<const DEst*>(lds + W.desc_off[kDescEsts]);
  L.loc_off = reinterpret_cast<const int*>(lds + W.desc_off[kDescLocOff]);
  L.loc_cells = reinterpret_cast<const int*>(lds + W.desc_off[kDescLocCells]);
  return L;
}

// Copies src[0, n) into LDS (block-wide; every thread of the block must call
// it).
__device__ __forceinline__ const double* StageHot(const double* src, double* lds, int n) {
  constexpr int kU = 8;
  for (int k0 = threadIdx.x; k0 < n; k0 += kU * blockDim.x) {
    double v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * blockDim.x;
      v[u] = k < n ? src[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * blockDim.x;
      if (k < n) lds[k] = v[u];
    }
  }
  __syncthreads();
  return lds;
}

// Copies the first grid's bucket hash into LDS (block-wide; every thread of
// the block must call it).  Returns the pointer lookups use, or nullptr when
// there is no hash to stage.
__device__ __forceinline__ const int* StageHash(const DevWorld& W, int* lds_hash) {
  if (W.energy_mode != 1 || W.grids[0].n_hash <= 0) return nullptr;
  const DGrid g0 = W.grids[0];
  const int* src = W.ipool + g0.hash_off;
  // several loads in flight per thread before the LDS stores
  constexpr int kU = 8;
  for (int k0 = threadIdx.x; k0 <= g0.n_hash; k0 += kU * blockDim.x) {
    int v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * blockDim.x;
      v[u] = k <= g0.n_hash ? src[k] : 0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int k = k0 + u * blockDim.x;
      if (k <= g0.n_hash) lds_hash[k] = v[u];
    }
  }
  __syncthreads();
  return lds_hash;
}

// Copies particle p of bank src to slot i of bank dst (bank compaction).
__device__ __forceinline__ void MoveParticle(const DevBank& src, uint32_t p, const DevBank& dst,
                                             uint32_t i) {
  dst.x[i] = src.x[p]; dst.y[i] = src.y[p]; dst.z[i] = src.z[p];
  dst.u[i] = src.u[p]; dst.v[i] = src.v[p]; dst.w[i] = src.w[p];
  dst.E[i] = src.E[p];
  dst.rng[i] = src.rng[p];
  dst.cell[i] = src.cell[p];
  dst.hist[i] = src.hist[p];
  dst.steps[i] = src.steps[p];
  dst.xs_grid[i] = src.xs_grid[p];
  dst.xs_gi[i] = src.xs_gi[p];
  dst.tslT[i] = src.tslT[p];
  dst.tslI[i] = src.tslI[p];
}
__device__ __forceinline__ void MoveIndirect(int n_pert, uint32_t cap, const DevBank& src,
                                             uint32_t p, const DevBank& dst, uint32_t i) {
  for (int j = 0; j < n_pert; j++) dst.ie[size_t(j) * cap + i] = src.ie[size_t(j) * cap + p];
}
// ---- surface crossing (TransportMethod.cpp:67-73, :98-105) ----------------
// Particle p, streamed to `pos` across surface `surf` out of cell `from`,
// enters the cell on the other side (`cell`, stored in the bank unless kMove:
// the caller stores it elsewhere) and scores the current estimators of the
// surface.  Returns whether it is still alive (surface_cross vs leak).
template <bool kLocate, bool kMove>
__device__ __forceinline__ bool CrossAt(const DevWorld& W, const RunState& R, uint32_t p,
                                        V3 pos, int surf, int from, V3 dir, double E,
                                        int& cell) {
  cell = (W.quirks & MMC_QUIRK_NUDGE_CROSSING)
             ? FindCell(W, pos)
             : FindCellAcross<kLocate>(W, pos, surf, 1 - FaceInside(W, from, surf));
  if (cell < 0) {
    atomicOr(R.misc, kErrLost);
    return false;
  }
  if (!kMove) R.B.cell[p] = cell;
  if (W.n_est_current) ScoreCurrent(W, R, p, R.B.hist[p], surf, dir, E);
  return W.cells[cell].material >= 0;
}

// ---- advance: XS lookup + distance to collision + CSG distance ------------
// SurfaceTracking/CellDeltaTracking::Transport loop body
// (TransportMethod.cpp:57-73, :93-118).  Next event of particle p: a surface
// crossing (to_x), done here at once from the registers that hold the
// particle (x_alive: it entered a cell with a material), a real collision
// (to_c) or, with delta tracking, a virtual collision (to_q: back to the
// live queue).
struct AdvanceOut {
  bool to_x, to_c, to_q, x_alive;
};
// The particle as a flight that ends in a collision leaves it: the values
// the advance writes to the bank, for a collision in the same kernel to take
// from registers instead of reading them back.
struct FlightEnd {
  V3 pos, dir;
  double E;
  uint32_t g;
  int cell;
  int grid, gi;  // the grid-index hint
  // the advance's last lookups at (E, T) that were not majorants: the
  // thermal-scattering total of table tsl (or tsl -1) and nuclide nuc's
  // total (or nuc -1); a collision at the same (E, T) takes them
  int tsl, nuc;
  double T, tsl_val, tot;
  double nuf;  // the material's nu Sigma_f at E (k-eigenvalue runs), or -1
};
// Particle::Stream -> TotalCrossSectionPerturbation::Stream
// (Particle.cpp:46-53, Perturbation.cpp:68-78): 1 / sigma - distance in a
// material that holds the perturbed nuclide, with sigma the collision
// probability density where the particle stands (the microscopic total, or
// the majorant with delta tracking: TransportMethod.cpp:79-82, :124-127).
// Kept out of line: problems without perturbations never call it, and their
// event code keeps its registers.
__device__ __noinline__ void StreamIndirect(const DevWorld& W, const RunState& R, uint32_t p,
                                            int mask, double sigma, double d) {
  const double inc = 1 / sigma - d;
  for (int j = 0; j < W.n_pert; j++)
    if (mask >> j & 1) R.B.ie[size_t(j) * R.capacity + p] += inc;
}

// hash: the energy hash (LDS) or nullptr; Ht: the thermal-scattering totals
// (DevWorld::tot or its LDS copy); kLocate: the crossing's cell search uses
// the cell locator (FindCellAcross).  kMove: a particle that lives on is
// written to slot i of bank D (every field a later event reads; the pending
// thermal collision fields are dead here), not back to slot p: the advance
// kernel compacts the bank as it goes.
// keep: where a flight that ends in a real collision leaves its particle
// (FlightEnd), or nullptr.
template <bool kLocate, bool kMove>
__device__ __forceinline__ AdvanceOut AdvanceEvent(const DevWorld& W, const RunState& R,
                                                   uint32_t p, const int* hash,
                                                   const double* Ht, const DevBank& D,
                                                   uint32_t i, FlightEnd* keep = nullptr) {
  AdvanceOut o{false, false, false, false};
  const bool delta = W.tracking == 1;
  V3 pos = LoadPos(R.B, p);
  const V3 dir = LoadDir(R.B, p);
  const double E = R.B.E[p];
  uint32_t g = R.B.rng[p];
  const int cell = R.B.cell[p];
  const DCell c = W.cells[cell];
  const DField fld = W.fields[c.field];
  const double aden = W.mats[c.material].aden;
  const double T = delta ? (fld.kind == 0 ? fld.c : fld.upper) : FieldAt(fld, pos);
  XsMemo memo;
  memo.hint_grid = R.B.xs_grid[p];
  memo.hint_gi = R.B.xs_gi[p];
  memo.tot = Ht;
  // the temperature, nuclide and value of the last total that is not the
  // majorant (for a collision that follows)
  double T_nt = T;
  int nt_nuc = -1;
  double nt_val = 0;
  const double sigma = MaterialTotal(W, c.material, E, T, delta, hash, &memo, &nt_nuc, &nt_val);
  const double d_coll = -log(1.0 - Canonical(g)) / (aden * sigma);
  int s;
  const double d_surf = NearestSurface(W, cell, pos, dir, s);
  const bool cross = delta ? (d_surf < d_coll) : !(d_coll < d_surf);
  double d;
  if (cross) {
    d = d_surf + kNudge;
    o.to_x = true;
  } else {
    bool real = true;
    if (delta) {
      T_nt = FieldAt(fld, pos);
      const double tot = MaterialTotal(W, c.material, E, T_nt, false, hash, &memo, &nt_nuc,
                                       &nt_val);
      real = Canonical(g) < tot / sigma;  // pre-flight position (quirk Q6)
    }
    d = d_coll;
    o.to_c = real;
    o.to_q = !real;
  }
  if (__builtin_expect(W.n_pert != 0, 0))
    StreamIndirect(W, R, p, W.mats[c.material].pert_mask, sigma, d);
  // k-eigenvalue sums: an increment of +0 (no fissile nuclide here) leaves
  // the non-negative sum as it is, so its read-modify-write is skipped
  double nuf_e = -1;
  if (R.keig) {
    const double nuf = MaterialNuFission(W, c.material, E, hash, &memo);
    nuf_e = nuf;
    if (nuf != 0) R.B.k_trk[R.B.hist[p]] += d * aden * nuf;
  }
  if (!kMove && memo.grid >= 0) {
    R.B.xs_grid[p] = memo.grid;
    R.B.xs_gi[p] = memo.gi;
  }
  pos = Add(pos, Mul(dir, d));
  const uint32_t st = R.B.steps[p] + 1;
  if (!kMove) {
    R.B.x[p] = pos.x; R.B.y[p] = pos.y; R.B.z[p] = pos.z;
    R.B.rng[p] = g;
    R.B.steps[p] = st;
  }
  if (st > R.max_steps) {
    atomicOr(R.misc, kErrSteps);
    o.to_x = o.to_c = o.to_q = false;
    return o;
  }
  int to_cell = cell;
  if (o.to_x) {
    o.x_alive = CrossAt<kLocate, kMove>(W, R, p, pos, s, cell, dir, E, to_cell);
  } else if (keep && o.to_c) {
    keep->pos = pos;
    keep->dir = dir;
    keep->E = E;
    keep->g = g;
    keep->cell = cell;
    keep->grid = memo.grid >= 0 ? memo.grid : memo.hint_grid;
    keep->gi = memo.grid >= 0 ? memo.gi : memo.hint_gi;
    keep->T = T_nt;
    keep->tsl = memo.tT == T_nt ? memo.tsl : -1;
    keep->tsl_val = memo.tval;
    keep->nuc = nt_nuc;
    keep->tot = nt_val;
    keep->nuf = nuf_e;
  }
  if (kMove && (o.to_c || o.to_q || o.x_alive)) {
    D.x[i] = pos.x; D.y[i] = pos.y; D.z[i] = pos.z;
    D.u[i] = dir.x; D.v[i] = dir.y; D.w[i] = dir.z;
    D.E[i] = E;
    D.rng[i] = g;
    D.cell[i] = to_cell;
    D.hist[i] = R.B.hist[p];
    D.steps[i] = st;
    D.xs_grid[i] = memo.grid >= 0 ? memo.grid : memo.hint_grid;
    D.xs_gi[i] = memo.grid >= 0 ? memo.gi : memo.hint_gi;
    MoveIndirect(W.n_pert, R.capacity, R.B, p, D, i);
  }
  return o;
}

#ifdef MMC_TRACKER_KERNELS
// kLds: the thermal-scattering totals and the World's descriptors are staged
// after the hash.  kMove: the kernel also compacts the bank, as KCompact does:
// the particle at queue position q is written to slot q of bank D (and the
// queues it joins hold q), and the particles parked for the POD sampler (fb,
// length misc[kMiscFallback]) move to the slots after the live ones; the host
// then swaps the banks.
template <bool kLds, bool kMove>
__global__ void __launch_bounds__(256) KAdvance(DevWorld Wg, RunState R, int iter, int qin,
                                                DevBank D, uint32_t* fb) {
  extern __shared__ int lds_hash[];
  if (blockIdx.x == 0 && threadIdx.x < kRingStride) *Ring(R, iter + 1, threadIdx.x) = 0;
  uint32_t n1;
  const uint32_t n = LiveLength(R, iter - 1, n1);
  const uint32_t nf = kMove && fb ? R.misc[kMiscFallback] : 0u;
  if (kMove && blockIdx.x == 0 && threadIdx.x == 0) R.misc[3] = n + nf;  // slots in use
  if (blockIdx.x * blockDim.x >= n + nf) return;  // no work: skip the staging too
  const int* hash = StageHash(Wg, lds_hash);
  double* lds_tot = reinterpret_cast<double*>(lds_hash + HashLdsInts(Wg));
  const double* Ht = kLds ? StageHot(Wg.tot, lds_tot, Wg.tot_n) : Wg.tot;
  const DevWorld W =
      kLds ? DescInLds(Wg, reinterpret_cast<unsigned char*>(lds_tot + Wg.tot_n)) : Wg;
  WaveTally t_adv, t_virt, t_cross;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n + nf;
       base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    const bool valid = q < n;
    AdvanceOut o{false, false, false, false};
    uint32_t p = 0;
    if (valid) {
      p = LiveAt(R, qin, n1, q);
      o = AdvanceEvent<true, kMove>(W, R, p, hash, Ht, D, q);
    } else if (kMove && q < n + nf) {
      const uint32_t f = fb[q - n];
      MoveParticle(R.B, f, D, q);
      MoveIndirect(W.n_pert, R.capacity, R.B, f, D, q);
      fb[q - n] = q;
    }
    const uint32_t slot = kMove ? q : p;  // the particle's slot after this kernel
    const bool live = o.to_q || o.x_alive;
    uint32_t iq, ic;
    // one atomic per block and pass: the block's collisions then sit together
    // in the collision queue, and the collision kernel reads neighbouring
    // particle slots (KCollide 7.86 -> 7.50 ms per cycle of the multi-zone
    // k-eigenvalue problem with 1e6 histories; KAdvance itself is unchanged)
    BlockPush2(Ring(R, iter, kQ), live, o.to_c, iq, ic);
    if (o.to_c) R.QC[ic] = slot;
    if (live) R.Q[qin ^ 1][iq] = slot;
    t_adv.add(valid);
    t_virt.add(o.to_q);
    t_cross.add(o.to_x);
  }
  t_adv.flush(Ev(R, kEvAdvance));
  t_virt.flush(Ev(R, kEvVirtual));
  t_cross.flush(Ev(R, kEvCross));
}
#endif  // MMC_TRACKER_KERNELS

// ---- collisions ------------------------------------------------------------
// site: for k-eigenvalue, the fission bank entry the caller claimed for this
// child (kClaimSite: claim one here).
constexpr uint32_t kClaimSite = 0xffffffffu;
__device__ __forceinline__ void BankChild(const DevWorld& W, const RunState& R, uint32_t p,
                                          uint32_t& g, V3 pos, V3 d, double E, int child,
                                          uint32_t* live_q, uint32_t* live_cnt,
                                          uint32_t site = kClaimSite) {
  const uint32_t seed = LcgNext(g);  // Particle::BankSecondaries (Particle.cpp:96-100)
  const uint32_t hist = R.B.hist[p];
  // The site key history * 64 + child holds 64 children per history.  The
  // World's nu-bar tables are checked to lie below 63 when it is built, so
  // a fission yields at most 63; this guards the key all the same.
  if (child >= 64) {
    atomicOr(R.misc, kErrOverflow);
    return;
  }
  if (R.keig) {
    const uint32_t k = site != kClaimSite ? site : atomicAdd(R.misc + 2, 1u);
    if (k >= R.site_cap) {
      atomicOr(R.misc, kErrOverflow);
      return;
    }
    mmc_site s;
    s.x = pos.x; s.y = pos.y; s.z = pos.z;
    s.u = d.x; s.v = d.y; s.w = d.z;
    s.E = E;
    s.seed = seed;
    s.pad = 0;
    s.key = uint64_t(hist) * 64u + uint64_t(child);
    R.sites[k] = s;
    return;
  }
  // fixed source: the child is a new particle of the same history, queued
  // with the survivors of this iteration (FixedSource.cpp:63-72)
  const uint32_t slot = atomicAdd(R.misc + 3, 1u);
  if (slot >= R.capacity) {
    atomicOr(R.misc, kErrOverflow);
    return;
  }
  R.B.x[slot] = pos.x; R.B.y[slot] = pos.y; R.B.z[slot] = pos.z;
  R.B.u[slot] = d.x; R.B.v[slot] = d.y; R.B.w[slot] = d.z;
  R.B.E[slot] = E;
  R.B.rng[slot] = LcgSeed(seed);
  R.B.cell[slot] = R.B.cell[p];
  R.B.hist[slot] = hist;
  R.B.steps[slot] = 0;
  R.B.xs_grid[slot] = -1;
  ClearIndirect(W, R, slot);
  if (live_q) live_q[atomicAdd(live_cnt, 1u)] = slot;
}

// Free-gas elastic scatter (ContinuousReaction.cpp:119-189).
__device__ __forceinline__ void FreeGasScatter(double awr, double T, uint32_t& g, V3& dir,
                                               double& E_io) {
  const double m_n = kNeutronMass;
  const double E = E_io;
  const double s_n = sqrt(2. * E / m_n);
  const V3 v_n = Mul(dir, s_n);
  const double beta = sqrt((awr * m_n) / (2. * kBoltzmann * T));
  const double y = beta * s_n;
  double x, mu;
  for (;;) {
    const double xi_1 = Canonical(g), xi_2 = Canonical(g);
    if (Canonical(g) < 2 / (sqrt(kPi) * y + 2)) {
      x = sqrt(-log(xi_1 * xi_2));
    } else {
      const double xi_3 = Canonical(g);
      const double z = cos(kPi * xi_3 / 2);
      x = sqrt(-log(xi_1) - log(xi_2) * z * z);
    }
    mu = 2 * Canonical(g) - 1;
    if (!(Canonical(g) >= sqrt(x * x + y * y - 2 * x * y * mu) / (x + y))) break;
  }
  const double s_T = x / beta;
  const double phi = 2 * kPi * Canonical(g);
  const V3 v_T = Mul(Rotate(dir, mu, phi), s_T);
  const V3 v_cm = Div(Add(v_n, Mul(v_T, awr)), (1 + awr));
  const V3 V_n = Sub(v_n, v_cm);
  const double mu_cm = 2 * Canonical(g) - 1;
  const double phi_cm = 2 * kPi * Canonical(g);
  const V3 Vp_dir = Rotate(Normalized(V_n), mu_cm, phi_cm);
  const V3 V_n_prime = Mul(Vp_dir, sqrt(Dot(V_n, V_n)));
  const V3 v_n_prime = Add(V_n_prime, v_cm);
  const double E_prime = 0.5 * m_n * Dot(v_n_prime, v_n_prime);
  const double mu_lab = Dot(dir, Normalized(v_n_prime));
  // Particle::Scatter (Particle.cpp:55-64)
  const double phi2 = 2 * kPi * Canonical(g);
  dir = Rotate(dir, mu_lab, phi2);
  E_io = E_prime;
}

// Outcome of one collision on one particle (shared by the tracker and the
// single-event parity kernel).
struct CollideOut {
  int event;      // mmc_event
  int children;
  bool tsl;       // deferred thermal scatter (tracker) / done (eval)
};

// Particle::SampleNuclide + Nuclide::Interact (Particle.cpp:110-124,
// Multigroup.cpp:49-73,249-287, Continuous.cpp:57-70,
// ContinuousReaction.cpp:60-62,119-189,252-265).  A thermal-scattering
// scatter stops after choosing the reaction (out.tsl); the caller samples
// (alpha, beta).  k_abs: the absorption estimator's sum of this collision,
// added to in k-eigenvalue runs (R->keig) only.  nt_nuc / nt_val: a nuclide
// and its microscopic total CeTotal at this (E, T), known to the caller, or -1.
__device__ CollideOut Collide(const DevWorld& W, const RunState* R, uint32_t p, V3 pos,
                              V3& dir, double& E, uint32_t& g, int cell, uint32_t* live_q,
                              uint32_t* live_cnt, const int* hash, double* k_abs,
                              int* tsl_index, XsMemo* memo = nullptr, int nt_nuc = -1,
                              double nt_val = 0) {
  CollideOut out{MMC_EV_SCATTER, 0, false};
  const DCell c = W.cells[cell];
  const DMaterial m = W.mats[c.material];
  const double T = FieldAt(W.fields[c.field], pos);
  // SampleNuclide.  With one nuclide the loop below returns it whatever the
  // totals are (the fallback is the last nuclide): only the draw is kept.
  int nuc = W.comps[m.comp_begin + m.comp_count - 1].nuclide;
  if (m.comp_count == 1) {
    (void)Canonical(g);
  } else {
    const double threshold_n =
        MaterialTotal(W, c.material, E, T, false, hash, memo) * Canonical(g);
    double acc = 0;
    for (int k = 0; k < m.comp_count; k++) {
      const DComp cp = W.comps[m.comp_begin + k];
      double tn;
      if (W.energy_mode == 0) {
        tn = W.mg_total[cp.nuclide * W.G + int(E) - 1];
      } else {
        const DCeNuclide n = W.ce[cp.nuclide];
        tn = CeTotal(W, n, NuclideGridIndex(W, n, hash, E, memo), E, T, false, memo);
      }
      acc += cp.afrac * tn;
      if (acc > threshold_n) {
        nuc = cp.nuclide;
        break;
      }
    }
  }
  if (W.energy_mode == 0) {
    const int G = W.G;
    const int gi = int(E) - 1;
    const double cap = W.mg_capture[nuc * G + gi];
    const double sc = W.mg_scatter_xs[nuc * G + gi];
    const double fi = W.mg_fission[nuc * G + gi];
    const double threshold = Canonical(g) * W.mg_total[nuc * G + gi];
    double acc = 0;
    acc += cap;
    if (acc > threshold) {
      out.event = MMC_EV_CAPTURE;
    } else {
      acc += sc;
      if (acc > threshold) {
        out.event = MMC_EV_SCATTER;
        const double t = Canonical(g);
        double a = 0;
        const double* probs = W.mg_scatter_probs + (size_t(nuc) * G + gi) * G;
        for (int go = 0; go < G; go++) {
          a += probs[go];
          if (a > t) {
            E = double(go + 1);
            dir = IsotropicDirection(g);
            break;
          }
        }
      } else {
        acc += fi;
        if (acc > threshold) {
          out.event = MMC_EV_FISSION;
          const int yield = int(W.mg_nubar[nuc * G + gi] + Canonical(g));
          const double* probs = W.mg_chi_probs + (size_t(nuc) * G + gi) * G;
          for (int i = 0; i < yield; i++) {
            const double t = Canonical(g);
            double a = 0;
            for (int go = 0; go < G; go++) {
              a += probs[go];
              if (a > t) {
                const V3 d = IsotropicDirection(g);
                if (R) BankChild(W, *R, p, g, pos, d, double(go + 1), out.children, live_q, live_cnt);
                else LcgNext(g);
                out.children++;
                break;
              }
            }
          }
        } else {
          out.event = -1;  // no reaction selected (Multigroup.cpp:70-72): unchanged
        }
      }
    }
    if (k_abs && R && R->keig &&
        (out.event == MMC_EV_CAPTURE || out.event == MMC_EV_FISSION) && cap + fi > 0)
      *k_abs += (W.mg_nubar[nuc * G + gi] * fi) / (cap + fi);
    return out;
  }
  const DCeNuclide n = W.ce[nuc];
  const int gi = NuclideGridIndex(W, n, hash, E, memo);
  for (int attempt = 0;; attempt++) {  // Continuous::Interact; tail recursion (Q4)
    if (attempt >= kMaxReactionResamples) {
      out.event = MMC_EV_CAPTURE;
      out.children = -1;  // flags kErrReaction to the caller
      return out;
    }
    const double threshold =
        Canonical(g) * (nuc == nt_nuc ? nt_val : CeTotal(W, n, gi, E, T, false, memo));
    double acc = 0;
    for (int k = 0; k < n.rx_count; k++) {
      const DCeReaction r = W.rx[n.rx_begin + k];
      acc += RxXs(W, n, r, gi, E, T, false, memo);
      if (!(acc > threshold)) continue;
      if (r.kind == 0) {
        out.event = MMC_EV_CAPTURE;
      } else if (r.kind == 1) {
        out.event = MMC_EV_SCATTER;
        if (r.tsl >= 0 && E < W.tsl[r.tsl].cutoff_energy) {
          out.tsl = true;
          *tsl_index = r.tsl;
          return out;  // the caller samples (alpha, beta)
        }
        double Eo = E;
        FreeGasScatter(n.awr, T, g, dir, Eo);
        E = Eo;
      } else {
        out.event = MMC_EV_FISSION;
        const int yield = int(TabAt(W, r.nubar, -1, E) + Canonical(g));
        // every child is banked: one atomic claims the fission's bank entries
        // (their order in the bank is fixed later by the site keys)
        const uint32_t site0 = R && R->keig && yield > 0
                                   ? atomicAdd(R->misc + 2, uint32_t(yield)) : kClaimSite;
        for (int i = 0; i < yield; i++) {
          const V3 d = IsotropicDirection(g);
          const double Ec = r.chi_kind == 1 ? SampleWatt(g, r.watt_a, r.watt_b) : E;
          if (R)
            BankChild(W, *R, p, g, pos, d, Ec, out.children, live_q, live_cnt,
                      R->keig ? site0 + uint32_t(i) : kClaimSite);
          else LcgNext(g);
          out.children++;
        }
      }
      if (k_abs && R && R->keig &&
          (out.event == MMC_EV_CAPTURE || out.event == MMC_EV_FISSION) && n.fissile) {
        double sa = 0, nusf = 0;
        for (int q = 0; q < n.rx_count; q++) {
          const DCeReaction rr = W.rx[n.rx_begin + q];
          if (rr.kind == 0) sa += TabAt(W, rr.xs, gi, E);
          if (rr.kind == 2) {
            const double sf = TabAt(W, rr.xs, gi, E);
            sa += sf;
            nusf += TableAt(W, rr.nubar, E) * sf;
          }
        }
        if (sa > 0) *k_abs += nusf / sa;
      }
      return out;
    }
  }
}

// Real collision of particle p (TransportMethod.cpp:74-90 / :106-118 with
// Particle::Scatter).  `alive`: back to the live queue; `to_t`: a bound
// thermal scattering whose (alpha, beta) is sampled by the TSL step.
// Children of a fixed-source fission go to live_q / live_cnt.
// The collision of particle p in state f: the values the advance that queued
// it left (in the bank, or in registers when the same kernel collides).
__device__ __forceinline__ void CollideEventAt(const DevWorld& W, const RunState& R, uint32_t p,
                                               const FlightEnd& f, const int* hash,
                                               const double* Ht, uint32_t* live_q,
                                               uint32_t* live_cnt, bool& alive, bool& to_t) {
  alive = to_t = false;
  const V3 pos = f.pos;
  V3 dir = f.dir;
  double E = f.E;
  uint32_t g = f.g;
  const int cell = f.cell;
  const DCell c = W.cells[cell];
  const double T = FieldAt(W.fields[c.field], pos);
  XsMemo memo;
  memo.hint_grid = f.grid;
  memo.hint_gi = f.gi;
  memo.hint_exact = true;  // left by the advance that ended in this collision
  memo.tot = Ht;
  // the advance's thermal-scattering total at this energy (used where T matches)
  memo.tsl = f.tsl;
  memo.tE = E;
  memo.tT = f.T;
  memo.tval = f.tsl_val;
  // and its nuclide total, where T matches
  const int nt_nuc = f.T == T ? f.nuc : -1;
  // the collision estimator of k: +0 increments (no fissile nuclide here) are
  // skipped, as for the track-length sum, and so is the total they would
  // divide when no flux estimator needs it either
  // (nu Sigma_f depends on the material and E only: the advance's value, where
  // it has one, is this one)
  const double nuf =
      R.keig ? (f.nuf >= 0 ? f.nuf : MaterialNuFission(W, c.material, E, hash, &memo)) : 0.0;
  if (W.n_est_flux || nuf != 0) {
    int nt_in = nt_nuc;
    double nt_v = f.tot;
    const double tot = MaterialTotal(W, c.material, E, T, false, hash, &memo, &nt_in, &nt_v);
    if (W.n_est_flux) ScoreFlux(W, R, R.B.hist[p], pos, W.mats[c.material].aden * tot);
    if (nuf != 0) R.B.k_col[R.B.hist[p]] += nuf / tot;
  }
  double kabs = 0;
  int tsl = -1;
  // (kabs is summed only in k-eigenvalue runs; its address is passed either
  // way, so that it stays in a register)
  const CollideOut o = Collide(W, &R, p, pos, dir, E, g, cell, live_q, live_cnt, hash, &kabs,
                               &tsl, &memo, nt_nuc, f.tot);
  if (o.children < 0) atomicOr(R.misc, kErrReaction);
  if (R.keig) {
    if (kabs != 0) R.B.k_abs[R.B.hist[p]] += kabs;
    if (o.event == MMC_EV_FISSION) R.fis_cnt[R.B.hist[p]] = uint32_t(o.children);
  }
  if (o.tsl) {
    to_t = true;
    R.B.tslT[p] = T;
    R.B.tslI[p] = tsl;
  } else {
    alive = o.event == MMC_EV_SCATTER || o.event < 0;
  }
  // a particle that died is never read again, and a deferred thermal
  // scattering keeps its direction and energy until the TSL step
  if (alive) {
    R.B.u[p] = dir.x; R.B.v[p] = dir.y; R.B.w[p] = dir.z;
    R.B.E[p] = E;
  }
  if (alive || to_t) R.B.rng[p] = g;
}
__device__ __forceinline__ void CollideEvent(const DevWorld& W, const RunState& R, uint32_t p,
                                             const int* hash, const double* Ht,
                                             uint32_t* live_q, uint32_t* live_cnt, bool& alive,
                                             bool& to_t) {
  FlightEnd f;
  f.pos = LoadPos(R.B, p);
  f.dir = LoadDir(R.B, p);
  f.E = R.B.E[p];
  f.g = R.B.rng[p];
  f.cell = R.B.cell[p];
  f.grid = R.B.xs_grid[p];
  f.gi = R.B.xs_gi[p];
  f.tsl = f.nuc = -1;
  f.T = f.tsl_val = f.tot = 0;
  f.nuf = -1;
  CollideEventAt(W, R, p, f, hash, Ht, live_q, live_cnt, alive, to_t);
}

#ifdef MMC_TRACKER_KERNELS
template <bool kLds>
__global__ void __launch_bounds__(256, 2) KCollide(DevWorld Wg, RunState R, int iter, int qin) {
  extern __shared__ int lds_hash[];
  const uint32_t n = *Ring(R, iter, kC);
  if (blockIdx.x * blockDim.x >= n) return;  // no work: skip the staging too
  const int* hash = StageHash(Wg, lds_hash);
  double* lds_tot = reinterpret_cast<double*>(lds_hash + HashLdsInts(Wg));
  const double* Ht = kLds ? StageHot(Wg.tot, lds_tot, Wg.tot_n) : Wg.tot;
  const DevWorld W =
      kLds ? DescInLds(Wg, reinterpret_cast<unsigned char*>(lds_tot + Wg.tot_n)) : Wg;
  WaveTally t_col;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    const bool valid = q < n;
    bool alive = false, to_t = false;
    uint32_t p = 0;
    if (valid) {
      p = R.QC[q];
      CollideEvent(W, R, p, hash, Ht, R.Q[qin ^ 1], Ring(R, iter, kQ), alive, to_t);
    }
    uint32_t it, iq;
    BlockPush2(Ring(R, iter, kT), to_t, alive, it, iq);
    if (to_t) R.QT[it] = p;
    if (alive) R.Q2[qin ^ 1][iq] = p;
    t_col.add(valid);
  }
  t_col.flush(Ev(R, kEvCollide));
}


// The advance and, where the flight ends in a real collision, the collision
// of the same particle, in one kernel: KAdvance's work, then KCollide's on
// the slot the advance just wrote (bank D when the kernel compacts).  In the
// multi-zone k-eigenvalue problem nine flights in ten end in a collision, so
// few lanes idle through the collision code, and the launch, the collision
// queue and the collision kernel's ramp go away.  k-eigenvalue iterations only: there a fission banks
// sites, while a fixed-source child claims a bank slot from misc[3], which
// block 0 of this kernel sets.  Queues: alive after the flight (virtual
// collision, crossing) -> Q (kQ); after the collision: thermal scattering ->
// QT (kT), alive -> Q2 (kQ2); kC stays 0.
template <bool kMove>
__global__ void __launch_bounds__(256, 2) KAdvCol(DevWorld Wg, RunState R, int iter, int qin,
                                                   DevBank D, uint32_t* fb) {
  extern __shared__ int lds_hash[];
  if (blockIdx.x == 0 && threadIdx.x < kRingStride) *Ring(R, iter + 1, threadIdx.x) = 0;
  uint32_t n1;
  const uint32_t n = LiveLength(R, iter - 1, n1);
  const uint32_t nf = kMove && fb ? R.misc[kMiscFallback] : 0u;
  if (kMove && blockIdx.x == 0 && threadIdx.x == 0) R.misc[3] = n + nf;  // slots in use
  if (blockIdx.x * blockDim.x >= n + nf) return;  // no work: skip the staging too
  const int* hash = StageHash(Wg, lds_hash);
  double* lds_tot = reinterpret_cast<double*>(lds_hash + HashLdsInts(Wg));
  const double* Ht = StageHot(Wg.tot, lds_tot, Wg.tot_n);
  const DevWorld W = DescInLds(Wg, reinterpret_cast<unsigned char*>(lds_tot + Wg.tot_n));
  RunState RD = R;  // the collision reads the bank the advance wrote
  if (kMove) RD.B = D;
  WaveTally t_adv, t_virt, t_cross, t_col;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n + nf;
       base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    const bool valid = q < n;
    AdvanceOut o{false, false, false, false};
    uint32_t p = 0;
    FlightEnd fe;
    if (valid) {
      p = LiveAt(R, qin, n1, q);
      o = AdvanceEvent<true, kMove>(W, R, p, hash, Ht, D, q, &fe);
    } else if (kMove && q < n + nf) {
      const uint32_t f = fb[q - n];
      MoveParticle(R.B, f, D, q);
      MoveIndirect(W.n_pert, R.capacity, R.B, f, D, q);
      fb[q - n] = q;
    }
    const uint32_t slot = kMove ? q : p;  // the particle's slot after this kernel
    const bool live = o.to_q || o.x_alive;
    bool alive = false, to_t = false;
    if (o.to_c) CollideEventAt(W, RD, slot, fe, hash, Ht, nullptr, nullptr, alive, to_t);
    const uint32_t iq = BlockPush(Ring(R, iter, kQ), live);
    if (live) R.Q[qin ^ 1][iq] = slot;
    uint32_t it, i2;
    BlockPush2(Ring(R, iter, kT), to_t, alive, it, i2);
    if (to_t) R.QT[it] = slot;
    if (alive) R.Q2[qin ^ 1][i2] = slot;
    t_adv.add(valid);
    t_virt.add(o.to_q);
    t_cross.add(o.to_x);
    t_col.add(o.to_c);
  }
  t_adv.flush(Ev(R, kEvAdvance));
  t_virt.flush(Ev(R, kEvVirtual));
  t_cross.flush(Ev(R, kEvCross));
  t_col.flush(Ev(R, kEvCollide));
}
#endif  // MMC_TRACKER_KERNELS

// ThermalScattering::Scatter (ThermalScattering.cpp:159-171) with POD sampling.
// H: the sampler's hot block (DevWorld::hot or its LDS copy).
__device__ __forceinline__ uint32_t TslScatterPod(const DevWorld& W, const double* H,
                                                  const DTsl& t, uint32_t& g, double T,
                                                  V3& dir, double& E, double& alpha,
                                                  double& beta) {
  const double E0 = E;
  if (!SampleBeta(W, H, t, g, E0, T, beta)) return kErrResample;
  const uint32_t e = SampleAlpha(W, H, t, g, beta, E0, T, alpha);
  if (e) return e;
  const double E_p = E0 + beta * kBoltzmann * T;
  const double mu = (E0 + E_p - alpha * t.awr * kBoltzmann * T) / (2 * sqrt(E0 * E_p));
  const double phi = 2 * kPi * Canonical(g);
  dir = Rotate(dir, mu, phi);
  E = E_p;
  return 0;
}

// Thermal scattering of particle p with the POD sampler; false on a
// sampling error (flagged in misc).
__device__ __forceinline__ bool TslPodEvent(const DevWorld& W, const RunState& R, uint32_t p,
                                            const double* H) {
  V3 dir = LoadDir(R.B, p);
  double E = R.B.E[p];
  uint32_t g = R.B.rng[p];
  double a = 0, b = 0;
  const uint32_t e = TslScatterPod(W, H, W.tsl[R.B.tslI[p]], g, R.B.tslT[p], dir, E, a, b);
  if (e) {
    atomicOr(R.misc, e);
    return false;
  }
  if (W.n_est_tsl) ScoreAlphaBeta(W, R, R.B.hist[p], a, b);
  R.B.u[p] = dir.x; R.B.v[p] = dir.y; R.B.w[p] = dir.z;
  R.B.E[p] = E;
  R.B.rng[p] = g;
  return true;
}

#ifdef MMC_TRACKER_KERNELS
// kLds: the hot block is staged in LDS (launched with hot_n doubles of
// dynamic LDS and only as many blocks as are resident at once).
// (The descriptor block stays in HBM here: a sample reads two descriptors,
// and the sampler's tables alone keep the LDS at two blocks per CU.)
// kThreads: threads per block (the launch option tsl_block): 384 fits three
// waves per SIMD in two blocks of a CU's LDS.
template <bool kLds, int kThreads = 256>
__global__ void __launch_bounds__(kThreads) KTslPod(DevWorld W, RunState R, int iter, int qin) {
  extern __shared__ double lds_hot[];
  const uint32_t n = *Ring(R, iter, kT);
  if (blockIdx.x * blockDim.x >= n) return;  // whole block idle: skip the staging
  const double* H = kLds ? StageHot(W.hot, lds_hot, W.hot_n) : W.hot;
  WaveTally t_tsl;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    const bool valid = q < n;
    bool alive = false;
    uint32_t p = 0;
    if (valid) {
      p = R.QT[q];
      alive = TslPodEvent(W, R, p, H);
    }
    // one atomic per block and pass (BlockPush): 5.92 -> 5.41 ms per cycle of
    // the multi-zone k-eigenvalue problem with 1e6 histories
    const uint32_t iq = BlockPush(Ring(R, iter, kQ), alive);
    if (alive) R.Q[qin ^ 1][iq] = p;
    t_tsl.add(valid);
  }
  t_tsl.flush(Ev(R, kEvTsl));
}

// ---- bank compaction ----------------------------------------------------------
// As histories end, the live particles spread over ever more of the bank, and
// a wave's loads of one field touch one cache line per lane.  This kernel
// moves the live queue's particles to slots 0..n-1 of the other bank, and
// the particles parked for the POD sampler by the surrogate step (fb, length
// misc[kMiscFallback]) to the slots after them; the host then swaps the
// banks, so that the following iterations read the particle state with unit
// stride.  The k-eigenvalue sums are indexed by history and stay where they
// are.
__global__ void __launch_bounds__(256) KCompact(RunState R, DevBank dst, int iter, int qin,
                                                uint32_t* fb, int n_pert) {
  uint32_t n1;
  const uint32_t n = LiveLength(R, iter - 1, n1);
  const uint32_t nf = fb ? R.misc[kMiscFallback] : 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n + nf;
       i += gridDim.x * blockDim.x) {
    if (i < n) {
      uint32_t& e = LiveAt(R, qin, n1, i);
      MoveParticle(R.B, e, dst, i);
      MoveIndirect(n_pert, R.capacity, R.B, e, dst, i);
      e = i;
    } else {
      MoveParticle(R.B, fb[i - n], dst, i);
      MoveIndirect(n_pert, R.capacity, R.B, fb[i - n], dst, i);
      fb[i - n] = i;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) R.misc[3] = n + nf;  // slots in use
}

// ---- tail: history-based completion of the last live particles ------------
// When few particles are left, one event per kernel launch leaves the GPU
// idle between short launches; this kernel instead runs every remaining
// particle to the end of its history, with the same event functions in the
// same order (so the same random numbers and results).  Each lane holds one
// particle and its next event; on every step the wave runs the event type
// that most of its lanes wait for, so a step costs one event's latency
// instead of the sum over the types present in the wave.  A lane whose
// history ends takes the next queued particle.  Thermal scatterings use the
// POD sampler here, unless kPark.  Used only where a collision cannot start a
// new particle (k-eigenvalue, or no fission).
// kThreads: threads per block (the launch option tail_block).  The tail
// stages more LDS than two blocks of a CU can hold, so its occupancy is one
// block per CU: 512 threads make that two waves per SIMD (and cap the
// registers at 256 per lane, arch and acc together).
// kPark: one round of the tail with the network sampler.  A thermal
// collision parks its particle in the thermal-scattering queue of iteration
// `iter` (QT, ring slot kT) and the lane takes the next particle; the
// network kernel then scatters the parked particles into the live queue of
// iteration `iter`, which the next round runs.  Like KAdvance, the round
// clears the ring row of iteration iter + 1 for the next round.
template <bool kLds, int kThreads = 256, bool kPark = false>
__global__ void __launch_bounds__(kThreads) KTail(DevWorld Wg, RunState R, int iter, int qin) {
  enum { kAdv = 0, kCol = 1, kTsl = 2, kDone = 3 };
  extern __shared__ int lds_hash[];
  if (kPark && blockIdx.x == 0 && threadIdx.x < kRingStride) *Ring(R, iter + 1, threadIdx.x) = 0;
  uint32_t n1;
  const uint32_t n = LiveLength(R, iter - 1, n1);
  if (blockIdx.x * blockDim.x >= n) return;
  // kLds: the hash, the totals, the sampler's tables and the World's
  // descriptors in LDS; a parking round does not sample, and its LDS holds
  // the hash, the totals and the descriptors only (the event kernels' LDS)
  const int* hash = kLds ? StageHash(Wg, lds_hash) : nullptr;
  double* lds_tot = reinterpret_cast<double*>(lds_hash + HashLdsInts(Wg));
  const double* Ht = kLds ? StageHot(Wg.tot, lds_tot, Wg.tot_n) : Wg.tot;
  const double* H = kLds && !kPark ? StageHot(Wg.hot, lds_tot + Wg.tot_n, Wg.hot_n) : Wg.hot;
  const DevWorld W = kLds ? DescInLds(Wg, reinterpret_cast<unsigned char*>(
                                              lds_tot + Wg.tot_n + (kPark ? 0 : Wg.hot_n)))
                          : Wg;
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  int st = q < n ? kAdv : kDone;
  uint32_t p = st == kAdv ? LiveAt(R, qin, n1, q) : 0u;
  unsigned long long c_adv = 0, c_virt = 0, c_cross = 0, c_coll = 0, c_tsl = 0, c_park = 0;
  unsigned long long c_tail = st == kAdv ? 1 : 0;
  FlightEnd fe;  // a lane in state kCol: its particle, as its advance left it
  for (;;) {
    int pick = -1, most = 0;
#pragma unroll
    for (int k = kAdv; k <= kTsl; k++) {
      const int c = __popcll(__ballot(st == k));
      if (c > most) {
        most = c;
        pick = k;
      }
    }
    if (pick < 0) break;  // every lane of the wave is done
    if (st != pick) continue;
    if (st == kAdv) {
      const AdvanceOut o = AdvanceEvent<false, false>(W, R, p, hash, Ht, R.B, p, &fe);
      c_adv++;
      c_virt += o.to_q ? 1 : 0;
      c_cross += o.to_x ? 1 : 0;
      st = (o.to_q || o.x_alive) ? kAdv : (o.to_c ? kCol : kDone);
    } else if (st == kCol) {
      c_coll++;
      bool alive, to_t;
      CollideEventAt(W, R, p, fe, hash, Ht, nullptr, nullptr, alive, to_t);
      if (kPark) {
        // one queue push for the wave's parked lanes (the lanes active in
        // this branch; the first of them adds the count)
        const uint64_t m = __ballot(to_t);
        if (m) {
          const int lead = __ffsll(static_cast<unsigned long long>(m)) - 1;
          uint32_t base = 0;
          if (__lane_id() == uint32_t(lead))
            base = atomicAdd(Ring(R, iter, kT), uint32_t(__popcll(m)));
          base = __shfl(base, lead);
          if (to_t) R.QT[base + __popcll(m & ((1ull << __lane_id()) - 1ull))] = p;
        }
        c_park += to_t ? 1 : 0;
        st = alive && !to_t ? kAdv : kDone;
      } else {
        st = to_t ? kTsl : (alive ? kAdv : kDone);
      }
    } else if constexpr (!kPark) {
      c_tsl++;
      st = TslPodEvent(W, R, p, H) ? kAdv : kDone;
    } else {
      st = kDone;  // (a parking round has no thermal step)
    }
    if (st == kDone) {
      q += stride;
      if (q < n) {
        p = LiveAt(R, qin, n1, q);
        st = kAdv;
        c_tail++;
      }
    }
  }
  if (c_tail) {
    atomicAdd(Ev(R, kEvAdvance), c_adv);
    atomicAdd(Ev(R, kEvVirtual), c_virt);
    atomicAdd(Ev(R, kEvCross), c_cross);
    atomicAdd(Ev(R, kEvCollide), c_coll);
    atomicAdd(Ev(R, kEvTsl), c_tsl);
    atomicAdd(Ev(R, kEvTail), c_tail);
    atomicAdd(Ev(R, kEvTailAdv), c_adv);
    atomicAdd(Ev(R, kEvTailCol), c_coll);
    atomicAdd(Ev(R, kEvTailTsl), c_tsl + c_park);
    atomicAdd(Ev(R, kEvTailCross), c_cross);
  }
}

#endif  // MMC_TRACKER_KERNELS

}  // namespace mmc
