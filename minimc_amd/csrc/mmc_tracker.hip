// C ABI of the tracker (include/minimc_amd.h): context creation (World
// flattening into device tables), the event loop driver, per-history tally
// commit, the k-eigenvalue cycle (fission bank sort and resampling), the
// single-event parity entry point and the TSL batch samplers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/minimc_amd.h"
#define MMC_TRACKER_KERNELS
#include "mmc_kernels.hpp"
#include "mmc_mlp.hpp"
#include "mmc_tally.hpp"

using namespace mmc;

#define HIPCHK(call)                                                   \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) {                                            \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return MMC_ERR_HIP;                                              \
    }                                                                  \
  } while (0)

static constexpr int kBlock = 256;
// Largest thermal-scattering hot block staged in LDS (of 160 KiB per CU),
// and the largest total-cross-section part staged by the event kernels.
static constexpr size_t kHotLdsMax = 128 * 1024;
static constexpr size_t kTotLdsMax = 32 * 1024;
static_assert(kMaxPert == MMC_MAX_PERTURBATIONS, "perturbation limit");
static constexpr size_t kDescLdsMax = 24 * 1024;  // the World's descriptor block
static constexpr size_t kLdsPerBlock = 160 * 1024;
// Largest nu-bar: a fission then yields at most 63 children, and the fission
// site key (history * 64 + child) keeps the children of one history apart.
static constexpr double kMaxNubar = 63.0;

struct mmc_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  mmc_options opt{};
  DevWorld W{};
  std::vector<void*> allocs;
  uint32_t* host_rb = nullptr;  // pinned copy of the ring and misc (one read per batch)
  bool has_tsl = false;
  int n_tsl = 0;               // thermal-scattering tables of the World
  bool fixed_fission = false;  // fixed-source problem with fission children
  int n_hash0 = 0;             // LDS entries staged by the CE kernels
  uint32_t lds_grid = 1024;    // grid cap of the LDS-staging kernels
  uint32_t tail_grid = 512;    // blocks of the tail kernel resident at once
  uint32_t tail_park_grid = 512;  // and of its parking rounds
  int tail_block = 256;        // threads per block of KTail (256 or 512)
  int tsl_block = 256;         // threads per block of KTslPod (256 or 384)
  size_t hot_lds = 0;          // LDS bytes of KTslPod<true> (0: it does not stage)
  uint32_t tsl_grid = 1024;    // grid cap of KTslPod when it stages the hot block
  size_t ev_lds = 0;           // LDS bytes of KAdvance / KCollide
  bool ev_tot_lds = false;     // ... which stage the totals and descriptors after the hash
  size_t tail_lds = 0;         // LDS bytes of KTail<true> (0: KTail<false>, no LDS)
  size_t tail_park_lds = 0;    // and of its parking rounds (hash, totals, descriptors)
  DevBank alt{};               // second bank: target of the compaction
  // bank
  uint32_t cap = 0;
  RunState R{};
  uint32_t* scan_tmp = nullptr;
  uint32_t scan_cap = 0;
  // tallies
  int NB = 0;
  double *S = nullptr, *S2 = nullptr;
  std::vector<int> est_off, est_size;
  std::vector<int> cell_material;  // host copy for argument checks
  double *bin_sum = nullptr, *bin_sq = nullptr;  // a batch's per-bin sums (TallyReduce)
  TallyWork tally{};
  // k-eigenvalue
  mmc_site* src = nullptr;
  mmc_site* sites_sorted = nullptr;
  uint32_t* fis_off = nullptr;
  uint64_t N = 0;
  int mesh_n[3] = {1, 1, 1};
  double mesh_lo[3] = {-1e300, -1e300, -1e300};
  double mesh_hi[3] = {1e300, 1e300, 1e300};
  uint32_t* mesh_cnt = nullptr;
  double* kpart = nullptr;
  uint32_t kpart_cap = 0;
  mmc_counters cnt{};
  unsigned long long ev_host[kEvN] = {};
  // surrogate
  MlpState mlp{};
  // kernel timing
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_kind;  // kind of the launch bracketed by pair i
  size_t ev_used = 0;
  double k_ms[MMC_K_N] = {};
  uint64_t k_launches[MMC_K_N] = {};
  // Captured batches of event iterations (RunEvents with opt.graphs), keyed
  // by everything the captured launch arguments depend on: ring phase, queue
  // parity, grid, the bank arrays in use, the record buffers, the network.
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    std::vector<int> kinds;  // profiling brackets recorded by the graph
    size_t ev_used = 0;
    int swaps = 0;           // bank swaps the batch makes
  };
  std::map<std::array<uint64_t, 8>, GraphEntry> graphs;
  uint64_t graph_captures = 0, graph_launches = 0;
};

template <class T>
static T* DevAlloc(mmc_ctx* ctx, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
  ctx->allocs.push_back(p);
  return static_cast<T*>(p);
}
template <class T>
static T* DevUpload(mmc_ctx* ctx, const std::vector<T>& v) {
  T* p = DevAlloc<T>(ctx, v.size());
  if (p && !v.empty()) hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return p;
}
static inline uint32_t Blocks(uint64_t n, uint32_t cap = 4096) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return uint32_t(std::min<uint64_t>(b, cap));
}

// ---------------------------------------------------------------------------
// World flattening
namespace {

struct Flat {
  std::vector<double> pool;
  std::vector<double> hot;  // the thermal-scattering sampler's tables (DevWorld::hot)
  std::vector<double> tot;  // the thermal-scattering totals (DevWorld::tot)
  std::vector<int> ipool;
  std::vector<DGrid> grids;
  std::vector<std::vector<double>> grid_keys;
  std::vector<DTable> tables;

  int AddDoubles(const double* p, size_t n) {
    const int off = int(pool.size());
    pool.insert(pool.end(), p, p + n);
    return off;
  }
  int AddHot(const double* p, size_t n) {
    const int off = int(hot.size());
    hot.insert(hot.end(), p, p + n);
    return off;
  }
  int AddTot(const double* p, size_t n) {
    const int off = int(tot.size());
    tot.insert(tot.end(), p, p + n);
    return off;
  }
  int AddGrid(const double* x, int n) {
    for (size_t g = 0; g < grids.size(); g++)
      if (grid_keys[g].size() == size_t(n) &&
          std::memcmp(grid_keys[g].data(), x, sizeof(double) * n) == 0)
        return int(g);
    DGrid g{};
    g.x_off = AddDoubles(x, n);
    g.n = n;
    g.n_hash = 0;
    g.hash_off = 0;
    g.x_first = n > 0 ? x[0] : 0.0;
    g.x_last = n > 0 ? x[n - 1] : 0.0;
    if (n >= 64 && x[0] > 0) {
      // log-spaced buckets; H[b] = upper_bound(x, edge_b) (host, exact)
      int nh = 1;
      while (nh < n / 4 && nh < 8192) nh <<= 1;
      const double l0 = std::log(x[0]);
      const double dl = (std::log(x[n - 1]) - l0) / nh;
      g.log_x0 = l0;
      g.inv_dlog = 1.0 / dl;
      g.n_hash = nh;
      g.hash_off = int(ipool.size());
      for (int b = 0; b <= nh; b++) {
        const double edge = b == 0 ? x[0] : (b == nh ? x[n - 1] : std::exp(l0 + b * dl));
        ipool.push_back(int(std::upper_bound(x, x + n, edge) - x));
      }
      // ensure the extreme buckets cover everything strictly inside
      ipool[g.hash_off] = 0;
      ipool[g.hash_off + nh] = n;
    }
    grids.push_back(g);
    grid_keys.emplace_back(x, x + n);
    return int(grids.size()) - 1;
  }
  int AddTable(const mmc_table& t) {
    if (t.n <= 0 || !t.x || !t.y) return -1;
    DTable d{};
    d.grid = AddGrid(t.x, t.n);
    d.y_off = AddDoubles(t.y, t.n);
    d.y_stride = 1;
    tables.push_back(d);
    return int(tables.size()) - 1;
  }
};

DBins MakeBins(Flat& F, const mmc_bins& b, bool present) {
  DBins d{};
  d.kind = present ? b.kind : MMC_BINS_NONE;
  switch (d.kind) {
    case MMC_BINS_LINSPACE:
    case MMC_BINS_LOGSPACE:
      d.n_bins = b.n + 2;
      d.lo = b.lo;
      d.hi = b.hi;
      d.width = (b.hi - b.lo) / double(d.n_bins - 2);
      d.base_log = std::log(d.kind == MMC_BINS_LOGSPACE && b.base > 0 ? b.base : 10.0);
      break;
    case MMC_BINS_BOUNDARIES:
      d.n_bins = b.n + 1;
      d.bounds_off = F.AddDoubles(b.bounds, b.n);
      d.n_bounds = b.n;
      break;
    default:
      d.n_bins = 1;
  }
  return d;
}

// Cell locator (DevWorld::loc_*): a uniform grid over the finite bounds of
// the cell boxes.  Along an axis where no box is bounded the grid has one
// voxel and loc_inv = 0.  Voxel v lists every cell whose box meets v widened
// by 1e-6 of a voxel on each side, in cell order; the tolerance covers a
// voxel index computed across a boundary by rounding.  About 8 voxels per
// cell, at most 4096, halved until off + list fit in `budget` bytes; no grid
// (loc_n[0] = 0) for a few cells, or when even 8 voxels do not fit.
void BuildLocator(const std::vector<DCell>& cells, size_t budget, DevWorld& W,
                  std::vector<int>& off, std::vector<int>& list) {
  W.loc_n[0] = W.loc_n[1] = W.loc_n[2] = 0;
  off.assign(1, 0);
  list.clear();
  const int nc = int(cells.size());
  if (nc <= 4) return;
  double lo[3], hi[3];
  int active = 0;
  for (int a = 0; a < 3; a++) {
    lo[a] = HUGE_VAL;
    hi[a] = -HUGE_VAL;
    for (const DCell& c : cells)
      for (double v : {c.lo[a], c.hi[a]})
        if (std::isfinite(v)) {
          lo[a] = std::min(lo[a], v);
          hi[a] = std::max(hi[a], v);
        }
    if (hi[a] > lo[a]) active++;
  }
  if (!active) return;
  for (int target = std::min(4096, 8 * nc); target >= 8; target /= 2) {
    const int per = std::max(1, int(std::lround(std::pow(double(target), 1.0 / active))));
    int n[3];
    double inv[3], w[3];
    for (int a = 0; a < 3; a++) {
      const bool on = hi[a] > lo[a];
      n[a] = on ? std::min(per, 64) : 1;
      inv[a] = on ? n[a] / (hi[a] - lo[a]) : 0.0;
      w[a] = on ? (hi[a] - lo[a]) / n[a] : 0.0;
    }
    off.assign(1, 0);
    list.clear();
    for (int i = 0; i < n[0]; i++)
      for (int j = 0; j < n[1]; j++)
        for (int k = 0; k < n[2]; k++) {
          const int idx[3] = {i, j, k};
          for (int c = 0; c < nc; c++) {
            bool meets = true;
            for (int a = 0; a < 3 && meets; a++) {
              if (inv[a] == 0) continue;
              const double vlo = lo[a] + idx[a] * w[a] - 1e-6 * w[a];
              const double vhi = lo[a] + (idx[a] + 1) * w[a] + 1e-6 * w[a];
              meets = cells[c].lo[a] <= vhi && cells[c].hi[a] >= vlo;
            }
            if (meets) list.push_back(c);
          }
          off.push_back(int(list.size()));
        }
    if ((off.size() + list.size()) * sizeof(int) + 32 <= budget) {
      for (int a = 0; a < 3; a++) {
        W.loc_n[a] = n[a];
        W.loc_lo[a] = inv[a] == 0 ? 0.0 : lo[a];
        W.loc_inv[a] = inv[a];
      }
      return;
    }
  }
  off.assign(1, 0);
  list.clear();
}

}  // namespace

static int Fail(mmc_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}
// A launch that fails at configuration time (more LDS than a CU holds, a bad
// grid) runs nothing and leaves its queue rows at zero, which would read as
// a finished batch.  Every batch of launches is followed by this check
// before its results are read.
static int LaunchCheck(mmc_ctx* ctx, const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return Fail(ctx, MMC_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
  return MMC_OK;
}
int MmcLaunchCheck(mmc_ctx* ctx, const char* where) { return LaunchCheck(ctx, where); }
// Static LDS of a kernel (its __shared__ arrays), which adds to the dynamic
// LDS a launch asks for.
static size_t StaticLds(const void* fn) {
  hipFuncAttributes a{};
  return hipFuncGetAttributes(&a, fn) == hipSuccess ? a.sharedSizeBytes : 0;
}

extern "C" int mmc_abi_version(void) { return MMC_ABI_VERSION; }

extern "C" int mmc_device_count(int* n) {
  if (!n) return MMC_ERR_ARG;
  if (hipGetDeviceCount(n) != hipSuccess) {
    *n = 0;
    return MMC_ERR_NODEV;
  }
  return MMC_OK;
}

static int BuildWorld(mmc_ctx* ctx, const mmc_world_desc* wd, const mmc_tally_desc* td) {
  Flat F;
  DevWorld& W = ctx->W;
  W.energy_mode = wd->energy_mode;
  W.G = wd->groups;
  W.tracking = ctx->opt.tracking;
  W.quirks = ctx->opt.quirks;
  W.n_cells = wd->n_cells;
  if (wd->n_cells <= 0 || wd->n_materials <= 0 || wd->n_nuclides <= 0)
    return Fail(ctx, MMC_ERR_ARG, "world needs cells, materials and nuclides");
  std::vector<DSurface> surf(wd->n_surfaces);
  for (int i = 0; i < wd->n_surfaces; i++) {
    const auto& s = wd->surfaces[i];
    if (s.kind < 0 || s.kind > 4) return Fail(ctx, MMC_ERR_ARG, "bad surface kind");
    surf[i] = DSurface{s.a, s.b, s.c, s.r, s.kind, 0};
  }
  std::vector<DCell> cells(wd->n_cells);
  for (int i = 0; i < wd->n_cells; i++) {
    const auto& c = wd->cells[i];
    if (c.material >= wd->n_materials || c.field < 0 || c.field >= wd->n_fields ||
        c.face_begin < 0 || c.face_begin + c.face_count > wd->n_faces)
      return Fail(ctx, MMC_ERR_ARG, "bad cell " + std::to_string(i));
    cells[i] = DCell{c.material, c.face_begin, c.face_count, c.field, {}, {}};
    ctx->cell_material.push_back(c.material);
  }
  // Bounding box of each cell, from the surfaces' Contains tests
  // (SurfContains): a plane's inside is coordinate < a, its outside
  // coordinate >= a; a sphere's inside lies within center +- r; a CylinderX's
  // inside (quirk Q1: sqrt(y^2 + z^2) < r^2) within |y|, |z| < r^2; an
  // outside of a sphere or cylinder bounds nothing.  Each bound is widened by
  // 1e-9 (1 + |bound|), far above the nudge and the rounding of a point that
  // has just crossed the surface, so a box never rejects a point the exact
  // tests (or the forced sense of a crossing) accept.
  for (int i = 0; i < wd->n_cells; i++) {
    const auto& c = wd->cells[i];
    double lo[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL}, hi[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
    for (int k = 0; k < c.face_count; k++) {
      const auto& f = wd->faces[c.face_begin + k];
      const auto& sf = wd->surfaces[f.surface];
      const int axis = sf.kind == MMC_PLANEX ? 0 : sf.kind == MMC_PLANEY ? 1
                       : sf.kind == MMC_PLANEZ ? 2 : -1;
      if (axis >= 0) {
        if (f.inside) hi[axis] = std::min(hi[axis], sf.a);
        else lo[axis] = std::max(lo[axis], sf.a);
      } else if (f.inside && sf.kind == MMC_SPHERE) {
        const double ctr[3] = {sf.a, sf.b, sf.c};
        for (int a = 0; a < 3; a++) {
          lo[a] = std::max(lo[a], ctr[a] - sf.r);
          hi[a] = std::min(hi[a], ctr[a] + sf.r);
        }
      } else if (f.inside && sf.kind == MMC_CYLINDERX) {
        const double h = sf.r * sf.r;
        for (int a = 1; a < 3; a++) {
          lo[a] = std::max(lo[a], -h);
          hi[a] = std::min(hi[a], h);
        }
      }
    }
    for (int a = 0; a < 3; a++) {
      cells[i].lo[a] = lo[a] == -HUGE_VAL ? lo[a] : lo[a] - 1e-9 * (1 + std::fabs(lo[a]));
      cells[i].hi[a] = hi[a] == HUGE_VAL ? hi[a] : hi[a] + 1e-9 * (1 + std::fabs(hi[a]));
    }
  }
  std::vector<DFace> faces(wd->n_faces);
  for (int i = 0; i < wd->n_faces; i++) {
    if (wd->faces[i].surface < 0 || wd->faces[i].surface >= wd->n_surfaces)
      return Fail(ctx, MMC_ERR_ARG, "bad face surface");
    faces[i] = DFace{wd->faces[i].surface, wd->faces[i].inside ? 1 : 0};
  }
  std::vector<DField> fields(wd->n_fields);
  for (int i = 0; i < wd->n_fields; i++) {
    const auto& f = wd->fields[i];
    fields[i] = DField{f.kind, 0, f.c, f.gx, f.gy, f.gz, f.b,
                       f.kind == MMC_FIELD_CONSTANT ? f.c : f.upper};
  }
  // total cross-section perturbations
  if (wd->n_perturbations < 0 || wd->n_perturbations > MMC_MAX_PERTURBATIONS ||
      (wd->n_perturbations > 0 && !wd->perturbation_nuclide))
    return Fail(ctx, MMC_ERR_ARG, "bad perturbations");
  for (int j = 0; j < wd->n_perturbations; j++)
    if (wd->perturbation_nuclide[j] < 0 || wd->perturbation_nuclide[j] >= wd->n_nuclides)
      return Fail(ctx, MMC_ERR_ARG, "bad perturbation nuclide");
  W.n_pert = wd->n_perturbations;
  // Material::AssignNuclides normalises atom fractions (Material.cpp:82-92)
  std::vector<DMaterial> mats(wd->n_materials);
  std::vector<DComp> comps;
  for (int i = 0; i < wd->n_materials; i++) {
    const auto& m = wd->materials[i];
    double sum = 0;
    for (int k = 0; k < m.comp_count; k++) sum += wd->components[m.comp_begin + k].afrac;
    mats[i] = DMaterial{m.aden, int(comps.size()), m.comp_count, 0, 0};
    for (int k = 0; k < m.comp_count; k++) {
      const auto& c = wd->components[m.comp_begin + k];
      if (c.nuclide < 0 || c.nuclide >= wd->n_nuclides)
        return Fail(ctx, MMC_ERR_ARG, "bad component nuclide");
      comps.push_back(DComp{c.afrac / sum, c.nuclide, 0});
      // perturbations of this nuclide act in this material (Perturbation.cpp:71-75)
      for (int j = 0; j < W.n_pert; j++)
        if (wd->perturbation_nuclide[j] == c.nuclide) mats[i].pert_mask |= 1 << j;
    }
  }
  const int G = W.G;
  const int NN = wd->n_nuclides;
  std::vector<double> cap, sxs, fis, nub, tot, sp, cp;
  std::vector<DCeNuclide> ce;
  std::vector<DCeReaction> rx;
  std::vector<DTsl> tsl;
  std::vector<DTslPart> parts;
  if (W.energy_mode == MMC_MULTIGROUP) {
    if (G <= 0) return Fail(ctx, MMC_ERR_ARG, "multigroup needs groups > 0");
    cap.assign(size_t(NN) * G, 0);
    sxs = fis = nub = tot = cap;
    sp.assign(size_t(NN) * G * G, 0);
    cp = sp;
    for (int n = 0; n < NN; n++) {
      const auto& d = wd->mg[n];
      for (int g = 0; g < G; g++) {
        if (d.capture) cap[n * G + g] = d.capture[g];
        if (d.fission_xs) fis[n * G + g] = d.fission_xs[g];
        if (d.nubar && d.fission_xs) nub[n * G + g] = d.nubar[g];
        // a fission yields int(nu-bar + U) children, at most 63 (the site
        // key holds 64 per history)
        if (!(nub[n * G + g] >= 0 && nub[n * G + g] < kMaxNubar))
          return Fail(ctx, MMC_ERR_ARG, "nubar must lie in [0, 63)");
      }
      // columns of the XML matrices: incoming group gp, outgoing g
      // (Multigroup.cpp:124-147); normalised per column (:168-186)
      auto column_probs = [&](const double* m, std::vector<double>& out, double* sums) {
        for (int gp = 0; gp < G; gp++) {
          double s = 0;
          for (int g = 0; g < G; g++) s += m ? m[G * g + gp] : 0.0;
          if (sums) sums[gp] = s;
          for (int g = 0; g < G; g++) {
            const double v = m ? m[G * g + gp] : 0.0;
            out[(size_t(n) * G + gp) * G + g] = s == 0. ? v : v / s;
          }
        }
      };
      column_probs(d.scatter, sp, &sxs[n * G]);
      column_probs(d.chi, cp, nullptr);
      for (int g = 0; g < G; g++) {
        double t = 0;  // CreateTotalXS (Multigroup.cpp:231-243), enum order
        if (d.capture) t = t + cap[n * G + g];
        if (d.scatter) t = t + sxs[n * G + g];
        if (d.fission_xs) t = t + fis[n * G + g];
        tot[n * G + g] = t;
      }
    }
  } else {
    for (int n = 0; n < NN; n++) {
      const auto& d = wd->ce[n];
      DCeNuclide dn{};
      dn.awr = d.awr;
      dn.total_temperature = d.total_temperature;
      dn.has_total = d.has_total && d.total.n > 0;
      dn.total = dn.has_total ? F.AddTable(d.total) : -1;
      dn.rx_begin = int(rx.size());
      dn.rx_count = d.n_reactions;
      dn.fissile = 0;
      int grid = dn.has_total ? F.tables[dn.total].grid : -2;
      for (int k = 0; k < d.n_reactions; k++) {
        const auto& r = d.reactions[k];
        DCeReaction dr{};
        dr.kind = r.kind;
        dr.tsl = r.kind == MMC_RX_SCATTER ? r.tsl : -1;
        if (dr.tsl >= wd->n_tsl) return Fail(ctx, MMC_ERR_ARG, "bad tsl index");
        dr.xs = F.AddTable(r.xs);
        if (dr.xs < 0) return Fail(ctx, MMC_ERR_ARG, "reaction without xs table");
        dr.nubar = r.kind == MMC_RX_FISSION ? F.AddTable(r.nubar) : -1;
        if (r.kind == MMC_RX_FISSION) {
          if (dr.nubar < 0) return Fail(ctx, MMC_ERR_ARG, "fission without nubar");
          for (int i = 0; i < r.nubar.n; i++)
            if (!(r.nubar.y[i] >= 0 && r.nubar.y[i] < kMaxNubar))
              return Fail(ctx, MMC_ERR_ARG, "nubar must lie in [0, 63)");
          dn.fissile = 1;
        }
        dr.chi_kind = r.chi_kind;
        dr.temperature = r.temperature;
        dr.watt_a = r.watt_a;
        dr.watt_b = r.watt_b;
        const int rg = F.tables[dr.xs].grid;
        grid = grid == -2 ? rg : (grid == rg ? grid : -1);
        rx.push_back(dr);
      }
      dn.grid = grid >= 0 ? grid : -1;
      // A collision reads every reaction's cross section at the same grid
      // index: interleave the reaction tables of a shared grid point by
      // point, so that one cache line serves all of them.
      if (dn.grid >= 0 && dn.rx_count >= 2) {
        const int c = dn.rx_count, n = F.grids[dn.grid].n;
        std::vector<double> block(size_t(n) * c);
        for (int k = 0; k < c; k++) {
          const DTable& tb = F.tables[rx[dn.rx_begin + k].xs];
          for (int i = 0; i < n; i++) block[size_t(i) * c + k] = F.pool[tb.y_off + i];
        }
        const int off = F.AddDoubles(block.data(), block.size());
        for (int k = 0; k < c; k++) {
          DTable& tb = F.tables[rx[dn.rx_begin + k].xs];
          tb.y_off = off + k;
          tb.y_stride = c;
        }
      }
      ce.push_back(dn);
    }
    for (int i = 0; i < wd->n_tsl; i++) {
      const auto& d = wd->tsl[i];
      DTsl t{};
      t.majorant = F.AddTable(d.majorant);
      if (t.majorant < 0 || d.n_E < 1 || d.n_T < 1 || d.rank < 1)
        return Fail(ctx, MMC_ERR_ARG, "bad tsl tables");
      t.E_axis_off = F.AddTot(d.total_E_axis, d.n_E);
      t.nE = d.n_E;
      t.E_modes_off = F.AddTot(d.total_E, size_t(d.n_E) * d.rank);
      t.T_axis_off = F.AddTot(d.total_T_axis, d.n_T);
      t.nT = d.n_T;
      t.T_modes_off = F.AddTot(d.total_T, size_t(d.n_T) * d.rank);
      t.S_off = F.AddTot(d.total_S, d.rank);
      t.rank = d.rank;
      // every partition has points on each axis: the samplers evaluate the
      // ends of a CDF interval at in-range indices
      bool bad_part = false;
      auto add_parts = [&](const mmc_tsl_partition* p, int np, int& begin, int& xs_off,
                           int& nxs, int& ends_off) {
        begin = int(parts.size());
        std::vector<double> xs;
        std::vector<int> ends;
        for (int k = 0; k < np; k++) {
          if (p[k].n_F < 1 || p[k].n_x < 1 || p[k].n_T < 1 || p[k].rank < 1) bad_part = true;
          DTslPart q{};
          q.F_off = F.AddHot(p[k].F, p[k].n_F);
          q.nF = p[k].n_F;
          q.cdf_off = F.AddHot(p[k].cdf_modes, size_t(p[k].n_F) * p[k].rank);
          q.S_off = F.AddHot(p[k].S, p[k].rank);
          q.x_off = F.AddDoubles(p[k].x, p[k].n_x);
          q.nx = p[k].n_x;
          q.T_off = F.AddHot(p[k].T, p[k].n_T);
          q.nT = p[k].n_T;
          q.xt_off = F.AddDoubles(p[k].x_T_modes, size_t(p[k].n_x) * p[k].n_T * p[k].rank);
          q.rank = p[k].rank;
          parts.push_back(q);
          xs.insert(xs.end(), p[k].x, p[k].x + p[k].n_x);
          ends.push_back(int(xs.size()));
        }
        xs_off = F.AddHot(xs.data(), xs.size());
        nxs = int(xs.size());
        const std::vector<double> ends_d(ends.begin(), ends.end());
        ends_off = F.AddHot(ends_d.data(), ends_d.size());
      };
      add_parts(d.beta_partitions, d.n_beta_partitions, t.bp_begin, t.Es_off, t.nEs,
                t.E_ends_off);
      t.n_bp = d.n_beta_partitions;
      add_parts(d.alpha_partitions, d.n_alpha_partitions, t.ap_begin, t.betas_off,
                t.nbetas, t.beta_ends_off);
      t.n_ap = d.n_alpha_partitions;
      if (bad_part) return Fail(ctx, MMC_ERR_ARG, "bad tsl partition (an empty axis)");
      t.beta_cutoff = d.beta_cutoff;
      t.alpha_cutoff = d.alpha_cutoff;
      t.cutoff_energy = d.total_E_axis[d.n_E - 1];  // ThermalScattering.hpp:125-126
      t.awr = 0;
      tsl.push_back(t);
    }
    for (int n = 0; n < NN; n++)
      for (int k = 0; k < ce[n].rx_count; k++)
        if (rx[ce[n].rx_begin + k].tsl >= 0) tsl[rx[ce[n].rx_begin + k].tsl].awr = ce[n].awr;
    ctx->has_tsl = !tsl.empty();
    ctx->n_tsl = int(tsl.size());
    if (!F.grids.empty()) ctx->n_hash0 = F.grids[0].n_hash;
  }
  // estimators
  std::vector<DEst> ests;
  int nb = 0;
  if (td) {
    for (int i = 0; i < td->n_estimators; i++) {
      const auto& e = td->estimators[i];
      DEst d{};
      d.kind = e.kind;
      d.surface = e.surface;
      if (e.sensitivity < 0 || e.sensitivity > W.n_pert ||
          (e.sensitivity && e.kind != MMC_EST_CURRENT))
        return Fail(ctx, MMC_ERR_ARG, "bad estimator sensitivity");
      d.sens = e.sensitivity;
      d.has_dir = e.has_direction ? 1 : 0;
      const double nrm = std::sqrt(e.u * e.u + e.v * e.v + e.w * e.w);
      d.dx = e.u / nrm;
      d.dy = e.v / nrm;
      d.dz = e.w / nrm;
      d.cosine = MakeBins(F, e.cosine, d.has_dir);
      d.energy = MakeBins(F, e.energy, true);
      int size = 0;
      if (e.kind == MMC_EST_FLUX_MESH) {
        size = 1;
        for (int a = 0; a < 3; a++) {
          d.mesh_n[a] = e.mesh_n[a] > 0 ? e.mesh_n[a] : 1;
          d.mesh_lo[a] = e.mesh_lo[a];
          d.mesh_hi[a] = e.mesh_hi[a];
          size *= d.mesh_n[a];
        }
        W.n_est_flux++;
      } else if (e.kind == MMC_EST_TSL_ALPHA || e.kind == MMC_EST_TSL_BETA) {
        size = d.energy.n_bins;
        W.n_est_tsl++;
      } else if (e.kind == MMC_EST_CURRENT) {
        if (e.surface < 0 || e.surface >= wd->n_surfaces)
          return Fail(ctx, MMC_ERR_ARG, "current estimator surface out of range");
        size = d.cosine.n_bins * d.energy.n_bins;
        W.n_est_current++;
      } else {
        return Fail(ctx, MMC_ERR_ARG, "unknown estimator kind");
      }
      d.bin_off = nb;
      ctx->est_off.push_back(nb);
      ctx->est_size.push_back(size);
      nb += size;
      ests.push_back(d);
    }
  }
  W.n_est = int(ests.size());
  ctx->NB = nb;
  if (F.pool.empty()) F.pool.push_back(0);
  if (F.ipool.empty()) F.ipool.push_back(0);
  W.mg_capture = DevUpload(ctx, cap);
  W.mg_scatter_xs = DevUpload(ctx, sxs);
  W.mg_fission = DevUpload(ctx, fis);
  W.mg_nubar = DevUpload(ctx, nub);
  W.mg_total = DevUpload(ctx, tot);
  W.mg_scatter_probs = DevUpload(ctx, sp);
  W.mg_chi_probs = DevUpload(ctx, cp);
  W.pool = DevUpload(ctx, F.pool);
  W.ipool = DevUpload(ctx, F.ipool);
  W.hot_n = int(F.hot.size());
  if (F.hot.empty()) F.hot.push_back(0);
  W.hot = DevUpload(ctx, F.hot);
  W.tot_n = int(F.tot.size());
  if (F.tot.empty()) F.tot.push_back(0);
  W.tot = DevUpload(ctx, F.tot);
  // LDS use (layout as HashLdsInts on the device): KTslPod stages the
  // sampler's tables; KAdvance and KCollide the energy hash, then the
  // thermal-scattering totals; KTail the hash, the totals, then the sampler's
  // tables; each also the descriptor block, below
  ctx->hot_lds = W.hot_n > 0 && size_t(W.hot_n) * sizeof(double) <= kHotLdsMax
                     ? size_t(W.hot_n) * sizeof(double) : 0;
  const size_t hash_lds =
      ctx->n_hash0 > 0 ? sizeof(int) * size_t((ctx->n_hash0 + 2) & ~1) : 0;
  const size_t tot_lds = size_t(W.tot_n) * sizeof(double);
  {
    // the descriptor arrays, in one block (DevWorld::desc)
    std::vector<unsigned char> blk;
    auto add = [&](int k, const void* p, size_t bytes) {
      blk.resize((blk.size() + 15) & ~size_t(15));
      W.desc_off[k] = int(blk.size());
      if (bytes)
        blk.insert(blk.end(), static_cast<const unsigned char*>(p),
                   static_cast<const unsigned char*>(p) + bytes);
    };
    auto addv = [&](int k, const auto& v) { add(k, v.data(), v.size() * sizeof(v[0])); };
    addv(kDescSurfaces, surf);
    addv(kDescCells, cells);
    addv(kDescFaces, faces);
    addv(kDescFields, fields);
    addv(kDescMats, mats);
    addv(kDescComps, comps);
    addv(kDescCe, ce);
    addv(kDescRx, rx);
    addv(kDescTables, F.tables);
    addv(kDescGrids, F.grids);
    addv(kDescTsl, tsl);
    addv(kDescParts, parts);
    addv(kDescEsts, ests);
    // the cell locator, within what keeps a block that fits in LDS there
    const size_t base = (blk.size() + 15) & ~size_t(15);
    const size_t budget = base + 32 <= kDescLdsMax ? kDescLdsMax - base - 16 : 64 * 1024;
    std::vector<int> loc_off, loc_cells;
    BuildLocator(cells, budget, W, loc_off, loc_cells);
    addv(kDescLocOff, loc_off);
    addv(kDescLocCells, loc_cells);
    blk.resize(std::max<size_t>(16, (blk.size() + 15) & ~size_t(15)));
    unsigned char* d = DevUpload(ctx, blk);
    if (!d) return Fail(ctx, MMC_ERR_HIP, "descriptor upload failed");
    W.desc = d;
    W.desc_bytes = int(blk.size());
    W.surfaces = reinterpret_cast<const DSurface*>(d + W.desc_off[kDescSurfaces]);
    W.cells = reinterpret_cast<const DCell*>(d + W.desc_off[kDescCells]);
    W.faces = reinterpret_cast<const DFace*>(d + W.desc_off[kDescFaces]);
    W.fields = reinterpret_cast<const DField*>(d + W.desc_off[kDescFields]);
    W.mats = reinterpret_cast<const DMaterial*>(d + W.desc_off[kDescMats]);
    W.comps = reinterpret_cast<const DComp*>(d + W.desc_off[kDescComps]);
    W.ce = reinterpret_cast<const DCeNuclide*>(d + W.desc_off[kDescCe]);
    W.rx = reinterpret_cast<const DCeReaction*>(d + W.desc_off[kDescRx]);
    W.tables = reinterpret_cast<const DTable*>(d + W.desc_off[kDescTables]);
    W.grids = reinterpret_cast<const DGrid*>(d + W.desc_off[kDescGrids]);
    W.tsl = reinterpret_cast<const DTsl*>(d + W.desc_off[kDescTsl]);
    W.parts = reinterpret_cast<const DTslPart*>(d + W.desc_off[kDescParts]);
    W.ests = reinterpret_cast<const DEst*>(d + W.desc_off[kDescEsts]);
    W.loc_off = reinterpret_cast<const int*>(d + W.desc_off[kDescLocOff]);
    W.loc_cells = reinterpret_cast<const int*>(d + W.desc_off[kDescLocCells]);
  }
  // The staging kernels put the descriptor block last in their LDS (KAdvance,
  // KCollide: hash, totals, descriptors; KTail: all of them).  Their LDS variants need it to fit.  KTslPod stages only the
  // sampler's tables (hot_lds above).
  // Each LDS variant must fit beside the kernel's static LDS.
  const size_t desc_lds = size_t(W.desc_bytes) <= kDescLdsMax ? size_t(W.desc_bytes) : 0;
  const size_t ev_static = std::max(
      {StaticLds(reinterpret_cast<const void*>(KAdvance<true, false>)),
       StaticLds(reinterpret_cast<const void*>(KAdvance<true, true>)),
       StaticLds(reinterpret_cast<const void*>(KCollide<true>)),
       StaticLds(reinterpret_cast<const void*>(KAdvCol<false>)),
       StaticLds(reinterpret_cast<const void*>(KAdvCol<true>))});
  const size_t ev_static0 = std::max(
      {StaticLds(reinterpret_cast<const void*>(KAdvance<false, false>)),
       StaticLds(reinterpret_cast<const void*>(KAdvance<false, true>)),
       StaticLds(reinterpret_cast<const void*>(KCollide<false>))});
  const size_t tail_static =
      std::max(StaticLds(reinterpret_cast<const void*>(KTail<true>)),
               StaticLds(reinterpret_cast<const void*>(KTail<true, 512>)));
  if (ctx->hot_lds &&
      ctx->hot_lds + StaticLds(reinterpret_cast<const void*>(KTslPod<true>)) > kLdsPerBlock)
    ctx->hot_lds = 0;
  if (hash_lds + ev_static0 > kLdsPerBlock)
    return Fail(ctx, MMC_ERR_UNSUPPORTED, "energy hash does not fit in LDS");
  ctx->ev_tot_lds = desc_lds && tot_lds <= kTotLdsMax &&
                    hash_lds + tot_lds + desc_lds + ev_static <= kLdsPerBlock;
  ctx->ev_lds = ctx->ev_tot_lds ? hash_lds + tot_lds + desc_lds : hash_lds;
  {
    const size_t hot = size_t(W.hot_n) * sizeof(double);
    const size_t need = hash_lds + tot_lds + hot + desc_lds;
    ctx->tail_lds =
        desc_lds && hot <= kHotLdsMax && need + tail_static <= kLdsPerBlock ? need : 0;
    const size_t park_static =
        std::max(StaticLds(reinterpret_cast<const void*>(KTail<true, 256, true>)),
                 StaticLds(reinterpret_cast<const void*>(KTail<true, 512, true>)));
    const size_t park = hash_lds + tot_lds + desc_lds;
    ctx->tail_park_lds = desc_lds && tot_lds <= kTotLdsMax &&
                                 park + park_static <= kLdsPerBlock ? park : 0;
  }
  ctx->S = DevAlloc<double>(ctx, nb);
  ctx->S2 = DevAlloc<double>(ctx, nb);
  ctx->bin_sum = DevAlloc<double>(ctx, nb);
  ctx->bin_sq = DevAlloc<double>(ctx, nb);
  HIPCHK(hipMemset(ctx->S, 0, sizeof(double) * std::max(nb, 1)));
  HIPCHK(hipMemset(ctx->S2, 0, sizeof(double) * std::max(nb, 1)));
  // fixed-source problems with fission keep half the slots for children
  if (W.energy_mode == MMC_MULTIGROUP) {
    for (double v : fis) ctx->fixed_fission |= v > 0;
  } else {
    for (const auto& n : ce) ctx->fixed_fission |= n.fissile != 0;
  }
  return MMC_OK;
}

// Score record buffers (RunState::rec_key, rec_val) and the reduction's work
// space for n records.  A k-eigenvalue cycle that scores more grows them and
// runs again.
static int EnsureRecords(mmc_ctx* ctx, uint32_t n) {
  RunState& R = ctx->R;
  if (R.rec_key && R.rec_cap >= n) return MMC_OK;
  if (R.rec_key) (void)hipFree(R.rec_key);
  if (R.rec_val) (void)hipFree(R.rec_val);
  R.rec_key = nullptr;
  R.rec_val = nullptr;
  R.rec_cap = 0;
  if (hipMalloc(reinterpret_cast<void**>(&R.rec_key), sizeof(uint64_t) * size_t(n)) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&R.rec_val), sizeof(double) * size_t(n)) !=
          hipSuccess ||
      TallyWorkEnsure(ctx->tally, n))
    return Fail(ctx, MMC_ERR_HIP, "out of device memory (score records)");
  R.rec_cap = n;
  return MMC_OK;
}

static int AllocBank(mmc_ctx* ctx, uint32_t cap) {
  if (ctx->cap >= cap) return MMC_OK;
  if (ctx->cap) return Fail(ctx, MMC_ERR_ARG, "bank already sized");
  ctx->cap = cap;
  RunState& R = ctx->R;
  DevBank& B = R.B;
  B.x = DevAlloc<double>(ctx, cap);
  B.y = DevAlloc<double>(ctx, cap);
  B.z = DevAlloc<double>(ctx, cap);
  B.u = DevAlloc<double>(ctx, cap);
  B.v = DevAlloc<double>(ctx, cap);
  B.w = DevAlloc<double>(ctx, cap);
  B.E = DevAlloc<double>(ctx, cap);
  B.rng = DevAlloc<uint32_t>(ctx, cap);
  B.cell = DevAlloc<int32_t>(ctx, cap);
  B.hist = DevAlloc<uint32_t>(ctx, cap);
  B.steps = DevAlloc<uint32_t>(ctx, cap);
  B.k_col = DevAlloc<double>(ctx, cap);
  B.k_trk = DevAlloc<double>(ctx, cap);
  B.k_abs = DevAlloc<double>(ctx, cap);
  B.tslT = DevAlloc<double>(ctx, cap);
  B.tslI = DevAlloc<int32_t>(ctx, cap);
  B.xs_grid = DevAlloc<int32_t>(ctx, cap);
  B.xs_gi = DevAlloc<int32_t>(ctx, cap);
  B.ie = ctx->W.n_pert ? DevAlloc<double>(ctx, size_t(cap) * ctx->W.n_pert) : nullptr;
  // the compaction target: the moved fields get their own arrays, the
  // history-indexed k sums are shared
  DevBank& A = ctx->alt;
  A = B;
  A.x = DevAlloc<double>(ctx, cap);
  A.y = DevAlloc<double>(ctx, cap);
  A.z = DevAlloc<double>(ctx, cap);
  A.u = DevAlloc<double>(ctx, cap);
  A.v = DevAlloc<double>(ctx, cap);
  A.w = DevAlloc<double>(ctx, cap);
  A.E = DevAlloc<double>(ctx, cap);
  A.rng = DevAlloc<uint32_t>(ctx, cap);
  A.cell = DevAlloc<int32_t>(ctx, cap);
  A.hist = DevAlloc<uint32_t>(ctx, cap);
  A.steps = DevAlloc<uint32_t>(ctx, cap);
  A.xs_grid = DevAlloc<int32_t>(ctx, cap);
  A.xs_gi = DevAlloc<int32_t>(ctx, cap);
  A.tslT = DevAlloc<double>(ctx, cap);
  A.tslI = DevAlloc<int32_t>(ctx, cap);
  A.ie = ctx->W.n_pert ? DevAlloc<double>(ctx, size_t(cap) * ctx->W.n_pert) : nullptr;
  R.Q[0] = DevAlloc<uint32_t>(ctx, cap);
  R.Q[1] = DevAlloc<uint32_t>(ctx, cap);
  R.Q2[0] = DevAlloc<uint32_t>(ctx, cap);
  R.Q2[1] = DevAlloc<uint32_t>(ctx, cap);
  R.QC = DevAlloc<uint32_t>(ctx, cap);
  R.QT = DevAlloc<uint32_t>(ctx, cap);
  // misc follows the ring, so that one copy reads both
  R.ring = DevAlloc<uint32_t>(ctx, kRing * kRingStride + 16);
  R.misc = R.ring ? R.ring + kRing * kRingStride : nullptr;
  if (hipHostMalloc(reinterpret_cast<void**>(&ctx->host_rb),
                    sizeof(uint32_t) * (kRing * kRingStride + 16)) != hipSuccess)
    ctx->host_rb = nullptr;
  R.ev = DevAlloc<unsigned long long>(ctx, kEvWords);
  R.capacity = cap;
  if (const int rr = EnsureRecords(ctx, cap * 4 + 1024)) return rr;
  R.fis_cnt = DevAlloc<uint32_t>(ctx, cap);
  ctx->fis_off = DevAlloc<uint32_t>(ctx, cap);
  ctx->scan_cap = (cap + 1023) / 1024 + 1;
  ctx->scan_tmp = DevAlloc<uint32_t>(ctx, ctx->scan_cap);
  ctx->kpart_cap = Blocks(cap, 1u << 30) * 3;
  ctx->kpart = DevAlloc<double>(ctx, ctx->kpart_cap);
  if (!ctx->kpart || !ctx->host_rb || !R.Q2[0] || !R.Q2[1] || !B.tslI || !B.xs_grid || !B.xs_gi || !A.x ||
      !A.y || !A.z || !A.u ||
      !A.v || !A.w || !A.E || !A.rng || !A.cell || !A.hist || !A.steps ||
      !A.xs_grid || !A.xs_gi || !A.tslT || !A.tslI)
    return Fail(ctx, MMC_ERR_HIP, "out of device memory");
  HIPCHK(hipMemset(R.ev, 0, sizeof(unsigned long long) * kEvWords));
  R.max_steps = uint32_t(ctx->opt.max_steps ? std::min<uint64_t>(ctx->opt.max_steps, 0xffffffffull)
                                            : 1000000ull);
  return MMC_OK;
}

extern "C" int mmc_create(const mmc_world_desc* world, const mmc_tally_desc* tallies,
                          const mmc_options* options, mmc_ctx** out) {
  if (!world || !out) return MMC_ERR_ARG;
  *out = nullptr;
  auto* ctx = new mmc_ctx();
  if (options) ctx->opt = *options;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= ctx->opt.device) {
    delete ctx;
    return MMC_ERR_NODEV;
  }
  ctx->device = ctx->opt.device;
  if (ctx->opt.tail_block != 0 && ctx->opt.tail_block != 256 && ctx->opt.tail_block != 512) {
    delete ctx;
    return MMC_ERR_ARG;
  }
  if (ctx->opt.tsl_block != 0 && ctx->opt.tsl_block != 256 && ctx->opt.tsl_block != 384) {
    delete ctx;
    return MMC_ERR_ARG;
  }
  ctx->tail_block = ctx->opt.tail_block ? ctx->opt.tail_block : 256;
  if (ctx->opt.tail_network < 0) {
    *out = ctx;
    return Fail(ctx, MMC_ERR_ARG, "tail_network must be 0 or a particle count");
  }
  if (ctx->opt.split_collide != 0 && ctx->opt.split_collide != 1) {
    *out = ctx;
    return Fail(ctx, MMC_ERR_ARG, "split_collide must be 0 or 1");
  }
  ctx->tsl_block = ctx->opt.tsl_block ? ctx->opt.tsl_block : 256;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return MMC_ERR_HIP;
  }
  const int rc = BuildWorld(ctx, world, tallies);
  if (rc) {
    *out = ctx;  // keep the message available through mmc_last_error
    return rc;
  }
  // the no-allocation sampler's static scratch: its address on this device
  if (MlpScratchInit(ctx->device)) {
    *out = ctx;
    return Fail(ctx, MMC_ERR_HIP, "no address for the static scratch");
  }
  // The CE kernels stage the energy hash (up to 32 KiB) into LDS once per
  // block: cap their grid at the blocks that are resident at once, so that
  // each block amortises the staging over several grid-stride passes.
  int n_cu = 0;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) ==
          hipSuccess && n_cu > 0) {
    const size_t lds = ctx->ev_lds;
    int occ_a = 0, occ_c = 0, occ_m = 0;
    hipError_t ea, ec, em;
    if (ctx->ev_tot_lds) {
      ea = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_a, KAdvance<true, false>, kBlock,
                                                        lds);
      em = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_m, KAdvance<true, true>, kBlock,
                                                        lds);
      ec = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_c, KCollide<true>, kBlock, lds);
      int occ_f = 0;
      if (ec == hipSuccess &&
          hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, KAdvCol<true>, kBlock, lds) ==
              hipSuccess)
        occ_c = std::min(occ_c, occ_f);
    } else {
      ea = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_a, KAdvance<false, false>, kBlock,
                                                        lds);
      em = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_m, KAdvance<false, true>, kBlock,
                                                        lds);
      ec = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_c, KCollide<false>, kBlock, lds);
    }
    if (ea != hipSuccess || ec != hipSuccess || em != hipSuccess) occ_a = occ_c = occ_m = 2;
    ctx->lds_grid =
        uint32_t(std::max(1, std::min(std::min(occ_a, occ_m), occ_c))) * uint32_t(n_cu);
    // the tail kernel's lanes take the next particle when one ends: launch
    // only the blocks that are resident at once
    int occ_t = 0;
    const int tb = ctx->tail_block;
    const hipError_t et =
        tb == 512 ? (ctx->tail_lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                         &occ_t, KTail<true, 512>, tb, ctx->tail_lds)
                                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                         &occ_t, KTail<false, 512>, tb, 0))
                  : (ctx->tail_lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                         &occ_t, KTail<true>, tb, ctx->tail_lds)
                                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                         &occ_t, KTail<false>, tb, 0));
    if (et != hipSuccess) occ_t = 1;
    ctx->tail_grid = uint32_t(std::max(1, occ_t)) * uint32_t(n_cu);
    int occ_p = 0;
    const size_t pl = ctx->tail_park_lds;
    const hipError_t ep =
        tb == 512 ? (pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &occ_p, KTail<true, 512, true>, tb, pl)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &occ_p, KTail<false, 512, true>, tb, 0))
                  : (pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &occ_p, KTail<true, 256, true>, tb, pl)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &occ_p, KTail<false, 256, true>, tb, 0));
    if (ep != hipSuccess) occ_p = 1;
    ctx->tail_park_grid = uint32_t(std::max(1, occ_p)) * uint32_t(n_cu);
    // KTslPod with the hot block in LDS: resident blocks only
    int occ_s = 0;
    if (ctx->hot_lds &&
        (ctx->tsl_block == 384
             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, KTslPod<true, 384>, 384,
                                                            ctx->hot_lds)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, KTslPod<true>, kBlock,
                                                            ctx->hot_lds)) != hipSuccess)
      occ_s = 1;
    ctx->tsl_grid = uint32_t(std::max(1, occ_s)) * uint32_t(n_cu);
  }
  uint64_t cap = ctx->opt.max_in_flight ? ctx->opt.max_in_flight : (1ull << 20);
  cap = std::min<uint64_t>(std::max<uint64_t>(cap, 1024), 1ull << 28);
  const int rb = AllocBank(ctx, uint32_t(cap));
  *out = ctx;
  return rb;
}

extern "C" void mmc_destroy(mmc_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto& g : ctx->graphs) hipGraphExecDestroy(g.second.exec);
  for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
  for (void* p : ctx->allocs) hipFree(p);
  if (ctx->R.rec_key) hipFree(ctx->R.rec_key);
  if (ctx->R.rec_val) hipFree(ctx->R.rec_val);
  TallyWorkFree(ctx->tally);
  if (ctx->host_rb) hipHostFree(ctx->host_rb);
  MlpFree(ctx->mlp);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* mmc_last_error(const mmc_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

// ---------------------------------------------------------------------------
// scan (exclusive, u32) and small utility kernels
__global__ void KScanBlock(const uint32_t* in, uint32_t* out, uint32_t* sums, uint32_t n) {
  __shared__ uint32_t s[1024];
  const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  const uint32_t v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 1024; off <<= 1) {
    const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}
__global__ void KScanTop(uint32_t* sums, uint32_t nb) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nb ? sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = s[threadIdx.x] - v + carry;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
}
__global__ void KScanAdd(uint32_t* out, const uint32_t* sums, uint32_t n) {
  const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
}
static int Scan(mmc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n) {
  const uint32_t nb = (n + 1023) / 1024;
  if (nb + 1 > ctx->scan_cap) return Fail(ctx, MMC_ERR_OVERFLOW, "scan too large");
  hipLaunchKernelGGL(KScanBlock, dim3(nb), dim3(1024), 0, ctx->stream, in, out,
                     ctx->scan_tmp, n);
  hipLaunchKernelGGL(KScanTop, dim3(1), dim3(1024), 0, ctx->stream, ctx->scan_tmp, nb);
  hipLaunchKernelGGL(KScanAdd, dim3(nb), dim3(1024), 0, ctx->stream, out, ctx->scan_tmp, n);
  return MMC_OK;
}
__global__ void KFillU32(uint32_t* p, uint32_t v, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void KFillF64(double* p, double v, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}
static void FillU32(mmc_ctx* ctx, uint32_t* p, uint32_t v, uint32_t n) {
  hipLaunchKernelGGL(KFillU32, dim3(Blocks(n, 2048)), dim3(kBlock), 0, ctx->stream, p, v, n);
}
static void FillF64(mmc_ctx* ctx, double* p, double v, uint32_t n) {
  hipLaunchKernelGGL(KFillF64, dim3(Blocks(n, 2048)), dim3(kBlock), 0, ctx->stream, p, v, n);
}

// ---- tally commit (ScorableProxy::CommitHistory, Scorable.cpp:101-106;
// EstimatorSet::operator+=, Estimator.cpp:215-231): a batch's per-bin sums of
// s_{h,b} and s_{h,b}^2 (TallyReduce, in a fixed order) join S and S2.
__global__ void KFoldFixed(double* S, double* S2, const double* sum, const double* sq,
                           uint32_t nb) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  S[b] += sum[b];
  S2[b] += sq[b];
}

// ---------------------------------------------------------------------------
// Event loop
static int CheckMisc(mmc_ctx* ctx, const uint32_t* misc) {
  const uint32_t e = misc[0];
  if (!e) return MMC_OK;
  if (e & kErrLost)
    return Fail(ctx, MMC_ERR_LOST,
                "Point does not belong to any Cell. Please checkthat all space is "
                "either assigned a material or void.");
  if (e & kErrResample)
    return Fail(ctx, MMC_ERR_RESAMPLE,
                "Number of thermal scattering resamples exceeded limit: 10");
  if (e & kErrOverflow) return Fail(ctx, MMC_ERR_OVERFLOW, "bank, queue or fission-site overflow");
  if (e & kErrRecords) return Fail(ctx, MMC_ERR_OVERFLOW, "score record overflow");
  if (e & kErrSteps) return Fail(ctx, MMC_ERR_STEPS, "history step limit exceeded");
  if (e & kErrReaction)
    return Fail(ctx, MMC_ERR_ARG, "no reaction selected after 1000 resamples");
  return Fail(ctx, MMC_ERR_ARG, "alpha sampling outside the beta grid");
}

// Kernel timing brackets (profiling mode only).
static void ProfBegin(mmc_ctx* ctx, int kind) {
  if (!ctx->prof) return;
  if (ctx->ev_used + 2 > ctx->ev_pool.size()) {
    for (int k = 0; k < 64; k++) {
      hipEvent_t e;
      hipEventCreate(&e);
      ctx->ev_pool.push_back(e);
    }
  }
  hipEventRecord(ctx->ev_pool[ctx->ev_used], ctx->stream);
  ctx->ev_kind.push_back(kind);
}
static void ProfEnd(mmc_ctx* ctx) {
  if (!ctx->prof) return;
  hipEventRecord(ctx->ev_pool[ctx->ev_used + 1], ctx->stream);
  ctx->ev_used += 2;
}
// Bracket helper for launch sequences owned by other files: kind >= 0 closes
// any open bracket and opens one of that kind, -1 closes.
static void ProfMark(void* p, int kind) {
  mmc_ctx* ctx = static_cast<mmc_ctx*>(p);
  if (!ctx->prof) return;
  if (ctx->ev_kind.size() * 2 > ctx->ev_used) ProfEnd(ctx);
  if (kind >= 0) ProfBegin(ctx, kind);
}
// After a stream synchronisation: fold the recorded pairs into the totals.
static void ProfCollect(mmc_ctx* ctx) {
  if (!ctx->prof) return;
  for (size_t i = 0; i < ctx->ev_kind.size(); i++) {
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev_pool[2 * i], ctx->ev_pool[2 * i + 1]);
    ctx->k_ms[ctx->ev_kind[i]] += ms;
    ctx->k_launches[ctx->ev_kind[i]]++;
  }
  ctx->ev_kind.clear();
  ctx->ev_used = 0;
}

// Exchanges the particle fields that KCompact moves between the two banks.
static void SwapBanks(mmc_ctx* ctx) {
  DevBank& B = ctx->R.B;
  DevBank& A = ctx->alt;
  std::swap(B.x, A.x); std::swap(B.y, A.y); std::swap(B.z, A.z);
  std::swap(B.u, A.u); std::swap(B.v, A.v); std::swap(B.w, A.w);
  std::swap(B.E, A.E);
  std::swap(B.rng, A.rng);
  std::swap(B.cell, A.cell);
  std::swap(B.hist, A.hist);
  std::swap(B.steps, A.steps);
  std::swap(B.xs_grid, A.xs_grid);
  std::swap(B.xs_gi, A.xs_gi);
  std::swap(B.tslT, A.tslT);
  std::swap(B.tslI, A.tslI);
  std::swap(B.ie, A.ie);
}

// move: KAdvance compacts the bank into ctx->alt (with the parked fallback
// particles fb), and the banks are swapped before the rest of the iteration.
// k-eigenvalue iterations with the tables in LDS run KAdvCol, which also
// collides its particles; the others run KAdvance, then KCollide.
static int LaunchTsl(mmc_ctx* ctx, int iter, int qin, uint32_t grid);
static int LaunchIteration(mmc_ctx* ctx, int iter, int qin, uint32_t grid, bool move,
                           uint32_t* fb) {
  const size_t lds = ctx->ev_lds;
  const uint32_t sgrid = lds ? std::min(grid, ctx->lds_grid) : grid;
  ProfBegin(ctx, MMC_K_ADVANCE);
  const dim3 ag(sgrid), ab(kBlock);
  if (ctx->R.keig && ctx->ev_tot_lds && !ctx->opt.split_collide) {
    // k-eigenvalue: the advance kernel collides its particles itself
    if (move)
      hipLaunchKernelGGL((KAdvCol<true>), ag, ab, lds, ctx->stream, ctx->W, ctx->R, iter, qin,
                         ctx->alt, fb);
    else
      hipLaunchKernelGGL((KAdvCol<false>), ag, ab, lds, ctx->stream, ctx->W, ctx->R, iter, qin,
                         ctx->R.B, nullptr);
    ProfEnd(ctx);
    if (move) SwapBanks(ctx);
    return LaunchTsl(ctx, iter, qin, grid);
  }
  if (ctx->ev_tot_lds) {
    if (move)
      hipLaunchKernelGGL((KAdvance<true, true>), ag, ab, lds, ctx->stream, ctx->W, ctx->R, iter,
                         qin, ctx->alt, fb);
    else
      hipLaunchKernelGGL((KAdvance<true, false>), ag, ab, lds, ctx->stream, ctx->W, ctx->R,
                         iter, qin, ctx->R.B, nullptr);
  } else {
    if (move)
      hipLaunchKernelGGL((KAdvance<false, true>), ag, ab, lds, ctx->stream, ctx->W, ctx->R,
                         iter, qin, ctx->alt, fb);
    else
      hipLaunchKernelGGL((KAdvance<false, false>), ag, ab, lds, ctx->stream, ctx->W, ctx->R,
                         iter, qin, ctx->R.B, nullptr);
  }
  ProfEnd(ctx);
  if (move) SwapBanks(ctx);
  ProfBegin(ctx, MMC_K_COLLIDE);
  if (ctx->ev_tot_lds)
    hipLaunchKernelGGL(KCollide<true>, dim3(sgrid), dim3(kBlock), lds, ctx->stream, ctx->W,
                       ctx->R, iter, qin);
  else
    hipLaunchKernelGGL(KCollide<false>, dim3(sgrid), dim3(kBlock), lds, ctx->stream, ctx->W,
                       ctx->R, iter, qin);
  ProfEnd(ctx);
  return LaunchTsl(ctx, iter, qin, grid);
}

// The thermal-scattering step of iteration `iter`: the network or the POD
// sampler for the collisions queued in QT.
static int LaunchTsl(mmc_ctx* ctx, int iter, int qin, uint32_t grid) {
  if (ctx->has_tsl) {
    if (ctx->opt.tsl_mode == MMC_TSL_SURROGATE && ctx->mlp.ready) {
      const int rc = MlpTslStep(ctx->mlp, ctx->W, ctx->R, iter, qin, grid, ctx->stream,
                                ProfMark, ctx);
      if (rc) return Fail(ctx, MMC_ERR_HIP, "surrogate step failed");
    } else {
      ProfBegin(ctx, MMC_K_TSL);
      if (ctx->hot_lds && ctx->tsl_block == 384)
        hipLaunchKernelGGL((KTslPod<true, 384>), dim3(std::min(grid, ctx->tsl_grid)), dim3(384),
                           ctx->hot_lds, ctx->stream, ctx->W, ctx->R, iter, qin);
      else if (ctx->hot_lds)
        hipLaunchKernelGGL(KTslPod<true>, dim3(std::min(grid, ctx->tsl_grid)), dim3(kBlock),
                           ctx->hot_lds, ctx->stream, ctx->W, ctx->R, iter, qin);
      else
        hipLaunchKernelGGL(KTslPod<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->W,
                           ctx->R, iter, qin);
      ProfEnd(ctx);
    }
  }
  return MMC_OK;
}

extern "C" int mmc_set_profiling(mmc_ctx* ctx, int enable) {
  if (!ctx) return MMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ProfCollect(ctx);
  ctx->prof = enable != 0;
  for (int k = 0; k < MMC_K_N; k++) {
    ctx->k_ms[k] = 0;
    ctx->k_launches[k] = 0;
  }
  return MMC_OK;
}
extern "C" int mmc_get_kernel_times(mmc_ctx* ctx, double* total_ms, uint64_t* launches) {
  if (!ctx || !total_ms || !launches) return MMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ProfCollect(ctx);
  for (int k = 0; k < MMC_K_N; k++) {
    total_ms[k] = ctx->k_ms[k];
    launches[k] = ctx->k_launches[k];
  }
  return MMC_OK;
}


// One batch of K event iterations from iteration `iter`, then the POD
// sampler for the collisions the surrogate parked, then the copy of the ring
// and counters into pinned host memory (the caller synchronises).  `swaps`
// counts the bank swaps of the batch's compactions.
static int LaunchBatch(mmc_ctx* ctx, int& iter, int& qin, uint32_t grid, int K,
                       bool compact_ok, int every, bool surrogate, int& swaps) {
  for (int k = 0; k < K; k++) {
    // compaction every iteration is done by the advance kernel itself
    const bool move = compact_ok && every == 1 && iter > 0;
    if (compact_ok && every > 1 && iter > 0 && iter % every == 0) {
      hipLaunchKernelGGL(KCompact, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->R,
                         ctx->alt, iter, qin, surrogate ? ctx->mlp.fb : nullptr,
                         ctx->W.n_pert);
      SwapBanks(ctx);
      swaps++;
    }
    const int rc =
        LaunchIteration(ctx, iter, qin, grid, move, surrogate ? ctx->mlp.fb : nullptr);
    if (rc) return rc;
    if (move) swaps++;
    iter++;
    qin ^= 1;
  }
  if (surrogate) {  // parked fallback particles rejoin before the count is read
    ProfBegin(ctx, MMC_K_TSL);
    const int rc =
        MlpFlushFallback(ctx->mlp, ctx->W, ctx->R, iter, qin, ctx->hot_lds, ctx->stream);
    ProfEnd(ctx);
    if (rc) return Fail(ctx, MMC_ERR_HIP, "surrogate fallback failed");
  }
  if (const int lc = LaunchCheck(ctx, "event kernels")) return lc;
  HIPCHK(hipMemcpyAsync(ctx->host_rb, ctx->R.ring,
                        sizeof(uint32_t) * (kRing * kRingStride + 16), hipMemcpyDeviceToHost,
                        ctx->stream));
  return MMC_OK;
}

// The grid of a captured batch: the next power of two, so that a few graphs
// serve every live count (the kernels grid-stride over their queues and
// blocks without work return at once).
static uint32_t GraphGrid(uint32_t grid) {
  uint32_t g = 1;
  while (g < grid) g <<= 1;
  return std::min<uint32_t>(g, 4096u);
}

// LaunchBatch as a HIP graph: captured the first time its key is seen, then
// replayed.  A replay makes the host-side changes the capture made (the bank
// swaps, iteration and queue parity) without launching anything itself.
static int GraphBatch(mmc_ctx* ctx, int& iter, int& qin, uint32_t grid, int K,
                      bool compact_ok, int every, bool surrogate) {
  const std::array<uint64_t, 8> key = {
      uint64_t(iter % kRing), uint64_t(qin), grid,
      uint64_t(reinterpret_cast<uintptr_t>(ctx->R.B.x)),
      uint64_t(reinterpret_cast<uintptr_t>(ctx->R.rec_key)), ctx->R.rec_cap,
      uint64_t(reinterpret_cast<uintptr_t>(ctx->mlp.packed)) ^
          (uint64_t(reinterpret_cast<uintptr_t>(ctx->mlp.fb)) << 1),
      uint64_t(surrogate) | uint64_t(ctx->prof) << 1 | uint64_t(every + 1) << 2};
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end()) {
    mmc_ctx::GraphEntry ent;
    const size_t kinds0 = ctx->ev_kind.size();
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int swaps = 0;
    const int rc = LaunchBatch(ctx, iter, qin, grid, K, compact_ok, every, surrogate, swaps);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      return rc ? rc : Fail(ctx, MMC_ERR_HIP, std::string("graph capture: ") +
                                                  hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&ent.exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess)
      return Fail(ctx, MMC_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    ent.kinds.assign(ctx->ev_kind.begin() + kinds0, ctx->ev_kind.end());
    ent.ev_used = ctx->ev_used;
    ent.swaps = swaps;
    it = ctx->graphs.emplace(key, std::move(ent)).first;
    ctx->graph_captures++;
  } else {
    for (int s = 0; s < it->second.swaps; s++) SwapBanks(ctx);
    iter += K;
    qin ^= K & 1;
    if (ctx->prof) {
      // the graph records the same events, from the start of the pool
      ctx->ev_kind = it->second.kinds;
      ctx->ev_used = it->second.ev_used;
    }
  }
  HIPCHK(hipGraphLaunch(it->second.exec, ctx->stream));
  ctx->graph_launches++;
  return MMC_OK;
}

// One launch of the tail kernel over the live queue of iteration iter - 1 (at
// most `bound` particles); park: a round that parks thermal collisions for
// the network instead of sampling them.
template <bool kLds, int kThreads, bool kPark>
static void TailLaunch(mmc_ctx* ctx, dim3 grid, int iter, int qin) {
  const size_t lds = !kLds ? 0 : kPark ? ctx->tail_park_lds : ctx->tail_lds;
  hipLaunchKernelGGL((KTail<kLds, kThreads, kPark>), grid, dim3(kThreads), lds, ctx->stream,
                     ctx->W, ctx->R, iter, qin);
}
template <int kThreads, bool kPark>
static void TailLaunchT(mmc_ctx* ctx, dim3 grid, int iter, int qin) {
  if (kPark ? ctx->tail_park_lds : ctx->tail_lds)
    TailLaunch<true, kThreads, kPark>(ctx, grid, iter, qin);
  else
    TailLaunch<false, kThreads, kPark>(ctx, grid, iter, qin);
}
static void LaunchTail(mmc_ctx* ctx, int iter, int qin, uint32_t bound, bool park) {
  const uint32_t tb = uint32_t(ctx->tail_block);
  const uint32_t resident = park ? ctx->tail_park_grid : ctx->tail_grid;
  const dim3 grid(std::max(1u, std::min(uint32_t((bound + tb - 1) / tb), resident)));
  ProfBegin(ctx, MMC_K_TAIL);
  if (tb == 512)
    park ? TailLaunchT<512, true>(ctx, grid, iter, qin)
         : TailLaunchT<512, false>(ctx, grid, iter, qin);
  else
    park ? TailLaunchT<256, true>(ctx, grid, iter, qin)
         : TailLaunchT<256, false>(ctx, grid, iter, qin);
  ProfEnd(ctx);
}

// Tail rounds with the network (surrogate mode): each round runs the live
// particles to their next thermal collision (parked) or to the end of their
// history, then the network kernel scatters the parked ones into the next
// round's live queue.  Rounds run in groups of K, each group ending with the
// POD sampler for the network's fallbacks and a read of the live count,
// until fewer than `until` particles are live; `bound` is then the live
// count.
static int TailRounds(mmc_ctx* ctx, int& iter, int& qin, uint32_t& bound, uint32_t until,
                      int K) {
  uint32_t* const ring = ctx->host_rb;
  uint32_t* const misc = ctx->host_rb + kRing * kRingStride;
  while (bound >= until && bound > 0) {
    for (int k = 0; k < K; k++) {
      LaunchTail(ctx, iter, qin, bound, true);
      if (MlpTslStep(ctx->mlp, ctx->W, ctx->R, iter, qin, 0, ctx->stream, ProfMark, ctx))
        return Fail(ctx, MMC_ERR_HIP, "surrogate step failed");
      iter++;
      qin ^= 1;
    }
    ProfBegin(ctx, MMC_K_TSL);
    const int rf =
        MlpFlushFallback(ctx->mlp, ctx->W, ctx->R, iter, qin, ctx->hot_lds, ctx->stream);
    ProfEnd(ctx);
    if (rf) return Fail(ctx, MMC_ERR_HIP, "surrogate fallback failed");
    if (const int lc = LaunchCheck(ctx, "tail rounds")) return lc;
    HIPCHK(hipMemcpyAsync(ring, ctx->R.ring, sizeof(uint32_t) * (kRing * kRingStride + 16),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ProfCollect(ctx);
    ctx->cnt.iterations += K;
    const int rc = CheckMisc(ctx, misc);
    if (rc) return rc;
    const uint32_t* row = ring + ((iter - 1 + kRing) % kRing) * kRingStride;
    bound = row[kQ] + row[kQ2];
  }
  return MMC_OK;
}

// Runs event iterations until no particle is alive.  The live-queue length
// is read back every `K` iterations; in between, kernels read queue lengths
// from device memory and grid-stride over them.
static int RunEvents(mmc_ctx* ctx, uint32_t n0) {
  const int K = ctx->opt.graph_iters > 0 ? ctx->opt.graph_iters : 8;
  const bool surrogate = ctx->has_tsl && ctx->opt.tsl_mode == MMC_TSL_SURROGATE;
  if (surrogate) {
    if (!ctx->mlp.ready)
      return Fail(ctx, MMC_ERR_UNSUPPORTED, "surrogate mode without a network (mmc_mlp_set)");
    const int rc = MlpEnsureStaging(ctx, ctx->cap);
    if (rc) return rc;
  }
  uint32_t bound = n0;
  int iter = 0, qin = 0;
  uint32_t* const ring = ctx->host_rb;  // pinned: the ring, then misc
  uint32_t* const misc = ctx->host_rb + kRing * kRingStride;
  constexpr size_t kMiscBytes = 16 * sizeof(uint32_t);
  // history-based tail (KTail) once few particles are left, where no
  // collision can start a new particle
  const bool tail_ok = ctx->opt.tail_threshold >= 0 && (ctx->R.keig || !ctx->fixed_fission);
  const uint32_t tail_at =
      ctx->opt.tail_threshold > 0
          ? uint32_t(ctx->opt.tail_threshold)
          : std::max<uint32_t>(1024u, std::min<uint32_t>(ctx->cap / 32u, 32768u));
  // compaction (KCompact): by default whenever the live particles fill less
  // than half of the slots in use, else every compact_interval iterations;
  // only where no collision allocates new slots
  const int every = ctx->opt.compact_interval;
  const bool compact_ok = every >= 0 && (ctx->R.keig || !ctx->fixed_fission);
  // Captured batches: k-eigenvalue cycles (the launch arguments repeat from
  // cycle to cycle), with compaction every iteration, at readbacks or never
  // (KCompact every k > 1 iterations would tie a batch to iter mod k).
  const bool graphs = ctx->opt.graphs > 0 && ctx->R.keig && every <= 1;
  // the tail's thermal collisions go to the network while at least
  // net_tail_at particles are live (an option: every round waits for its
  // slowest particle, and the rounds cost more than the network saves on
  // the benchmark problems)
  const bool net_tail = surrogate && ctx->opt.tail_network > 0;
  const uint32_t net_tail_at = uint32_t(std::max(ctx->opt.tail_network, 1));
  uint32_t in_use = n0;
  for (;;) {
    if (tail_ok && bound < tail_at && iter > 0) {
      if (net_tail) {
        const int rc = TailRounds(ctx, iter, qin, bound, net_tail_at, K);
        if (rc) return rc;
        if (bound == 0) break;
      }
      LaunchTail(ctx, iter, qin, bound, false);
      if (const int lc = LaunchCheck(ctx, "tail kernel")) return lc;
      HIPCHK(hipMemcpyAsync(misc, ctx->R.misc, kMiscBytes, hipMemcpyDeviceToHost,
                            ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ProfCollect(ctx);
      const int rc = CheckMisc(ctx, misc);
      if (rc) return rc;
      break;
    }
    if (graphs && iter > 0) {
      const int rc = GraphBatch(ctx, iter, qin, GraphGrid(Blocks(bound)), K, compact_ok,
                                every, surrogate);
      if (rc) return rc;
    } else {
      int swaps = 0;
      const int rc =
          LaunchBatch(ctx, iter, qin, Blocks(bound), K, compact_ok, every, surrogate, swaps);
      if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ProfCollect(ctx);
    ctx->cnt.iterations += K;
    const int rc = CheckMisc(ctx, misc);
    if (rc) return rc;
    const uint32_t* row = ring + ((iter - 1 + kRing) % kRing) * kRingStride;
    const uint32_t live = row[kQ] + row[kQ2];  // live queue: Q, then Q2
    if (live == 0) break;
    bound = ctx->fixed_fission && !ctx->R.keig ? ctx->cap : live;
    if (compact_ok && every == 0 && 2 * live < in_use) {
      hipLaunchKernelGGL(KCompact, dim3(Blocks(live)), dim3(kBlock), 0, ctx->stream, ctx->R,
                         ctx->alt, iter, qin, surrogate ? ctx->mlp.fb : nullptr,
                         ctx->W.n_pert);
      SwapBanks(ctx);
      in_use = live;
    }
  }
  // (the next run resets the ring: ResetRun)
  return MMC_OK;
}

static int ResetRun(mmc_ctx* ctx) {
  HIPCHK(hipMemsetAsync(ctx->R.ring, 0, sizeof(uint32_t) * kRing * kRingStride, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->R.misc, 0, sizeof(uint32_t) * 16, ctx->stream));
  return MMC_OK;
}

static int ReadEvents(mmc_ctx* ctx) {
  unsigned long long copies[kEvWords];
  HIPCHK(hipMemcpyAsync(copies, ctx->R.ev, sizeof(copies), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < kEvN; k++) {
    unsigned long long s = 0;
    for (int c = 0; c < kEvSlots; c++) s += copies[c * kEvStride + k];
    ctx->ev_host[k] = s;
  }
  ctx->cnt.advances = ctx->ev_host[kEvAdvance];
  ctx->cnt.xs_lookups = ctx->ev_host[kEvAdvance];
  ctx->cnt.collisions = ctx->ev_host[kEvCollide];
  ctx->cnt.crossings = ctx->ev_host[kEvCross];
  ctx->cnt.virtual_collisions = ctx->ev_host[kEvVirtual];
  ctx->cnt.tsl_collisions = ctx->ev_host[kEvTsl];
  ctx->cnt.tsl_fallbacks = ctx->ev_host[kEvFallback];
  ctx->cnt.tail_particles = ctx->ev_host[kEvTail];
  ctx->cnt.tail_advances = ctx->ev_host[kEvTailAdv];
  ctx->cnt.tail_collisions = ctx->ev_host[kEvTailCol];
  ctx->cnt.tail_tsl = ctx->ev_host[kEvTailTsl];
  ctx->cnt.tail_crossings = ctx->ev_host[kEvTailCross];
  ctx->cnt.mlp_rows = ctx->ev_host[kEvMlp];
  ctx->cnt.graph_launches = ctx->graph_launches;
  ctx->cnt.graph_captures = ctx->graph_captures;
  return MMC_OK;
}

extern "C" int mmc_run_fixed_source(mmc_ctx* ctx, const mmc_source_desc* source,
                                    uint64_t seed0, uint64_t first, uint64_t n) {
  if (!ctx || !source) return MMC_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  ctx->R.keig = 0;
  uint32_t chunk_cap = ctx->fixed_fission ? ctx->cap / 2 : ctx->cap;
  uint64_t done = 0;
  unsigned long long ev_snap[kEvWords];
  while (done < n) {
    const uint32_t m = uint32_t(std::min<uint64_t>(n - done, chunk_cap));
    const uint64_t f = first + done;
    if (f + m > 0xffffffffull) return Fail(ctx, MMC_ERR_ARG, "history index exceeds 2^32");
    ctx->R.chunk_first = uint32_t(f);
    int rc = ResetRun(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ev_snap, ctx->R.ev, sizeof(ev_snap), hipMemcpyDeviceToHost,
                          ctx->stream));
    hipLaunchKernelGGL(KSourceFixed, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       ctx->stream, ctx->W, ctx->R, *source, seed0, uint32_t(f), m);
    rc = RunEvents(ctx, m);
    if (rc == MMC_ERR_OVERFLOW && m > 1) {
      // The chunk's secondaries or score records did not fit in the bank:
      // restore the event counters and run the same histories again in
      // chunks half as large.  A history's random numbers depend only on its
      // index, and its scores are committed only when its chunk completes,
      // so the results are those of an uninterrupted run.
      HIPCHK(hipMemcpyAsync(ctx->R.ev, ev_snap, sizeof(ev_snap), hipMemcpyHostToDevice,
                            ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ctx->err.clear();
      chunk_cap = m / 2;
      continue;
    }
    if (rc) return rc;
    // commit this chunk's per-history scores
    uint32_t misc[16];
    HIPCHK(hipMemcpyAsync(misc, ctx->R.misc, sizeof(misc), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const uint32_t nrec = misc[1];
    if (nrec > 0 && ctx->NB) {
      if (TallyReduce(ctx->tally, ctx->R.rec_key, reinterpret_cast<uint64_t*>(ctx->R.rec_val),
                      nrec, ctx->NB, ctx->bin_sum, ctx->bin_sq, ctx->stream))
        return Fail(ctx, MMC_ERR_HIP, "tally reduction failed");
      hipLaunchKernelGGL(KFoldFixed, dim3(Blocks(ctx->NB, 1u << 30)), dim3(kBlock), 0,
                         ctx->stream, ctx->S, ctx->S2, ctx->bin_sum, ctx->bin_sq,
                         uint32_t(ctx->NB));
    }
    if (const int lc = LaunchCheck(ctx, "tally commit")) return lc;
    done += m;
    ctx->cnt.histories += m;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int rc = ReadEvents(ctx);
  ctx->cnt.transport_seconds +=
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// ---------------------------------------------------------------------------
// k-eigenvalue
__global__ void KInitSites(const mmc_source_desc s, mmc_site* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t g = LcgSeed(uint64_t(i) + 1);  // seeds 1..N (KEigenvalue.cpp:31-33)
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
  mmc_site site;
  site.x = q[0]; site.y = q[1]; site.z = q[2];
  site.u = d.x; site.v = d.y; site.w = d.z;
  site.E = s.energy_kind == 1 ? Uniform(g, s.E, s.E_hi) : s.E;
  site.seed = LcgNext(g);
  site.pad = 0;
  site.key = i;
  out[i] = site;
}

extern "C" int mmc_set_entropy_mesh(mmc_ctx* ctx, const int32_t n[3], const double lo[3],
                                    const double hi[3]) {
  if (!ctx || !n || !lo || !hi) return MMC_ERR_ARG;
  for (int a = 0; a < 3; a++) {
    if (n[a] < 1) return Fail(ctx, MMC_ERR_ARG, "entropy mesh needs n >= 1");
    ctx->mesh_n[a] = n[a];
    ctx->mesh_lo[a] = lo[a];
    ctx->mesh_hi[a] = hi[a];
  }
  return MMC_OK;
}

static int EnsureKeig(mmc_ctx* ctx, uint64_t n) {
  if (n > ctx->cap) return Fail(ctx, MMC_ERR_ARG, "k-eigenvalue source larger than max_in_flight");
  if (!ctx->src) {
    ctx->src = DevAlloc<mmc_site>(ctx, ctx->cap);
    ctx->sites_sorted = DevAlloc<mmc_site>(ctx, size_t(ctx->cap) * 8);
    ctx->R.sites = DevAlloc<mmc_site>(ctx, size_t(ctx->cap) * 8);
    ctx->R.site_cap = ctx->cap * 8;
    ctx->mesh_cnt = DevAlloc<uint32_t>(ctx, 1u << 20);
    if (!ctx->mesh_cnt) return Fail(ctx, MMC_ERR_HIP, "out of device memory");
  }
  return MMC_OK;
}

extern "C" int mmc_keig_init(mmc_ctx* ctx, const mmc_source_desc* source, uint64_t n) {
  if (!ctx || !source || n == 0) return MMC_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  int rc = EnsureKeig(ctx, n);
  if (rc) return rc;
  ctx->N = n;
  hipLaunchKernelGGL(KInitSites, dim3(Blocks(n, 1u << 30)), dim3(kBlock), 0, ctx->stream,
                     *source, ctx->src, uint32_t(n));
  if (const int lc = LaunchCheck(ctx, "initial sites")) return lc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MMC_OK;
}

__global__ void KSitePlace(const mmc_site* in, uint32_t m, const uint32_t* off, mmc_site* out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const mmc_site s = in[k];
  out[off[s.key / 64] + uint32_t(s.key % 64)] = s;
}
__device__ __forceinline__ uint64_t Remix(uint64_t s, uint64_t c) {
  uint64_t z = s + c * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return z % 2147483647ULL;
}
// source j <- bank[(j M) / N]; extra copies of a site get Remix(seed, copy).
__global__ void KResample(const mmc_site* bank, uint64_t M, uint64_t N, mmc_site* src) {
  const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint64_t idx = (j * M) / N;
  const uint64_t j0 = (idx * N + M - 1) / M;
  const uint64_t copy = j - j0;
  mmc_site s = bank[idx];
  if (copy) s.seed = uint32_t(Remix(s.seed, copy));
  s.key = j;
  src[j] = s;
}
// Entropy-mesh counts of the banked sites.  Meshes of up to kMeshLds bins
// are counted in a per-block LDS histogram (the bank is ordered by history,
// so neighbouring sites hit the same bins and global atomics would serialise
// on them), merged with one global atomic per non-empty bin and block; the
// grid is kept small (kMeshBlocks) so that the merge stays cheap.
constexpr int kMeshLds = 8192;
constexpr uint32_t kMeshBlocks = 128;
__global__ void KMeshCount(const mmc_site* s, uint32_t m, int nx, int ny, int nz, double lx,
                           double ly, double lz, double hx, double hy, double hz,
                           uint32_t* cnt, int nmesh) {
  extern __shared__ uint32_t hist_lds[];
  const bool lds = nmesh <= kMeshLds;
  if (lds) {
    for (int i = threadIdx.x; i < nmesh; i += blockDim.x) hist_lds[i] = 0;
    __syncthreads();
  }
  const double lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
  const int nn[3] = {nx, ny, nz};
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
    const double q[3] = {s[k].x, s[k].y, s[k].z};
    int idx = 0;
    bool in = true;
    for (int a = 0; a < 3 && in; a++) {
      const double f = (q[a] - lo[a]) / (hi[a] - lo[a]);
      if (!(f >= 0 && f < 1)) {
        in = false;
      } else {
        int i = int(f * nn[a]);
        if (i >= nn[a]) i = nn[a] - 1;
        idx = idx * nn[a] + i;
      }
    }
    if (!in) continue;
    if (lds)
      atomicAdd(hist_lds + idx, 1u);
    else
      atomicAdd(cnt + idx, 1u);
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < nmesh; i += blockDim.x)
      if (hist_lds[i]) atomicAdd(cnt + i, hist_lds[i]);
  }
}
// deterministic per-block sums of the per-history k accumulators
__global__ void KKSum(const double* a, const double* b, const double* c, uint32_t n,
                      double* part) {
  __shared__ double s[3][kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  s[0][threadIdx.x] = i < n ? a[i] : 0;
  s[1][threadIdx.x] = i < n ? b[i] : 0;
  s[2][threadIdx.x] = i < n ? c[i] : 0;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int q = 0; q < 3; q++) s[q][threadIdx.x] += s[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = s[threadIdx.x][0];
}
__global__ void KFoldCycle(double* S, double* S2, const double* cyc, uint32_t nb, double inv_n) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double v = cyc[b] * inv_n;
  S[b] += v;
  S2[b] += v * v;
}

extern "C" int mmc_run_cycle(mmc_ctx* ctx, int active, mmc_cycle_stats* out) {
  if (!ctx || ctx->N == 0) return MMC_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t N = uint32_t(ctx->N);
  RunState& R = ctx->R;
  R.keig = 1;
  R.chunk_first = 0;  // score records carry the source site index
  int rc;
  // RunEvents ends with a read of the counters after its last kernel
  const uint32_t* misc = ctx->host_rb + kRing * kRingStride;
  const size_t src_lds = size_t(ctx->W.desc_bytes) <= kDescLdsMax ? size_t(ctx->W.desc_bytes) : 0;
  const mmc_counters cnt_snap = ctx->cnt;
  unsigned long long ev_snap[kEvWords];
  for (;;) {
    rc = ResetRun(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ev_snap, R.ev, sizeof(ev_snap), hipMemcpyDeviceToHost, ctx->stream));
    if (src_lds)
      hipLaunchKernelGGL(KSourceSites<true>, dim3(Blocks(N, 1u << 30)), dim3(kBlock), src_lds,
                         ctx->stream, ctx->W, R, ctx->src, N);
    else
      hipLaunchKernelGGL(KSourceSites<false>, dim3(Blocks(N, 1u << 30)), dim3(kBlock), 0,
                         ctx->stream, ctx->W, R, ctx->src, N);
    rc = RunEvents(ctx, N);
    if (rc == MMC_ERR_OVERFLOW && (misc[0] & ~uint32_t(kErrRecords)) == 0) {
      // Only the score records overflowed.  The cycle is a function of its
      // source alone, so it runs again, unchanged, with room for the records
      // (misc[1] counts the records asked for until the overflow was seen).
      HIPCHK(hipMemcpyAsync(R.ev, ev_snap, sizeof(ev_snap), hipMemcpyHostToDevice,
                            ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ctx->cnt = cnt_snap;
      ctx->err.clear();
      const uint64_t want = std::max<uint64_t>(2ull * R.rec_cap, uint64_t(misc[1]) * 5 / 4);
      rc = EnsureRecords(ctx, uint32_t(std::min<uint64_t>(want, 0xfffffff0ull)));
      if (rc) return rc;
      continue;
    }
    if (rc) return rc;
    break;
  }
  const uint32_t M = misc[2];
  const uint32_t nrec = misc[1];
  // sort the bank by key = history * 64 + child
  rc = Scan(ctx, R.fis_cnt, ctx->fis_off, N);
  if (rc) return rc;
  if (M) hipLaunchKernelGGL(KSitePlace, dim3(Blocks(M, 1u << 30)), dim3(kBlock), 0,
                            ctx->stream, R.sites, M, ctx->fis_off, ctx->sites_sorted);
  // entropy
  const uint32_t nmesh = uint32_t(ctx->mesh_n[0]) * ctx->mesh_n[1] * ctx->mesh_n[2];
  if (nmesh > (1u << 20)) return Fail(ctx, MMC_ERR_ARG, "entropy mesh too large");
  FillU32(ctx, ctx->mesh_cnt, 0, nmesh);
  if (M) {
    const size_t lds = nmesh <= uint32_t(kMeshLds) ? sizeof(uint32_t) * nmesh : 0;
    hipLaunchKernelGGL(KMeshCount, dim3(Blocks(M, lds ? kMeshBlocks : 1u << 30)), dim3(kBlock),
                       lds, ctx->stream, ctx->sites_sorted, M, ctx->mesh_n[0], ctx->mesh_n[1],
                       ctx->mesh_n[2], ctx->mesh_lo[0], ctx->mesh_lo[1], ctx->mesh_lo[2],
                       ctx->mesh_hi[0], ctx->mesh_hi[1], ctx->mesh_hi[2], ctx->mesh_cnt,
                       int(nmesh));
  }
  const uint32_t nbk = Blocks(N, 1u << 30);
  hipLaunchKernelGGL(KKSum, dim3(nbk), dim3(kBlock), 0, ctx->stream, R.B.k_col, R.B.k_trk,
                     R.B.k_abs, N, ctx->kpart);
  // the cycle's tallies: per-bin sums of the scores in a fixed order, then
  // S += sum / N and S2 += (sum / N)^2 once per active cycle
  if (active && ctx->NB) {
    if (TallyReduce(ctx->tally, R.rec_key, reinterpret_cast<uint64_t*>(R.rec_val), nrec,
                    ctx->NB, ctx->bin_sum, nullptr, ctx->stream))
      return Fail(ctx, MMC_ERR_HIP, "tally reduction failed");
    hipLaunchKernelGGL(KFoldCycle, dim3(Blocks(ctx->NB, 1u << 30)), dim3(kBlock), 0,
                       ctx->stream, ctx->S, ctx->S2, ctx->bin_sum, uint32_t(ctx->NB), 1.0 / N);
  }
  if (const int lc = LaunchCheck(ctx, "cycle bank and sums")) return lc;
  std::vector<uint32_t> mc(nmesh);
  std::vector<double> kp(size_t(nbk) * 3);
  HIPCHK(hipMemcpyAsync(mc.data(), ctx->mesh_cnt, sizeof(uint32_t) * nmesh,
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(kp.data(), ctx->kpart, sizeof(double) * kp.size(),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  double kc = 0, kt = 0, ka = 0;
  for (uint32_t b = 0; b < nbk; b++) {
    kc += kp[b * 3];
    kt += kp[b * 3 + 1];
    ka += kp[b * 3 + 2];
  }
  uint64_t tot = 0;
  for (uint32_t c : mc) tot += c;
  double H = 0;
  for (uint32_t c : mc)
    if (c) {
      const double p = double(c) / double(tot);
      H -= p * std::log2(p);
    }
  ctx->cnt.histories += N;
  ctx->cnt.fission_sites += M;
  if (out) {
    out->k_analog = double(M) / double(N);
    out->k_collision = kc / N;
    out->k_track = kt / N;
    out->k_absorption = ka / N;
    out->entropy = H;
    out->n_sites = M;
  }
  if (M == 0) return Fail(ctx, MMC_ERR_ARG, "no fission sites banked");
  hipLaunchKernelGGL(KResample, dim3(Blocks(N, 1u << 30)), dim3(kBlock), 0, ctx->stream,
                     ctx->sites_sorted, uint64_t(M), uint64_t(N), ctx->src);
  if (const int lc = LaunchCheck(ctx, "resample")) return lc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  rc = ReadEvents(ctx);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  ctx->cnt.transport_seconds += secs;
  if (out) out->seconds = secs;
  return rc;
}

extern "C" int mmc_get_source(mmc_ctx* ctx, mmc_site* out, uint64_t cap, uint64_t* n) {
  if (!ctx || !n) return MMC_ERR_ARG;
  *n = ctx->N;
  if (out && ctx->N) {
    HIPCHK(hipMemcpy(out, ctx->src, sizeof(mmc_site) * std::min<uint64_t>(cap, ctx->N),
                     hipMemcpyDeviceToHost));
  }
  return MMC_OK;
}
extern "C" int mmc_set_source(mmc_ctx* ctx, const mmc_site* sites, uint64_t n) {
  if (!ctx || !sites || n == 0) return MMC_ERR_ARG;
  int rc = EnsureKeig(ctx, n);
  if (rc) return rc;
  ctx->N = n;
  HIPCHK(hipMemcpy(ctx->src, sites, sizeof(mmc_site) * n, hipMemcpyHostToDevice));
  return MMC_OK;
}

// ---------------------------------------------------------------------------
// tallies and counters
extern "C" int mmc_tally_size(mmc_ctx* ctx, int e, uint64_t* n) {
  if (!ctx || !n || e < 0 || e >= int(ctx->est_off.size())) return MMC_ERR_ARG;
  *n = uint64_t(ctx->est_size[e]);
  return MMC_OK;
}
extern "C" int mmc_get_tallies(mmc_ctx* ctx, int e, double* S, double* S2, uint64_t n) {
  if (!ctx || e < 0 || e >= int(ctx->est_off.size()) || n != uint64_t(ctx->est_size[e]))
    return MMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (S)
    HIPCHK(hipMemcpy(S, ctx->S + ctx->est_off[e], sizeof(double) * n, hipMemcpyDeviceToHost));
  if (S2)
    HIPCHK(hipMemcpy(S2, ctx->S2 + ctx->est_off[e], sizeof(double) * n, hipMemcpyDeviceToHost));
  return MMC_OK;
}
extern "C" int mmc_reset_tallies(mmc_ctx* ctx) {
  if (!ctx) return MMC_ERR_ARG;
  if (ctx->NB) {
    FillF64(ctx, ctx->S, 0.0, ctx->NB);
    FillF64(ctx, ctx->S2, 0.0, ctx->NB);
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MMC_OK;
}
extern "C" int mmc_get_counters(mmc_ctx* ctx, mmc_counters* out) {
  if (!ctx || !out) return MMC_ERR_ARG;
  *out = ctx->cnt;
  return MMC_OK;
}

// ---------------------------------------------------------------------------
// Single-event parity entry point
struct EvalSoA {
  double *x, *y, *z, *u, *v, *w, *E, *out0, *out1;
  uint32_t* rng;
  int32_t *cell, *surface, *event, *iout;
};

__global__ void KEval(DevWorld W, int kind, EvalSoA s, uint32_t n, uint32_t* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 pos{s.x[i], s.y[i], s.z[i]};
  V3 dir{s.u[i], s.v[i], s.w[i]};
  double E = s.E[i];
  uint32_t g = s.rng[i];
  const int cell = s.cell[i];
  switch (kind) {
    case MMC_EVAL_FIND_CELL:
      s.cell[i] = FindCell(W, pos);
      break;
    case MMC_EVAL_NEAREST_SURFACE: {
      int sf;
      s.out0[i] = NearestSurface(W, cell, pos, dir, sf);
      s.iout[i] = sf;
      break;
    }
    case MMC_EVAL_XS_TOTAL: {
      const DCell c = W.cells[cell];
      const DField f = W.fields[c.field];
      s.out0[i] = MaterialTotal(W, c.material, E, FieldAt(f, pos), false, nullptr);
      s.out1[i] = MaterialTotal(W, c.material, E, f.kind == 0 ? f.c : f.upper, true, nullptr);
      break;
    }
    case MMC_EVAL_COLLISION_DISTANCE: {
      const DCell c = W.cells[cell];
      const DField f = W.fields[c.field];
      const bool delta = W.tracking == 1;
      const double sig = MaterialTotal(W, c.material, E,
                                       delta ? (f.kind == 0 ? f.c : f.upper) : FieldAt(f, pos),
                                       delta, nullptr);
      s.out0[i] = -log(1.0 - Canonical(g)) / (W.mats[c.material].aden * sig);
      break;
    }
    case MMC_EVAL_COLLIDE: {
      int tsl = -1;
      CollideOut o = Collide(W, nullptr, 0, pos, dir, E, g, cell, nullptr, nullptr, nullptr,
                             nullptr, &tsl);
      if (o.tsl) {
        double a, b;
        const double T = FieldAt(W.fields[W.cells[cell].field], pos);
        const uint32_t e = TslScatterPod(W, W.hot, W.tsl[tsl], g, T, dir, E, a, b);
        if (e) atomicOr(err, e);
      }
      if (o.event >= 0) s.event[i] = o.event;
      if (o.children < 0) atomicOr(err, kErrReaction);
      s.iout[i] = o.children < 0 ? 0 : o.children;
      break;
    }
    case MMC_EVAL_SCATTER_ROTATE: {
      const double phi = 2 * kPi * Canonical(g);
      dir = Rotate(dir, s.out0[i], phi);
      E = s.out1[i];
      break;
    }
    case MMC_EVAL_TSL_SAMPLE: {
      double a = 0, b = 0;
      const DTsl t = W.tsl[0];
      uint32_t e =
          SampleBeta(W, W.hot, t, g, E, s.out0[i], b) ? 0u : uint32_t(kErrResample);
      if (!e) e = SampleAlpha(W, W.hot, t, g, b, E, s.out0[i], a);
      if (e) {
        atomicOr(err, e);
        a = b = __builtin_nan("");
      }
      s.out0[i] = a;
      s.out1[i] = b;
      break;
    }
    default:
      break;
  }
  s.u[i] = dir.x;
  s.v[i] = dir.y;
  s.w[i] = dir.z;
  s.E[i] = E;
  s.rng[i] = g;
}

extern "C" int mmc_eval_event(mmc_ctx* ctx, int kind, mmc_particle_soa* p, uint64_t n) {
  if (!ctx || !p || kind < 0 || kind > MMC_EVAL_TSL_SAMPLE) return MMC_ERR_ARG;
  if (n == 0) return MMC_OK;
  if (n > (1u << 26)) return Fail(ctx, MMC_ERR_ARG, "batch too large");
  if (kind == MMC_EVAL_TSL_SAMPLE && !ctx->has_tsl) return Fail(ctx, MMC_ERR_ARG, "no tsl");
  HIPCHK(hipSetDevice(ctx->device));
  // host-side range checks of the cell indices the kernels will use
  if (kind != MMC_EVAL_FIND_CELL && kind != MMC_EVAL_SCATTER_ROTATE &&
      kind != MMC_EVAL_TSL_SAMPLE) {
    const bool needs_material = kind != MMC_EVAL_NEAREST_SURFACE;
    for (uint64_t i = 0; i < n; i++) {
      const int c = p->cell[i];
      if (c < 0 || c >= ctx->W.n_cells) return Fail(ctx, MMC_ERR_ARG, "cell out of range");
      if (needs_material && ctx->cell_material[c] < 0)
        return Fail(ctx, MMC_ERR_ARG, "event needs a material cell");
    }
  }
  double* dbuf = nullptr;
  uint32_t* ibuf = nullptr;
  HIPCHK(hipMalloc(&dbuf, sizeof(double) * n * 9));
  HIPCHK(hipMalloc(&ibuf, sizeof(uint32_t) * (n * 5 + 1)));
  double* hd[9] = {p->x, p->y, p->z, p->u, p->v, p->w, p->E, p->out0, p->out1};
  EvalSoA s;
  double** sd[9] = {&s.x, &s.y, &s.z, &s.u, &s.v, &s.w, &s.E, &s.out0, &s.out1};
  for (int k = 0; k < 9; k++) {
    *sd[k] = dbuf + n * k;
    if (hd[k]) hipMemcpy(*sd[k], hd[k], sizeof(double) * n, hipMemcpyHostToDevice);
    else hipMemset(*sd[k], 0, sizeof(double) * n);
  }
  s.rng = ibuf;
  s.cell = reinterpret_cast<int32_t*>(ibuf + n);
  s.surface = reinterpret_cast<int32_t*>(ibuf + 2 * n);
  s.event = reinterpret_cast<int32_t*>(ibuf + 3 * n);
  s.iout = reinterpret_cast<int32_t*>(ibuf + 4 * n);
  uint32_t* err = ibuf + 5 * n;
  const void* hi[4] = {p->rng, p->cell, p->surface, p->event};
  void* di[4] = {s.rng, s.cell, s.surface, s.event};
  for (int k = 0; k < 4; k++) {
    if (hi[k]) hipMemcpy(di[k], hi[k], sizeof(uint32_t) * n, hipMemcpyHostToDevice);
    else hipMemset(di[k], 0, sizeof(uint32_t) * n);
  }
  hipMemset(s.iout, 0, sizeof(uint32_t) * n);
  hipMemset(err, 0, sizeof(uint32_t));
  hipLaunchKernelGGL(KEval, dim3(Blocks(n, 1u << 30)), dim3(kBlock), 0, ctx->stream, ctx->W,
                     kind, s, uint32_t(n), err);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < 9; k++)
    if (hd[k]) hipMemcpy(hd[k], *sd[k], sizeof(double) * n, hipMemcpyDeviceToHost);
  void* ho[5] = {p->rng, p->cell, p->surface, p->event, p->iout};
  void* dd[5] = {s.rng, s.cell, s.surface, s.event, s.iout};
  for (int k = 0; k < 5; k++)
    if (ho[k]) hipMemcpy(ho[k], dd[k], sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
  uint32_t herr = 0;
  hipMemcpy(&herr, err, sizeof(uint32_t), hipMemcpyDeviceToHost);
  hipFree(dbuf);
  hipFree(ibuf);
  if (e != hipSuccess) return Fail(ctx, MMC_ERR_HIP, hipGetErrorString(e));
  uint32_t misc[16] = {herr};
  return CheckMisc(ctx, misc);
}

// Source::Sample for histories seed0 + i, i < n, into host SoA.
__global__ void KEvalSource(DevWorld W, mmc_source_desc s, uint64_t seed0, EvalSoA o,
                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t g = LcgSeed(seed0 + i);
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
  o.x[i] = q[0]; o.y[i] = q[1]; o.z[i] = q[2];
  o.u[i] = d.x; o.v[i] = d.y; o.w[i] = d.z;
  o.E[i] = s.energy_kind == 1 ? Uniform(g, s.E, s.E_hi) : s.E;
  o.rng[i] = LcgSeed(LcgNext(g));
  o.cell[i] = FindCell(W, V3{q[0], q[1], q[2]});
}

extern "C" int mmc_sample_source(mmc_ctx* ctx, const mmc_source_desc* src, uint64_t seed0,
                                 uint64_t n, mmc_particle_soa* p) {
  if (!ctx || !src || !p || n == 0 || n > (1u << 26)) return MMC_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  double* dbuf = nullptr;
  uint32_t* ibuf = nullptr;
  HIPCHK(hipMalloc(&dbuf, sizeof(double) * n * 7));
  HIPCHK(hipMalloc(&ibuf, sizeof(uint32_t) * n * 2));
  EvalSoA s{};
  double** sd[7] = {&s.x, &s.y, &s.z, &s.u, &s.v, &s.w, &s.E};
  for (int k = 0; k < 7; k++) *sd[k] = dbuf + n * k;
  s.rng = ibuf;
  s.cell = reinterpret_cast<int32_t*>(ibuf + n);
  hipLaunchKernelGGL(KEvalSource, dim3(Blocks(n, 1u << 30)), dim3(kBlock), 0, ctx->stream,
                     ctx->W, *src, seed0, s, uint32_t(n));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  double* hd[7] = {p->x, p->y, p->z, p->u, p->v, p->w, p->E};
  for (int k = 0; k < 7; k++)
    if (hd[k]) hipMemcpy(hd[k], *sd[k], sizeof(double) * n, hipMemcpyDeviceToHost);
  if (p->rng) hipMemcpy(p->rng, s.rng, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
  if (p->cell) hipMemcpy(p->cell, s.cell, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
  hipFree(dbuf);
  hipFree(ibuf);
  if (e != hipSuccess) return Fail(ctx, MMC_ERR_HIP, hipGetErrorString(e));
  return MMC_OK;
}

// ---------------------------------------------------------------------------
// Batch POD sampling of (alpha, beta) on device-resident inputs.
template <bool kLds>
__global__ void __launch_bounds__(256)
    KTslBatch(DevWorld W, int tsl, uint32_t n, const double* E, const double* T, uint32_t* rng,
              double* alpha, double* beta, uint32_t* err) {
  extern __shared__ double lds_hot[];
  if (blockIdx.x * blockDim.x >= n) return;
  const double* H = kLds ? StageHot(W.hot, lds_hot, W.hot_n) : W.hot;
  const DTsl t = W.tsl[tsl];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    uint32_t g = rng[i];
    double a = 0, b = 0;
    uint32_t e = SampleBeta(W, H, t, g, E[i], T[i], b) ? 0u : uint32_t(kErrResample);
    if (!e) e = SampleAlpha(W, H, t, g, b, E[i], T[i], a);
    if (e) {
      atomicOr(err, e);
      a = b = __builtin_nan("");
    }
    alpha[i] = a;
    beta[i] = b;
    rng[i] = g;
  }
}

extern "C" int mmc_tsl_sample_pod(mmc_ctx* ctx, int tsl, uint64_t n, const double* E,
                                  const double* T, uint32_t* rng, double* alpha,
                                  double* beta) {
  if (!ctx || !E || !T || !rng || !alpha || !beta) return MMC_ERR_ARG;
  if (tsl < 0 || tsl >= ctx->n_tsl) return Fail(ctx, MMC_ERR_ARG, "no such tsl table");
  if (n == 0) return MMC_OK;
  if (n > 0xffffffffull) return Fail(ctx, MMC_ERR_ARG, "batch too large");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMemsetAsync(ctx->R.misc, 0, sizeof(uint32_t), ctx->stream));
  if (ctx->hot_lds)
    hipLaunchKernelGGL(KTslBatch<true>, dim3(std::min(Blocks(n, 1u << 30), ctx->tsl_grid)),
                       dim3(kBlock), ctx->hot_lds, ctx->stream, ctx->W, tsl, uint32_t(n), E, T,
                       rng, alpha, beta, ctx->R.misc);
  else
    hipLaunchKernelGGL(KTslBatch<false>, dim3(Blocks(n, 1u << 30)), dim3(kBlock), 0,
                       ctx->stream, ctx->W, tsl, uint32_t(n), E, T, rng, alpha, beta,
                       ctx->R.misc);
  if (const int lc = LaunchCheck(ctx, "POD sampler")) return lc;
  uint32_t misc[16] = {0};
  HIPCHK(hipMemcpyAsync(misc, ctx->R.misc, sizeof(uint32_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CheckMisc(ctx, misc);
}

extern "C" int mmc_device_malloc(mmc_ctx* ctx, uint64_t bytes, void** ptr) {
  if (!ctx || !ptr) return MMC_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return MMC_OK;
}
extern "C" int mmc_device_free(mmc_ctx* ctx, void* ptr) {
  if (!ctx) return MMC_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipFree(ptr));
  return MMC_OK;
}
extern "C" int mmc_memcpy_h2d(mmc_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return MMC_ERR_ARG;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MMC_OK;
}
extern "C" int mmc_memcpy_d2h(mmc_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return MMC_ERR_ARG;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return MMC_OK;
}

hipStream_t MmcStream(mmc_ctx* ctx) { return ctx->stream; }
MlpState& MmcMlp(mmc_ctx* ctx) { return ctx->mlp; }
const DevWorld& MmcWorld(mmc_ctx* ctx) { return ctx->W; }
int MmcFail(mmc_ctx* ctx, int code, const char* msg) { return Fail(ctx, code, msg); }
int MmcDevice(mmc_ctx* ctx) { return ctx->device; }
int MmcTslCount(mmc_ctx* ctx) { return ctx->n_tsl; }
