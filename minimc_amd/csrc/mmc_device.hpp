// Device-side World tables and per-particle physics for gfx950.
//
// Every function here evaluates one reference function on one particle
// state; the arithmetic keeps the reference's operation order and is
// compiled with contraction off, so results equal the CPU's up to the
// last-ulp differences of the libm calls (log, exp, sin, cos, erf, pow).
// References are to the minimc sources (file:line).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mmc {

// ---- constants (src/Constants.hpp:13-35) ---------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kNudge = 10 * 2.220446049250313080847e-16;
constexpr double kOnAxis = 0.9;
constexpr double kBoltzmann = 8.617333262145e-11;
constexpr double kNeutronMass = 1.045354912280858e-18;
constexpr double kRelTempTol = 0.01;
constexpr int kBetaResampleLimit = 10;
constexpr int kAlphaResampleLimit = 10;
constexpr uint32_t kLcgM = 2147483647u;

// error bits (device flag word)
enum : uint32_t {
  kErrLost = 1u,
  kErrResample = 2u,
  kErrSteps = 4u,
  kErrAlphaRange = 8u,
  kErrOverflow = 16u,
  kErrReaction = 32u,  // no reaction selected after kMaxReactionResamples
  kErrRecords = 64u    // more score records than the record buffer holds
};
constexpr int kMaxReactionResamples = 1000;

// ---- tables (flattened World) ---------------------------------------------
struct DSurface {
  double a, b, c, r;
  int kind;
  int pad;
};
struct DCell {
  int material, face_begin, face_count, field;
  // A box that holds every point the cell's face tests accept, widened by a
  // small margin (CellBox on the host): the cell searches skip a cell whose
  // box misses the point without reading its faces.  The margin keeps a
  // point that sits on a face just across its surface (a crossing) inside.
  double lo[3], hi[3];
};
struct DFace {
  int surface, inside;
};
struct DField {
  int kind, pad;
  double c, gx, gy, gz, b, upper;
};
struct DMaterial {
  double aden;
  int comp_begin, comp_count;
  int pert_mask, pad;  // bit j: the material holds perturbation j's nuclide
};
struct DComp {
  double afrac;
  int nuclide, pad;
};
// A ContinuousMap on a (possibly shared) key grid.
struct DTable {
  int grid;      // index into grids
  int y_off;     // values in pool: value i at pool[y_off + i * y_stride]
  int y_stride;  // 1, or the column count of an interleaved block
};
struct DGrid {
  int x_off, n;         // keys in pool
  int hash_off, n_hash; // bucket bounds in ipool (n_hash + 1 entries), 0 = none
  double log_x0, inv_dlog;
  double x_first, x_last;  // x[0] and x[n - 1], kept with the descriptor
};
struct DCeReaction {
  int kind, tsl, xs, nubar;  // xs / nubar: table index (-1 none)
  int chi_kind, pad;
  double temperature, watt_a, watt_b;
};
struct DCeNuclide {
  double awr, total_temperature;
  int has_total, total;  // table
  int rx_begin, rx_count;
  int grid;              // the grid shared by every table of this nuclide, or -1
  int fissile;
};
// Two small blocks that kernels copy into LDS: DevWorld::tot holds every
// table's total cross-section tables (E_axis_off, E_modes_off, T_axis_off,
// T_modes_off, S_off of DTsl), DevWorld::hot the sampler's axes and CDF mode
// tables (F_off, cdf_off, S_off, T_off of a partition; Es_off, E_ends_off,
// betas_off, beta_ends_off of a table).  x_off and xt_off are offsets into
// the pool.
struct DTslPart {
  int F_off, nF, cdf_off, S_off, x_off, nx, T_off, nT, xt_off, rank;
};
struct DTsl {
  int majorant;          // table
  int E_axis_off, nE, E_modes_off, T_axis_off, nT, T_modes_off, S_off, rank;
  int Es_off, nEs, E_ends_off, n_bp, bp_begin;       // beta partitions
  int betas_off, nbetas, beta_ends_off, n_ap, ap_begin;  // alpha partitions
  // (E_ends and beta_ends are stored as doubles; they are small integers)
  int pad;
  double beta_cutoff, alpha_cutoff, awr, cutoff_energy;
};
struct DBins {
  int kind, n_bins;  // total bins including overflow bins
  double lo, hi, base_log, width;
  int bounds_off, n_bounds;
};
struct DEst {
  int kind, surface, has_dir, sens;  // sens: 0, or 1 + perturbation index
  double dx, dy, dz;
  DBins cosine, energy;
  int mesh_n[3], bin_off;  // bin_off: first global tally bin
  double mesh_lo[3], mesh_hi[3];
};

struct DevWorld {
  int energy_mode, G, tracking, n_cells;
  int quirks, pad_q;
  int n_est, n_est_current, n_est_flux, n_est_tsl;
  const DSurface* surfaces;
  const DCell* cells;
  const DFace* faces;
  const DField* fields;
  const DMaterial* mats;
  const DComp* comps;
  // multigroup: [nuc*G] and [nuc*G*G] ([nuc][g_in][g_out])
  const double *mg_capture, *mg_scatter_xs, *mg_fission, *mg_nubar, *mg_total;
  const double *mg_scatter_probs, *mg_chi_probs;
  // continuous
  const DCeNuclide* ce;
  const DCeReaction* rx;
  const DTable* tables;
  const DGrid* grids;
  const DTsl* tsl;
  const DTslPart* parts;
  const double* pool;
  const int* ipool;
  const DEst* ests;
  // Cell locator: a uniform grid of loc_n[0] x loc_n[1] x loc_n[2] voxels
  // from loc_lo, loc_inv voxels per cm (loc_n[0] == 0: no grid).  Voxel v
  // lists, in cell order, the cells loc_cells[loc_off[v] .. loc_off[v+1])
  // whose boxes reach into it (widened by a small tolerance).
  const int* loc_off;
  const int* loc_cells;
  double loc_lo[3], loc_inv[3];
  int loc_n[3], pad_loc;
  const double* hot;  // thermal-scattering sampler tables (see DTsl)
  const double* tot;  // thermal-scattering total cross-section tables
  int hot_n, tot_n;   // doubles in each
  // Every descriptor array above (surfaces .. ests) lives in one block,
  // desc, at the byte offsets desc_off (order of kDesc*), so that a kernel
  // can copy the whole block into LDS at once (DescInLds).
  const unsigned char* desc;
  int desc_bytes;  // a multiple of 16
  int desc_off[15];
  int n_pert, pad_pert;  // total cross-section perturbations (at most kMaxPert)
};
constexpr int kMaxPert = 4;
enum DescArray {
  kDescSurfaces, kDescCells, kDescFaces, kDescFields, kDescMats, kDescComps, kDescCe,
  kDescRx, kDescTables, kDescGrids, kDescTsl, kDescParts, kDescEsts, kDescLocOff,
  kDescLocCells, kDescN
};
static_assert(kDescN == sizeof(DevWorld::desc_off) / sizeof(int), "one offset per array");

// ---- RNG: std::minstd_rand with libstdc++ canonical (see oracle/rng.hpp) --
__device__ __forceinline__ uint32_t LcgNext(uint32_t& x) {
  const uint64_t p = uint64_t(x) * 48271u;
  uint32_t r = uint32_t(p & 0x7fffffffu) + uint32_t(p >> 31);
  if (r >= kLcgM) r -= kLcgM;
  x = r;
  return r;
}
__host__ __device__ __forceinline__ uint32_t LcgSeed(uint64_t s) {
  uint32_t x = uint32_t(s % kLcgM);
  return x == 0 ? 1u : x;
}
__device__ __forceinline__ double Canonical(uint32_t& x) {
  const double r = 2147483646.0;
  const double r2 = 2147483646.0 * 2147483646.0;
  double sum = double(LcgNext(x) - 1u);
  sum = sum + double(LcgNext(x) - 1u) * r;
  double ret = sum / r2;
  if (ret >= 1.0) ret = 0x1.fffffffffffffp-1;  // nextafter(1, 0)
  return ret;
}
__device__ __forceinline__ double Uniform(uint32_t& x, double a, double b) {
  return Canonical(x) * (b - a) + a;
}

// ---- vectors (src/Point.cpp) ----------------------------------------------
struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 Add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 Mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 Div(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 Cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 Normalized(V3 p) {
  const double n = sqrt(Dot(p, p));
  return {p.x / n, p.y / n, p.z / n};
}
// Direction(RNG&), Point.cpp:90-96
__device__ __forceinline__ V3 IsotropicDirection(uint32_t& g) {
  V3 d;
  d.x = Uniform(g, -1., +1.);
  const double sin_theta = sqrt(1 - d.x * d.x);
  const double phi = Uniform(g, 0., 2 * kPi);
  double s, c;
  sincos(phi, &s, &c);
  d.y = sin_theta * c;
  d.z = sin_theta * s;
  return d;
}
// Direction(d, mu, phi) then Normalize, Point.cpp:98-129
__device__ __forceinline__ V3 Rotate(V3 d, double mu, double phi) {
  const bool off = d.x <= kOnAxis && d.x > -kOnAxis;
  const V3 ax = off ? V3{1, 0, 0} : V3{0, 1, 0};
  const V3 u = Normalized(Cross(d, ax));
  const V3 v = Normalized(Cross(d, u));
  double s, c;
  sincos(phi, &s, &c);
  const double st = sqrt(1 - mu * mu);
  const V3 uc = Mul(u, st * c);
  const V3 vc = Mul(v, st * s);
  const V3 dc = Mul(d, mu);
  return Normalized(Add(Add(uc, vc), dc));
}

// ---- geometry (src/CSGSurface.cpp, src/Cell.cpp, src/World.cpp) ---------
__device__ __forceinline__ double SolveQuadratic(double a, double b, double c) {
  const double disc = b * b - 4 * a * c;
  if (disc <= 0) return __builtin_huge_val();
  const double sq = sqrt(disc);
  const double lesser = b > 0 ? (-b - sq) / (2 * a) : (2 * c) / (-b + sq);
  const double greater = b > 0 ? (2 * c) / (-b - sq) : (-b + sq) / (2 * a);
  if (lesser > 0) return lesser;
  if (greater > 0) return greater;
  return __builtin_huge_val();
}
__device__ __forceinline__ double SurfDistance(const DSurface& s, V3 o, V3 d) {
  switch (s.kind) {
    case 0: {  // sphere
      const V3 oc = Sub(o, V3{s.a, s.b, s.c});
      return SolveQuadratic(1, 2 * Dot(oc, d), Dot(oc, oc) - s.r * s.r);
    }
    case 2: {  // cylinder along x
      const double pp = Dot(o, o);
      const double pw = Dot(o, V3{1, 0, 0});
      const double dw = Dot(d, V3{1, 0, 0});
      const double pd = Dot(o, d);
      return SolveQuadratic(1 - dw * dw, 2 * (pd - pw * dw), pp - pw * pw - s.r * s.r);
    }
    default: {  // planes x (1), y (3), z (4)
      const V3 ax = s.kind == 1 ? V3{1, 0, 0} : s.kind == 3 ? V3{0, 1, 0} : V3{0, 0, 1};
      const double t = (s.a - Dot(o, ax)) / Dot(d, ax);
      return t > 0 ? t : __builtin_huge_val();
    }
  }
}
__device__ __forceinline__ bool SurfContains(const DSurface& s, V3 p) {
  switch (s.kind) {
    case 0: {
      const V3 pc = Sub(p, V3{s.a, s.b, s.c});
      return Dot(pc, pc) < s.r * s.r;
    }
    case 2: {  // quirk Q1: sqrt(rho^2) < r^2 (CSGSurface.cpp:164-174)
      const double pp = Dot(p, p);
      const double pw = Dot(p, V3{1, 0, 0});
      return sqrt(pp - pw * pw) < s.r * s.r;
    }
    case 1:
      return Dot(p, V3{1, 0, 0}) < s.a;
    case 3:
      return Dot(p, V3{0, 1, 0}) < s.a;
    default:
      return Dot(p, V3{0, 0, 1}) < s.a;
  }
}
__device__ __forceinline__ bool InBox(const DCell& c, V3 p) {
  return p.x >= c.lo[0] && p.x <= c.hi[0] && p.y >= c.lo[1] && p.y <= c.hi[1] &&
         p.z >= c.lo[2] && p.z <= c.hi[2];
}
// The cells a search must test for point p, in cell order: the list of p's
// voxel in the cell locator ([*b, *e) of loc_cells), or every cell (b = 0,
// e = n_cells, list = false) where there is no grid or p lies outside it.
// A cell whose face tests accept p has a box that holds p, and that box
// reaches into p's voxel, so the first accepting cell of the list is the
// first accepting cell of the World.  A voxel index rounded across a voxel
// boundary is covered by the tolerance the lists are built with.
__device__ __forceinline__ bool LocCandidates(const DevWorld& W, V3 p, int& b, int& e) {
  b = 0;
  e = W.n_cells;
  if (W.loc_n[0] == 0) return false;
  const double t0 = (p.x - W.loc_lo[0]) * W.loc_inv[0];
  const double t1 = (p.y - W.loc_lo[1]) * W.loc_inv[1];
  const double t2 = (p.z - W.loc_lo[2]) * W.loc_inv[2];
  // NaN fails every comparison and falls back to the full search
  if (!(t0 >= 0 && t0 < W.loc_n[0] && t1 >= 0 && t1 < W.loc_n[1] && t2 >= 0 &&
        t2 < W.loc_n[2]))
    return false;
  const int v = (int(t0) * W.loc_n[1] + int(t1)) * W.loc_n[2] + int(t2);
  b = W.loc_off[v];
  e = W.loc_off[v + 1];
  return true;
}
__device__ __forceinline__ int FindCell(const DevWorld& W, V3 p) {
  int b, e;
  const bool list = LocCandidates(W, p, b, e);
  for (int k0 = b; k0 < e; k0++) {
    const int c = list ? W.loc_cells[k0] : k0;
    const DCell cell = W.cells[c];
    if (!InBox(cell, p)) continue;  // no face test can accept p
    bool in = true;
    for (int k = 0; k < cell.face_count && in; k++) {
      const DFace f = W.faces[cell.face_begin + k];
      in = SurfContains(W.surfaces[f.surface], p) == (f.inside != 0);
    }
    if (in) return c;
  }
  return -1;
}
// Crossing without the nudge quirk: the crossed surface's sense is flipped
// explicitly, every other face is tested on the point.  kLocate: search the
// locator's candidates; else every cell in order, a loop whose cell index is
// the same in every lane (faster for the tail kernel, whose single wave per
// SIMD waits on each dependent load).
template <bool kLocate>
__device__ __forceinline__ int FindCellAcross(const DevWorld& W, V3 p, int surf,
                                              int new_inside) {
  int b = 0, e = W.n_cells;
  const bool list = kLocate && LocCandidates(W, p, b, e);
  for (int k0 = b; k0 < e; k0++) {
    const int c = list ? W.loc_cells[k0] : k0;
    const DCell cell = W.cells[c];
    if (!InBox(cell, p)) continue;  // no face test can accept p
    bool in = true;
    for (int k = 0; k < cell.face_count && in; k++) {
      const DFace f = W.faces[cell.face_begin + k];
      in = f.surface == surf ? f.inside == new_inside
                             : SurfContains(W.surfaces[f.surface], p) == (f.inside != 0);
    }
    if (in) return c;
  }
  return -1;
}
__device__ __forceinline__ int FaceInside(const DevWorld& W, int cell, int surf) {
  const DCell c = W.cells[cell];
  for (int k = 0; k < c.face_count; k++) {
    const DFace f = W.faces[c.face_begin + k];
    if (f.surface == surf) return f.inside;
  }
  return 0;
}

// Cell::NearestSurface (Cell.cpp:37-51): first minimum wins.
__device__ __forceinline__ double NearestSurface(const DevWorld& W, int cell, V3 p,
                                                 V3 d, int& surf) {
  const DCell c = W.cells[cell];
  double best = __builtin_huge_val();
  surf = -1;
  for (int k = 0; k < c.face_count; k++) {
    const int s = W.faces[c.face_begin + k].surface;
    const double t = SurfDistance(W.surfaces[s], p, d);
    if (surf < 0 || t < best) {
      best = t;
      surf = s;
    }
  }
  return best;
}

__device__ __forceinline__ double FieldAt(const DField& f, V3 p) {
  return f.kind == 0 ? f.c : Dot(V3{f.gx, f.gy, f.gz}, p) + f.b;
}

// ---- grid search and ContinuousMap::at (ContinuousMap.hpp:20-56) ---------
// upper_bound(x[lo..hi), key) on ascending keys.
__device__ __forceinline__ int UpperBoundRange(const double* x, int lo, int hi,
                                               double key) {
  int len = hi - lo;
  while (len > 0) {
    const int half = len >> 1;
    const int mid = lo + half;
    if (key < x[mid]) {
      len = half;
    } else {
      lo = mid + 1;
      len = len - half - 1;
    }
  }
  return lo;
}
// The grid's log-spaced bucket bounds (staged in LDS by the caller when
// `hash` points there) narrow the search to at most three buckets.
__device__ __forceinline__ int GridUpperBound(const DevWorld& W, const DGrid& g,
                                              const int* hash, double key) {
  const double* x = W.pool + g.x_off;
  int lo = 0, hi = g.n;
  if (g.n_hash > 0 && key > g.x_first && key < g.x_last) {
    const double f = (log(key) - g.log_x0) * g.inv_dlog;
    int b = int(f);
    b = b < 0 ? 0 : (b >= g.n_hash ? g.n_hash - 1 : b);
    const int b0 = b > 0 ? b - 1 : 0;
    const int b1 = b + 2 < g.n_hash ? b + 2 : g.n_hash;
    lo = hash[b0];
    hi = hash[b1];
  }
  return UpperBoundRange(x, lo, hi, key);
}
__device__ __forceinline__ double TableAtIndex(const DevWorld& W, const DGrid& g,
                                               const DTable& tb, int hi, double k) {
  const double* x = W.pool + g.x_off;
  const double* y = W.pool + tb.y_off;
  const int s = tb.y_stride;
  if (hi == g.n) return y[(hi - 1) * s];
  if (hi == 0) return y[0];
  const double k_hi = x[hi], v_hi = y[hi * s], k_lo = x[hi - 1], v_lo = y[(hi - 1) * s];
  return v_lo + (v_hi - v_lo) / (k_hi - k_lo) * (k - k_lo);
}
__device__ __forceinline__ double TableAt(const DevWorld& W, int t, double k) {
  const DTable tb = W.tables[t];
  const DGrid g = W.grids[tb.grid];
  const int hi = GridUpperBound(W, g, W.ipool + g.hash_off, k);
  return TableAtIndex(W, g, tb, hi, k);
}
// With a precomputed grid index (same grid).
__device__ __forceinline__ double TableAtI(const DevWorld& W, int t, int hi, double k) {
  const DTable tb = W.tables[t];
  return TableAtIndex(W, W.grids[tb.grid], tb, hi, k);
}

// ---- thermal scattering POD (src/ThermalScattering.cpp) -------------------
__device__ __forceinline__ int UpperBoundArr(const double* x, int n, double key) {
  return UpperBoundRange(x, 0, n, key);
}
// The temperature bracket of BetaPartition/AlphaPartition::Evaluate
// (ThermalScattering.cpp:183-256): the same for every evaluation of one
// partition at one temperature, so a sample computes it once.
struct TBracket {
  int t_lo, t_hi;
};
__device__ __forceinline__ TBracket PartTBracket(const double* H, const DTslPart& P,
                                                 double T) {
  const double* Ts = H + P.T_off;
  const int cand = UpperBoundArr(Ts, P.nT, T);
  const bool above = cand == P.nT;
  const int t_hi = above ? cand - 1 : cand;
  const bool below = t_hi == 0;
  return {below ? t_hi : t_hi - 1, t_hi};
}
// The two rows of POD coefficients (grid point x_i at the bracket's two
// temperatures) that every evaluation of one sample multiplies.  A sample
// loads them once, into registers when the rank is kRowRegs: the CDF
// bisection then waits on the CDF modes in LDS only, not on a global load per
// probe, and the sums are unrolled with no branch per mode, so that an
// evaluation's LDS reads issue together.  Other ranks read the rows in
// memory at each evaluation.
constexpr int kRowRegs = 8;
struct PartRow {
  TBracket br;
  const double* xh;  // row at t_hi
  const double* xl;  // row at t_lo
  double h[kRowRegs], l[kRowRegs];
};
__device__ __forceinline__ PartRow LoadPartRow(const DevWorld& W, const DTslPart& P, int x_i,
                                               TBracket br) {
  PartRow r;
  r.br = br;
  r.xh = W.pool + P.xt_off + (x_i * P.nT + br.t_hi) * P.rank;
  r.xl = W.pool + P.xt_off + (x_i * P.nT + br.t_lo) * P.rank;
  // (other ranks load mode 0 into every register, in bounds and unused: no
  // branch, which keeps the rows in registers)
  const bool regs = P.rank == kRowRegs;
#pragma unroll
  for (int o = 0; o < kRowRegs; o++) {
    r.h[o] = r.xh[regs ? o : 0];
    r.l[o] = r.xl[regs ? o : 0];
  }
  return r;
}
// BetaPartition/AlphaPartition::Evaluate (ThermalScattering.cpp:183-256) at
// CDF point cdf_i, for the grid point and temperature bracket of row r.  The
// sums run over the modes in order, whichever copy of the rows they read.
// kRegs: the partition's rank is kRowRegs (the caller checked it once per
// sample), so that a sampler's evaluations hold no branch and overlap.
template <bool kRegs>
__device__ __forceinline__ double PartEvaluateAt(const double* H, const DTslPart& P, int cdf_i,
                                                 double T, const PartRow& r) {
  const double* Ts = H + P.T_off;
  const double* S = H + P.S_off;
  const double* cdf = H + P.cdf_off + cdf_i * P.rank;
  double v_hi = 0, v_lo = 0;
  if constexpr (kRegs) {
#pragma unroll
    for (int o = 0; o < kRowRegs; o++) {
      const double sc = S[o] * cdf[o];
      v_hi = v_hi + sc * r.h[o];
      v_lo = v_lo + sc * r.l[o];
    }
  } else {
    for (int o = 0; o < P.rank; o++) {
      const double sc = S[o] * cdf[o];
      v_hi = v_hi + sc * r.xh[o];
      v_lo = v_lo + sc * r.xl[o];
    }
  }
  const double T_hi = Ts[r.br.t_hi], T_lo = Ts[r.br.t_lo];
  return v_lo + (v_hi - v_lo) / (T_hi - T_lo) * (T - T_lo);
}
// The values at the two ends of CDF interval h of an axis of nF > 0 points:
// lo at h - 1 (fallback lo_out when h == 0) and hi at h (fallback hi_out when
// h == nF).  Both evaluations run at in-range indices whatever h is, so their
// reads and sums overlap; the fallbacks replace what they computed.
template <bool kRegs>
__device__ __forceinline__ void PartEnds(const double* H, const DTslPart& P, int h, double T,
                                         const PartRow& r, double lo_out, double hi_out,
                                         double& lo, double& hi) {
  const double e_lo = PartEvaluateAt<kRegs>(H, P, h != 0 ? h - 1 : 0, T, r);
  const double e_hi = PartEvaluateAt<kRegs>(H, P, h != P.nF ? h : P.nF - 1, T, r);
  lo = h != 0 ? e_lo : lo_out;
  hi = h != P.nF ? e_hi : hi_out;
}
// ThermalScattering::GetTotal (ThermalScattering.cpp:111-157)
__device__ __forceinline__ double TslTotal(const double* H, const DTsl& t, double E,
                                           double T) {
  const double* Ex = H + t.E_axis_off;
  const double* Tx = H + t.T_axis_off;
  const int e_hi = UpperBoundArr(Ex, t.nE, E);
  const bool below_E = e_hi == 0;
  const int e_lo = below_E ? e_hi : e_hi - 1;
  const int cand = UpperBoundArr(Tx, t.nT, T);
  const bool above_T = cand == t.nT;
  const int t_hi = above_T ? cand - 1 : cand;
  const bool below_T = t_hi == 0;
  const int t_lo = below_T ? t_hi : t_hi - 1;
  const double* S = H + t.S_off;
  const double* Em = H + t.E_modes_off;
  const double* Tm = H + t.T_modes_off;
  double ll = 0, lh = 0, hl = 0, hh = 0;
  if (t.rank == kRowRegs) {
    // (unrolled with no branch per mode: the LDS reads issue together)
#pragma unroll
    for (int o = 0; o < kRowRegs; o++) {
      const double sel = S[o] * Em[e_lo * kRowRegs + o];
      const double seh = S[o] * Em[e_hi * kRowRegs + o];
      ll = ll + sel * Tm[t_lo * kRowRegs + o];
      lh = lh + sel * Tm[t_hi * kRowRegs + o];
      hl = hl + seh * Tm[t_lo * kRowRegs + o];
      hh = hh + seh * Tm[t_hi * kRowRegs + o];
    }
  } else {
    for (int o = 0; o < t.rank; o++) {
      const double sel = S[o] * Em[e_lo * t.rank + o];
      const double seh = S[o] * Em[e_hi * t.rank + o];
      ll = ll + sel * Tm[t_lo * t.rank + o];
      lh = lh + sel * Tm[t_hi * t.rank + o];
      hl = hl + seh * Tm[t_lo * t.rank + o];
      hh = hh + seh * Tm[t_hi * t.rank + o];
    }
  }
  const double E_lo = Ex[e_lo], E_hi = Ex[e_hi];
  const double r_E = below_E ? 1 : (E - E_lo) / (E_hi - E_lo);
  const double xs_T_lo = ll + r_E * (hl - ll);
  const double xs_T_hi = lh + r_E * (hh - lh);
  const double T_lo = Tx[t_lo], T_hi = Tx[t_hi];
  const double r_T = below_T ? 1 : above_T ? 0 : (T - T_lo) / (T_hi - T_lo);
  return xs_T_lo + r_T * (xs_T_hi - xs_T_lo);
}
// ThermalScattering::SampleBeta (ThermalScattering.cpp:271-338)
// Partition of a concatenated grid index (beta_partition_E_ends /
// alpha_partition_beta_ends) and the index local to it.
__device__ __forceinline__ int PartitionOf(const double* ends, int n_parts, int idx,
                                           int& local) {
  int p_i = 0;
  while (p_i < n_parts && ends[p_i] <= idx) p_i++;
  local = p_i == 0 ? idx : idx - int(ends[p_i - 1]);
  return p_i;
}
// The draws of SampleBeta once the partition and its rows are known.
template <bool kRegs>
__device__ __forceinline__ bool BetaDraws(const double* H, const DTsl& t, const DTslPart& P,
                                          const PartRow& row, uint32_t& g, double E, double E_s,
                                          double T, double& beta) {
  const double* F = H + P.F_off;
  for (int k = 0; k < kBetaResampleLimit; k++) {
    const double u = Canonical(g);
    const int f_hi = UpperBoundArr(F, P.nF, u);
    const double F_lo = f_hi != 0 ? F[f_hi - 1] : 0;
    const double F_hi = f_hi != P.nF ? F[f_hi] : 1;
    double b_lo, b_hi;
    PartEnds<kRegs>(H, P, f_hi, T, row, -E_s / (kBoltzmann * T), t.beta_cutoff, b_lo, b_hi);
    const double b = b_lo + (u - F_lo) / (F_hi - F_lo) * (b_hi - b_lo);
    const double b_min = -E / (kBoltzmann * T);
    if (b_min <= b) {
      beta = b;
      return true;
    }
  }
  return false;
}
// H: the hot block (DevWorld::hot or its copy in LDS).
__device__ __forceinline__ bool SampleBeta(const DevWorld& W, const double* H, const DTsl& t,
                                           uint32_t& g, double E, double T, double& beta) {
  const double* Es = H + t.Es_off;
  const int e_hi = UpperBoundArr(Es, t.nEs, E);
  if (e_hi >= t.nEs) return false;  // E above the beta grid (assert at :278)
  const double r = e_hi != 0 ? (E - Es[e_hi - 1]) / (Es[e_hi] - Es[e_hi - 1]) : 1;
  const int e_s = r <= Canonical(g) ? e_hi - 1 : e_hi;
  const double E_s = Es[e_s];
  int local;
  const int p_i = PartitionOf(H + t.E_ends_off, t.n_bp, e_s, local);
  const DTslPart P = W.parts[t.bp_begin + p_i];
  const PartRow row = LoadPartRow(W, P, local, PartTBracket(H, P, T));
  return P.rank == kRowRegs ? BetaDraws<true>(H, t, P, row, g, E, E_s, T, beta)
                            : BetaDraws<false>(H, t, P, row, g, E, E_s, T, beta);
}
// find_cdf of SampleAlpha (ThermalScattering.cpp:395-418): upper_bound over
// the CDF axis with a POD evaluation inside the comparator (the reconstructed
// values need not ascend, so this stays a bisection), then linear
// interpolation of F.  SampleAlpha needs it at alpha_min and alpha_max; the
// two bisections are independent and run in lockstep here so that their
// memory round trips overlap.  Each follows exactly the probe sequence of a
// single std::upper_bound.
template <bool kRegs>
__device__ __forceinline__ void FindCdf2(const double* H, const DTslPart& P, double T,
                                         const PartRow& row, double a0, double a1,
                                         double alpha_cutoff, double& F0, double& F1) {
  const double* F = H + P.F_off;
  const int nF = P.nF;
  int first0 = 0, len0 = nF, first1 = 0, len1 = nF;
  while (len0 > 0 || len1 > 0) {
    const int half0 = len0 >> 1, half1 = len1 >> 1;
    const int mid0 = first0 + half0, mid1 = first1 + half1;
    // both probes run, at an index inside the axis, whether or not their
    // search is still open, so that their reads and sums overlap
    const double v0 = PartEvaluateAt<kRegs>(H, P, len0 > 0 ? mid0 : 0, T, row);
    const double v1 = PartEvaluateAt<kRegs>(H, P, len1 > 0 ? mid1 : 0, T, row);
    if (len0 > 0) {
      if (a0 < v0) {
        len0 = half0;
      } else {
        first0 = mid0 + 1;
        len0 = len0 - half0 - 1;
      }
    }
    if (len1 > 0) {
      if (a1 < v1) {
        len1 = half1;
      } else {
        first1 = mid1 + 1;
        len1 = len1 - half1 - 1;
      }
    }
  }
  const int h0 = first0, h1 = first1;
  double a_lo0, a_hi0, a_lo1, a_hi1;
  PartEnds<kRegs>(H, P, h0, T, row, 0, alpha_cutoff, a_lo0, a_hi0);
  PartEnds<kRegs>(H, P, h1, T, row, 0, alpha_cutoff, a_lo1, a_hi1);
  const double Fl0 = h0 != 0 ? F[h0 - 1] : 0, Fh0 = h0 != nF ? F[h0] : 1;
  const double Fl1 = h1 != 0 ? F[h1 - 1] : 0, Fh1 = h1 != nF ? F[h1] : 1;
  F0 = Fl0 + (a0 - a_lo0) * (Fh0 - Fl0) / (a_hi0 - a_lo0);
  F1 = Fl1 + (a1 - a_lo1) * (Fh1 - Fl1) / (a_hi1 - a_lo1);
}

// The CDF bounds and draws of SampleAlpha once the partition and its rows
// are known.
template <bool kRegs>
__device__ __forceinline__ uint32_t AlphaDraws(const double* H, const DTsl& t,
                                               const DTslPart& P, const PartRow& row,
                                               uint32_t& g, double b, double E, double T,
                                               double sqrt_E, double akT, double a_min,
                                               double a_max, double& alpha) {
  const double* F = H + P.F_off;
  const int nF = P.nF;
  double F_min, F_max;
  FindCdf2<kRegs>(H, P, T, row, a_min, a_max, t.alpha_cutoff, F_min, F_max);
  for (int k = 0; k < kAlphaResampleLimit; k++) {
    const double u = F_min + Canonical(g) * (F_max - F_min);
    const int hi = UpperBoundArr(F, nF, u);
    const double F_lo = hi != 0 ? F[hi - 1] : 0;
    const double F_hi = hi != nF ? F[hi] : 1;
    double a_lo, a_hi;
    PartEnds<kRegs>(H, P, hi, T, row, 0, t.alpha_cutoff, a_lo, a_hi);
    const double a = a_lo + (u - F_lo) / (F_hi - F_lo) * (a_hi - a_lo);
    if (a_min < a && a < a_max) {
      const double b_root = sqrt(E + b * kBoltzmann * T);
      const double em = sqrt_E - b_root, ep = sqrt_E + b_root;
      const double ba_min = em * em / akT;
      const double ba_max = ep * ep / akT;
      alpha = ba_min + (a - a_min) * (ba_max - ba_min) / (a_max - a_min);
      return 0;
    }
  }
  return kErrResample;
}
// ThermalScattering::SampleAlpha (ThermalScattering.cpp:340-463), quirk Q2
// (operator precedence at :361-365) reproduced.
__device__ __forceinline__ uint32_t SampleAlpha(const DevWorld& W, const double* H,
                                                const DTsl& t, uint32_t& g, double b,
                                                double E, double T, double& alpha) {
  const double* betas = H + t.betas_off;
  const double abs_b = fabs(b);
  const int sgn = (0 < b) - (b < 0);
  const int b_hi = UpperBoundArr(betas, t.nbetas, abs_b);
  if (sgn != 0 && b_hi >= t.nbetas) return kErrAlphaRange;
  const bool snap_lower = (sgn == 1 && t.beta_cutoff <= betas[b_hi]) ||
                          (sgn == -1 && -betas[b_hi] < -E / (kBoltzmann * T));
  const bool snap_min = b_hi == 0;
  const double r = snap_lower ? 0
                   : snap_min ? 1
                              : (abs_b - betas[b_hi - 1] / (betas[b_hi] - betas[b_hi - 1]));
  // |b| below the first grid beta with the lower-snap rule: the reference
  // reads betas.at(-1) (ThermalScattering.cpp:366-368); use the first beta.
  int b_s_i = r <= Canonical(g) ? b_hi - 1 : b_hi;
  if (b_hi == 0) b_s_i = 0;
  double b_s = sgn * betas[b_s_i];
  // ... and if that grid beta would make E' negative, the kinematic bounds
  // are those of the sampled beta itself.
  if (E + b_s * kBoltzmann * T < 0) b_s = b;
  const double sqrt_E = sqrt(E);
  const double bs_root = sqrt(E + b_s * kBoltzmann * T);
  const double akT = t.awr * kBoltzmann * T;
  const double dm = sqrt_E - bs_root, dp = sqrt_E + bs_root;
  const double a_min = dm * dm / akT;
  const double a_max = dp * dp / akT;
  int local;
  const int p_i = PartitionOf(H + t.beta_ends_off, t.n_ap, b_s_i, local);
  const DTslPart P = W.parts[t.ap_begin + p_i];
  const PartRow row = LoadPartRow(W, P, local, PartTBracket(H, P, T));
  return P.rank == kRowRegs
             ? AlphaDraws<true>(H, t, P, row, g, b, E, T, sqrt_E, akT, a_min, a_max, alpha)
             : AlphaDraws<false>(H, t, P, row, g, b, E, T, sqrt_E, akT, a_min, a_max, alpha);
}

// ---- continuous-energy cross sections --------------------------------------
__device__ __forceinline__ bool EvalValid(double T_eval, double T_req) {
  return (T_req - T_eval) / T_eval < kRelTempTol;  // ContinuousEvaluation.cpp:17-20
}
__device__ __forceinline__ bool FreeGasValid(double awr, double E, double T) {
  if (awr <= 1.0) return true;  // ContinuousReaction.cpp:207-223
  return E < 500 * kBoltzmann * T / awr;
}
__device__ __forceinline__ double FreeGasAdjust(double awr, double E, double T) {
  if (T == 0) return 1;  // ContinuousReaction.cpp:225-237
  const double x = sqrt(E / (kBoltzmann * T));
  const double arg = awr * x * x;
  return (1 + 1 / (2 * arg)) * erf(sqrt(arg)) + exp(-arg) / sqrt(kPi * arg);
}

// Lookups repeated within one event reuse their results.  An event keeps
// the last grid search (grid, E) -> index and the last thermal-scattering
// total (table, E, T) -> value; the grid index found by the previous event
// of the same particle comes in as a hint and is used only after checking
// that it is the upper bound of E (x[i-1] <= E < x[i]), so every value
// equals a fresh lookup's.
struct XsMemo {
  int grid = -1, gi = 0;
  double E = 0;
  int hint_grid = -1, hint_gi = 0;
  // hint_exact: the hint is the grid index of this very energy (the caller
  // knows it: a collision follows the advance that computed it, and nothing
  // between them changes the energy), so it is taken without reading the grid
  bool hint_exact = false;
  int tsl = -1;
  double tE = 0, tT = 0, tval = 0;
  const double* tot = nullptr;  // DevWorld::tot or its copy in LDS (never null in use)
};

// Grid index of E for a nuclide whose tables share one grid, or -1.
// lds_hash: grid 0's bucket hash staged in LDS, or nullptr.
__device__ __forceinline__ int NuclideGridIndex(const DevWorld& W, const DCeNuclide& n,
                                                const int* lds_hash, double E,
                                                XsMemo* memo = nullptr) {
  if (n.grid < 0) return -1;
  if (memo && memo->grid == n.grid && memo->E == E) return memo->gi;
  const DGrid g = W.grids[n.grid];
  int gi = -1;
  if (memo && memo->hint_grid == n.grid) {
    const double* x = W.pool + g.x_off;
    const int i = memo->hint_gi;
    if (i >= 0 && i <= g.n &&
        (memo->hint_exact || ((i == 0 || x[i - 1] <= E) && (i == g.n || x[i] > E))))
      gi = i;
  }
  // lds_hash is the LDS copy of grid 0's hash; other grids use their own
  if (gi < 0)
    gi = GridUpperBound(W, g, lds_hash && n.grid == 0 ? lds_hash : W.ipool + g.hash_off, E);
  if (memo) {
    memo->grid = n.grid;
    memo->E = E;
    memo->gi = gi;
  }
  return gi;
}
__device__ __forceinline__ double TabAt(const DevWorld& W, int t, int gi, double E) {
  return gi >= 0 ? TableAtI(W, t, gi, E) : TableAt(W, t, E);
}

// ContinuousScatter::GetCrossSection / GetMajorant (ContinuousReaction.cpp:77-117)
__device__ __forceinline__ double RxXs(const DevWorld& W, const DCeNuclide& n,
                                       const DCeReaction& r, int gi, double E, double T,
                                       bool majorant, XsMemo* memo = nullptr) {
  if (r.kind != 1) return TabAt(W, r.xs, gi, E);
  if (r.tsl >= 0) {
    const DTsl& t = W.tsl[r.tsl];
    if (E < t.cutoff_energy) {
      if (majorant) return TableAt(W, t.majorant, E);
      if (memo && memo->tsl == r.tsl && memo->tE == E && memo->tT == T) return memo->tval;
      const double v = TslTotal(memo ? memo->tot : W.tot, t, E, T);
      if (memo) {
        memo->tsl = r.tsl;
        memo->tE = E;
        memo->tT = T;
        memo->tval = v;
      }
      return v;
    }
  }
  if (EvalValid(r.temperature, T)) return TabAt(W, r.xs, gi, E);
  if (FreeGasValid(n.awr, E, T))
    return TabAt(W, r.xs, gi, E) / FreeGasAdjust(n.awr, E, r.temperature) *
           FreeGasAdjust(n.awr, E, T);
  return TabAt(W, r.xs, gi, E);
}
__device__ __forceinline__ bool ModifiesTotal(const DevWorld& W, const DCeNuclide& n,
                                              double E) {
  for (int k = 0; k < n.rx_count; k++) {
    const DCeReaction r = W.rx[n.rx_begin + k];
    if (r.kind == 1 && r.tsl >= 0 && E < W.tsl[r.tsl].cutoff_energy) return true;
  }
  return false;
}
// Continuous::GetTotal / GetMajorant (Continuous.cpp:26-55)
__device__ __forceinline__ double CeTotal(const DevWorld& W, const DCeNuclide& n, int gi,
                                          double E, double T, bool majorant,
                                          XsMemo* memo = nullptr) {
  if (n.has_total && !ModifiesTotal(W, n, E) && EvalValid(n.total_temperature, T))
    return TabAt(W, n.total, gi, E);
  double acc = 0;
  for (int k = 0; k < n.rx_count; k++)
    acc = acc + RxXs(W, n, W.rx[n.rx_begin + k], gi, E, T, majorant, memo);
  return acc;
}

// Microscopic material total (or majorant): Material.cpp:41-61.
// nt_nuc / nt_val (a total that is not the majorant, optional): on entry a
// nuclide whose microscopic total CeTotal at this (E, T) the caller knows
// (nt_val), or -1; on return the last nuclide summed and its total.
__device__ __forceinline__ double MaterialTotal(const DevWorld& W, int mat, double E,
                                                double T, bool majorant,
                                                const int* lds_hash, XsMemo* memo = nullptr,
                                                int* nt_nuc = nullptr,
                                                double* nt_val = nullptr) {
  const DMaterial m = W.mats[mat];
  double acc = 0;
  if (W.energy_mode == 0) {
    const int g = int(E) - 1;
    for (int k = 0; k < m.comp_count; k++) {
      const DComp c = W.comps[m.comp_begin + k];
      acc = acc + c.afrac * W.mg_total[c.nuclide * W.G + g];
    }
    return acc;
  }
  for (int k = 0; k < m.comp_count; k++) {
    const DComp c = W.comps[m.comp_begin + k];
    const DCeNuclide n = W.ce[c.nuclide];
    const int gi = NuclideGridIndex(W, n, lds_hash, E, memo);
    const double v = nt_nuc && !majorant && *nt_nuc == c.nuclide
                         ? *nt_val
                         : CeTotal(W, n, gi, E, T, majorant, memo);
    if (nt_nuc && !majorant) {
      *nt_nuc = c.nuclide;
      *nt_val = v;
    }
    acc = acc + c.afrac * v;
  }
  return acc;
}
// nu Sigma_f (microscopic, afrac-weighted) for the k estimators.
__device__ __forceinline__ double MaterialNuFission(const DevWorld& W, int mat, double E,
                                                    const int* lds_hash = nullptr,
                                                    XsMemo* memo = nullptr) {
  const DMaterial m = W.mats[mat];
  double acc = 0;
  for (int k = 0; k < m.comp_count; k++) {
    const DComp c = W.comps[m.comp_begin + k];
    if (W.energy_mode == 0) {
      const int g = int(E) - 1;
      acc += c.afrac * (W.mg_nubar[c.nuclide * W.G + g] * W.mg_fission[c.nuclide * W.G + g]);
    } else {
      const DCeNuclide n = W.ce[c.nuclide];
      if (!n.fissile) continue;
      const int gi = NuclideGridIndex(W, n, lds_hash, E, memo);
      double v = 0;
      for (int q = 0; q < n.rx_count; q++) {
        const DCeReaction r = W.rx[n.rx_begin + q];
        if (r.kind == 2) v += TableAt(W, r.nubar, E) * TabAt(W, r.xs, gi, E);
      }
      acc += c.afrac * v;
    }
  }
  return acc;
}

// Watt spectrum sampling (additive extension; LA-9721 rule C64 Maxwellian).
__device__ __forceinline__ double SampleWatt(uint32_t& g, double a, double b) {
  const double xi1 = Canonical(g), xi2 = Canonical(g), xi3 = Canonical(g);
  const double c = cos(kPi * xi3 / 2);
  const double Em = -a * (log(xi1) + log(xi2) * c * c);
  const double xi4 = Canonical(g);
  return Em + a * a * b / 4 + (2 * xi4 - 1) * sqrt(a * a * b * Em);
}

// ---- tally bins (src/Bins.cpp) ---------------------------------------------
__device__ __forceinline__ int BinIndex(const DevWorld& W, const DBins& b, double v) {
  switch (b.kind) {
    case 1:
      if (v < b.lo) return 0;
      if (v >= b.hi) return b.n_bins - 1;
      return int((v - b.lo) / b.width + 1);
    case 2: {
      const double lv = log(v) / b.base_log;
      if (lv < b.lo) return 0;
      if (lv >= b.hi) return b.n_bins - 1;
      return int((lv - b.lo) / b.width + 1);
    }
    case 3:
      return UpperBoundArr(W.pool + b.bounds_off, b.n_bounds, v);
    default:
      return 0;
  }
}

}  // namespace mmc
