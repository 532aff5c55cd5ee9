// ORACLE — test infrastructure only (see oracle.h).  Built with
// g++ -O3 -ffp-contract=off; the product never links it.
//
// A CPU restatement of the reference's particle-history hot path.  Each
// function cites the reference file:line it follows (paths relative to the
// reference's repository root).  Where the reference is silent (k-eigenvalue
// cycles, the flux/alpha/beta estimators, Watt fission spectra, plane-y/z
// surfaces, uniform-energy sources) this file defines the behaviour and the
// product follows the same definition.
//
// Parity pinning (what checks each part, in tests/):
//   RNG            C++ standard known answer (10000th minstd_rand output from
//                  seed 1 is 399268537) + canonical closed form.
//   geometry       reference test vectors (test/test_CSGSurface.cpp:23-56,
//                  test/test_Cell.cpp:22-27, test/test_World.cpp:46-53,
//                  test/test_Point.cpp:49-58) + closed-form roots.
//   multigroup     reference totals (test/test_Nuclide.cpp:68-73),
//                  infinite-medium k_inf closed form, point-source absorber
//                  leakage e^-1 (test/test_FixedSource.cpp:13-26).
//   continuous     lin-lin closed form on synthetic grids; free-gas cross
//                  section adjustment closed form; free-gas kernel: the
//                  stationary collision-energy law (mean 7/4 kT at A = 1, by
//                  quadrature) and the stationary-target limit.
//   fission        Watt spectrum: mean 3a/2 + a^2 b/4 and a KS test against
//                  the integrated Watt CDF on a k-eigenvalue bank.
//   TSL POD        rank-full reconstruction equals the tabulated inverse
//                  CDFs; sampled alpha/beta marginals against the analytic
//                  S(alpha,beta) by KS/W1; alpha bounds closed form.
//   tallies        mean / std-dev formula (src/Scorable.cpp:51-70) on
//                  hand-made score streams.
//   k-eigenvalue   no reference behaviour (src/KEigenvalue.cpp:36-62 is a
//                  stub): k_inf closed form in an infinite medium; Shannon
//                  entropy recomputed from the bank in the test.  The
//                  systematic resampling is parity unpinned.
//   perturbations  indirect effects (src/Perturbation.cpp:68-78): in a pure
//                  absorber around a point source every leaking history
//                  streams once, so its indirect effect is exactly
//                  1/sigma - (R + nudge) and the sensitivity tally equals
//                  that times the leakage tally.
//   output text    orc_format_* write the .out blocks with libstdc++'s own
//                  std::scientific streams, the formatter the reference
//                  uses; nothing else pins them (the reference ships no
//                  .out files).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "rng.hpp"

namespace orc {

// ---- constants (src/Constants.hpp:13-35) ---------------------------------
constexpr double kPi = M_PI;
constexpr double kNudge = 10 * std::numeric_limits<double>::epsilon();
constexpr double kOnAxis = 0.9;
constexpr double kBoltzmann = 8.617333262145e-11;
constexpr double kNeutronMass = 1.045354912280858e-18;
constexpr double kRelTempTol = 0.01;
constexpr int kBetaResampleLimit = 10;
constexpr int kAlphaResampleLimit = 10;
constexpr double kInf = std::numeric_limits<double>::infinity();

enum ErrBits : int {
  kErrLost = 1,
  kErrResample = 2,
  kErrSteps = 4,
  kErrAlphaRange = 8,
  kErrOverflow = 16,  // a 65th child of one history (site keys hold 64)
  kErrReaction = 32
};
constexpr int kMaxReactionResamples = 1000;

// ---- Point / Direction (src/Point.cpp) -----------------------------------
struct Pt {
  double x = 0, y = 0, z = 0;
};
inline Pt operator+(const Pt& l, const Pt& r) {
  return {l.x + r.x, l.y + r.y, l.z + r.z};
}
inline Pt operator-(const Pt& l, const Pt& r) {
  return {l.x - r.x, l.y - r.y, l.z - r.z};
}
// Point * Real and Real * Point both evaluate component * scalar
// (Point.cpp:14-22).
inline Pt operator*(const Pt& l, double r) { return {l.x * r, l.y * r, l.z * r}; }
inline Pt operator*(double l, const Pt& r) { return r * l; }
inline Pt operator/(const Pt& l, double r) { return {l.x / r, l.y / r, l.z / r}; }
inline double Dot(const Pt& a, const Pt& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;  // Point::Dot, Point.cpp:42-44
}
inline Pt Cross(const Pt& a, const Pt& b) {  // Point::Cross, Point.cpp:46-49
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Pt Normalized(Pt p) {  // Point::Normalize, Point.cpp:37-40
  const double n = std::sqrt(Dot(p, p));
  p.x /= n;
  p.y /= n;
  p.z /= n;
  return p;
}
// Direction(RNG&), Point.cpp:90-96: not normalised afterwards.
inline Pt IsotropicDirection(Lcg& g) {
  Pt d;
  d.x = Uniform(g, -1., +1.);
  const double sin_theta = std::sqrt(1 - d.x * d.x);
  const double phi = Uniform(g, 0., 2 * kPi);
  d.y = sin_theta * std::cos(phi);
  d.z = sin_theta * std::sin(phi);
  return d;
}
// Direction(d, mu, phi), Point.cpp:98-121, result normalised through
// Direction(Point&&) (Point.cpp:127-129).
inline Pt Rotate(const Pt& d, double mu, double phi) {
  const bool is_off_xaxis = d.x <= kOnAxis && d.x > -kOnAxis;
  const Pt xaxis = is_off_xaxis ? Pt{1, 0, 0} : Pt{0, 1, 0};
  const Pt u = Normalized(Cross(d, xaxis));
  const Pt v = Normalized(Cross(d, u));
  const Pt u_comp = std::sqrt(1 - mu * mu) * std::cos(phi) * u;
  const Pt v_comp = std::sqrt(1 - mu * mu) * std::sin(phi) * v;
  const Pt d_comp = mu * d;
  return Normalized(u_comp + v_comp + d_comp);
}

// ---- ContinuousMap::at (src/ContinuousMap.hpp:20-56) ----------------------
struct Table {
  std::vector<double> x, y;
  bool empty() const { return x.empty(); }
  double at(double k) const {
    const size_t hi = std::upper_bound(x.begin(), x.end(), k) - x.begin();
    if (hi == x.size()) return y[hi - 1];
    if (hi == 0) return y[0];
    const double k_hi = x[hi], v_hi = y[hi];
    const double k_lo = x[hi - 1], v_lo = y[hi - 1];
    return v_lo + (v_hi - v_lo) / (k_hi - k_lo) * (k - k_lo);
  }
};
Table MakeTable(const mmc_table& t) {
  Table r;
  if (t.n > 0 && t.x && t.y) {
    r.x.assign(t.x, t.x + t.n);
    r.y.assign(t.y, t.y + t.n);
  }
  return r;
}

// libstdc++ std::upper_bound with a value comparator (bits/stl_algo.h).
template <class Less>
size_t UpperBoundIdx(size_t n, Less value_less_than_elem) {
  size_t first = 0, len = n;
  while (len > 0) {
    const size_t half = len >> 1;
    const size_t mid = first + half;
    if (value_less_than_elem(mid)) {
      len = half;
    } else {
      first = mid + 1;
      len = len - half - 1;
    }
  }
  return first;
}
inline size_t UpperBound(const std::vector<double>& v, double x) {
  return std::upper_bound(v.begin(), v.end(), x) - v.begin();
}

// ---- World ---------------------------------------------------------------
struct Surface {
  int kind;
  double a, b, c, r;
};
// SolveQuadratic, CSGSurface.cpp:71-97.
double SolveQuadratic(double a, double b, double c) {
  const double disc = b * b - 4 * a * c;
  if (disc <= 0) return kInf;
  const double lesser = b > 0 ? (-b - std::sqrt(disc)) / (2 * a)
                              : (2 * c) / (-b + std::sqrt(disc));
  const double greater = b > 0 ? (2 * c) / (-b - std::sqrt(disc))
                               : (-b + std::sqrt(disc)) / (2 * a);
  if (lesser > 0) return lesser;
  if (greater > 0) return greater;
  return kInf;
}
// Sphere (CSGSurface.cpp:107-116), PlaneX (:125-138), CylinderX (:148-174);
// plane-y and plane-z follow PlaneX with the other axes.
double Distance(const Surface& s, const Pt& o, const Pt& d) {
  switch (s.kind) {
    case MMC_SPHERE: {
      const Pt oc = o - Pt{s.a, s.b, s.c};
      return SolveQuadratic(1, 2 * Dot(oc, d), Dot(oc, oc) - s.r * s.r);
    }
    case MMC_PLANEX:
    case MMC_PLANEY:
    case MMC_PLANEZ: {
      const Pt axis = s.kind == MMC_PLANEX   ? Pt{1, 0, 0}
                      : s.kind == MMC_PLANEY ? Pt{0, 1, 0}
                                             : Pt{0, 0, 1};
      const double v = s.a - Dot(o, axis);
      const double dd = Dot(d, axis);
      const double t = v / dd;
      return t > 0 ? t : kInf;
    }
    case MMC_CYLINDERX: {
      const Pt w{1, 0, 0};
      const double pp = Dot(o, o);
      const double pw = Dot(o, w);
      const double dw = Dot(d, w);
      const double pd = Dot(o, d);
      return SolveQuadratic(1 - dw * dw, 2 * (pd - pw * dw),
                            pp - pw * pw - s.r * s.r);
    }
  }
  return kInf;
}
bool Contains(const Surface& s, const Pt& p) {
  switch (s.kind) {
    case MMC_SPHERE: {
      const Pt pc = p - Pt{s.a, s.b, s.c};
      return Dot(pc, pc) < s.r * s.r;
    }
    case MMC_PLANEX:
      return Dot(p, Pt{1, 0, 0}) < s.a;
    case MMC_PLANEY:
      return Dot(p, Pt{0, 1, 0}) < s.a;
    case MMC_PLANEZ:
      return Dot(p, Pt{0, 0, 1}) < s.a;
    case MMC_CYLINDERX: {
      // Quirk Q1 reproduced: sqrt(rho^2) compared with r^2
      // (CSGSurface.cpp:164-174).
      const Pt w{1, 0, 0};
      const double pp = Dot(p, p);
      const double pw = Dot(p, w);
      return std::sqrt(pp - pw * pw) < s.r * s.r;
    }
  }
  return false;
}

struct Field {  // ScalarField (ScalarField.cpp:36-58)
  int kind;
  double c, gx, gy, gz, b, lower, upper;
  double at(const Pt& p) const {
    if (kind == MMC_FIELD_CONSTANT) return c;
    return Dot(Pt{gx, gy, gz}, p) + b;
  }
  double upper_bound() const { return kind == MMC_FIELD_CONSTANT ? c : upper; }
};

struct Cell {
  int material;
  std::vector<std::pair<int, bool>> faces;  // (surface, inside)
  int field;
};
struct Material {
  double aden;
  std::vector<std::pair<int, double>> comps;  // normalised afrac
};

struct MgNuclide {
  std::vector<double> capture, scatter_xs, fission_xs, nubar, total;
  // per incoming group g (0-based): probabilities over outgoing groups
  std::vector<std::vector<double>> scatter_probs, chi_probs;
  bool has_scatter = false, has_fission = false;
};

struct TslPartition {
  int nF, nx, nT, rank;
  std::vector<double> F, cdf, S, x, T, xt;
  // BetaPartition::Evaluate / AlphaPartition::Evaluate
  // (ThermalScattering.cpp:183-215, :225-256)
  double Evaluate(size_t cdf_i, size_t x_i, double Tq) const {
    const size_t cand = UpperBound(T, Tq);
    const bool above = cand == T.size();
    const size_t t_hi = above ? cand - 1 : cand;
    const bool below = t_hi == 0;
    const size_t t_lo = below ? t_hi : t_hi - 1;
    double v_hi = 0, v_lo = 0;
    for (int o = 0; o < rank; o++) {
      v_hi += S[o] * cdf[cdf_i * rank + o] * xt[(x_i * nT + t_hi) * rank + o];
      v_lo += S[o] * cdf[cdf_i * rank + o] * xt[(x_i * nT + t_lo) * rank + o];
    }
    const double T_hi = T[t_hi], T_lo = T[t_lo];
    return v_lo + (v_hi - v_lo) / (T_hi - T_lo) * (Tq - T_lo);
  }
};

struct Tsl {
  Table majorant;
  int nE, nT, rank;
  std::vector<double> E_axis, E_modes, T_axis, T_modes, S;
  std::vector<TslPartition> beta_parts, alpha_parts;
  std::vector<double> Es, betas;
  std::vector<size_t> E_ends, beta_ends;
  double beta_cutoff, alpha_cutoff, awr, cutoff_energy;
};

struct CeReaction {
  int kind;
  int tsl;
  Table xs;
  double temperature;
  Table nubar;
  int chi_kind;
  double watt_a, watt_b;
};
struct CeNuclide {
  double awr;
  bool has_total;
  Table total;
  double total_temperature;
  std::vector<CeReaction> rx;
};

struct Bins {
  int kind = MMC_BINS_NONE;
  size_t n_bins = 1;
  double lo = 0, hi = 0, base = 10, width = 0;
  std::vector<double> bounds;
  size_t size() const { return n_bins; }
  // NoBins / LinspaceBins / LogspaceBins / BoundaryBins ::GetIndex
  // (Bins.cpp:44-48, :72-82, :112-123, :158-162)
  size_t index(double v) const {
    switch (kind) {
      case MMC_BINS_LINSPACE:
        if (v < lo) return 0;
        if (v >= hi) return n_bins - 1;
        return size_t((v - lo) / width + 1);
      case MMC_BINS_LOGSPACE: {
        const double lv = std::log(v) / std::log(base);
        if (lv < lo) return 0;
        if (lv >= hi) return n_bins - 1;
        return size_t((lv - lo) / width + 1);
      }
      case MMC_BINS_BOUNDARIES:
        return UpperBound(bounds, v);
      default:
        return 0;
    }
  }
};
Bins MakeBins(const mmc_bins& b) {
  Bins r;
  r.kind = b.kind;
  switch (b.kind) {
    case MMC_BINS_LINSPACE:
    case MMC_BINS_LOGSPACE:
      r.n_bins = size_t(b.n) + 2;
      r.lo = b.lo;
      r.hi = b.hi;
      r.base = b.kind == MMC_BINS_LOGSPACE ? (b.base > 0 ? b.base : 10.0) : 10;
      r.width = (r.hi - r.lo) / double(r.n_bins - 2);
      break;
    case MMC_BINS_BOUNDARIES:
      r.bounds.assign(b.bounds, b.bounds + b.n);
      r.n_bins = size_t(b.n) + 1;
      break;
    default:
      r.n_bins = 1;
  }
  return r;
}

struct Estimator {
  int kind;
  int surface;
  int sens = 0;  // 0, or 1 + perturbation: CurrentTotalCrossSectionSensitivity
  bool has_dir;
  Pt dir;
  Bins cosine, energy;
  int mesh_n[3];
  double mesh_lo[3], mesh_hi[3];
  size_t size() const {
    if (kind == MMC_EST_FLUX_MESH)
      return size_t(mesh_n[0]) * mesh_n[1] * mesh_n[2];
    if (kind == MMC_EST_TSL_ALPHA || kind == MMC_EST_TSL_BETA)
      return energy.size();
    return cosine.size() * energy.size();
  }
};

struct Particle {
  Pt pos, dir;
  double E = 0;  // MeV, or group number for multigroup
  int cell = -1;
  Lcg rng;
  int event = MMC_EV_BIRTH;
  int surface = -1;
  uint64_t hist = 0;
  // indirect effect of each total cross-section perturbation
  // (Particle::indirect_effects, Particle.hpp:105-106); zero for a new
  // particle, secondaries included (the reference gives secondaries none)
  double ie[MMC_MAX_PERTURBATIONS] = {0, 0, 0, 0};
  std::vector<Particle> secondaries;
};

struct Source {
  mmc_source_desc d;
};

// Per-thread accumulators.
struct Acc {
  std::vector<std::vector<double>> S, S2;
  mmc_counters cnt{};
  double k_col = 0, k_trk = 0, k_abs = 0;
  int err = 0;
  void init(const std::vector<Estimator>& ests) {
    S.resize(ests.size());
    S2.resize(ests.size());
    for (size_t e = 0; e < ests.size(); e++) {
      S[e].assign(ests[e].size(), 0.0);
      S2[e].assign(ests[e].size(), 0.0);
    }
  }
};

struct World {
  int energy_mode = MMC_MULTIGROUP;
  int G = 0;
  int tracking = MMC_SURFACE_TRACKING;
  int quirks = 0;
  uint64_t max_steps = 1000000;
  std::vector<Surface> surfaces;
  std::vector<Field> fields;
  std::vector<Cell> cells;
  std::vector<Material> materials;
  std::vector<MgNuclide> mg;
  std::vector<CeNuclide> ce;
  std::vector<Tsl> tsl;
  std::vector<Estimator> ests;
  std::vector<int> pert_nuclide;  // perturbed nuclide of each perturbation
  Acc total;  // merged tallies
  std::string error;
  // k-eigenvalue state
  std::vector<mmc_site> source;
  int mesh_n[3] = {1, 1, 1};
  double mesh_lo[3] = {-1e300, -1e300, -1e300};
  double mesh_hi[3] = {1e300, 1e300, 1e300};
  std::vector<std::vector<double>> cycle_S;  // per estimator
};

// ---- World queries --------------------------------------------------------
// Cell::Contains (Cell.cpp:27-35)
bool CellContains(const World& w, const Cell& c, const Pt& p) {
  for (const auto& f : c.faces)
    if (Contains(w.surfaces[f.first], p) != f.second) return false;
  return true;
}
// World::FindCellContaining (World.cpp:26-37): first cell in order.
int FindCell(const World& w, const Pt& p) {
  for (size_t i = 0; i < w.cells.size(); i++)
    if (CellContains(w, w.cells[i], p)) return int(i);
  return -1;
}
// Crossing without the nudge quirk: the first cell whose face on the crossed
// surface has the flipped sense and whose other faces contain the point.
int FindCellAcross(const World& w, const Pt& p, int surf, bool new_inside) {
  for (size_t i = 0; i < w.cells.size(); i++) {
    bool ok = true;
    for (const auto& f : w.cells[i].faces) {
      ok = f.first == surf ? f.second == new_inside
                           : Contains(w.surfaces[f.first], p) == f.second;
      if (!ok) break;
    }
    if (ok) return int(i);
  }
  return -1;
}
bool FaceInside(const World& w, int cell, int surf) {
  for (const auto& f : w.cells[cell].faces)
    if (f.first == surf) return f.second;
  return false;
}

// Cell::NearestSurface (Cell.cpp:37-51): min_element keeps the first of
// equal distances; the winner's distance is computed again.
std::pair<int, double> NearestSurface(const World& w, const Cell& c,
                                      const Pt& p, const Pt& d) {
  if (c.faces.empty()) return {-1, kInf};
  size_t best = 0;
  for (size_t i = 1; i < c.faces.size(); i++) {
    if (Distance(w.surfaces[c.faces[i].first], p, d) <
        Distance(w.surfaces[c.faces[best].first], p, d))
      best = i;
  }
  const int s = c.faces[best].first;
  return {s, Distance(w.surfaces[s], p, d)};
}

// ---- continuous-energy cross sections -------------------------------------
// ContinuousEvaluation::IsValid (ContinuousEvaluation.cpp:17-20), quirk Q7.
inline bool EvalValid(double T_eval, double T_req) {
  const double rel = (T_req - T_eval) / T_eval;
  return rel < kRelTempTol;
}
// ThermalScattering::IsValid (ThermalScattering.cpp:101-104)
inline bool TslValid(const Tsl& t, double E) { return E < t.cutoff_energy; }

// ThermalScattering::EvaluateInelastic (ThermalScattering.cpp:260-269)
double EvaluateInelastic(const Tsl& t, size_t ei, size_t ti) {
  double r = 0;
  for (int o = 0; o < t.rank; o++)
    r += t.S[o] * t.E_modes[ei * t.rank + o] * t.T_modes[ti * t.rank + o];
  return r;
}
// ThermalScattering::GetTotal (ThermalScattering.cpp:111-157)
double TslTotal(const Tsl& t, double E, double T) {
  const size_t e_hi = UpperBound(t.E_axis, E);
  const bool below_E = e_hi == 0;
  const size_t e_lo = below_E ? e_hi : e_hi - 1;
  const size_t cand = UpperBound(t.T_axis, T);
  const bool above_T = cand == t.T_axis.size();
  const size_t t_hi = above_T ? cand - 1 : cand;
  const bool below_T = t_hi == 0;
  const size_t t_lo = below_T ? t_hi : t_hi - 1;
  const double ll = EvaluateInelastic(t, e_lo, t_lo);
  const double lh = EvaluateInelastic(t, e_lo, t_hi);
  const double hl = EvaluateInelastic(t, e_hi, t_lo);
  const double hh = EvaluateInelastic(t, e_hi, t_hi);
  const double E_lo = t.E_axis[e_lo], E_hi = t.E_axis[e_hi];
  const double r_E = below_E ? 1 : (E - E_lo) / (E_hi - E_lo);
  const double xs_T_lo = ll + r_E * (hl - ll);
  const double xs_T_hi = lh + r_E * (hh - lh);
  const double T_lo = t.T_axis[t_lo], T_hi = t.T_axis[t_hi];
  const double r_T = below_T ? 1 : above_T ? 0 : (T - T_lo) / (T_hi - T_lo);
  return xs_T_lo + r_T * (xs_T_hi - xs_T_lo);
}

// ContinuousScatter::IsFreeGasScatteringValid (ContinuousReaction.cpp:207-223)
inline bool FreeGasValid(double awr, double E, double T) {
  if (awr <= 1.0) return true;
  return E < 500 * kBoltzmann * T / awr;
}
// ContinuousScatter::GetFreeGasScatterAdjustment (ContinuousReaction.cpp:225-237)
inline double FreeGasAdjust(double awr, double E, double T) {
  if (T == 0) return 1;
  const double x = std::sqrt(E / (kBoltzmann * T));
  const double arg = awr * x * x;
  return (1 + 1 / (2 * arg)) * std::erf(std::sqrt(arg)) +
         std::exp(-arg) / std::sqrt(kPi * arg);
}

// ContinuousScatter::GetCrossSection / GetMajorant (ContinuousReaction.cpp:77-117);
// ContinuousReaction::GetCrossSection (:52-55) for capture / fission.
double RxXs(const World& w, const CeNuclide& n, const CeReaction& r, double E,
            double T, bool majorant) {
  if (r.kind != MMC_RX_SCATTER) return r.xs.at(E);
  if (r.tsl >= 0 && TslValid(w.tsl[r.tsl], E))
    return majorant ? w.tsl[r.tsl].majorant.at(E) : TslTotal(w.tsl[r.tsl], E, T);
  if (EvalValid(r.temperature, T)) return r.xs.at(E);
  if (FreeGasValid(n.awr, E, T))
    return r.xs.at(E) / FreeGasAdjust(n.awr, E, r.temperature) *
           FreeGasAdjust(n.awr, E, T);
  return r.xs.at(E);
}
bool ModifiesTotal(const World& w, const CeNuclide& n, double E) {
  for (const auto& r : n.rx)
    if (r.kind == MMC_RX_SCATTER && r.tsl >= 0 && TslValid(w.tsl[r.tsl], E))
      return true;
  return false;
}
// Continuous::GetTotal / GetMajorant (Continuous.cpp:26-55); T is the
// requested temperature (cell field at the position) or, for the majorant,
// the field's upper bound.
double CeTotal(const World& w, const CeNuclide& n, double E, double T,
               bool majorant) {
  if (n.has_total && !ModifiesTotal(w, n, E) &&
      EvalValid(n.total_temperature, T))
    return n.total.at(E);
  double acc = 0;
  for (const auto& r : n.rx) acc = acc + RxXs(w, n, r, E, T, majorant);
  return acc;
}

double NuclideTotal(const World& w, int nuc, const Particle& p, double T,
                    bool majorant) {
  if (w.energy_mode == MMC_MULTIGROUP) {
    // Multigroup::GetTotal / GetMajorant (Multigroup.cpp:40-47)
    return w.mg[nuc].total[size_t(p.E) - 1];
  }
  return CeTotal(w, w.ce[nuc], p.E, T, majorant);
}
double CellT(const World& w, const Particle& p) {
  return w.fields[w.cells[p.cell].field].at(p.pos);
}
double CellTmax(const World& w, const Particle& p) {
  return w.fields[w.cells[p.cell].field].upper_bound();
}
// Material::GetMicroscopicTotal / GetMicroscopicMajorant (Material.cpp:41-61)
double MaterialTotal(const World& w, const Particle& p, bool majorant) {
  const Material& m = w.materials[w.cells[p.cell].material];
  const double T = majorant ? CellTmax(w, p) : CellT(w, p);
  double acc = 0;
  for (const auto& c : m.comps)
    acc = acc + c.second * NuclideTotal(w, c.first, p, T, majorant);
  return acc;
}
// nu Sigma_f / sigma pieces for the k estimators (new behaviour).
double NuclideNuFission(const World& w, int nuc, const Particle& p) {
  if (w.energy_mode == MMC_MULTIGROUP) {
    const auto& n = w.mg[nuc];
    if (!n.has_fission) return 0;
    const size_t g = size_t(p.E) - 1;
    return n.nubar[g] * n.fission_xs[g];
  }
  double acc = 0;
  for (const auto& r : w.ce[nuc].rx)
    if (r.kind == MMC_RX_FISSION) acc += r.nubar.at(p.E) * r.xs.at(p.E);
  return acc;
}
double MaterialNuFission(const World& w, const Particle& p) {
  const Material& m = w.materials[w.cells[p.cell].material];
  double acc = 0;
  for (const auto& c : m.comps) acc += c.second * NuclideNuFission(w, c.first, p);
  return acc;
}

// ---- physics --------------------------------------------------------------
struct Ctx {
  World* w;
  Acc* acc;
  bool keig;
  std::vector<mmc_site>* bank;  // k-eigenvalue fission bank (per thread)
  uint32_t child_count = 0;
  // per-history pending scores (ScorableProxy::pending_scores)
  std::vector<std::map<size_t, double>> pending;
};

// Particle::BankSecondaries (Particle.cpp:96-100): the child's seed is the
// parent's next raw draw, taken after the child's direction (and energy).
void BankSecondary(Ctx& c, Particle& p, const Pt& dir, double E) {
  const uint32_t seed = uint32_t(p.rng());
  if (c.keig && c.child_count >= 64) {  // nu-bar < 63 is checked at orc_create
    c.acc->err |= kErrOverflow;
    return;
  }
  if (c.keig) {
    mmc_site s{};
    s.x = p.pos.x;
    s.y = p.pos.y;
    s.z = p.pos.z;
    s.u = dir.x;
    s.v = dir.y;
    s.w = dir.z;
    s.E = E;
    s.seed = seed;
    s.key = p.hist * 64 + c.child_count++;
    c.bank->push_back(s);
  } else {
    Particle ch;
    ch.pos = p.pos;
    ch.dir = dir;
    ch.E = E;
    ch.cell = p.cell;
    ch.rng.Seed(seed);
    ch.hist = p.hist;
    p.secondaries.push_back(ch);
  }
  c.acc->cnt.fission_sites++;
}

// Particle::Scatter (Particle.cpp:55-64)
void ScatterTo(Particle& p, double mu, double E) {
  const double phi = 2 * kPi * Canonical(p.rng);
  p.dir = Rotate(p.dir, mu, phi);
  p.E = E;
}

// Watt fission spectrum (extension): Maxwellian with temperature a by
// LA-9721 rule C64, then E = E_m + a^2 b / 4 + (2 xi - 1) sqrt(a^2 b E_m).
double SampleWatt(Lcg& g, double a, double b) {
  const double xi1 = Canonical(g), xi2 = Canonical(g), xi3 = Canonical(g);
  const double c = std::cos(kPi * xi3 / 2);
  const double Em = -a * (std::log(xi1) + std::log(xi2) * c * c);
  const double xi4 = Canonical(g);
  return Em + a * a * b / 4 + (2 * xi4 - 1) * std::sqrt(a * a * b * Em);
}

// Multigroup::Interact, Scatter, Fission (Multigroup.cpp:49-73, :249-287)
void MgInteract(Ctx& c, const MgNuclide& n, Particle& p) {
  const size_t gi = size_t(p.E) - 1;
  const double threshold = Canonical(p.rng) * n.total[gi];
  double acc = 0;
  acc += n.capture[gi];
  if (acc > threshold) {
    p.event = MMC_EV_CAPTURE;
    return;
  }
  acc += n.scatter_xs[gi];
  if (acc > threshold) {
    p.event = MMC_EV_SCATTER;
    const double t = Canonical(p.rng);
    double a = 0;
    const auto& probs = n.scatter_probs[gi];
    for (int g = 0; g < c.w->G; g++) {
      a += probs[g];
      if (a > t) {
        p.E = double(g + 1);
        p.dir = IsotropicDirection(p.rng);
        return;
      }
    }
    return;  // unreachable for normalised columns (Multigroup.cpp:263-264)
  }
  acc += n.fission_xs[gi];
  if (acc > threshold) {
    p.event = MMC_EV_FISSION;
    const size_t yield = size_t(n.nubar[gi] + Canonical(p.rng));
    const auto& probs = n.chi_probs[gi];
    for (size_t i = 0; i < yield; i++) {
      const double t = Canonical(p.rng);
      double a = 0;
      for (int g = 0; g < c.w->G; g++) {
        a += probs[g];
        if (a > t) {
          const Pt d = IsotropicDirection(p.rng);
          BankSecondary(c, p, d, double(g + 1));
          break;
        }
      }
    }
    return;
  }
  // Multigroup.cpp:70-72: unreachable when total = sum of reactions.
}

// ThermalScattering::SampleBeta (ThermalScattering.cpp:271-338)
bool SampleBeta(const Tsl& t, Lcg& g, double E, double T, double& beta) {
  const size_t e_hi = UpperBound(t.Es, E);
  if (e_hi >= t.Es.size()) return false;  // E above the beta grid (:278 asserts)
  const double r =
      e_hi != 0 ? (E - t.Es[e_hi - 1]) / (t.Es[e_hi] - t.Es[e_hi - 1]) : 1;
  const size_t e_s = r <= Canonical(g) ? e_hi - 1 : e_hi;
  const double E_s = t.Es[e_s];
  const size_t p_i =
      std::upper_bound(t.E_ends.begin(), t.E_ends.end(), e_s) - t.E_ends.begin();
  const TslPartition& P = t.beta_parts[p_i];
  const size_t local = p_i == 0 ? e_s : e_s - t.E_ends[p_i - 1];
  for (int k = 0; k < kBetaResampleLimit; k++) {
    const double F = Canonical(g);
    const size_t f_hi = UpperBound(P.F, F);
    const double F_lo = f_hi != 0 ? P.F[f_hi - 1] : 0;
    const double F_hi = f_hi != P.F.size() ? P.F[f_hi] : 1;
    const double b_lo =
        f_hi != 0 ? P.Evaluate(f_hi - 1, local, T) : -E_s / (kBoltzmann * T);
    const double b_hi =
        f_hi != P.F.size() ? P.Evaluate(f_hi, local, T) : t.beta_cutoff;
    const double b = b_lo + (F - F_lo) / (F_hi - F_lo) * (b_hi - b_lo);
    const double b_min = -E / (kBoltzmann * T);
    if (b_min <= b) {
      beta = b;
      return true;
    }
  }
  return false;
}

// ThermalScattering::SampleAlpha (ThermalScattering.cpp:340-463), with the
// operator-precedence quirk Q2 of :361-365 reproduced.
int SampleAlpha(const Tsl& t, Lcg& g, double b, double E, double T,
                double& alpha) {
  const double abs_b = std::abs(b);
  const int sgn = (0 < b) - (b < 0);
  const size_t b_hi = UpperBound(t.betas, abs_b);
  if (sgn != 0 && b_hi >= t.betas.size()) return kErrAlphaRange;
  const bool snap_lower =
      (sgn == 1 && t.beta_cutoff <= t.betas[b_hi]) ||
      (sgn == -1 && -t.betas[b_hi] < -E / (kBoltzmann * T));
  const bool snap_min = b_hi == 0;
  const double r =
      snap_lower ? 0
      : snap_min ? 1
                 : (abs_b - t.betas[b_hi - 1] / (t.betas[b_hi] - t.betas[b_hi - 1]));
  // With |b| below the first grid beta and the lower-snap rule firing, the
  // reference reads betas.at(-1) and throws (ThermalScattering.cpp:366-368);
  // the first grid beta is used instead.
  size_t b_s_i = r <= Canonical(g) ? b_hi - 1 : b_hi;
  if (b_hi == 0) b_s_i = 0;
  double b_s = sgn * t.betas[b_s_i];
  // ... and if that grid beta would make E' negative, the kinematic bounds
  // are those of the sampled beta itself.
  if (E + b_s * kBoltzmann * T < 0) b_s = b;
  const double sqrt_E = std::sqrt(E);
  const double bs_root = std::sqrt(E + b_s * kBoltzmann * T);
  const double akT = t.awr * kBoltzmann * T;
  const double a_min = std::pow(sqrt_E - bs_root, 2) / akT;
  const double a_max = std::pow(sqrt_E + bs_root, 2) / akT;
  const size_t p_i =
      std::upper_bound(t.beta_ends.begin(), t.beta_ends.end(), b_s_i) -
      t.beta_ends.begin();
  const TslPartition& P = t.alpha_parts[p_i];
  const size_t local = p_i == 0 ? b_s_i : b_s_i - t.beta_ends[p_i - 1];
  const size_t nF = P.F.size();
  auto find_cdf = [&](double a) {
    const size_t hi =
        UpperBoundIdx(nF, [&](size_t i) { return a < P.Evaluate(i, local, T); });
    const double F_lo = hi != 0 ? P.F[hi - 1] : 0;
    const double F_hi = hi != nF ? P.F[hi] : 1;
    const double a_lo = hi != 0 ? P.Evaluate(hi - 1, local, T) : 0;
    const double a_hi = hi != nF ? P.Evaluate(hi, local, T) : t.alpha_cutoff;
    return F_lo + (a - a_lo) * (F_hi - F_lo) / (a_hi - a_lo);
  };
  const double F_min = find_cdf(a_min);
  const double F_max = find_cdf(a_max);
  for (int k = 0; k < kAlphaResampleLimit; k++) {
    const double F = F_min + Canonical(g) * (F_max - F_min);
    const size_t hi = UpperBound(P.F, F);
    const double F_lo = hi != 0 ? P.F[hi - 1] : 0;
    const double F_hi = hi != nF ? P.F[hi] : 1;
    const double a_lo = hi != 0 ? P.Evaluate(hi - 1, local, T) : 0;
    const double a_hi = hi != nF ? P.Evaluate(hi, local, T) : t.alpha_cutoff;
    const double a = a_lo + (F - F_lo) / (F_hi - F_lo) * (a_hi - a_lo);
    if (a_min < a && a < a_max) {
      const double b_root = std::sqrt(E + b * kBoltzmann * T);
      const double ba_min = std::pow(sqrt_E - b_root, 2) / akT;
      const double ba_max = std::pow(sqrt_E + b_root, 2) / akT;
      alpha = ba_min + (a - a_min) * (ba_max - ba_min) / (a_max - a_min);
      return 0;
    }
  }
  return kErrResample;
}

int TslSample(const Tsl& t, Lcg& g, double E, double T, double& alpha,
              double& beta) {
  if (!SampleBeta(t, g, E, T, beta)) return kErrResample;
  return SampleAlpha(t, g, beta, E, T, alpha);
}

void ScoreTslHist(Ctx& c, double alpha, double beta) {
  for (size_t e = 0; e < c.w->ests.size(); e++) {
    const Estimator& est = c.w->ests[e];
    if (est.kind == MMC_EST_TSL_ALPHA)
      c.pending[e][est.energy.index(alpha)] += 1.0;
    else if (est.kind == MMC_EST_TSL_BETA)
      c.pending[e][est.energy.index(beta)] += 1.0;
  }
}

// ContinuousScatter::Interact (ContinuousReaction.cpp:119-189)
int CeScatter(Ctx& c, const CeNuclide& n, const CeReaction& r, Particle& p) {
  World& w = *c.w;
  p.event = MMC_EV_SCATTER;
  const double T = CellT(w, p);
  if (r.tsl >= 0 && TslValid(w.tsl[r.tsl], p.E)) {
    // ThermalScattering::Scatter (ThermalScattering.cpp:159-171)
    const Tsl& t = w.tsl[r.tsl];
    const double E = p.E;
    double alpha = 0, beta = 0;
    const int err = TslSample(t, p.rng, E, T, alpha, beta);
    c.acc->cnt.tsl_collisions++;
    if (err) return err;
    ScoreTslHist(c, alpha, beta);
    const double E_p = E + beta * kBoltzmann * T;
    const double mu =
        (E + E_p - alpha * t.awr * kBoltzmann * T) / (2 * std::sqrt(E * E_p));
    ScatterTo(p, mu, E_p);
    return 0;
  }
  const double m_n = kNeutronMass;
  const double awr = n.awr;
  const double E = p.E;
  const double s_n = std::sqrt(2. * E / m_n);
  const Pt v_n = s_n * p.dir;
  const double beta = std::sqrt((awr * m_n) / (2. * kBoltzmann * T));
  const double y = beta * s_n;
  double x, mu;
  do {
    const double xi_1 = Canonical(p.rng), xi_2 = Canonical(p.rng);
    if (Canonical(p.rng) < 2 / (std::sqrt(kPi) * y + 2)) {
      x = std::sqrt(-std::log(xi_1 * xi_2));
    } else {
      const double xi_3 = Canonical(p.rng);
      const double z = std::cos(kPi * xi_3 / 2);
      x = std::sqrt(-std::log(xi_1) - std::log(xi_2) * z * z);
    }
    mu = 2 * Canonical(p.rng) - 1;
  } while (Canonical(p.rng) >=
           std::sqrt(x * x + y * y - 2 * x * y * mu) / (x + y));
  const double s_T = x / beta;
  const double phi = 2 * kPi * Canonical(p.rng);
  const Pt v_T = s_T * Rotate(p.dir, mu, phi);
  const Pt v_cm = (v_n + awr * v_T) / (1 + awr);
  const Pt V_n = v_n - v_cm;
  const double mu_cm = 2 * Canonical(p.rng) - 1;
  const double phi_cm = 2 * kPi * Canonical(p.rng);
  const Pt Vp_dir = Rotate(Normalized(V_n), mu_cm, phi_cm);
  const Pt V_n_prime = std::sqrt(Dot(V_n, V_n)) * Vp_dir;
  const Pt v_n_prime = V_n_prime + v_cm;
  const double E_prime = 0.5 * m_n * Dot(v_n_prime, v_n_prime);
  ScatterTo(p, Dot(p.dir, Normalized(v_n_prime)), E_prime);
  return 0;
}

// Continuous::Interact (Continuous.cpp:57-70) with the reactions of
// ContinuousReaction.cpp; the tail-recursive resample (quirk Q4) is a loop.
int CeInteract(Ctx& c, const CeNuclide& n, Particle& p) {
  World& w = *c.w;
  const double T = CellT(w, p);
  for (int attempt = 0;; attempt++) {
    if (attempt >= kMaxReactionResamples) return kErrReaction;
    const double threshold = Canonical(p.rng) * CeTotal(w, n, p.E, T, false);
    double acc = 0;
    for (const auto& r : n.rx) {
      acc += RxXs(w, n, r, p.E, T, false);
      if (acc > threshold) {
        if (r.kind == MMC_RX_CAPTURE) {
          p.event = MMC_EV_CAPTURE;  // ContinuousCapture::Interact (:60-62)
          return 0;
        }
        if (r.kind == MMC_RX_SCATTER) return CeScatter(c, n, r, p);
        // ContinuousFission::Interact (:252-265)
        p.event = MMC_EV_FISSION;
        const size_t yield = size_t(r.nubar.at(p.E) + Canonical(p.rng));
        for (size_t i = 0; i < yield; i++) {
          const Pt d = IsotropicDirection(p.rng);
          const double E = r.chi_kind == MMC_CHI_WATT
                               ? SampleWatt(p.rng, r.watt_a, r.watt_b)
                               : p.E;
          BankSecondary(c, p, d, E);
        }
        return 0;
      }
    }
  }
}

// Particle::SampleNuclide (Particle.cpp:110-124)
int SampleNuclide(const World& w, Particle& p) {
  const Material& m = w.materials[w.cells[p.cell].material];
  const double threshold = MaterialTotal(w, p, false) * Canonical(p.rng);
  const double T = CellT(w, p);
  double acc = 0;
  for (const auto& c : m.comps) {
    acc += c.second * NuclideTotal(w, c.first, p, T, false);
    if (acc > threshold) return c.first;
  }
  return m.comps.back().first;  // Particle.cpp:121-123: unreachable
}

// Nuclide::Interact (Nuclide.cpp:24-26) after SampleNuclide.
int Collide(Ctx& c, Particle& p) {
  World& w = *c.w;
  const int nuc = SampleNuclide(w, p);
  // absorption k estimator (new): nu sigma_f / sigma_a of the sampled
  // nuclide at an absorption (capture or fission).
  double nusf = 0, sa = 0;
  if (c.keig) {
    if (w.energy_mode == MMC_MULTIGROUP) {
      const auto& n = w.mg[nuc];
      const size_t g = size_t(p.E) - 1;
      sa = n.capture[g] + n.fission_xs[g];
      nusf = n.has_fission ? n.nubar[g] * n.fission_xs[g] : 0;
    } else {
      for (const auto& r : w.ce[nuc].rx) {
        if (r.kind == MMC_RX_CAPTURE) sa += r.xs.at(p.E);
        if (r.kind == MMC_RX_FISSION) {
          sa += r.xs.at(p.E);
          nusf += r.nubar.at(p.E) * r.xs.at(p.E);
        }
      }
    }
  }
  int err = 0;
  if (w.energy_mode == MMC_MULTIGROUP)
    MgInteract(c, w.mg[nuc], p);
  else
    err = CeInteract(c, w.ce[nuc], p);
  if (c.keig && (p.event == MMC_EV_CAPTURE || p.event == MMC_EV_FISSION) &&
      sa > 0)
    c.acc->k_abs += nusf / sa;
  return err;
}

inline bool IsAlive(const Particle& p) {  // Particle::IsAlive (:126-129)
  return p.event != MMC_EV_CAPTURE && p.event != MMC_EV_LEAK &&
         p.event != MMC_EV_FISSION;
}

// EstimatorSetProxy::Score -> CurrentEstimator::GetScore -> ParticleBins
// (Estimator.cpp:142-151, :247-251; Scorable.cpp:81-99; Bins.cpp:196-204)
void ScoreCurrent(Ctx& c, const Particle& p) {
  if (!(p.event == MMC_EV_SURFACE_CROSS || p.event == MMC_EV_LEAK)) return;
  for (size_t e = 0; e < c.w->ests.size(); e++) {
    const Estimator& est = c.w->ests[e];
    if (est.kind != MMC_EST_CURRENT || est.surface != p.surface) continue;
    const size_t ci = est.has_dir ? est.cosine.index(Dot(est.dir, p.dir)) : 0;
    const size_t ei = est.energy.index(p.E);
    // CurrentTotalCrossSectionSensitivity::GetScore (Sensitivity.cpp:55-58)
    const double score = est.sens ? p.ie[est.sens - 1] * 1.0 : 1.0;
    c.pending[e][est.energy.size() * ci + ei] += score;
  }
}
void ScoreFluxMesh(Ctx& c, const Particle& p, double sigma_t_macro) {
  for (size_t e = 0; e < c.w->ests.size(); e++) {
    const Estimator& est = c.w->ests[e];
    if (est.kind != MMC_EST_FLUX_MESH) continue;
    const double q[3] = {p.pos.x, p.pos.y, p.pos.z};
    size_t idx = 0;
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
    if (in) c.pending[e][idx] += 1.0 / sigma_t_macro;
  }
}

// TotalCrossSectionPerturbation::Stream (Perturbation.cpp:68-78), called by
// Particle::Stream (Particle.cpp:46-53) before the move: in a material that
// holds the perturbed nuclide, 1 / (collision probability density) -
// distance.  The density is the microscopic total, or the majorant with
// delta tracking (TransportMethod.cpp:79-82, :124-127).
void StreamIndirect(const World& w, const Material& m, double sigma, double d, Particle& p) {
  for (size_t j = 0; j < w.pert_nuclide.size(); j++) {
    bool holds = false;
    for (const auto& comp : m.comps) holds = holds || comp.first == w.pert_nuclide[j];
    if (holds) p.ie[j] += 1 / sigma - d;
  }
}

// SurfaceTracking::Transport / CellDeltaTracking::Transport
// (TransportMethod.cpp:52-77, :88-122)
int Transport(Ctx& c, Particle& p) {
  World& w = *c.w;
  p.cell = FindCell(w, p.pos);
  if (p.cell < 0) return kErrLost;
  if (w.cells[p.cell].material < 0) {  // born in a void cell: leaks at once
    p.event = MMC_EV_LEAK;
    return 0;
  }
  uint64_t steps = 0;
  while (IsAlive(p)) {
    if (++steps > w.max_steps) return kErrSteps;
    const Material& m = w.materials[w.cells[p.cell].material];
    const bool delta = w.tracking == MMC_CELL_DELTA_TRACKING;
    const double sigma = MaterialTotal(w, p, delta);
    const double d_coll = Exponential(p.rng, m.aden * sigma);
    c.acc->cnt.xs_lookups++;
    c.acc->cnt.advances++;
    const auto ns = NearestSurface(w, w.cells[p.cell], p.pos, p.dir);
    bool cross = delta ? (ns.second < d_coll) : !(d_coll < ns.second);
    if (cross) {
      const double d = ns.second + kNudge;
      if (c.keig) c.acc->k_trk += d * m.aden * MaterialNuFission(w, p);
      StreamIndirect(w, m, sigma, d, p);
      p.pos = p.pos + p.dir * d;  // Particle::Stream (Particle.cpp:46-53)
      p.cell = (w.quirks & MMC_QUIRK_NUDGE_CROSSING)
                   ? FindCell(w, p.pos)
                   : FindCellAcross(w, p.pos, ns.first, !FaceInside(w, p.cell, ns.first));
      if (p.cell < 0) return kErrLost;
      p.surface = ns.first;
      p.event = w.cells[p.cell].material >= 0 ? MMC_EV_SURFACE_CROSS
                                              : MMC_EV_LEAK;
      c.acc->cnt.crossings++;
      ScoreCurrent(c, p);
      continue;
    }
    bool real = true;
    if (delta) {
      const double tot = MaterialTotal(w, p, false);
      real = Bernoulli(p.rng, tot / sigma);  // pre-flight position (Q6)
    }
    if (c.keig) c.acc->k_trk += d_coll * m.aden * MaterialNuFission(w, p);
    StreamIndirect(w, m, sigma, d_coll, p);
    p.pos = p.pos + p.dir * d_coll;
    if (!real) {
      p.event = MMC_EV_VIRTUAL;
      c.acc->cnt.virtual_collisions++;
      continue;
    }
    c.acc->cnt.collisions++;
    const double tot = MaterialTotal(w, p, false);
    ScoreFluxMesh(c, p, m.aden * tot);
    if (c.keig) c.acc->k_col += MaterialNuFission(w, p) / tot;
    const int err = Collide(c, p);
    if (err) return err;
  }
  return 0;
}

// Source::Sample (Source.cpp:143-154), IsotropicFlux::Sample (:124-129)
Particle SampleSource(const mmc_source_desc& s, uint64_t seed) {
  Lcg g(seed);
  Particle p;
  const double lo[3] = {s.box_lo[0], s.box_lo[1], s.box_lo[2]};
  const double hi[3] = {s.box_hi[0], s.box_hi[1], s.box_hi[2]};
  double q[3] = {s.x, s.y, s.z};
  for (int a = 0; a < 3; a++)
    if (hi[a] > lo[a]) q[a] = Uniform(g, lo[a], hi[a]);
  p.pos = {q[0], q[1], q[2]};
  if (s.dir_kind == MMC_DIR_ISOTROPIC) {
    p.dir = IsotropicDirection(g);
  } else if (s.dir_kind == MMC_DIR_ISOTROPIC_FLUX) {
    const double mu = std::sqrt(Canonical(g));
    const double phi = Uniform(g, 0., 2 * kPi);
    p.dir = Rotate(Normalized(Pt{s.u, s.v, s.w}), mu, phi);
  } else {
    p.dir = Normalized(Pt{s.u, s.v, s.w});
  }
  p.E = s.energy_kind == MMC_E_UNIFORM ? Uniform(g, s.E, s.E_hi) : s.E;
  p.rng.Seed(g());
  return p;
}

// splitmix64 finaliser for extra copies of a fission site (new behaviour).
uint64_t Remix(uint64_t s, uint64_t c) {
  uint64_t z = s + c * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return z % kLcgM;
}

void Commit(Ctx& c) {  // ScorableProxy::CommitHistory (Scorable.cpp:101-106)
  for (size_t e = 0; e < c.pending.size(); e++) {
    for (const auto& kv : c.pending[e]) {
      c.acc->S[e][kv.first] += kv.second;
      c.acc->S2[e][kv.first] += kv.second * kv.second;
    }
    c.pending[e].clear();
  }
}

void Merge(Acc& into, const Acc& from) {  // EstimatorSet::operator+=
  for (size_t e = 0; e < into.S.size(); e++)
    for (size_t b = 0; b < into.S[e].size(); b++) {
      into.S[e][b] += from.S[e][b];
      into.S2[e][b] += from.S2[e][b];
    }
  uint64_t* a = &into.cnt.histories;
  const uint64_t* f = &from.cnt.histories;
  for (int i = 0; i < 11; i++) a[i] += f[i];
  into.err |= from.err;
}

}  // namespace orc

using namespace orc;

struct orc_world {
  World w;
};

static int Fail(orc_world* w, int code, const std::string& msg) {
  if (w) w->w.error = msg;
  return code;
}

extern "C" {

int orc_rng_draws(uint64_t seed, uint64_t n, uint64_t* out) {
  Lcg g(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = g();
  return 0;
}
int orc_rng_canonicals(uint64_t seed, uint64_t n, double* out) {
  Lcg g(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = Canonical(g);
  return 0;
}
int orc_rng_exponentials(uint64_t seed, double lambda, uint64_t n,
                         double* out) {
  Lcg g(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = Exponential(g, lambda);
  return 0;
}
int orc_rng_directions(uint64_t seed, uint64_t n, double* out3) {
  Lcg g(seed);
  for (uint64_t i = 0; i < n; i++) {
    const Pt d = IsotropicDirection(g);
    out3[3 * i] = d.x;
    out3[3 * i + 1] = d.y;
    out3[3 * i + 2] = d.z;
  }
  return 0;
}

int orc_create(const mmc_world_desc* wd, const mmc_tally_desc* td,
               const mmc_options* opt, orc_world** out) {
  if (!wd || !out) return MMC_ERR_ARG;
  auto* ow = new orc_world();
  World& w = ow->w;
  w.energy_mode = wd->energy_mode;
  w.G = wd->groups;
  if (opt) {
    w.tracking = opt->tracking;
    w.quirks = opt->quirks;
    if (opt->max_steps) w.max_steps = opt->max_steps;
  }
  for (int i = 0; i < wd->n_surfaces; i++) {
    const auto& s = wd->surfaces[i];
    w.surfaces.push_back({s.kind, s.a, s.b, s.c, s.r});
  }
  for (int i = 0; i < wd->n_fields; i++) {
    const auto& f = wd->fields[i];
    w.fields.push_back({f.kind, f.c, f.gx, f.gy, f.gz, f.b, f.lower, f.upper});
  }
  for (int i = 0; i < wd->n_cells; i++) {
    const auto& c = wd->cells[i];
    Cell cell;
    cell.material = c.material;
    cell.field = c.field;
    for (int k = 0; k < c.face_count; k++) {
      const auto& f = wd->faces[c.face_begin + k];
      cell.faces.push_back({f.surface, f.inside != 0});
    }
    w.cells.push_back(cell);
  }
  // Material::AssignNuclides normalises atom fractions (Material.cpp:82-92).
  for (int i = 0; i < wd->n_materials; i++) {
    const auto& m = wd->materials[i];
    Material mat;
    mat.aden = m.aden;
    double sum = 0;
    for (int k = 0; k < m.comp_count; k++) sum += wd->components[m.comp_begin + k].afrac;
    for (int k = 0; k < m.comp_count; k++) {
      const auto& c = wd->components[m.comp_begin + k];
      mat.comps.push_back({c.nuclide, c.afrac / sum});
    }
    w.materials.push_back(mat);
  }
  if (wd->n_perturbations < 0 || wd->n_perturbations > MMC_MAX_PERTURBATIONS) {
    delete ow;
    return MMC_ERR_ARG;  // bad perturbations
  }
  for (int j = 0; j < wd->n_perturbations; j++)
    w.pert_nuclide.push_back(wd->perturbation_nuclide[j]);
  const int G = w.G;
  if (w.energy_mode == MMC_MULTIGROUP) {
    for (int i = 0; i < wd->n_nuclides; i++) {
      const auto& d = wd->mg[i];
      MgNuclide n;
      auto vec = [&](const double* p) {
        return p ? std::vector<double>(p, p + G) : std::vector<double>(G, 0.0);
      };
      n.capture = vec(d.capture);
      n.fission_xs = vec(d.fission_xs);
      n.nubar = vec(d.nubar);
      n.has_fission = d.fission_xs != nullptr;
      // site keys hold 64 children per history: nu-bar below 63
      if (n.has_fission)
        for (double v : n.nubar)
          if (!(v >= 0 && v < 63.0)) { *out = ow; return Fail(ow, MMC_ERR_ARG, "nubar must lie in [0, 63)"); }
      n.has_scatter = d.scatter != nullptr;
      // TwoDimensional column construction and NormalizedTwoDimensional
      // (Multigroup.cpp:124-186); CreateScatterXS column sums (:201-210).
      auto columns = [&](const double* m) {
        std::vector<std::vector<double>> cols(G, std::vector<double>(G, 0.0));
        if (!m) return cols;
        for (int gp = 0; gp < G; gp++)
          for (int g = 0; g < G; g++) cols[gp][g] = m[G * g + gp];
        return cols;
      };
      auto normalise = [&](std::vector<std::vector<double>> cols) {
        for (auto& col : cols) {
          double s = 0;
          for (double v : col) s += v;
          if (s == 0.) continue;
          for (double& v : col) v /= s;
        }
        return cols;
      };
      const auto sc = columns(d.scatter);
      n.scatter_xs.assign(G, 0.0);
      for (int gp = 0; gp < G; gp++) {
        double s = 0;
        for (int g = 0; g < G; g++) s += sc[gp][g];
        n.scatter_xs[gp] = s;
      }
      n.scatter_probs = normalise(sc);
      n.chi_probs = normalise(columns(d.chi));
      // CreateTotalXS (Multigroup.cpp:231-243): reactions in enum order.
      n.total.assign(G, 0.0);
      for (int g = 0; g < G; g++) {
        double t = 0;
        if (d.capture) t = t + n.capture[g];
        if (d.scatter) t = t + n.scatter_xs[g];
        if (d.fission_xs) t = t + n.fission_xs[g];
        n.total[g] = t;
      }
      w.mg.push_back(n);
    }
  } else {
    for (int i = 0; i < wd->n_nuclides; i++) {
      const auto& d = wd->ce[i];
      CeNuclide n;
      n.awr = d.awr;
      n.has_total = d.has_total != 0;
      n.total = MakeTable(d.total);
      n.total_temperature = d.total_temperature;
      for (int k = 0; k < d.n_reactions; k++) {
        const auto& r = d.reactions[k];
        CeReaction rx;
        rx.kind = r.kind;
        rx.tsl = r.kind == MMC_RX_SCATTER ? r.tsl : -1;
        rx.xs = MakeTable(r.xs);
        rx.temperature = r.temperature;
        rx.nubar = MakeTable(r.nubar);
        if (r.kind == MMC_RX_FISSION)
          for (int i = 0; i < r.nubar.n; i++)
            if (!(r.nubar.y[i] >= 0 && r.nubar.y[i] < 63.0))
              { *out = ow; return Fail(ow, MMC_ERR_ARG, "nubar must lie in [0, 63)"); }
        rx.chi_kind = r.chi_kind;
        rx.watt_a = r.watt_a;
        rx.watt_b = r.watt_b;
        n.rx.push_back(rx);
      }
      w.ce.push_back(n);
    }
    // ThermalScattering constructor (ThermalScattering.cpp:25-99)
    for (int i = 0; i < wd->n_tsl; i++) {
      const auto& d = wd->tsl[i];
      Tsl t;
      t.majorant = MakeTable(d.majorant);
      t.nE = d.n_E;
      t.nT = d.n_T;
      t.rank = d.rank;
      t.E_axis.assign(d.total_E_axis, d.total_E_axis + d.n_E);
      t.E_modes.assign(d.total_E, d.total_E + size_t(d.n_E) * d.rank);
      t.T_axis.assign(d.total_T_axis, d.total_T_axis + d.n_T);
      t.T_modes.assign(d.total_T, d.total_T + size_t(d.n_T) * d.rank);
      t.S.assign(d.total_S, d.total_S + d.rank);
      auto part = [](const mmc_tsl_partition& p) {
        TslPartition q;
        q.nF = p.n_F;
        q.nx = p.n_x;
        q.nT = p.n_T;
        q.rank = p.rank;
        q.F.assign(p.F, p.F + p.n_F);
        q.cdf.assign(p.cdf_modes, p.cdf_modes + size_t(p.n_F) * p.rank);
        q.S.assign(p.S, p.S + p.rank);
        q.x.assign(p.x, p.x + p.n_x);
        q.T.assign(p.T, p.T + p.n_T);
        q.xt.assign(p.x_T_modes, p.x_T_modes + size_t(p.n_x) * p.n_T * p.rank);
        return q;
      };
      size_t end = 0;
      for (int k = 0; k < d.n_beta_partitions; k++) {
        t.beta_parts.push_back(part(d.beta_partitions[k]));
        const auto& q = t.beta_parts.back();
        t.Es.insert(t.Es.end(), q.x.begin(), q.x.end());
        end += q.x.size();
        t.E_ends.push_back(end);
      }
      end = 0;
      for (int k = 0; k < d.n_alpha_partitions; k++) {
        t.alpha_parts.push_back(part(d.alpha_partitions[k]));
        const auto& q = t.alpha_parts.back();
        t.betas.insert(t.betas.end(), q.x.begin(), q.x.end());
        end += q.x.size();
        t.beta_ends.push_back(end);
      }
      t.beta_cutoff = d.beta_cutoff;
      t.alpha_cutoff = d.alpha_cutoff;
      t.awr = 0;
      t.cutoff_energy = t.E_axis.back();  // ThermalScattering.hpp:125-126
      w.tsl.push_back(t);
    }
    // awr comes from the owning nuclide (ThermalScattering.cpp:99)
    for (const auto& n : w.ce)
      for (const auto& r : n.rx)
        if (r.tsl >= 0) w.tsl[r.tsl].awr = n.awr;
  }
  if (td) {
    for (int i = 0; i < td->n_estimators; i++) {
      const auto& e = td->estimators[i];
      Estimator est;
      est.kind = e.kind;
      est.surface = e.surface;
      est.sens = e.sensitivity;
      est.has_dir = e.has_direction != 0;
      est.dir = Normalized(Pt{e.u, e.v, e.w});
      est.cosine = MakeBins(est.has_dir ? e.cosine : mmc_bins{});
      est.energy = MakeBins(e.energy);
      for (int a = 0; a < 3; a++) {
        est.mesh_n[a] = e.mesh_n[a] > 0 ? e.mesh_n[a] : 1;
        est.mesh_lo[a] = e.mesh_lo[a];
        est.mesh_hi[a] = e.mesh_hi[a];
      }
      w.ests.push_back(est);
    }
  }
  w.total.init(w.ests);
  w.cycle_S.resize(w.ests.size());
  for (size_t e = 0; e < w.ests.size(); e++) w.cycle_S[e].assign(w.ests[e].size(), 0.0);
  *out = ow;
  return 0;
}

void orc_destroy(orc_world* w) { delete w; }
const char* orc_last_error(const orc_world* w) {
  return w ? w->w.error.c_str() : "null world";
}

static void ToSoA(const Particle& p, mmc_particle_soa* s, uint64_t i) {
  s->x[i] = p.pos.x;
  s->y[i] = p.pos.y;
  s->z[i] = p.pos.z;
  s->u[i] = p.dir.x;
  s->v[i] = p.dir.y;
  s->w[i] = p.dir.z;
  s->E[i] = p.E;
  s->rng[i] = uint32_t(p.rng.x);
  if (s->cell) s->cell[i] = p.cell;
  if (s->surface) s->surface[i] = p.surface;
  if (s->event) s->event[i] = p.event;
}
static Particle FromSoA(const mmc_particle_soa* s, uint64_t i) {
  Particle p;
  p.pos = {s->x[i], s->y[i], s->z[i]};
  p.dir = {s->u[i], s->v[i], s->w[i]};
  p.E = s->E[i];
  p.rng.x = s->rng[i];
  p.cell = s->cell ? s->cell[i] : -1;
  p.surface = s->surface ? s->surface[i] : -1;
  p.event = s->event ? s->event[i] : MMC_EV_BIRTH;
  return p;
}

int orc_sample_source(orc_world* w, const mmc_source_desc* src,
                      uint64_t seed0, uint64_t n, mmc_particle_soa* out) {
  if (!w || !src || !out) return MMC_ERR_ARG;
  for (uint64_t i = 0; i < n; i++) {
    Particle p = SampleSource(*src, seed0 + i);
    p.cell = FindCell(w->w, p.pos);
    ToSoA(p, out, i);
  }
  return 0;
}

int orc_eval_event(orc_world* ow, int kind, mmc_particle_soa* s, uint64_t n) {
  if (!ow || !s) return MMC_ERR_ARG;
  World& w = ow->w;
  Acc acc;
  acc.init(w.ests);
  std::vector<mmc_site> bank;
  Ctx c{&w, &acc, false, &bank, 0, {}};
  c.pending.resize(w.ests.size());
  int err = 0;
  for (uint64_t i = 0; i < n; i++) {
    Particle p = FromSoA(s, i);
    switch (kind) {
      case MMC_EVAL_FIND_CELL:
        p.cell = FindCell(w, p.pos);
        break;
      case MMC_EVAL_NEAREST_SURFACE: {
        const auto ns = NearestSurface(w, w.cells[p.cell], p.pos, p.dir);
        s->iout[i] = ns.first;
        s->out0[i] = ns.second;
        break;
      }
      case MMC_EVAL_XS_TOTAL:
        s->out0[i] = MaterialTotal(w, p, false);
        s->out1[i] = MaterialTotal(w, p, true);
        break;
      case MMC_EVAL_COLLISION_DISTANCE: {
        const Material& m = w.materials[w.cells[p.cell].material];
        const double sig =
            MaterialTotal(w, p, w.tracking == MMC_CELL_DELTA_TRACKING);
        s->out0[i] = Exponential(p.rng, m.aden * sig);
        break;
      }
      case MMC_EVAL_COLLIDE: {
        p.secondaries.clear();
        const int e = Collide(c, p);
        if (e) err = e;
        s->iout[i] = int32_t(p.secondaries.size());
        break;
      }
      case MMC_EVAL_SCATTER_ROTATE:
        ScatterTo(p, s->out0[i], s->out1[i]);
        break;
      case MMC_EVAL_TSL_SAMPLE: {
        double a = 0, b = 0;
        const int e = TslSample(w.tsl[0], p.rng, p.E, s->out0[i], a, b);
        if (e) {
          err = e;
          a = b = std::nan("");
        }
        s->out0[i] = a;
        s->out1[i] = b;
        break;
      }
      default:
        return MMC_ERR_ARG;
    }
    ToSoA(p, s, i);
  }
  if (err & kErrResample) return Fail(ow, MMC_ERR_RESAMPLE, "TSL resample limit exceeded");
  if (err) return Fail(ow, MMC_ERR_ARG, "event error");
  return 0;
}

static int ErrToStatus(orc_world* ow, int err) {
  if (!err) return 0;
  if (err & kErrLost)
    return Fail(ow, MMC_ERR_LOST,
                "Point does not belong to any Cell. Please checkthat all space "
                "is either assigned a material or void.");
  if (err & kErrResample)
    return Fail(ow, MMC_ERR_RESAMPLE,
                "Number of thermal scattering resamples exceeded limit: 10");
  if (err & kErrSteps) return Fail(ow, MMC_ERR_STEPS, "history step limit exceeded");
  if (err & kErrReaction) return Fail(ow, MMC_ERR_ARG, "no reaction selected after 1000 resamples");
  if (err & kErrOverflow) return Fail(ow, MMC_ERR_OVERFLOW, "bank, queue or record overflow");
  return Fail(ow, MMC_ERR_ARG, "alpha sampling outside the beta grid");
}

// FixedSource::Solve / StartWorker (FixedSource.cpp:22-77): workers pull
// history indices from an atomic counter, transport the source particle and
// its secondaries (back of the bank first, secondaries spliced to the
// front), commit per-history scores; per-thread tallies merge serially.
int orc_run_fixed_source(orc_world* ow, const mmc_source_desc* src,
                         uint64_t seed0, uint64_t first, uint64_t n,
                         int threads) {
  if (!ow || !src) return MMC_ERR_ARG;
  World& w = ow->w;
  if (threads < 1) threads = 1;
  std::atomic<uint64_t> next{0};
  std::vector<Acc> accs(threads);
  for (auto& a : accs) a.init(w.ests);
  auto worker = [&](int t) {
    Acc& acc = accs[t];
    std::vector<mmc_site> bank_unused;
    Ctx c{&w, &acc, false, &bank_unused, 0, {}};
    c.pending.resize(w.ests.size());
    for (;;) {
      const uint64_t k = next++;
      if (k >= n) break;
      const uint64_t hist = first + k;
      std::vector<Particle> bank;
      bank.push_back(SampleSource(*src, seed0 + hist));
      bank.back().hist = hist;
      while (!bank.empty()) {
        Particle p = std::move(bank.back());
        bank.pop_back();
        const int err = Transport(c, p);
        if (err) {
          acc.err |= err;
          break;
        }
        // secondaries go to the front: processed after everything queued
        bank.insert(bank.begin(), p.secondaries.begin(), p.secondaries.end());
      }
      Commit(c);
      acc.cnt.histories++;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(worker, t);
  for (auto& th : pool) th.join();
  for (auto& a : accs) Merge(w.total, a);
  return ErrToStatus(ow, w.total.err);
}

int orc_tsl_evaluate(orc_world* ow, int tsl, int which, int part, uint64_t cdf_i,
                     uint64_t x_i, double T, double* out) {
  if (!ow || !out || tsl < 0 || tsl >= int(ow->w.tsl.size())) return MMC_ERR_ARG;
  const Tsl& t = ow->w.tsl[tsl];
  const auto& parts = which == 0 ? t.beta_parts : t.alpha_parts;
  if (part < 0 || part >= int(parts.size())) return MMC_ERR_ARG;
  const TslPartition& P = parts[part];
  if (cdf_i >= P.F.size() || x_i >= P.x.size()) return MMC_ERR_ARG;
  *out = P.Evaluate(cdf_i, x_i, T);
  return 0;
}
int orc_tsl_total(orc_world* ow, int tsl, double E, double T, double* out) {
  if (!ow || !out || tsl < 0 || tsl >= int(ow->w.tsl.size())) return MMC_ERR_ARG;
  *out = TslTotal(ow->w.tsl[tsl], E, T);
  return 0;
}

int orc_set_entropy_mesh(orc_world* ow, const int32_t n[3], const double lo[3],
                         const double hi[3]) {
  if (!ow) return MMC_ERR_ARG;
  for (int a = 0; a < 3; a++) {
    ow->w.mesh_n[a] = n[a];
    ow->w.mesh_lo[a] = lo[a];
    ow->w.mesh_hi[a] = hi[a];
  }
  return 0;
}

// KEigenvalue constructor: initial sites use seeds 1..N (KEigenvalue.cpp:31-33).
int orc_keig_init(orc_world* ow, const mmc_source_desc* src, uint64_t n) {
  if (!ow || !src || n == 0) return MMC_ERR_ARG;
  World& w = ow->w;
  w.source.clear();
  for (uint64_t s = 1; s <= n; s++) {
    Particle p = SampleSource(*src, s);
    mmc_site site{};
    site.x = p.pos.x;
    site.y = p.pos.y;
    site.z = p.pos.z;
    site.u = p.dir.x;
    site.v = p.dir.y;
    site.w = p.dir.z;
    site.E = p.E;
    site.seed = uint32_t(p.rng.x);
    site.key = s - 1;
    w.source.push_back(site);
  }
  return 0;
}

static double Entropy(const World& w, const std::vector<mmc_site>& bank) {
  const int nx = w.mesh_n[0], ny = w.mesh_n[1], nz = w.mesh_n[2];
  std::vector<uint64_t> cnt(size_t(nx) * ny * nz, 0);
  uint64_t tot = 0;
  for (const auto& s : bank) {
    const double q[3] = {s.x, s.y, s.z};
    size_t idx = 0;
    bool in = true;
    for (int a = 0; a < 3; a++) {
      const double f = (q[a] - w.mesh_lo[a]) / (w.mesh_hi[a] - w.mesh_lo[a]);
      if (!(f >= 0 && f < 1)) {
        in = false;
        break;
      }
      int i = int(f * w.mesh_n[a]);
      if (i >= w.mesh_n[a]) i = w.mesh_n[a] - 1;
      idx = idx * w.mesh_n[a] + i;
    }
    if (in) {
      cnt[idx]++;
      tot++;
    }
  }
  double h = 0;
  for (uint64_t c : cnt)
    if (c) {
      const double p = double(c) / double(tot);
      h -= p * std::log2(p);
    }
  return h;
}

// One k-eigenvalue cycle (new behaviour; the reference's Solve is a stub,
// KEigenvalue.cpp:36-62): transport every source site as history j, bank
// fission children instead of transporting them, sort the bank by
// key = history * 64 + child index, then resample N sites systematically:
// source j <- bank[(j * M) / N]; the first copy keeps the banked seed, later
// copies take Remix(seed, copy).
int orc_run_cycle(orc_world* ow, int active, int threads, mmc_cycle_stats* out) {
  if (!ow) return MMC_ERR_ARG;
  World& w = ow->w;
  const uint64_t N = w.source.size();
  if (N == 0) return Fail(ow, MMC_ERR_ARG, "empty source");
  if (threads < 1) threads = 1;
  const auto t0 = std::chrono::steady_clock::now();
  std::atomic<uint64_t> next{0};
  std::vector<Acc> accs(threads);
  std::vector<std::vector<mmc_site>> banks(threads);
  for (auto& a : accs) a.init(w.ests);
  auto worker = [&](int t) {
    Acc& acc = accs[t];
    Ctx c{&w, &acc, true, &banks[t], 0, {}};
    c.pending.resize(w.ests.size());
    for (;;) {
      const uint64_t j = next++;
      if (j >= N) break;
      const mmc_site& s = w.source[j];
      Particle p;
      p.pos = {s.x, s.y, s.z};
      p.dir = {s.u, s.v, s.w};
      p.E = s.E;
      p.rng.Seed(s.seed);
      p.hist = j;
      c.child_count = 0;
      const int err = Transport(c, p);
      if (err) acc.err |= err;
      // cycle tallies: plain sums, folded into S/S2 per cycle below
      for (size_t e = 0; e < c.pending.size(); e++) {
        for (const auto& kv : c.pending[e]) acc.S[e][kv.first] += kv.second;
        c.pending[e].clear();
      }
      acc.cnt.histories++;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(worker, t);
  for (auto& th : pool) th.join();
  Acc cyc;
  cyc.init(w.ests);
  double kc = 0, kt = 0, ka = 0;
  std::vector<mmc_site> bank;
  for (int t = 0; t < threads; t++) {
    Merge(cyc, accs[t]);
    kc += accs[t].k_col;
    kt += accs[t].k_trk;
    ka += accs[t].k_abs;
    bank.insert(bank.end(), banks[t].begin(), banks[t].end());
  }
  if (cyc.err) return ErrToStatus(ow, cyc.err);
  std::sort(bank.begin(), bank.end(),
            [](const mmc_site& a, const mmc_site& b) { return a.key < b.key; });
  const uint64_t M = bank.size();
  uint64_t* a = &w.total.cnt.histories;
  const uint64_t* f = &cyc.cnt.histories;
  for (int i = 0; i < 11; i++) a[i] += f[i];
  if (active) {
    for (size_t e = 0; e < w.ests.size(); e++)
      for (size_t b = 0; b < w.ests[e].size(); b++) {
        const double v = cyc.S[e][b] / double(N);
        w.total.S[e][b] += v;
        w.total.S2[e][b] += v * v;
      }
  }
  if (out) {
    out->k_analog = double(M) / double(N);
    out->k_collision = kc / double(N);
    out->k_track = kt / double(N);
    out->k_absorption = ka / double(N);
    out->entropy = Entropy(w, bank);
    out->n_sites = M;
  }
  if (M == 0) return Fail(ow, MMC_ERR_ARG, "no fission sites banked");
  std::vector<mmc_site> next_src(N);
  for (uint64_t j = 0; j < N; j++) {
    const uint64_t idx = (j * M) / N;
    const uint64_t j0 = (idx * N + M - 1) / M;
    const uint64_t copy = j - j0;
    mmc_site s = bank[idx];
    if (copy) s.seed = uint32_t(Remix(s.seed, copy));
    s.key = j;
    next_src[j] = s;
  }
  w.source.swap(next_src);
  const auto t1 = std::chrono::steady_clock::now();
  if (out) out->seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

int orc_get_source(orc_world* ow, mmc_site* out, uint64_t cap, uint64_t* n) {
  if (!ow || !n) return MMC_ERR_ARG;
  *n = ow->w.source.size();
  if (out) std::memcpy(out, ow->w.source.data(),
                       sizeof(mmc_site) * std::min<uint64_t>(cap, *n));
  return 0;
}
int orc_set_source(orc_world* ow, const mmc_site* sites, uint64_t n) {
  if (!ow || (!sites && n)) return MMC_ERR_ARG;
  ow->w.source.assign(sites, sites + n);
  return 0;
}

int orc_tally_size(orc_world* ow, int e, uint64_t* n) {
  if (!ow || !n || e < 0 || e >= int(ow->w.ests.size())) return MMC_ERR_ARG;
  *n = ow->w.ests[e].size();
  return 0;
}
int orc_get_tallies(orc_world* ow, int e, double* S, double* S2, uint64_t n) {
  if (!ow || e < 0 || e >= int(ow->w.ests.size())) return MMC_ERR_ARG;
  const auto& s = ow->w.total.S[e];
  const auto& s2 = ow->w.total.S2[e];
  if (n != s.size()) return MMC_ERR_ARG;
  if (S) std::memcpy(S, s.data(), n * sizeof(double));
  if (S2) std::memcpy(S2, s2.data(), n * sizeof(double));
  return 0;
}
int orc_reset_tallies(orc_world* ow) {
  if (!ow) return MMC_ERR_ARG;
  ow->w.total = Acc{};
  ow->w.total.init(ow->w.ests);
  return 0;
}
int orc_get_counters(orc_world* ow, mmc_counters* out) {
  if (!ow || !out) return MMC_ERR_ARG;
  *out = ow->w.total.cnt;
  return 0;
}

int orc_tsl_sample(orc_world* ow, int tsl, uint64_t n, const double* E,
                   const double* T, uint32_t* rng, double* alpha,
                   double* beta, int threads) {
  if (!ow || tsl < 0 || tsl >= int(ow->w.tsl.size())) return MMC_ERR_ARG;
  if (threads < 1) threads = 1;
  const Tsl& t = ow->w.tsl[tsl];
  std::atomic<int> err{0};
  auto worker = [&](int k) {
    const uint64_t lo = n * k / threads, hi = n * (k + 1) / threads;
    for (uint64_t i = lo; i < hi; i++) {
      Lcg g;
      g.x = rng[i];
      double a = 0, b = 0;
      const int e = TslSample(t, g, E[i], T[i], a, b);
      if (e) {
        err |= e;
        a = b = std::nan("");
      }
      alpha[i] = a;
      beta[i] = b;
      rng[i] = uint32_t(g.x);
    }
  };
  std::vector<std::thread> pool;
  for (int k = 0; k < threads; k++) pool.emplace_back(worker, k);
  for (auto& th : pool) th.join();
  return ErrToStatus(ow, err.load());
}

// ---- output text (the reference's .out file) --------------------------------
static uint64_t CopyOut(const std::string& t, char* buf, uint64_t cap) {
  if (buf && cap > 0) {
    const uint64_t k = std::min<uint64_t>(t.size(), cap - 1);
    std::memcpy(buf, t.data(), k);
    buf[k] = 0;
  }
  return t.size();
}
uint64_t orc_format_scientific(double v, char* buf, uint64_t cap) {
  std::ostringstream os;
  os << std::scientific << v;
  return CopyOut(os.str(), buf, cap);
}
// Scorable::GetScoreAsString (Scorable.cpp:51-70)
uint64_t orc_format_scores(const double* S, const double* S2, uint64_t n,
                           double total_weight, char* buf, uint64_t cap) {
  std::ostringstream os;
  os << "mean\n----\n" << std::scientific;
  for (uint64_t i = 0; i < n; i++) os << S[i] / total_weight << ", ";
  os << "\n\n" << "std dev\n-------\n";
  for (uint64_t i = 0; i < n; i++)
    os << std::sqrt(S2[i] - S[i] * S[i] / total_weight) / total_weight << ", ";
  os << "\n\n";
  return CopyOut(os.str(), buf, cap);
}
// NoBins / LinspaceBins / LogspaceBins / BoundaryBins::to_string
// (Bins.cpp:51, :84-90, :125-131, :164-170); n is the user's bin count
// (the reference adds the two overflow bins).
uint64_t orc_format_bins(int kind, int n, double lo, double hi, double base,
                         const double* bounds, char* buf, uint64_t cap) {
  std::ostringstream os;
  if (kind == MMC_BINS_NONE) {
    os << "none";
  } else if (kind == MMC_BINS_BOUNDARIES) {
    for (int i = 0; i < n; i++) os << std::scientific << bounds[i] << ", ";
  } else {
    const size_t n_bins = size_t(n) + 2;
    const double width = (hi - lo) / double(n_bins - 2);
    for (size_t i = 0; i < n_bins - 1; i++) {
      if (kind == MMC_BINS_LINSPACE)
        os << std::scientific << lo + double(i) * width << ", ";
      else
        os << std::scientific << std::pow(base, lo + double(i) * width) << ", ";
    }
  }
  return CopyOut(os.str(), buf, cap);
}

}  // extern "C"
