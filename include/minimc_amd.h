/*
 * minimc_amd.h — C ABI of the MI355X particle-history tracker.
 *
 * The tracker replaces the reference's per-particle history loop:
 *   FixedSource::StartWorker            (src/FixedSource.cpp:40-77)
 *   Particle::Transport -> TransportMethod::Transport
 *                                        (src/Particle.cpp:42-44,
 *                                         src/TransportMethod.cpp:52-122)
 *   the (unimplemented) KEigenvalue cycle body (src/KEigenvalue.cpp:36-84)
 * One virtual call per particle is the wrong granularity for a GPU, so the
 * boundary sits one level up, at the history loop: a host driver hands a
 * flattened World (surfaces, cells, materials, nuclides, tables) to
 * mmc_create once, then asks for whole batches of histories.
 *
 * Only plain C types cross this boundary.  Every descriptor pointer is read
 * during the call that receives it and never retained; output buffers are
 * caller-allocated.  Calls on one context are serialised (not re-entrant).
 * Every entry point returns MMC_OK (0) or an error code; mmc_last_error()
 * gives the message.
 *
 * Units follow the reference (src/BasicTypes.hpp:10-27): MeV, cm, barn, K.
 */
#ifndef MINIMC_AMD_H
#define MINIMC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMC_ABI_VERSION 3

/* ---- status codes ------------------------------------------------------ */
enum mmc_status {
  MMC_OK = 0,
  MMC_ERR_ARG = 1,         /* malformed descriptor or argument            */
  MMC_ERR_HIP = 2,         /* a HIP runtime call failed                    */
  MMC_ERR_LOST = 3,        /* a particle left every cell
                              (World::FindCellContaining, World.cpp:33-36) */
  MMC_ERR_RESAMPLE = 4,    /* TSL resample limit exceeded
                              (ThermalScattering.cpp:335-337,460-462)       */
  MMC_ERR_OVERFLOW = 5,    /* a queue or bank capacity was exceeded        */
  MMC_ERR_NODEV = 6,       /* no usable gfx950 device                      */
  MMC_ERR_UNSUPPORTED = 7, /* feature not available or not configured      */
  MMC_ERR_STEPS = 8        /* a history exceeded the step limit            */
};

/* ---- geometry (src/CSGSurface.hpp:31-118, src/Cell.hpp:30-68) ----------- */
enum mmc_surface_kind {
  MMC_SPHERE = 0,    /* center (a,b,c), radius r                          */
  MMC_PLANEX = 1,    /* x = a                                              */
  MMC_CYLINDERX = 2, /* y^2 + z^2 = r^2 (Distance), see Contains quirk Q1  */
  MMC_PLANEY = 3,    /* y = a   (additive extension)                       */
  MMC_PLANEZ = 4     /* z = a   (additive extension)                       */
};
typedef struct mmc_surface {
  int32_t kind;
  int32_t pad;
  double a, b, c, r;
} mmc_surface;

/* One (surface, sense) pair of a cell; inside = 1 for sense "-1"
 * (Cell::AssignSurfaceSenses, Cell.cpp:53-80). */
typedef struct mmc_face {
  int32_t surface;
  int32_t inside;
} mmc_face;

/* ScalarField (src/ScalarField.cpp:36-58): constant c, or g.p + b. */
enum mmc_field_kind { MMC_FIELD_CONSTANT = 0, MMC_FIELD_LINEAR = 1 };
typedef struct mmc_field {
  int32_t kind;
  int32_t pad;
  double c;
  double gx, gy, gz, b;
  double lower, upper;
} mmc_field;

/* material = -1 for a void cell (Cell::AssignMaterial, Cell.cpp:82-97). */
typedef struct mmc_cell {
  int32_t material;
  int32_t face_begin;
  int32_t face_count;
  int32_t field;
} mmc_cell;

/* ---- materials (src/Material.cpp:30-94) --------------------------------- */
typedef struct mmc_component {
  int32_t nuclide;
  int32_t pad;
  double afrac; /* as given; mmc_create normalises per material */
} mmc_component;
typedef struct mmc_material {
  double aden; /* number density, atoms / (barn cm) */
  int32_t comp_begin;
  int32_t comp_count;
} mmc_material;

/* ---- multigroup nuclide (src/Multigroup.cpp) ----------------------------
 * Matrices are G x G in XML order: row = outgoing group, column = incoming
 * group (Multigroup.cpp:124-147).  Absent reactions pass NULL. */
typedef struct mmc_mg_nuclide {
  const double* capture;    /* [G] or NULL                              */
  const double* scatter;    /* [G*G] or NULL                            */
  const double* fission_xs; /* [G] or NULL                              */
  const double* nubar;      /* [G] (with fission)                       */
  const double* chi;        /* [G*G] (with fission)                     */
} mmc_mg_nuclide;

/* ---- continuous-energy nuclide (src/Continuous*.cpp) ---------------------
 * A table is a ContinuousMap (ContinuousMap.hpp:20-56): ascending keys,
 * lin-lin interpolation, snapping to the end values outside. */
typedef struct mmc_table {
  int32_t n;
  int32_t pad;
  const double* x;
  const double* y;
} mmc_table;

enum mmc_reaction_kind {
  MMC_RX_CAPTURE = 0,
  MMC_RX_SCATTER = 1,
  MMC_RX_FISSION = 2
};
/* Fission-neutron energy: the reference gives children the parent's energy
 * (ContinuousReaction.cpp:258-264); a Watt spectrum is an additive extension
 * p(E) ~ exp(-E/a) sinh(sqrt(b E)). */
enum mmc_chi_kind { MMC_CHI_INHERIT = 0, MMC_CHI_WATT = 1 };
typedef struct mmc_ce_reaction {
  int32_t kind;
  int32_t tsl; /* index into mmc_world_desc.tsl, or -1 (scatter only)     */
  mmc_table xs;
  double temperature; /* evaluation temperature                          */
  mmc_table nubar;    /* fission only                                    */
  int32_t chi_kind;
  int32_t pad;
  double watt_a, watt_b;
} mmc_ce_reaction;
typedef struct mmc_ce_nuclide {
  double awr;
  int32_t has_total;
  int32_t n_reactions; /* in XML order, "total" excluded               */
  mmc_table total;
  double total_temperature;
  const mmc_ce_reaction* reactions;
} mmc_ce_nuclide;

/* ---- thermal scattering law, POD form (src/ThermalScattering.hpp) --------
 * A partition holds rank-r modes so that
 *   value(F_j; x_i, T_k) = sum_r S_r * cdf_modes[j][r] * x_T_modes[i][k][r]
 * with x = incident energy E for beta partitions and x = beta for alpha
 * partitions (BetaPartition/AlphaPartition::Evaluate,
 * ThermalScattering.cpp:183-256).  Layout is the reference's pandas-HDF5
 * block layout: row-major, last axis fastest (HDF5DataSet.hpp:118-167).
 * Every count is at least 1: mmc_create refuses a partition with an empty
 * axis (MMC_ERR_ARG). */
typedef struct mmc_tsl_partition {
  int32_t n_F, n_x, n_T, rank;
  const double* F;         /* [n_F]            CDF axis                 */
  const double* cdf_modes; /* [n_F * rank]                              */
  const double* S;         /* [rank]           singular values          */
  const double* x;         /* [n_x]            E (beta) or beta (alpha) */
  const double* T;         /* [n_T]                                     */
  const double* x_T_modes; /* [n_x * n_T * rank]                        */
} mmc_tsl_partition;
typedef struct mmc_tsl {
  mmc_table majorant;
  int32_t n_E, n_T, rank, pad;
  const double* total_E_axis; /* [n_E]        scatter_xs_E axis 0         */
  const double* total_E;      /* [n_E * rank]                            */
  const double* total_T_axis; /* [n_T]                                   */
  const double* total_T;      /* [n_T * rank]                            */
  const double* total_S;      /* [rank]                                  */
  int32_t n_beta_partitions, n_alpha_partitions;
  const mmc_tsl_partition* beta_partitions;
  const mmc_tsl_partition* alpha_partitions;
  double beta_cutoff, alpha_cutoff;
} mmc_tsl;

/* ---- the World --------------------------------------------------------- */
enum mmc_energy_mode { MMC_MULTIGROUP = 0, MMC_CONTINUOUS = 1 };
typedef struct mmc_world_desc {
  int32_t energy_mode; /* mmc_energy_mode                                */
  int32_t groups;      /* G (multigroup)                                 */
  int32_t n_surfaces, n_faces, n_fields, n_cells;
  int32_t n_materials, n_components, n_nuclides, n_tsl;
  const mmc_surface* surfaces;
  const mmc_face* faces;
  const mmc_field* fields;
  const mmc_cell* cells; /* FindCellContaining scans them in this order */
  const mmc_material* materials;
  const mmc_component* components;
  const mmc_mg_nuclide* mg; /* [n_nuclides] when multigroup             */
  const mmc_ce_nuclide* ce; /* [n_nuclides] when continuous             */
  const mmc_tsl* tsl;       /* [n_tsl]                                  */
  /* Total cross-section perturbations (TotalCrossSectionPerturbation,
   * src/Perturbation.cpp:44-83): the perturbed nuclide of each, in input
   * order, at most MMC_MAX_PERTURBATIONS.  Every particle carries one
   * indirect effect per perturbation (Particle::indirect_effects,
   * src/Particle.hpp:105-106). */
  int32_t n_perturbations;
  int32_t pad_p;
  const int32_t* perturbation_nuclide;
} mmc_world_desc;
#define MMC_MAX_PERTURBATIONS 4

/* ---- estimators (src/Estimator.cpp:142-151, src/Bins.cpp) ---------------- */
enum mmc_bins_kind {
  MMC_BINS_NONE = 0,
  MMC_BINS_LINSPACE = 1,
  MMC_BINS_LOGSPACE = 2,
  MMC_BINS_BOUNDARIES = 3
};
typedef struct mmc_bins {
  int32_t kind;
  int32_t n;            /* "bins" attribute, or number of boundaries     */
  double lo, hi, base;  /* min, max, logspace base                       */
  const double* bounds; /* boundaries                                    */
} mmc_bins;
/* MMC_EST_CURRENT is the reference's CurrentEstimator.  The others are
 * additive extensions for the k-eigenvalue and thermal-scattering configs:
 * a collision-estimator flux on a Cartesian mesh (score 1/Sigma_t at each
 * real collision, binned by position), and histograms of the (alpha, beta)
 * pairs sampled in thermal scattering collisions. */
enum mmc_estimator_kind {
  MMC_EST_CURRENT = 0,
  MMC_EST_FLUX_MESH = 1,
  MMC_EST_TSL_ALPHA = 2,
  MMC_EST_TSL_BETA = 3
};
/* A current estimator with sensitivity = j + 1 is the sensitivity of that
 * estimator to perturbation j (CurrentTotalCrossSectionSensitivity,
 * src/Sensitivity.cpp:45-58): same surface and bins, and each score is the
 * particle's indirect effect times the current's score. */
typedef struct mmc_estimator {
  int32_t kind;
  int32_t surface;       /* MMC_EST_CURRENT                             */
  int32_t has_direction; /* cosine bins present                         */
  int32_t sensitivity;   /* 0, or 1 + perturbation index                */
  double u, v, w;        /* cosine reference direction (unnormalised)   */
  mmc_bins cosine;
  mmc_bins energy;       /* alpha / beta bins for the TSL histograms    */
  int32_t mesh_n[3];     /* MMC_EST_FLUX_MESH                           */
  int32_t pad2;
  double mesh_lo[3], mesh_hi[3];
} mmc_estimator;
typedef struct mmc_tally_desc {
  int32_t n_estimators;
  int32_t pad;
  const mmc_estimator* estimators;
} mmc_tally_desc;

/* ---- sources (src/Source.cpp:26-154) ------------------------------------ */
enum mmc_dir_kind {
  MMC_DIR_CONSTANT = 0,
  MMC_DIR_ISOTROPIC = 1,
  MMC_DIR_ISOTROPIC_FLUX = 2
};
enum mmc_energy_kind {
  MMC_E_CONSTANT = 0,
  MMC_E_UNIFORM = 1 /* additive extension: uniform in [E, E_hi)         */
};
typedef struct mmc_source_desc {
  double x, y, z;
  int32_t dir_kind;
  int32_t energy_kind;
  double u, v, w; /* constant direction / isotropic-flux reference    */
  double E, E_hi; /* MeV, or group number for multigroup             */
  /* optional position box: when box_hi > box_lo on an axis, x (y, z) is
   * uniform in [lo, hi) on that axis (additive extension, k-eig start) */
  double box_lo[3], box_hi[3];
} mmc_source_desc;

/* ---- options ------------------------------------------------------------ */
enum mmc_tracking { MMC_SURFACE_TRACKING = 0, MMC_CELL_DELTA_TRACKING = 1 };
enum mmc_tsl_mode { MMC_TSL_POD = 0, MMC_TSL_SURROGATE = 1 };
typedef struct mmc_options {
  int32_t tracking;     /* mmc_tracking (TransportMethod.cpp:27-45)      */
  int32_t tsl_mode;     /* mmc_tsl_mode                                 */
  int32_t device;       /* HIP device ordinal                           */
  int32_t block;        /* threads per workgroup (0 = default 256)      */
  uint64_t max_in_flight; /* particles resident per chunk (0 = default)  */
  int32_t graph_iters;  /* event iterations per host read of the queue
                           lengths, and per captured graph when graphs is
                           on (0 = 8)                                     */
  int32_t tail_threshold; /* live particles below which the fused tail
                             kernel finishes a batch (0 = default,
                             -1 = never)                                */
  uint64_t max_steps;   /* per-history event limit (0 = 1e6)            */
  int32_t quirks;       /* mmc_quirk bits (0 = corrected behaviour)      */
  int32_t compact_interval; /* bank compaction: 0 = when the live particles
                             fill less than half of the slots in use (at a
                             read of the queue lengths), k > 0 = before
                             every k-th iteration, -1 = never             */
  int32_t graphs;       /* 1: k-eigenvalue cycles replay each batch of
                           graph_iters iterations (and its read of the
                           queue lengths) as a captured HIP graph, one per
                           ring phase and launch shape; 0: launch the
                           kernels one by one                             */
  int32_t tail_block;   /* threads per block of the tail kernel: 256 or 512
                           (0 = 256)                                      */
  int32_t tsl_block;    /* threads per block of the POD thermal-scattering
                           kernel: 256 or 384 (0 = 256)                   */
  int32_t tail_network; /* surrogate mode, k > 0: while at least k
                           particles are live, the tail runs in rounds that
                           park thermal collisions for the network; below
                           k the tail kernel samples them with POD.  0 (the
                           default): the tail samples with POD throughout */
  int32_t split_collide; /* k-eigenvalue runs with the tables in LDS: 0 (the
                           default): the advance kernel also does the
                           collision that ends a flight; 1: separate
                           advance and collision kernels, as in fixed-
                           source runs.  Results are the same bit for bit. */
} mmc_options;
/* Reference behaviours that are reproduced only on request.
 * MMC_QUIRK_NUDGE_CROSSING: after a surface crossing the new cell is the
 *   first cell containing the point nudged by 10 epsilon (absolute;
 *   TransportMethod.cpp:68, Constants.hpp:15).  At |x| >~ 1 cm the nudge can
 *   vanish in rounding and leave the particle in the old cell, beyond its
 *   surface.  Without the bit the crossed surface's sense is flipped
 *   explicitly when the new cell is searched. */
enum mmc_quirk { MMC_QUIRK_NUDGE_CROSSING = 1 };

/* ---- particle state for single-event parity checks ----------------------
 * Structure of arrays, n entries each.  For multigroup, E holds the group
 * number as a double.  rng holds the minstd_rand state (< 2^31). */
typedef struct mmc_particle_soa {
  double *x, *y, *z, *u, *v, *w, *E;
  uint32_t* rng;
  int32_t* cell;
  int32_t* surface; /* last crossed surface, -1 none                   */
  int32_t* event;   /* mmc_event                                        */
  double* out0;     /* event-specific scalar outputs                    */
  double* out1;
  int32_t* iout;    /* event-specific integer output                    */
} mmc_particle_soa;

/* Particle::Event (Particle.hpp:36-44). */
enum mmc_event {
  MMC_EV_BIRTH = 0,
  MMC_EV_SCATTER = 1,
  MMC_EV_CAPTURE = 2,
  MMC_EV_FISSION = 3,
  MMC_EV_SURFACE_CROSS = 4,
  MMC_EV_LEAK = 5,
  MMC_EV_VIRTUAL = 6
};

/* Single-event kinds for mmc_eval_event (each one names the reference
 * function it evaluates on identical particle states). */
enum mmc_event_kind {
  MMC_EVAL_FIND_CELL = 0,       /* World::FindCellContaining -> cell      */
  MMC_EVAL_NEAREST_SURFACE = 1, /* Cell::NearestSurface -> iout, out0    */
  MMC_EVAL_XS_TOTAL = 2,        /* Material::GetMicroscopicTotal -> out0,
                                   Material::GetMicroscopicMajorant -> out1 */
  MMC_EVAL_COLLISION_DISTANCE = 3, /* exponential(aden*total)(rng) -> out0 */
  MMC_EVAL_COLLIDE = 4,         /* collision at the current position:
                                   Particle::SampleNuclide +
                                   Nuclide::Interact; iout = number of
                                   secondaries banked                   */
  MMC_EVAL_SCATTER_ROTATE = 5,  /* Particle::Scatter(mu=out0, E'=out1)    */
  MMC_EVAL_TSL_SAMPLE = 6       /* ThermalScattering::SampleBeta then
                                   SampleAlpha: in E, T = out0;
                                   out0 = alpha, out1 = beta (tsl 0)     */
};

/* ---- counters ----------------------------------------------------------- */
typedef struct mmc_counters {
  uint64_t histories;
  uint64_t xs_lookups;   /* cross-section lookup events                 */
  uint64_t advances;     /* distance-to-collision + CSG distance        */
  uint64_t collisions;   /* real collisions                             */
  uint64_t tsl_collisions;
  uint64_t tsl_fallbacks; /* surrogate inputs outside the trained box    */
  uint64_t crossings;    /* surface crossings (incl. leaks)             */
  uint64_t virtual_collisions;
  uint64_t fission_sites;
  uint64_t iterations;   /* event-loop iterations                       */
  uint64_t tail_particles; /* particles finished by the fused tail kernel */
  double transport_seconds;
  /* events run inside the tail kernel (also counted above) */
  uint64_t tail_advances, tail_collisions, tail_tsl, tail_crossings;
  uint64_t mlp_rows;     /* thermal collisions sampled by the network     */
  uint64_t graph_launches, graph_captures; /* batches replayed as graphs,
                                              and graphs captured         */
} mmc_counters;

/* ---- k-eigenvalue cycle statistics -------------------------------------- */
typedef struct mmc_cycle_stats {
  double k_analog;      /* fission sites banked / source sites          */
  double k_collision;   /* sum over collisions of nu Sigma_f / Sigma_t  */
  double k_track;       /* sum over tracks of d * nu Sigma_f            */
  double k_absorption;  /* sum over absorptions of nu sigma_f / sigma_a */
  double entropy;       /* Shannon entropy of the fission sites (bits)  */
  uint64_t n_sites;     /* fission sites banked this cycle              */
  double seconds;
} mmc_cycle_stats;

/* A fission site: what the next cycle's source needs.  key orders the bank
 * deterministically: parent history index * 64 + child index. */
typedef struct mmc_site {
  double x, y, z, u, v, w, E;
  uint32_t seed;
  uint32_t pad;
  uint64_t key;
} mmc_site;

/* ---- context ------------------------------------------------------------ */
typedef struct mmc_ctx mmc_ctx;

int mmc_abi_version(void);
int mmc_device_count(int* n);

int mmc_create(const mmc_world_desc* world, const mmc_tally_desc* tallies,
               const mmc_options* options, mmc_ctx** out);
void mmc_destroy(mmc_ctx* ctx);
const char* mmc_last_error(const mmc_ctx* ctx);

/* Fixed source: histories seed0 + i for i in [first, first + n)
 * (FixedSource.cpp:61).  Tallies accumulate across calls. */
int mmc_run_fixed_source(mmc_ctx* ctx, const mmc_source_desc* source,
                         uint64_t seed0, uint64_t first, uint64_t n);

/* k-eigenvalue: the initial source is sampled with seeds 1..n
 * (KEigenvalue.cpp:31-33).  Each call to mmc_run_cycle transports the
 * current source, banks fission sites, sorts them by key and resamples n
 * sites into the next source. */
int mmc_keig_init(mmc_ctx* ctx, const mmc_source_desc* source, uint64_t n);
int mmc_run_cycle(mmc_ctx* ctx, int active, mmc_cycle_stats* out);
/* Copy the current source (n entries) to host memory. */
int mmc_get_source(mmc_ctx* ctx, mmc_site* out, uint64_t cap, uint64_t* n);
/* Replace the current source with n host sites. */
int mmc_set_source(mmc_ctx* ctx, const mmc_site* sites, uint64_t n);
/* Shannon entropy mesh (default: the bounding box of the non-void cells'
 * planes split 8x8x8, or 100x1x1 for slabs). */
int mmc_set_entropy_mesh(mmc_ctx* ctx, const int32_t n[3], const double lo[3],
                         const double hi[3]);

/* Tally results: S and S2 per bin of estimator e (Scorable.cpp:101-106).
 * For fixed-source runs S2 holds per-history squares; for k-eigenvalue
 * flux tallies it holds per-active-cycle squares. */
int mmc_tally_size(mmc_ctx* ctx, int estimator, uint64_t* n);
int mmc_get_tallies(mmc_ctx* ctx, int estimator, double* S, double* S2,
                    uint64_t n);
int mmc_reset_tallies(mmc_ctx* ctx);
int mmc_get_counters(mmc_ctx* ctx, mmc_counters* out);

/* Kernel timing: with profiling on, every event-kernel launch is bracketed
 * by HIP events on the context's stream and its duration accumulated per
 * kernel kind (mmc_kernel_kind).  Off by default. */
enum mmc_kernel_kind {
  MMC_K_ADVANCE = 0, /* XS lookup + collision distance + CSG distance,
                        and the surface crossing that follows (cell
                        search + current tally); in k-eigenvalue runs
                        with the tables in LDS also the collision that
                        ends the flight                                */
  MMC_K_CROSS = 1,   /* reserved: crossings run in the advance kernel,
                        so this kind stays at 0 launches               */
  MMC_K_COLLIDE = 2, /* nuclide + reaction sampling, collision physics
                        (its own kernel where MMC_K_ADVANCE does not
                        collide)                                       */
  MMC_K_TSL = 3,     /* thermal-scattering (alpha, beta) sampling: the
                        POD sampler, or the surrogate's gather + apply */
  MMC_K_MLP = 4,     /* the surrogate's fused BF16-MFMA network        */
  MMC_K_TAIL = 5,    /* history-based completion of the last particles */
  MMC_K_N = 6
};
int mmc_set_profiling(mmc_ctx* ctx, int enable);
/* total_ms[k] and launches[k] for k < MMC_K_N (accumulated since enable). */
int mmc_get_kernel_times(mmc_ctx* ctx, double* total_ms, uint64_t* launches);

/* Single-event parity entry point: evaluates kind on n particle states
 * already in host memory (in place). */
int mmc_eval_event(mmc_ctx* ctx, int kind, mmc_particle_soa* p, uint64_t n);

/* Source::Sample (Source.cpp:143-154) for seeds seed0 + i, i < n, into host
 * particle state (x..E, rng = the particle's seed, cell). */
int mmc_sample_source(mmc_ctx* ctx, const mmc_source_desc* source,
                      uint64_t seed0, uint64_t n, mmc_particle_soa* out);

/* ---- device memory (for callers without a HIP runtime of their own) ------
 * Allocations are on the context's device; copies are synchronous. */
int mmc_device_malloc(mmc_ctx* ctx, uint64_t bytes, void** ptr);
int mmc_device_free(mmc_ctx* ctx, void* ptr);
int mmc_memcpy_h2d(mmc_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int mmc_memcpy_d2h(mmc_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- thermal-scattering samplers on device-resident batches -------------
 * Pointers are device pointers.  E in MeV, T in K, rng minstd states
 * (updated in place).  alpha, beta outputs.  tsl selects the table. */
int mmc_tsl_sample_pod(mmc_ctx* ctx, int tsl, uint64_t n, const double* E,
                       const double* T, uint32_t* rng, double* alpha,
                       double* beta);

/* ---- thermal-scattering surrogate (no counterpart in the reference) -----
 * MLP (log E, T, u1, u2) -> (alpha, beta), widths w[0..n_layers], SiLU
 * between layers.  Parameters are FP32 masters; inference runs BF16 MFMA
 * with FP32 accumulation.  Inputs outside [e_min <= E <= e_max,
 * t_lo <= T <= t_hi], and network outputs whose outgoing energy leaves the
 * box (E' < e_min, including E' <= 0), fall back to the POD sampler. */
typedef struct mmc_mlp_desc {
  int32_t n_layers; /* number of linear layers                         */
  int32_t widths[8];
  float in_shift[4], in_scale[4];   /* x_norm = (x - shift) * scale        */
  float out_shift[2], out_scale[2]; /* y = y_norm * scale + shift          */
  double e_max, t_lo, t_hi;         /* trained box                         */
  double e_min;
} mmc_mlp_desc;
/* params: host FP32, layer l: W_l [widths[l+1] x widths[l]] row-major then
 * b_l [widths[l+1]], layers consecutive. */
int mmc_mlp_set(mmc_ctx* ctx, int tsl, const mmc_mlp_desc* desc,
                const float* params, uint64_t n_params);
int mmc_mlp_get(mmc_ctx* ctx, float* params, uint64_t n_params);
/* Device batch inference of the network: x [n x 4] raw inputs (ln E[MeV],
 * T[K], u1, u2) -> y [n x 2] raw outputs (the scaled energy transfer
 * beta kT / (E + kT) and the position of alpha inside its kinematic range
 * [alpha_min(beta), alpha_max(beta)]); FP32 in and out, BF16 MFMA with
 * FP32 accumulation inside (widths 4-256-256-256-2 only). */
int mmc_mlp_infer(mmc_ctx* ctx, uint64_t n, const float* x, float* y);
/* The deterministic POD transport map the network is trained to imitate:
 * the quantile of beta at level u1 and then of alpha at level u2 under the
 * POD sampler's own laws, i.e. the mixtures over the two bracketing
 * tabulated energies (betas) that its stochastic interpolation draws from,
 * truncated where it rejects.  Device x [n x 4] raw inputs -> y [n x 2]
 * raw outputs in the network's output parametrisation. */
int mmc_tsl_pod_map(mmc_ctx* ctx, int tsl, uint64_t n, const float* x, float* y);
/* Surrogate sampling of (alpha, beta) for device batches with fallback. */
int mmc_tsl_sample_surrogate(mmc_ctx* ctx, int tsl, uint64_t n,
                             const double* E, const double* T, uint32_t* rng,
                             double* alpha, double* beta);
/* mmc_tsl_sample_surrogate without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The network's inputs, outputs and error word live in a static device
 * array of 6 MiB per GPU (held until the process ends), so n is at most
 * 262144 rows; a larger n returns MMC_ERR_ARG before anything runs.  The
 * results equal mmc_tsl_sample_surrogate's bit for bit.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other host
 * threads, other contexts) race. */
int mmc_tsl_sample_surrogate_noalloc(mmc_ctx* ctx, int tsl, uint64_t n,
                                     const double* E, const double* T, uint32_t* rng,
                                     double* alpha, double* beta);
/* Training on device samples: x [n x 4] normalised inputs, y [n x 2]
 * normalised targets (device pointers).  Runs `steps` Adam steps over
 * minibatches of `batch` rows drawn in order (wrapping), returns the mean
 * loss of the last step.  Forward and backward are this library's own
 * FP32 kernels. */
typedef struct mmc_train_opts {
  uint64_t batch;
  uint64_t steps;
  float lr, beta1, beta2, eps;
  uint64_t offset; /* first row of the first minibatch                 */
  const float* weights; /* optional device [n x n_out] loss weights: the
                           minimised loss is sum w (f - y)^2 / (2 b n_out) */
} mmc_train_opts;
int mmc_mlp_train(mmc_ctx* ctx, uint64_t n, const float* x, const float* y,
                  const mmc_train_opts* opts, float* last_loss);
/* Normalises raw device training data in place with the network's shifts
 * and scales: x <- (x - in_shift) * in_scale, y <- (y - out_shift) /
 * out_scale (y may be NULL). */
int mmc_mlp_normalize(mmc_ctx* ctx, uint64_t n, float* x, float* y);
/* One forward+backward on a host batch (normalised units): returns the
 * loss L = sum over rows and outputs of (f(x) - y)^2 / (2 n n_out) and, if
 * grads is not NULL, dL/dparams in the params layout.  For gradient checks
 * against the CPU. */
int mmc_mlp_grad(mmc_ctx* ctx, uint64_t n, const float* x_host,
                 const float* y_host, float* loss, float* grads);

#ifdef __cplusplus
}
#endif
#endif /* MINIMC_AMD_H */
