"""ctypes mirror of include/minimc_amd.h (the C ABI of the tracker).

Field order and types follow the header exactly; tests check the struct
sizes against the library's own sizeof values.
"""
import ctypes as C

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_u32_p = C.POINTER(C.c_uint32)
c_i32_p = C.POINTER(C.c_int32)

MMC_OK = 0
MMC_ERR_ARG, MMC_ERR_HIP, MMC_ERR_LOST, MMC_ERR_RESAMPLE = 1, 2, 3, 4
MMC_ERR_OVERFLOW, MMC_ERR_NODEV, MMC_ERR_UNSUPPORTED, MMC_ERR_STEPS = 5, 6, 7, 8

MMC_SPHERE, MMC_PLANEX, MMC_CYLINDERX, MMC_PLANEY, MMC_PLANEZ = 0, 1, 2, 3, 4
MMC_FIELD_CONSTANT, MMC_FIELD_LINEAR = 0, 1
MMC_RX_CAPTURE, MMC_RX_SCATTER, MMC_RX_FISSION = 0, 1, 2
MMC_CHI_INHERIT, MMC_CHI_WATT = 0, 1
MMC_MULTIGROUP, MMC_CONTINUOUS = 0, 1
MMC_BINS_NONE, MMC_BINS_LINSPACE, MMC_BINS_LOGSPACE, MMC_BINS_BOUNDARIES = 0, 1, 2, 3
MMC_EST_CURRENT, MMC_EST_FLUX_MESH, MMC_EST_TSL_ALPHA, MMC_EST_TSL_BETA = 0, 1, 2, 3
MMC_DIR_CONSTANT, MMC_DIR_ISOTROPIC, MMC_DIR_ISOTROPIC_FLUX = 0, 1, 2
MMC_E_CONSTANT, MMC_E_UNIFORM = 0, 1
MMC_SURFACE_TRACKING, MMC_CELL_DELTA_TRACKING = 0, 1
MMC_TSL_POD, MMC_TSL_SURROGATE = 0, 1
MMC_QUIRK_NUDGE_CROSSING = 1
(MMC_EV_BIRTH, MMC_EV_SCATTER, MMC_EV_CAPTURE, MMC_EV_FISSION,
 MMC_EV_SURFACE_CROSS, MMC_EV_LEAK, MMC_EV_VIRTUAL) = range(7)
(MMC_EVAL_FIND_CELL, MMC_EVAL_NEAREST_SURFACE, MMC_EVAL_XS_TOTAL,
 MMC_EVAL_COLLISION_DISTANCE, MMC_EVAL_COLLIDE, MMC_EVAL_SCATTER_ROTATE,
 MMC_EVAL_TSL_SAMPLE, MMC_EVAL_SOURCE) = range(8)


class Surface(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("a", C.c_double),
                ("b", C.c_double), ("c", C.c_double), ("r", C.c_double)]


class Face(C.Structure):
    _fields_ = [("surface", C.c_int32), ("inside", C.c_int32)]


class Field(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("c", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double),
                ("b", C.c_double), ("lower", C.c_double), ("upper", C.c_double)]


class Cell(C.Structure):
    _fields_ = [("material", C.c_int32), ("face_begin", C.c_int32),
                ("face_count", C.c_int32), ("field", C.c_int32)]


class Component(C.Structure):
    _fields_ = [("nuclide", C.c_int32), ("pad", C.c_int32), ("afrac", C.c_double)]


class Material(C.Structure):
    _fields_ = [("aden", C.c_double), ("comp_begin", C.c_int32),
                ("comp_count", C.c_int32)]


class MgNuclide(C.Structure):
    _fields_ = [("capture", c_double_p), ("scatter", c_double_p),
                ("fission_xs", c_double_p), ("nubar", c_double_p),
                ("chi", c_double_p)]


class Table(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad", C.c_int32), ("x", c_double_p),
                ("y", c_double_p)]


class CeReaction(C.Structure):
    _fields_ = [("kind", C.c_int32), ("tsl", C.c_int32), ("xs", Table),
                ("temperature", C.c_double), ("nubar", Table),
                ("chi_kind", C.c_int32), ("pad", C.c_int32),
                ("watt_a", C.c_double), ("watt_b", C.c_double)]


class CeNuclide(C.Structure):
    _fields_ = [("awr", C.c_double), ("has_total", C.c_int32),
                ("n_reactions", C.c_int32), ("total", Table),
                ("total_temperature", C.c_double),
                ("reactions", C.POINTER(CeReaction))]


class TslPartition(C.Structure):
    _fields_ = [("n_F", C.c_int32), ("n_x", C.c_int32), ("n_T", C.c_int32),
                ("rank", C.c_int32), ("F", c_double_p), ("cdf_modes", c_double_p),
                ("S", c_double_p), ("x", c_double_p), ("T", c_double_p),
                ("x_T_modes", c_double_p)]


class Tsl(C.Structure):
    _fields_ = [("majorant", Table), ("n_E", C.c_int32), ("n_T", C.c_int32),
                ("rank", C.c_int32), ("pad", C.c_int32),
                ("total_E_axis", c_double_p), ("total_E", c_double_p),
                ("total_T_axis", c_double_p), ("total_T", c_double_p),
                ("total_S", c_double_p),
                ("n_beta_partitions", C.c_int32), ("n_alpha_partitions", C.c_int32),
                ("beta_partitions", C.POINTER(TslPartition)),
                ("alpha_partitions", C.POINTER(TslPartition)),
                ("beta_cutoff", C.c_double), ("alpha_cutoff", C.c_double)]


class WorldDesc(C.Structure):
    _fields_ = [("energy_mode", C.c_int32), ("groups", C.c_int32),
                ("n_surfaces", C.c_int32), ("n_faces", C.c_int32),
                ("n_fields", C.c_int32), ("n_cells", C.c_int32),
                ("n_materials", C.c_int32), ("n_components", C.c_int32),
                ("n_nuclides", C.c_int32), ("n_tsl", C.c_int32),
                ("surfaces", C.POINTER(Surface)), ("faces", C.POINTER(Face)),
                ("fields", C.POINTER(Field)), ("cells", C.POINTER(Cell)),
                ("materials", C.POINTER(Material)),
                ("components", C.POINTER(Component)),
                ("mg", C.POINTER(MgNuclide)), ("ce", C.POINTER(CeNuclide)),
                ("tsl", C.POINTER(Tsl)),
                ("n_perturbations", C.c_int32), ("pad_p", C.c_int32),
                ("perturbation_nuclide", C.POINTER(C.c_int32))]


MMC_MAX_PERTURBATIONS = 4


class Bins(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("lo", C.c_double),
                ("hi", C.c_double), ("base", C.c_double), ("bounds", c_double_p)]


class Estimator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("surface", C.c_int32),
                ("has_direction", C.c_int32), ("sensitivity", C.c_int32),
                ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("cosine", Bins), ("energy", Bins),
                ("mesh_n", C.c_int32 * 3), ("pad2", C.c_int32),
                ("mesh_lo", C.c_double * 3), ("mesh_hi", C.c_double * 3)]


class TallyDesc(C.Structure):
    _fields_ = [("n_estimators", C.c_int32), ("pad", C.c_int32),
                ("estimators", C.POINTER(Estimator))]


class SourceDesc(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("dir_kind", C.c_int32), ("energy_kind", C.c_int32),
                ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("E", C.c_double), ("E_hi", C.c_double),
                ("box_lo", C.c_double * 3), ("box_hi", C.c_double * 3)]


class Options(C.Structure):
    _fields_ = [("tracking", C.c_int32), ("tsl_mode", C.c_int32),
                ("device", C.c_int32), ("block", C.c_int32),
                ("max_in_flight", C.c_uint64), ("graph_iters", C.c_int32),
                ("tail_threshold", C.c_int32), ("max_steps", C.c_uint64),
                ("quirks", C.c_int32), ("compact_interval", C.c_int32),
                ("graphs", C.c_int32), ("tail_block", C.c_int32),
                ("tsl_block", C.c_int32), ("tail_network", C.c_int32),
                ("split_collide", C.c_int32)]


class ParticleSoA(C.Structure):
    _fields_ = [("x", c_double_p), ("y", c_double_p), ("z", c_double_p),
                ("u", c_double_p), ("v", c_double_p), ("w", c_double_p),
                ("E", c_double_p), ("rng", c_u32_p), ("cell", c_i32_p),
                ("surface", c_i32_p), ("event", c_i32_p), ("out0", c_double_p),
                ("out1", c_double_p), ("iout", c_i32_p)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "histories", "xs_lookups", "advances", "collisions", "tsl_collisions",
        "tsl_fallbacks", "crossings", "virtual_collisions", "fission_sites",
        "iterations", "tail_particles")] + [("transport_seconds", C.c_double)] + [
        (n, C.c_uint64) for n in ("tail_advances", "tail_collisions", "tail_tsl",
                                  "tail_crossings", "mlp_rows", "graph_launches",
                                  "graph_captures")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CycleStats(C.Structure):
    _fields_ = [("k_analog", C.c_double), ("k_collision", C.c_double),
                ("k_track", C.c_double), ("k_absorption", C.c_double),
                ("entropy", C.c_double), ("n_sites", C.c_uint64),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Site(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("E", C.c_double), ("seed", C.c_uint32), ("pad", C.c_uint32),
                ("key", C.c_uint64)]


class MlpDesc(C.Structure):
    _fields_ = [("n_layers", C.c_int32), ("widths", C.c_int32 * 8),
                ("in_shift", C.c_float * 4), ("in_scale", C.c_float * 4),
                ("out_shift", C.c_float * 2), ("out_scale", C.c_float * 2),
                ("e_max", C.c_double), ("t_lo", C.c_double), ("t_hi", C.c_double),
                ("e_min", C.c_double)]


class TrainOpts(C.Structure):
    _fields_ = [("batch", C.c_uint64), ("steps", C.c_uint64), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("offset", C.c_uint64), ("weights", C.c_void_p)]


STRUCTS = {
    "mmc_surface": Surface, "mmc_face": Face, "mmc_field": Field,
    "mmc_cell": Cell, "mmc_component": Component, "mmc_material": Material,
    "mmc_mg_nuclide": MgNuclide, "mmc_table": Table,
    "mmc_ce_reaction": CeReaction, "mmc_ce_nuclide": CeNuclide,
    "mmc_tsl_partition": TslPartition, "mmc_tsl": Tsl,
    "mmc_world_desc": WorldDesc, "mmc_bins": Bins, "mmc_estimator": Estimator,
    "mmc_tally_desc": TallyDesc, "mmc_source_desc": SourceDesc,
    "mmc_options": Options, "mmc_particle_soa": ParticleSoA,
    "mmc_counters": Counters, "mmc_cycle_stats": CycleStats, "mmc_site": Site,
    "mmc_mlp_desc": MlpDesc, "mmc_train_opts": TrainOpts,
}


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(c_double_p)
