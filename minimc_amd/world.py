"""Host-side World model and its flattening into the C-ABI descriptor.

`Problem` mirrors what the reference builds in its constructors
(World, Material, Nuclide, Cell, EstimatorSet, Source; src/World.cpp:20-189)
and `Problem.descriptors()` flattens it into the plain-C structs of
include/minimc_amd.h.  `load_xml` reads the reference's XML input schema
(test/minimc.xsd) with the same creation order as World.cpp:
surfaces, nuclides and materials in order of first use by the cells, cells in
document order.
"""
import ctypes as C
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import abi

ROOM_TEMPERATURE = 293.6  # Constants.hpp:22


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Problem:
    """A flattened World plus tallies, source and run settings."""

    def __init__(self, energy_mode=abi.MMC_MULTIGROUP, groups=0):
        self.energy_mode = energy_mode
        self.groups = groups
        self.surfaces = []      # dicts: name, kind, a, b, c, r
        self.fields = []        # dicts: kind, c, g(3), b, lower, upper
        self.cells = []         # dicts: name, material, faces[(s, inside)], field
        self.materials = []     # dicts: name, aden, comps[(nuclide, afrac)]
        self.nuclides = []      # dicts (mg or ce)
        self.tsl = []           # dicts of numpy arrays
        self.estimators = []    # dicts
        # total cross-section perturbations: dicts name, nuclide (index)
        self.perturbations = []
        self.source = None      # dict for SourceDesc
        self.histories = 0
        self.seed = 1
        self.threads = 1
        self.tracking = abi.MMC_SURFACE_TRACKING
        self.problem = "fixedsource"
        self.inactive = 0
        self.active = 0

    # -- builders -----------------------------------------------------------
    def add_field(self, kind=abi.MMC_FIELD_CONSTANT, c=ROOM_TEMPERATURE,
                  g=(0.0, 0.0, 0.0), b=0.0, lower=None, upper=None):
        self.fields.append(dict(kind=kind, c=c, g=tuple(g), b=b,
                                lower=c if lower is None else lower,
                                upper=c if upper is None else upper))
        return len(self.fields) - 1

    def add_surface(self, name, kind, a=0.0, b=0.0, c=0.0, r=0.0):
        self.surfaces.append(dict(name=name, kind=kind, a=a, b=b, c=c, r=r))
        return len(self.surfaces) - 1

    def add_cell(self, name, material, faces, field=0):
        self.cells.append(dict(name=name, material=material,
                               faces=list(faces), field=field))
        return len(self.cells) - 1

    def add_material(self, name, aden, comps):
        self.materials.append(dict(name=name, aden=aden, comps=list(comps)))
        return len(self.materials) - 1

    def surface_index(self, name):
        for i, s in enumerate(self.surfaces):
            if s["name"] == name:
                return i
        raise RuntimeError(
            'Surface "%s" not found. Must be one of: [%s]'
            % (name, "".join('"%s", ' % s["name"] for s in self.surfaces)))

    # -- flattening ---------------------------------------------------------
    def descriptors(self):
        """Returns (WorldDesc, TallyDesc, keepalive)."""
        keep = []

        def arr(struct, items):
            a = (struct * max(1, len(items)))(*items)
            keep.append(a)
            return a

        def dp(x):
            if x is None:
                return None
            x = _f64(x)
            keep.append(x)
            return abi.dptr(x)

        def table(t):
            if t is None:
                return abi.Table(0, 0, None, None)
            x, y = _f64(t[0]), _f64(t[1])
            keep.extend([x, y])
            return abi.Table(len(x), 0, abi.dptr(x), abi.dptr(y))

        w = abi.WorldDesc()
        w.energy_mode = self.energy_mode
        w.groups = self.groups
        surfs = [abi.Surface(s["kind"], 0, s["a"], s["b"], s["c"], s["r"])
                 for s in self.surfaces]
        faces, cells = [], []
        for c in self.cells:
            cells.append(abi.Cell(c["material"], len(faces), len(c["faces"]),
                                  c["field"]))
            faces += [abi.Face(s, 1 if inside else 0) for s, inside in c["faces"]]
        fields = [abi.Field(f["kind"], 0, f["c"], f["g"][0], f["g"][1], f["g"][2],
                            f["b"], f["lower"], f["upper"]) for f in self.fields]
        comps, mats = [], []
        for m in self.materials:
            mats.append(abi.Material(m["aden"], len(comps), len(m["comps"])))
            comps += [abi.Component(n, 0, a) for n, a in m["comps"]]
        w.n_surfaces, w.surfaces = len(surfs), arr(abi.Surface, surfs)
        w.n_faces, w.faces = len(faces), arr(abi.Face, faces)
        w.n_fields, w.fields = len(fields), arr(abi.Field, fields)
        w.n_cells, w.cells = len(cells), arr(abi.Cell, cells)
        w.n_materials, w.materials = len(mats), arr(abi.Material, mats)
        w.n_components, w.components = len(comps), arr(abi.Component, comps)
        w.n_nuclides = len(self.nuclides)
        if self.energy_mode == abi.MMC_MULTIGROUP:
            mg = [abi.MgNuclide(dp(n.get("capture")), dp(n.get("scatter")),
                                dp(n.get("fission_xs")), dp(n.get("nubar")),
                                dp(n.get("chi"))) for n in self.nuclides]
            w.mg = arr(abi.MgNuclide, mg)
        else:
            ce = []
            for n in self.nuclides:
                rxs = []
                for r in n["reactions"]:
                    rxs.append(abi.CeReaction(
                        r["kind"], r.get("tsl", -1), table(r["xs"]),
                        r.get("temperature", ROOM_TEMPERATURE),
                        table(r.get("nubar")), r.get("chi_kind", abi.MMC_CHI_INHERIT),
                        0, r.get("watt_a", 0.0), r.get("watt_b", 0.0)))
                has_total = n.get("total") is not None
                ce.append(abi.CeNuclide(
                    n["awr"], 1 if has_total else 0, len(rxs),
                    table(n.get("total")),
                    n.get("total_temperature", ROOM_TEMPERATURE),
                    arr(abi.CeReaction, rxs)))
            w.ce = arr(abi.CeNuclide, ce)
            tsls = []
            for t in self.tsl:
                def parts(lst):
                    out = []
                    for p in lst:
                        out.append(abi.TslPartition(
                            len(p["F"]), len(p["x"]), len(p["T"]), len(p["S"]),
                            dp(p["F"]), dp(p["cdf_modes"]), dp(p["S"]), dp(p["x"]),
                            dp(p["T"]), dp(p["x_T_modes"])))
                    return arr(abi.TslPartition, out), len(out)
                bp, nbp = parts(t["beta_partitions"])
                ap, nap = parts(t["alpha_partitions"])
                tsls.append(abi.Tsl(
                    table(t["majorant"]), len(t["total_E_axis"]),
                    len(t["total_T_axis"]), len(t["total_S"]), 0,
                    dp(t["total_E_axis"]), dp(t["total_E"]), dp(t["total_T_axis"]),
                    dp(t["total_T"]), dp(t["total_S"]), nbp, nap, bp, ap,
                    t["beta_cutoff"], t["alpha_cutoff"]))
            w.n_tsl = len(tsls)
            w.tsl = arr(abi.Tsl, tsls)
        if self.perturbations:
            pn = np.array([p["nuclide"] for p in self.perturbations], dtype=np.int32)
            keep.append(pn)
            w.n_perturbations = len(pn)
            w.perturbation_nuclide = pn.ctypes.data_as(C.POINTER(C.c_int32))
        ests = [self._estimator(e, keep) for e in self.estimators]
        td = abi.TallyDesc(len(ests), 0, arr(abi.Estimator, ests))
        keep.append(w)
        return w, td, keep

    @staticmethod
    def _bins(b, keep):
        if not b:
            return abi.Bins(abi.MMC_BINS_NONE, 0, 0.0, 0.0, 0.0, None)
        kind = b["kind"]
        if kind == abi.MMC_BINS_BOUNDARIES:
            bd = _f64(b["bounds"])
            keep.append(bd)
            return abi.Bins(kind, len(bd), 0.0, 0.0, 0.0, abi.dptr(bd))
        return abi.Bins(kind, int(b["bins"]), b["min"], b["max"],
                        b.get("base", 10.0), None)

    def _estimator(self, e, keep):
        est = abi.Estimator()
        est.kind = e.get("kind", abi.MMC_EST_CURRENT)
        est.surface = e.get("surface", -1)
        est.sensitivity = e.get("sensitivity", 0)
        cos = e.get("cosine")
        est.has_direction = 1 if cos else 0
        if cos:
            est.u, est.v, est.w = cos["direction"]
            est.cosine = self._bins(cos["bins"], keep)
        est.energy = self._bins(e.get("energy"), keep)
        mesh = e.get("mesh")
        if mesh:
            for a in range(3):
                est.mesh_n[a] = mesh["n"][a]
                est.mesh_lo[a] = mesh["lo"][a]
                est.mesh_hi[a] = mesh["hi"][a]
        return est

    def estimator_size(self, i):
        e = self.estimators[i]

        def bsize(b):
            if not b:
                return 1
            if b["kind"] == abi.MMC_BINS_BOUNDARIES:
                return len(b["bounds"]) + 1
            return int(b["bins"]) + 2
        kind = e.get("kind", abi.MMC_EST_CURRENT)
        if kind == abi.MMC_EST_FLUX_MESH:
            return int(np.prod(e["mesh"]["n"]))
        if kind in (abi.MMC_EST_TSL_ALPHA, abi.MMC_EST_TSL_BETA):
            return bsize(e.get("energy"))
        cos = e.get("cosine")
        return (bsize(cos["bins"]) if cos else 1) * bsize(e.get("energy"))

    def source_desc(self, src=None):
        s = dict(self.source if src is None else src)
        d = abi.SourceDesc()
        d.x, d.y, d.z = s.get("position", (0.0, 0.0, 0.0))
        d.dir_kind = s.get("dir_kind", abi.MMC_DIR_ISOTROPIC)
        d.u, d.v, d.w = s.get("direction", (1.0, 0.0, 0.0))
        d.energy_kind = s.get("energy_kind", abi.MMC_E_CONSTANT)
        d.E = s["E"]
        d.E_hi = s.get("E_hi", s["E"])
        box = s.get("box")
        if box:
            for a in range(3):
                d.box_lo[a], d.box_hi[a] = box[0][a], box[1][a]
        return d


# ---------------------------------------------------------------------------
# XML input (test/minimc.xsd).  Creation order follows World.cpp:85-189.

def _floats(node, path, expected):
    vals = [float(t) for t in (node.text or "").split()]
    if len(vals) != expected:
        raise RuntimeError("%s: Expected %d entries but got %d"
                           % (path, expected, len(vals)))
    return np.array(vals)


def _resolve(path, base):
    return path if os.path.isabs(path) else os.path.join(base, path)


def _load_table(path, base):
    """A ContinuousMap table: a pandas-HDF5 file of one level (the
    reference's format, HDF5DataSet<1>::ToContinuousMap), or a .npy (n x 2:
    key, value) or .npz (x, y)."""
    p = _resolve(path, base)
    if p.endswith((".hdf5", ".h5")):
        from . import pandas_h5
        x, y = pandas_h5.read_table(p)
        return _f64(x), _f64(y)
    if not os.path.exists(p):
        raise RuntimeError("File not found: " + path)
    if p.endswith(".npz"):
        z = np.load(p)
        return _f64(z["x"]), _f64(z["y"])
    a = np.load(p)
    return _f64(a[:, 0]), _f64(a[:, 1])


_SURF = {"sphere": abi.MMC_SPHERE, "planex": abi.MMC_PLANEX,
         "cylinderx": abi.MMC_CYLINDERX, "planey": abi.MMC_PLANEY,
         "planez": abi.MMC_PLANEZ}


def _bins_from_xml(node):
    if node is None:
        return None
    child = list(node)
    if not child:
        return None
    b = child[0]
    if b.tag == "linspace":
        return dict(kind=abi.MMC_BINS_LINSPACE, min=float(b.get("min")),
                    max=float(b.get("max")), bins=int(b.get("bins")))
    if b.tag == "logspace":
        return dict(kind=abi.MMC_BINS_LOGSPACE, min=float(b.get("min")),
                    max=float(b.get("max")), bins=int(b.get("bins")),
                    base=float(b.get("base", 10.0)))
    if b.tag == "boundaries":
        return dict(kind=abi.MMC_BINS_BOUNDARIES,
                    bounds=[float(t) for t in (b.text or "").split()])
    raise RuntimeError("unknown bins " + b.tag)


def _source_from_xml(node, multigroup):
    s = {}
    pos = node.find("position")[0]
    s["position"] = (float(pos.get("x")), float(pos.get("y")), float(pos.get("z")))
    box = node.find("position").find("box")
    if box is not None:
        s["box"] = ([float(box.get(k)) for k in ("xmin", "ymin", "zmin")],
                    [float(box.get(k)) for k in ("xmax", "ymax", "zmax")])
    d = node.find("direction")[0]
    if d.tag == "constant":
        s["dir_kind"] = abi.MMC_DIR_CONSTANT
        s["direction"] = (float(d.get("x")), float(d.get("y")), float(d.get("z")))
    elif d.tag == "isotropic":
        s["dir_kind"] = abi.MMC_DIR_ISOTROPIC
    else:
        s["dir_kind"] = abi.MMC_DIR_ISOTROPIC_FLUX
        s["direction"] = (float(d.get("x")), float(d.get("y")), float(d.get("z")))
    e = node.find("energy")[0]
    if e.tag == "uniform":
        s["energy_kind"] = abi.MMC_E_UNIFORM
        s["E"], s["E_hi"] = float(e.get("min")), float(e.get("max"))
    else:
        s["E"] = float(int(e.get("energy"))) if multigroup else float(e.get("energy"))
    return s


def load_xml(path):
    """Build a Problem from a minimc XML input."""
    base = os.path.dirname(os.path.abspath(path))
    root = ET.parse(path).getroot()
    nuc_root = root.find("nuclides")
    etype = list(nuc_root)[0]
    multigroup = etype.tag == "multigroup"
    G = int(etype.get("groups")) if multigroup else 0
    pr = Problem(abi.MMC_MULTIGROUP if multigroup else abi.MMC_CONTINUOUS, G)
    gen = root.find("general")
    pr.histories = int(gen.find("histories").text)
    pr.threads = int(gen.find("threads").text)
    pr.seed = int(gen.find("seed").text) if gen.find("seed") is not None else 1
    tr = gen.find("tracking")
    tracking = (tr.text or "").strip() if tr is not None else ""
    pr.tracking = (abi.MMC_CELL_DELTA_TRACKING if tracking == "cell delta"
                   else abi.MMC_SURFACE_TRACKING)
    # global temperature (World::CreateTemperature, World.cpp:165-173)
    tnode = root.find("temperature")
    if tnode is None:
        pr.add_field(c=ROOM_TEMPERATURE)
    else:
        f = list(tnode)[0]
        if f.tag == "constant":
            pr.add_field(c=float(f.get("c")))
        else:
            bnd, g = f.find("bounds"), f.find("gradient")
            pr.add_field(kind=abi.MMC_FIELD_LINEAR, c=0.0,
                         g=(float(g.get("x")), float(g.get("y")), float(g.get("z"))),
                         b=float(f.find("intercept").get("b")),
                         lower=float(bnd.get("lower")), upper=float(bnd.get("upper")))
    if pr.tracking == abi.MMC_SURFACE_TRACKING:
        f0 = pr.fields[0]
        constant = f0["kind"] == abi.MMC_FIELD_CONSTANT or f0["g"] == (0.0, 0.0, 0.0)
        if not constant:
            raise RuntimeError(
                "Surface tracking with continuous global temperature not allowed")
    cells_node = root.find("cells")
    surfaces_node = root.find("surfaces")
    # surfaces in order of first use (World::CreateCSGSurfaces)
    for cell in cells_node:
        for sn in cell:
            name = sn.get("name")
            if any(s["name"] == name for s in pr.surfaces):
                continue
            node = None
            for cand in surfaces_node:
                if cand.get("name") == name:
                    node = cand
            if node is None:
                raise RuntimeError(
                    'Surface node "%s" not found. Must be one of: [%s]' % (
                        name, "".join('"%s", ' % c.get("name") for c in surfaces_node)))
            kind = _SURF[node.tag]
            if kind == abi.MMC_SPHERE:
                c = node.find("center")
                pr.add_surface(name, kind, float(c.get("x")), float(c.get("y")),
                               float(c.get("z")), float(node.find("radius").get("r")))
            elif kind == abi.MMC_CYLINDERX:
                pr.add_surface(name, kind, r=float(node.get("r")))
            else:
                axis = {"planex": "x", "planey": "y", "planez": "z"}[node.tag]
                pr.add_surface(name, kind, a=float(node.get(axis)))
    materials_node = root.find("materials")

    def find_material(name):
        for m in materials_node:
            if m.get("name") == name:
                return m
        raise RuntimeError('Material node "%s" not found. Must be one of: [%s]' % (
            name, "".join('"%s", ' % m.get("name") for m in materials_node)))

    nuc_names = []
    for cell in cells_node:
        if cell.tag == "void":
            continue
        for nn in find_material(cell.get("material")):
            name = nn.get("name")
            if name in nuc_names:
                continue
            node = None
            for cand in etype:
                if cand.get("name") == name:
                    node = cand
            if node is None:
                raise RuntimeError(
                    'Nuclide node "%s" not found. Must be one of: [%s]' % (
                        name, "".join('"%s", ' % c.get("name") for c in etype)))
            nuc_names.append(name)
            pr.nuclides.append(_nuclide_from_xml(node, multigroup, G, base, pr))
    mat_names = []
    for cell in cells_node:
        if cell.tag == "void":
            continue
        name = cell.get("material")
        if name in mat_names:
            continue
        mat_names.append(name)
        node = find_material(name)
        pr.add_material(name, float(node.get("aden")),
                        [(nuc_names.index(n.get("name")), float(n.get("afrac")))
                         for n in node])
    for cell in cells_node:
        faces = [(pr.surface_index(s.get("name")), s.get("sense") == "-1")
                 for s in cell]
        if cell.tag == "void":
            pr.add_cell("", -1, faces, 0)
            continue
        field = 0
        if cell.get("temperature") is not None:
            field = pr.add_field(c=float(cell.get("temperature")))
        pr.add_cell(cell.get("name"), mat_names.index(cell.get("material")),
                    faces, field)
    pt = list(root.find("problemtype"))[0]
    if pt.tag == "fixedsource":
        pr.problem = "fixedsource"
        pr.source = _source_from_xml(pt, multigroup)
    else:
        pr.problem = "keigenvalue"
        pr.inactive = int(pt.get("inactive"))
        pr.active = int(pt.get("active"))
        pr.source = _source_from_xml(pt.find("initialsource"), multigroup)
    # perturbations (PerturbationSet, Perturbation.cpp:87-118), before the
    # estimators that name them
    pert_names = []
    pert_root = root.find("perturbations")
    if pert_root is not None:
        for pnode in pert_root:
            if pnode.tag != "total":
                raise RuntimeError("unknown perturbation " + pnode.tag)
            nname = pnode.get("nuclide")
            if nname not in nuc_names:  # World::FindNuclideByName (World.cpp:63-81)
                raise RuntimeError('Nuclide "%s" not found. Must be one of: [%s]'
                                   % (nname, "".join('"%s", ' % n for n in nuc_names)))
            pr.perturbations.append(dict(name=pnode.get("name"),
                                         nuclide=nuc_names.index(nname)))
            pert_names.append(pnode.get("name"))
    if len(pr.perturbations) > abi.MMC_MAX_PERTURBATIONS:
        raise RuntimeError("at most %d perturbations" % abi.MMC_MAX_PERTURBATIONS)
    est_root = root.find("estimators")
    if est_root is not None:
        for e in est_root:
            if e.tag == "current":
                bins = e.find("bins")
                est = dict(kind=abi.MMC_EST_CURRENT, name=e.get("name"),
                           surface=pr.surface_index(e.get("surface")))
                if bins is not None:
                    est["energy"] = _bins_from_xml(bins.find("energy"))
                    cos = bins.find("cosine")
                    if cos is not None:
                        est["cosine"] = dict(
                            direction=(float(cos.get("u")), float(cos.get("v")),
                                       float(cos.get("w"))),
                            bins=_bins_from_xml(cos))
                pr.estimators.append(est)
                # sensitivities of this current (Estimator.cpp:83-90); each is
                # named "<estimator>::<perturbation>" (Scorable.cpp:22-28)
                sens = e.find("sensitivities")
                if sens is not None:
                    parent = len(pr.estimators) - 1
                    for sp in sens:
                        pname = sp.get("name")
                        if pname not in pert_names:  # Perturbation.cpp:99-118
                            raise RuntimeError(
                                'Perturbation "%s" not found. Must be one of: [%s]'
                                % (pname, "".join('"%s", ' % n for n in pert_names)))
                        pr.estimators.append(dict(
                            est, name=est["name"] + "::" + pname,
                            sensitivity=pert_names.index(pname) + 1, parent=parent))
            elif e.tag == "flux":
                m = e.find("mesh")
                pr.estimators.append(dict(
                    kind=abi.MMC_EST_FLUX_MESH, name=e.get("name"),
                    mesh=dict(n=[int(m.get(k)) for k in ("nx", "ny", "nz")],
                              lo=[float(m.get(k)) for k in ("xmin", "ymin", "zmin")],
                              hi=[float(m.get(k)) for k in ("xmax", "ymax", "zmax")])))
            elif e.tag in ("alpha", "beta"):
                pr.estimators.append(dict(
                    kind=abi.MMC_EST_TSL_ALPHA if e.tag == "alpha" else abi.MMC_EST_TSL_BETA,
                    name=e.get("name"), energy=_bins_from_xml(e)))
    return pr


def _nuclide_from_xml(node, multigroup, G, base, pr):
    path = "/minimc/nuclides/%s/nuclide" % ("multigroup" if multigroup else "continuous")
    neutron = node.find("neutron")
    if neutron is None:
        raise RuntimeError('%s: "neutron" node not found' % path)
    npath = path + "/neutron"
    if multigroup:
        n = {}
        # Multigroup::CreateReactions (Multigroup.cpp:212-229)
        for r in neutron:
            if r.tag == "capture":
                n["capture"] = _floats(r, npath + "/capture", G)
            elif r.tag == "scatter":
                n["scatter"] = _floats(r, npath + "/scatter", G * G)
            elif r.tag == "fission":
                n["fission_xs"] = _floats(r.find("xs"), npath + "/fission/xs", G)
                n["nubar"] = _floats(r.find("nubar"), npath + "/fission/nubar", G)
                n["chi"] = _floats(r.find("chi"), npath + "/fission/chi", G * G)
        return n
    n = {"awr": float(node.get("awr")), "reactions": []}
    for r in neutron:
        if r.tag == "total":
            n["total"] = _load_table(r.get("file"), base)
            n["total_temperature"] = float(r.get("temperature"))
            continue
        kind = {"capture": abi.MMC_RX_CAPTURE, "scatter": abi.MMC_RX_SCATTER,
                "fission": abi.MMC_RX_FISSION}[r.tag]
        ev = r if r.tag == "capture" else r.find("xs")
        rx = dict(kind=kind, xs=_load_table(ev.get("file"), base),
                  temperature=float(ev.get("temperature")))
        if r.tag == "fission":
            rx["nubar"] = _load_table(r.find("nubar").get("file"), base)
            watt = r.find("watt")
            if watt is not None:
                rx["chi_kind"] = abi.MMC_CHI_WATT
                rx["watt_a"], rx["watt_b"] = float(watt.get("a")), float(watt.get("b"))
        if r.tag == "scatter" and r.find("tsl") is not None:
            pr.tsl.append(_tsl_from_xml(r.find("tsl"), base))
            rx["tsl"] = len(pr.tsl) - 1
        n["reactions"].append(rx)
    return n


def _tsl_from_xml(t, base):
    """The thermal-scattering law of a <tsl> element (minimc_basictypes.xsd:
    96-133; ThermalScattering::ThermalScattering, ThermalScattering.cpp:24-99):
    pandas-HDF5 files for the majorant (E), total_E (E, order), total_T
    (T, order) and total_S (order), and per partition CDF (F, order), S
    (order) and E_T (E, T, order) for beta or beta_T (beta, T, order) for
    alpha.  A `file` attribute instead names one .npz holding all of them
    (tsl_data.save_tsl_npz)."""
    if t.get("file") is not None:
        from .tsl_data import load_tsl_npz
        return load_tsl_npz(_resolve(t.get("file"), base))
    from . import pandas_h5

    def rd(node, attr, D):
        v = node.get(attr)
        if v is None:
            raise RuntimeError('tsl: attribute "%s" missing' % attr)
        return pandas_h5.read(_resolve(v, base), D)

    (mE,), mv = rd(t, "majorant", 1)
    (E_axis, _), tot_E = rd(t, "total_E", 2)
    (T_axis, _), tot_T = rd(t, "total_T", 2)
    _, tot_S = rd(t, "total_S", 1)
    rank = len(tot_S)
    if tot_E.shape[1] != rank or tot_T.shape[1] != rank:
        raise RuntimeError("tsl: total_E, total_T and total_S disagree on the rank")

    def parts(tag, x_attr):
        node = t.find(tag)
        if node is None or len(node) == 0:
            raise RuntimeError('tsl: "%s" needs at least one partition' % tag)
        out = []
        for p in node:
            (F, _), cdf = rd(p, "CDF", 2)
            _, S = rd(p, "S", 1)
            (x, T, _), xt = rd(p, x_attr, 3)
            if cdf.shape[1] != len(S) or xt.shape[2] != len(S):
                raise RuntimeError("tsl %s: CDF, S and %s disagree on the rank"
                                   % (tag, x_attr))
            out.append(dict(F=F, cdf_modes=np.ascontiguousarray(cdf), S=S, x=x, T=T,
                            x_T_modes=np.ascontiguousarray(xt)))
        return out

    return dict(majorant=(mE, mv), total_E_axis=E_axis, total_E=np.ascontiguousarray(tot_E),
                total_T_axis=T_axis, total_T=np.ascontiguousarray(tot_T), total_S=tot_S,
                beta_partitions=parts("beta_partitions", "E_T"),
                alpha_partitions=parts("alpha_partitions", "beta_T"),
                beta_cutoff=float(t.get("beta_cutoff")),
                alpha_cutoff=float(t.get("alpha_cutoff")))
