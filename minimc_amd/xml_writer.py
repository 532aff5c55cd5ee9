"""Writes a Problem as a minimc XML input with its tables in pandas-HDF5
files: the input format of the reference (test/minimc.xsd,
test/minimc_basictypes.xsd; tables as HDF5DataSet<D> reads them,
src/HDF5DataSet.hpp:18-167), plus the additive extensions world.load_xml
reads (planey/planez, a <watt> fission spectrum, a source <box>, flux and
alpha/beta estimators).

write_xml(pr, path) puts the tables in a directory next to the XML file
("<stem>_data/") and names them by relative path, so that load_xml(path)
rebuilds the same World.  Nuclides are named "nuclide-<i>"; surfaces,
materials and cells keep their names (nameless material cells get
"cell-<i>").
"""
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import abi, pandas_h5


def _fmt(v):
    return repr(float(v))


def _table(data_dir, rel_dir, name, t):
    x, y = np.asarray(t[0], dtype=np.float64), np.asarray(t[1], dtype=np.float64)
    pandas_h5.write(os.path.join(data_dir, name), [x], y)
    return os.path.join(rel_dir, name)


def write_tsl(t, data_dir, rel_dir, prefix):
    """The thermal-scattering law t (the mmc_tsl layout of Problem.tsl) as
    the reference's <tsl> element and its pandas-HDF5 files."""
    def put(name, axes, values):
        pandas_h5.write(os.path.join(data_dir, prefix + name), axes, values)
        return os.path.join(rel_dir, prefix + name)

    rank = len(t["total_S"])
    order = np.arange(rank, dtype=np.float64)
    el = ET.Element("tsl")
    el.set("majorant", put("majorant.hdf5", [t["majorant"][0]], t["majorant"][1]))
    el.set("total_T", put("total_T.hdf5", [t["total_T_axis"], order], t["total_T"]))
    el.set("total_S", put("total_S.hdf5", [order], t["total_S"]))
    el.set("total_E", put("total_E.hdf5", [t["total_E_axis"], order], t["total_E"]))
    el.set("beta_cutoff", _fmt(t["beta_cutoff"]))
    el.set("alpha_cutoff", _fmt(t["alpha_cutoff"]))
    for tag, x_attr, key in (("beta_partitions", "E_T", "beta"),
                             ("alpha_partitions", "beta_T", "alpha")):
        node = ET.SubElement(el, tag)
        for i, p in enumerate(t[tag]):
            r = np.arange(len(p["S"]), dtype=np.float64)
            q = ET.SubElement(node, "partition")
            q.set("CDF", put("%s_%d_CDF.hdf5" % (key, i), [p["F"], r], p["cdf_modes"]))
            q.set("S", put("%s_%d_S.hdf5" % (key, i), [r], p["S"]))
            q.set(x_attr, put("%s_%d_%s.hdf5" % (key, i, x_attr), [p["x"], p["T"], r],
                              p["x_T_modes"]))
    return el


def _bins(parent, tag, b):
    node = ET.SubElement(parent, tag)
    kind = b["kind"]
    if kind == abi.MMC_BINS_LINSPACE:
        ET.SubElement(node, "linspace", min=_fmt(b["min"]), max=_fmt(b["max"]),
                      bins=str(b["bins"]))
    elif kind == abi.MMC_BINS_LOGSPACE:
        e = ET.SubElement(node, "logspace", min=_fmt(b["min"]), max=_fmt(b["max"]),
                          bins=str(b["bins"]))
        if "base" in b:
            e.set("base", _fmt(b["base"]))
    elif kind == abi.MMC_BINS_BOUNDARIES:
        ET.SubElement(node, "boundaries").text = " ".join(_fmt(v) for v in b["bounds"])
    return node


def _source(parent, s, multigroup):
    pos = ET.SubElement(parent, "position")
    x, y, z = s.get("position", (0.0, 0.0, 0.0))
    ET.SubElement(pos, "constant", x=_fmt(x), y=_fmt(y), z=_fmt(z))
    if s.get("box"):
        lo, hi = s["box"]
        ET.SubElement(pos, "box", xmin=_fmt(lo[0]), ymin=_fmt(lo[1]), zmin=_fmt(lo[2]),
                      xmax=_fmt(hi[0]), ymax=_fmt(hi[1]), zmax=_fmt(hi[2]))
    d = ET.SubElement(parent, "direction")
    kind = s.get("dir_kind", abi.MMC_DIR_ISOTROPIC)
    if kind == abi.MMC_DIR_ISOTROPIC:
        ET.SubElement(d, "isotropic")
    else:
        u, v, w = s.get("direction", (1.0, 0.0, 0.0))
        ET.SubElement(d, "constant" if kind == abi.MMC_DIR_CONSTANT else "isotropic-flux",
                      x=_fmt(u), y=_fmt(v), z=_fmt(w))
    e = ET.SubElement(parent, "energy")
    if s.get("energy_kind", abi.MMC_E_CONSTANT) == abi.MMC_E_UNIFORM:
        ET.SubElement(e, "uniform", min=_fmt(s["E"]), max=_fmt(s["E_hi"]))
    else:
        ET.SubElement(e, "constant",
                      energy=str(int(s["E"])) if multigroup else _fmt(s["E"]))
    pt = ET.SubElement(parent, "particletype")
    ET.SubElement(pt, "constant", type="neutron")


def write_xml(pr, path, data_dir=None, threads=1):
    """Writes pr to `path`; returns the path.  Tables go to data_dir
    (default "<stem>_data" beside the XML file)."""
    base = os.path.dirname(os.path.abspath(path))
    stem = os.path.splitext(os.path.basename(path))[0]
    data_dir = data_dir or os.path.join(base, stem + "_data")
    os.makedirs(data_dir, exist_ok=True)
    rel_dir = os.path.relpath(data_dir, base)
    mg = pr.energy_mode == abi.MMC_MULTIGROUP
    root = ET.Element("minimc")
    gen = ET.SubElement(root, "general")
    ET.SubElement(gen, "particles").text = "neutron"
    ET.SubElement(gen, "histories").text = str(int(pr.histories))
    ET.SubElement(gen, "threads").text = str(threads)
    ET.SubElement(gen, "seed").text = str(int(pr.seed))
    ET.SubElement(gen, "tracking").text = (
        "cell delta" if pr.tracking == abi.MMC_CELL_DELTA_TRACKING else "surface")
    # nuclides
    nucs = ET.SubElement(root, "nuclides")
    group = ET.SubElement(nucs, "multigroup" if mg else "continuous")
    if mg:
        group.set("groups", str(pr.groups))
    names = ["nuclide-%d" % i for i in range(len(pr.nuclides))]
    for i, n in enumerate(pr.nuclides):
        node = ET.SubElement(group, "nuclide", name=names[i])
        neu = ET.SubElement(node, "neutron")
        if mg:
            txt = lambda a: " ".join(_fmt(v) for v in np.asarray(a).reshape(-1))
            if n.get("capture") is not None:
                ET.SubElement(neu, "capture").text = txt(n["capture"])
            if n.get("scatter") is not None:
                ET.SubElement(neu, "scatter").text = txt(n["scatter"])
            if n.get("fission_xs") is not None:
                f = ET.SubElement(neu, "fission")
                ET.SubElement(f, "xs").text = txt(n["fission_xs"])
                ET.SubElement(f, "nubar").text = txt(n["nubar"])
                ET.SubElement(f, "chi").text = txt(n["chi"])
            continue
        node.set("awr", _fmt(n["awr"]))
        if n.get("total") is not None:
            ET.SubElement(neu, "total", file=_table(data_dir, rel_dir, "%s-total.hdf5" %
                                                    names[i], n["total"]),
                          temperature=_fmt(n.get("total_temperature", 293.6)))
        for k, r in enumerate(n["reactions"]):
            T = _fmt(r.get("temperature", 293.6))
            f = lambda what: _table(data_dir, rel_dir, "%s-%d-%s.hdf5" % (names[i], k, what),
                                    r[what])
            if r["kind"] == abi.MMC_RX_CAPTURE:
                ET.SubElement(neu, "capture", file=f("xs"), temperature=T)
            elif r["kind"] == abi.MMC_RX_SCATTER:
                sc = ET.SubElement(neu, "scatter")
                ET.SubElement(sc, "xs", file=f("xs"), temperature=T)
                if r.get("tsl", -1) >= 0:
                    sc.append(write_tsl(pr.tsl[r["tsl"]], data_dir, rel_dir,
                                        "%s-tsl%d-" % (names[i], r["tsl"])))
            else:
                fi = ET.SubElement(neu, "fission")
                ET.SubElement(fi, "xs", file=f("xs"), temperature=T)
                ET.SubElement(fi, "nubar", file=f("nubar"), temperature=T)
                if r.get("chi_kind", abi.MMC_CHI_INHERIT) == abi.MMC_CHI_WATT:
                    ET.SubElement(fi, "watt", a=_fmt(r["watt_a"]), b=_fmt(r["watt_b"]))
    # materials
    mats = ET.SubElement(root, "materials")
    for m in pr.materials:
        node = ET.SubElement(mats, "material", name=m["name"], aden=_fmt(m["aden"]))
        for nuc, afrac in m["comps"]:
            ET.SubElement(node, "nuclide", name=names[nuc], afrac=_fmt(afrac))
    # surfaces
    surfs = ET.SubElement(root, "surfaces")
    for s in pr.surfaces:
        k = s["kind"]
        if k == abi.MMC_SPHERE:
            node = ET.SubElement(surfs, "sphere", name=s["name"])
            ET.SubElement(node, "center", x=_fmt(s["a"]), y=_fmt(s["b"]), z=_fmt(s["c"]))
            ET.SubElement(node, "radius", r=_fmt(s["r"]))
        elif k == abi.MMC_CYLINDERX:
            ET.SubElement(surfs, "cylinderx", name=s["name"], r=_fmt(s["r"]))
        else:
            tag, ax = {abi.MMC_PLANEX: ("planex", "x"), abi.MMC_PLANEY: ("planey", "y"),
                       abi.MMC_PLANEZ: ("planez", "z")}[k]
            ET.SubElement(surfs, tag, {"name": s["name"], ax: _fmt(s["a"])})
    # cells; a cell's own constant temperature is its temperature attribute,
    # field 0 is the global <temperature>
    cells = ET.SubElement(root, "cells")
    for i, c in enumerate(pr.cells):
        if c["material"] < 0:
            node = ET.SubElement(cells, "void")
        else:
            node = ET.SubElement(cells, "cell", name=c["name"] or "cell-%d" % i,
                                 material=pr.materials[c["material"]]["name"])
            if c["field"] != 0:
                f = pr.fields[c["field"]]
                if f["kind"] != abi.MMC_FIELD_CONSTANT:
                    raise ValueError("a cell's own temperature must be constant")
                node.set("temperature", _fmt(f["c"]))
        for s, inside in c["faces"]:
            ET.SubElement(node, "surface", name=pr.surfaces[s]["name"],
                          sense="-1" if inside else "+1")
    # problem type
    pt = ET.SubElement(root, "problemtype")
    if pr.problem == "keigenvalue":
        k = ET.SubElement(pt, "keigenvalue", inactive=str(pr.inactive), active=str(pr.active))
        _source(ET.SubElement(k, "initialsource"), pr.source, mg)
    else:
        _source(ET.SubElement(pt, "fixedsource"), pr.source, mg)
    f0 = pr.fields[0]
    temp = ET.SubElement(root, "temperature")
    if f0["kind"] == abi.MMC_FIELD_CONSTANT:
        ET.SubElement(temp, "constant", c=_fmt(f0["c"]))
    else:
        lin = ET.SubElement(temp, "linear")
        ET.SubElement(lin, "bounds", lower=_fmt(f0["lower"]), upper=_fmt(f0["upper"]))
        ET.SubElement(lin, "intercept", b=_fmt(f0["b"]))
        ET.SubElement(lin, "gradient", x=_fmt(f0["g"][0]), y=_fmt(f0["g"][1]),
                      z=_fmt(f0["g"][2]))
    if pr.perturbations:
        pn = ET.SubElement(root, "perturbations")
        for p in pr.perturbations:
            ET.SubElement(pn, "total", name=p["name"], nuclide=names[p["nuclide"]])
    if pr.estimators:
        ests = ET.SubElement(root, "estimators")
        last = {}
        for i, e in enumerate(pr.estimators):
            kind = e["kind"]
            if kind == abi.MMC_EST_CURRENT and e.get("sensitivity"):
                sens = last[e["parent"]].find("sensitivities")
                if sens is None:
                    sens = ET.SubElement(last[e["parent"]], "sensitivities")
                ET.SubElement(sens, "perturbation",
                              name=pr.perturbations[e["sensitivity"] - 1]["name"])
                continue
            if kind == abi.MMC_EST_CURRENT:
                node = ET.SubElement(ests, "current", name=e["name"],
                                     surface=pr.surfaces[e["surface"]]["name"])
                if e.get("energy") or e.get("cosine"):
                    bins = ET.SubElement(node, "bins")
                    if e.get("energy"):
                        _bins(bins, "energy", e["energy"])
                    if e.get("cosine"):
                        cos = _bins(bins, "cosine", e["cosine"]["bins"])
                        u, v, w = e["cosine"]["direction"]
                        cos.set("u", _fmt(u))
                        cos.set("v", _fmt(v))
                        cos.set("w", _fmt(w))
                last[i] = node
            elif kind == abi.MMC_EST_FLUX_MESH:
                node = ET.SubElement(ests, "flux", name=e["name"])
                m = e["mesh"]
                ET.SubElement(node, "mesh", nx=str(m["n"][0]), ny=str(m["n"][1]),
                              nz=str(m["n"][2]), xmin=_fmt(m["lo"][0]),
                              ymin=_fmt(m["lo"][1]), zmin=_fmt(m["lo"][2]),
                              xmax=_fmt(m["hi"][0]), ymax=_fmt(m["hi"][1]),
                              zmax=_fmt(m["hi"][2]))
            else:
                node = ET.SubElement(ests, "alpha" if kind == abi.MMC_EST_TSL_ALPHA
                                     else "beta", name=e["name"])
                for child in list(_bins(node, "x", e["energy"])):
                    node.append(child)
                node.remove(node.find("x"))
    ET.indent(root)
    ET.ElementTree(root).write(path, encoding="unicode", xml_declaration=False)
    return path
