"""Scene builders used by the tests (independent of the product's importers).

* load_fixture(model, config): one of the reference's scene fixtures copied
  under tests/golden/models/.  Follows the reference's import rules:
  config XML (src/IO/ReadXML.cpp:109-168, 598-622, 848-871, 950-1010),
  OBJ faces in file order with (inner, outer) ids from `o name` / `g N` /
  `g N_name` lines (src/IO/ImportExport.cpp:952-1053), cell normals from
  the face winding, and the emitter model of src/Model/Emitter.cpp:186-339.
* slab(layers, ...): a planar layered slab in the reference's mesh
  convention, one surface per interface (like models/model3).
"""
import gzip
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")
f32 = np.float32


def _num(el, tag, default=None):
    x = el.find(tag)
    return default if x is None else float(x.text.strip())


def parse_config(path):
    root = ET.parse(path).getroot()
    model = root.find("model")
    dims = model.find("dimensions")
    tissues = {}
    for t in model.find("tissueList").findall("tissue"):
        tid = int(t.find("ID").text)
        tissues[tid] = dict(name=t.find("name").text.strip(), g=_num(t, "anisotropy"),
                            mua=_num(t, "mua"), mus=_num(t, "mus"), n=_num(t, "refraction"))
    em = root.find("Emittermodel")
    sp = root.find("SimulationParams")
    logs = root.find("Logs")
    dg = logs.find("DetectionGrid")
    return dict(
        modelfile=model.find("modelfile").text.strip(),
        ppu=_num(model, "PixelPerUnit", 1.0),
        dims=[_num(dims, a) for a in "xyz"],
        tissues=[tissues[i] for i in range(len(tissues))],
        position=[_num(em.find("position"), a) for a in "xyz"],
        distance=_num(em, "distance"),
        heading=[_num(em.find("heading"), a) for a in "xyz"],
        nphotons=int(_num(sp, "numberOfPhotons")),
        chance=_num(sp, "chance"), w_th=_num(sp, "weightThreshold"),
        fetal_id=int(_num(logs, "fetalID", 1)),
        det_grid=[_num(dg, a) for a in ("x", "y", "z", "r", "phi", "rho")],
    )


def parse_obj(path, tissue_names):
    opener = gzip.open if path.endswith(".gz") else open
    verts, vn, faces, med, fvn = [], [], [], [], []
    inner = outer = 0
    with opener(path, "rt") as f:
        for line in f:
            line = line.rstrip("\n").rstrip("\r")
            if line.startswith("v "):
                verts.append([float(x) for x in line.split()[1:4]])
            elif line.startswith("vn "):
                vn.append([float(x) for x in line.split()[1:4]])
            elif re.match(r"^o (\w+)", line):
                name = re.match(r"^o (\w+)", line).group(1)
                if name in tissue_names:
                    inner = tissue_names.index(name)
            elif re.match(r"^f (.+)", line):
                toks = line.split()[1:]
                idx = [int(t.split("/")[0]) - 1 for t in toks]
                faces.append(idx[:3])
                parts = toks[0].split("/")
                fvn.append(int(parts[2]) - 1 if len(parts) > 2 and parts[2] else -1)
                med.append([inner, outer])
            elif re.match(r"^g (\d+)$", line):
                outer = int(re.match(r"^g (\d+)$", line).group(1))
            elif re.match(r"^g (\d+)_(\w+)", line):
                outer = int(re.match(r"^g (\d+)_(\w+)", line).group(1))
    # vertices as float32 (vtkOBJReader's points), normals from them in double
    V = np.array(verts, np.float64).astype(f32).astype(np.float64)
    F = np.array(faces, np.int64)
    P0, P1, P2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(P1 - P0, P2 - P0)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    tri = np.concatenate([P0, P1, P2], axis=1).astype(f32)
    file_vn = np.array([vn[i] if i >= 0 else [np.nan] * 3 for i in fvn], np.float64)
    return tri, n.astype(f32), np.array(med, np.int32), file_vn


def _normalize_f(a):
    a = [f32(x) for x in a]
    s = f32(math.sqrt(float(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])))
    return [f32(x / s) for x in a]


def _cross_f(a, b):
    return [f32(a[1] * b[2] - a[2] * b[1]), f32(a[2] * b[0] - a[0] * b[2]), f32(a[0] * b[1] - a[1] * b[0])]


def emitter_basis(heading):
    """CalculateEmitterModel (Emitter.cpp:299-339), float arithmetic."""
    ez = [f32(x) for x in heading]
    if ez[2] != 0:
        a = [f32(1), f32(1), f32(-(ez[0] + ez[1]) / ez[2])]
    elif ez[1] != 0:
        a = [f32(1), f32(-ez[0] / ez[1]), f32(0)]
    elif ez[0] != 0:
        a = [f32(0), f32(1), f32(0)]
    else:
        a = [f32(0)] * 3
    a = _normalize_f(a)
    ez = _normalize_f(ez)
    ey = _normalize_f(_cross_f(ez, a))
    ex = _normalize_f(_cross_f(ey, ez))
    return ex, ey, ez


def make_scene(tri, nrm, med, tissues, position, distance, heading, dims, det_grid, chance, w_th,
               fetal_id=1, ppu=1.0):
    ex, ey, ez = emitter_basis(heading)
    h = [f32(x) for x in heading]
    p0 = [f32(x) for x in position]
    d = f32(distance)
    pstart = [f32(-h[i] * d + p0[i]) for i in range(3)]  # Emitter.cpp:186-205
    return dict(
        tri=np.asarray(tri, f32), nrm=np.asarray(nrm, f32), med=np.asarray(med, np.int32),
        nmuasg=np.array([[t["n"], t["mua"], t["mus"], t["g"]] for t in tissues], f32),
        fetal_id=fetal_id, pstart=np.array(pstart, f32), vstart=np.array([h], f32),
        vocc=np.array([1], np.int32), poi_basis=np.array(p0 + ex + ey + ez, f32),
        model_dims=np.array(dims, f32), det_grid=np.array(det_grid, f32),
        factor=f32(ppu), chance=f32(chance), w_th=f32(w_th))


def load_fixture(model, config="config.xml"):
    d = os.path.join(MODELS, model)
    cfg = parse_config(os.path.join(d, config))
    objname = os.path.basename(cfg["modelfile"])
    path = os.path.join(d, objname)
    if not os.path.exists(path):
        path += ".gz"
    tri, nrm, med, file_vn = parse_obj(path, [t["name"] for t in cfg["tissues"]])
    sc = make_scene(tri, nrm, med, cfg["tissues"], cfg["position"], cfg["distance"], cfg["heading"],
                    cfg["dims"], cfg["det_grid"], cfg["chance"], cfg["w_th"], cfg["fetal_id"], cfg["ppu"])
    sc["file_vn"] = file_vn
    sc["config"] = cfg
    return sc


def _quad(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d) of a planar quad, winding kept."""
    return [a + b + c, a + c + d]


def slab(layers, half=5.0, ambient_n=1.0, distance=0.002, det_grid=(0.05, 0.05, 0.01, 0.05, 1, 1),
         chance=0.1, w_th=1e-4):
    """Planar layered slab, x, y in [-half, half], layer i between z_i and
    z_{i+1}, one surface per interface: normals point from inner (deeper)
    tissue to outer.  layers = [(thickness, n, mua, mus, g), ...]."""
    tissues = [dict(n=ambient_n, mua=0.0, mus=0.0, g=0.0)]
    for (_, n, mua, mus, g) in layers:
        tissues.append(dict(n=n, mua=mua, mus=mus, g=g))
    zs = [0.0]
    for L in layers:
        zs.append(zs[-1] + L[0])
    h = half
    tri, med = [], []
    # top surface z=0, normal -z (towards the ambient), inner = layer 1
    for t in _quad([-h, -h, 0.0], [-h, h, 0.0], [h, h, 0.0], [h, -h, 0.0]):
        tri.append(t); med.append([1, 0])
    # interfaces z_i between layer i (outer, above) and i+1 (inner, below): normal -z
    for i in range(1, len(layers)):
        z = zs[i]
        for t in _quad([-h, -h, z], [-h, h, z], [h, h, z], [h, -h, z]):
            tri.append(t); med.append([i + 1, i])
    # bottom z = z_n: normal +z, inner = last layer
    z = zs[-1]
    for t in _quad([-h, -h, z], [h, -h, z], [h, h, z], [-h, h, z]):
        tri.append(t); med.append([len(layers), 0])
    # side walls per layer, outward normals, outer = ambient
    for i in range(len(layers)):
        za, zb = zs[i], zs[i + 1]
        for t in _quad([h, -h, za], [h, h, za], [h, h, zb], [h, -h, zb]):      # +x
            tri.append(t); med.append([i + 1, 0])
        for t in _quad([-h, -h, za], [-h, -h, zb], [-h, h, zb], [-h, h, za]):  # -x
            tri.append(t); med.append([i + 1, 0])
        for t in _quad([-h, h, za], [-h, h, zb], [h, h, zb], [h, h, za]):      # +y
            tri.append(t); med.append([i + 1, 0])
        for t in _quad([-h, -h, za], [h, -h, za], [h, -h, zb], [-h, -h, zb]):  # -y
            tri.append(t); med.append([i + 1, 0])
    tri = np.array(tri, np.float64)
    P0, P1, P2 = tri[:, 0:3], tri[:, 3:6], tri[:, 6:9]
    n = np.cross(P1 - P0, P2 - P0)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    dims = [2 * h, 2 * h, zs[-1]]
    return make_scene(tri.astype(f32), n.astype(f32), np.array(med, np.int32), tissues,
                      [0.0, 0.0, 0.0], distance, [0.0, 0.0, 1.0], dims, det_grid, chance, w_th)


C0_LAYERS = [(0.15, 1.40, 0.4, 200.0, 0.9), (1.0, 1.44, 0.1, 120.0, 0.9), (2.0, 1.37, 0.3, 80.0, 0.9)]
# ten layers (ten regions: 20 (mua, mus) parameters for the adjoint), 1 cm,
# one index (under the signed-cosine Fresnel an inward crossing between
# unequal indices almost always reflects, which would keep the photons in
# the top layers)
SLAB10_LAYERS = [(0.1, 1.40, 0.05 + 0.08 * i, 15.0 + 3.0 * i, 0.80 + 0.015 * i) for i in range(10)]


def escape_scene(n_up):
    """models/model3 (open mesh of eight 10 x 10 cm faces) with two start
    directions (GetvStart's vStartOcc counts, MCLightPropagation.cl:423-467):
    the first n_up photons head +z into the layers as the fixture's emitter
    does; the rest head away from the mesh (0.6, 0, -0.8), find no face ahead
    and leave it (the reference's mediumIDs[-1] read, .cl:1475; counted as
    escaped here)."""
    sc = load_fixture("model3")
    sc["vstart"] = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, -0.8]], f32)
    sc["vocc"] = np.array([n_up, 1 << 30], np.int32)
    return sc


# oblique photon through a weakly scattering slab at mismatched n
OBLIQUE_D = 1.0
OBLIQUE_N = (1.0, 1.5, 1.2)
OBLIQUE_THETA = math.radians(40.0)
OBLIQUE_X0, OBLIQUE_Z0 = 2.0, -0.5


def oblique_slab():
    """A 20 x 20 cm slab z in [0, OBLIQUE_D] (two quads, normals +z, so both
    crossings of a photon heading down have cos_in > 0), tissue 1 weakly
    scattering (n 1.5), behind it tissue 2 (n 1.2, mua = mus = 0: the photon
    leaves there).  The source sits at (OBLIQUE_X0, 0, OBLIQUE_Z0), heading at
    OBLIQUE_THETA from +z in the x-z plane."""
    h, d = 10.0, OBLIQUE_D
    tri, med = [], []
    for z, pair in ((0.0, (0, 1)), (d, (1, 2))):
        a, b, c, e = [-h, -h, z], [h, -h, z], [h, h, z], [-h, h, z]
        tri += [a + b + c, a + c + e]
        med += [pair, pair]
    n0, n1, n2 = OBLIQUE_N
    nm = [[n0, 0, 0, 0], [n1, 1e-6, 1e-6, 0.0], [n2, 0, 0, 0]]
    v = [math.sin(OBLIQUE_THETA), 0.0, math.cos(OBLIQUE_THETA)]
    return dict(tri=np.array(tri, f32).reshape(-1, 9), nrm=np.tile(np.array([0, 0, 1], f32), (4, 1)),
                med=np.array(med, np.int32), nmuasg=np.array(nm, f32), fetal_id=-1,
                pstart=np.array([OBLIQUE_X0, 0.0, OBLIQUE_Z0], f32), vstart=np.array([v], f32),
                vocc=np.array([1], np.int32), poi_basis=np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], f32),
                model_dims=np.array([20.0, 20.0, 2 * d], f32), det_grid=np.array([0.1, 0.1, 0.1, 0.1, 1.0, 1.0], f32),
                factor=f32(1.0), chance=f32(0.1), w_th=f32(1e-4))
