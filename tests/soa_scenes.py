"""Scene descriptions in the reference's FrameScene form (for the oracle's blob serializer), and
the same scenes built as bf_scene_desc through the package's Scene builder, for the blob tests."""
import importlib.util
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent


def _load_serializer():
    name = "oracle_scene_soa"
    if name not in sys.modules:
        spec = importlib.util.spec_from_file_location(name, ROOT / "oracle" / "scene_soa.py")
        module = importlib.util.module_from_spec(spec)
        sys.modules[name] = module
        spec.loader.exec_module(module)
    return sys.modules[name]


soa = _load_serializer()
F32 = np.float32


def matrix16(linear, translation):
    """Row-major 4x4 float32 affine matrix."""
    m = np.eye(4)
    m[:3, :3] = linear
    m[:3, 3] = translation
    return [F32(v) for v in m.reshape(-1)]


def cornell_description(bf, lane_constant=False, u=0.375):
    """The Cornell box of the package's scene generator (same walls, boxes, light and camera), with
    every spectrum sampled at the uniform visible packet of `u`. lane_constant=False gives each
    lane its own value, a rough-conductor mixture on the back wall, a parent chain (the two boxes
    are children of an invisible, translated group instance) and an environment; lane_constant=True keeps every
    spectrum flat so the scene equals the constant-spectrum Cornell box."""
    sc = bf.scenes
    wl = soa.uniform_visible_wavelengths(u)
    lanes = (lambda v: [v] * 4) if lane_constant else (lambda v: [v * (0.85 + 0.1 * k) for k in range(4)])

    def material(mid, closures, probabilities, emission=None, frame_mode=0, rotation=0.0):
        return {"id": mid, "spectral": {"wavelengths": wl, "closures": closures, "probabilities": probabilities,
                                        "frame_mode": frame_mode, "tangent_rotation": rotation,
                                        "emission": emission if emission is not None else [0.0] * 4}}

    materials = [
        material(21, [soa.lambertian(lanes(0.73))], [1.0]),
        material(22, [soa.lambertian(lanes(0.12))], [1.0]),
        material(23, [soa.lambertian(lanes(0.45))], [1.0]),
        material(24, [soa.lambertian([0.0] * 4)], [1.0], emission=lanes(17.0)),
    ]
    back = 21
    if not lane_constant:
        materials.append(material(28, [soa.lambertian(lanes(0.5)),
                                       soa.rough_conductor(lanes(0.8), [0.27, 0.35, 0.62, 1.1], [3.6, 2.9, 2.6, 2.2], 0.15, 0.4)],
                                  [0.7, 0.3], frame_mode=1, rotation=0.5))
        back = 28
    geometries, instances = [], []

    def mesh(gid, positions, normals, uvs, indices):
        geometries.append({"id": gid, "positions": np.asarray(positions, np.float32), "normals": np.asarray(normals, np.float32),
                           "uvs": np.asarray(uvs, np.float32), "triangles": np.asarray(indices, np.uint32)})

    walls = [
        (11, sc.quad([[-1, -1, 0], [-1, -1, 2], [1, -1, 2], [1, -1, 0]], [0, 1, 0]), 21),
        (12, sc.quad([[-1, 1, 0], [1, 1, 0], [1, 1, 2], [-1, 1, 2]], [0, -1, 0]), 21),
        (13, sc.quad([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [0, 0, 1]), back),
        (14, sc.quad([[-1, -1, 0], [-1, 1, 0], [-1, 1, 2], [-1, -1, 2]], [1, 0, 0]), 22),
        (15, sc.quad([[1, -1, 0], [1, -1, 2], [1, 1, 2], [1, 1, 0]], [-1, 0, 0]), 23),
        (16, sc.quad([[-0.32, 0.985, 0.62], [0.32, 0.985, 0.62], [0.32, 0.985, 1.18], [-0.32, 0.985, 1.18]],
                     [0, -1, 0]), 24),
    ]
    identity = matrix16(np.eye(3), (0.0, 0.0, 0.0))
    for k, (gid, shape, mid) in enumerate(walls):
        mesh(gid, *shape)
        instances.append({"id": 31 + k, "parent": None, "object": 1 + k, "geometry": gid, "material": mid,
                          "local_to_parent": identity, "visibility_mask": 0xFFFFFFFF})
    mesh(17, *sc.unit_box())
    short = matrix16(sc.rotation_y(math.radians(15.0)) @ np.diag([0.55, 0.55, 0.55]), (0.36, -0.725, 1.30))
    tall = matrix16(sc.rotation_y(math.radians(-18.0)) @ np.diag([0.55, 1.15, 0.55]), (-0.38, -0.425, 0.72))
    parent = None
    if not lane_constant:
        # an invisible group instance (visibility mask 0) that parents and shifts both boxes
        instances.append({"id": 29, "parent": None, "object": 9, "geometry": 17, "material": 21,
                          "local_to_parent": matrix16(np.eye(3), (0.05, 0.0, -0.02)), "visibility_mask": 0x0})
        parent = 29
    objects = sorted({i["object"] for i in instances} | {7, 8})
    instances.append({"id": 37, "parent": parent, "object": 7, "geometry": 17, "material": 21,
                      "local_to_parent": short, "visibility_mask": 0xFFFFFFFF})
    instances.append({"id": 38, "parent": parent, "object": 8, "geometry": 17, "material": 21,
                      "local_to_parent": tall, "visibility_mask": 0xFFFFFFFF})
    d = {"objects": [{"id": o} for o in objects], "geometries": geometries, "materials": materials,
         "instances": instances, "punctual_lights": []}
    if not lane_constant:
        d["environment"] = {"wavelengths": wl, "radiance": [0.02, 0.03, 0.05, 0.04]}
    return d, wl


def scene_from_canonical(bf, canon, constant_spectra=False):
    """The canonical scene (soa.frame_scene output) as a bf_scene_desc through the package's Scene
    builder: point-sampled spectra at the packet (or, with constant_spectra, one-value spectra for
    flat lanes), world matrices as serialized."""
    abi = bf.abi
    s = bf.scenes.Scene()
    sampled = s.sampled_spectrum

    def spectrum(packet, values):
        if constant_spectra and all(F32(v) == F32(values[0]) for v in values):
            return s.spectrum(float(F32(values[0])))
        return sampled(packet, values)

    packet = None
    gindex = {}
    for g in canon["geometries"]:
        gindex[g["id"]] = s.add_geometry(g["id"], np.asarray(g["positions"], np.float64), np.asarray(g["normals"], np.float64),
                                         np.asarray(g["uvs"], np.float64), np.asarray(g["triangles"]))
    mindex = {}
    for m in canon["materials"]:
        sp = m["spectral"]
        packet = [w[0] for w in sp["wavelengths"]]
        closures = []
        for c, p in zip(sp["closures"], sp["probabilities"]):
            rec = {"kind": c["kind"], "lobes": c["lobes"], "weight": bf.scenes.SpectrumId(spectrum(packet, c["weight"])),
                   "params": [float(v) for v in c["parameters"]], "probability": float(F32(p))}
            if c["kind"] == abi.CLOSURE_ROUGH_CONDUCTOR:
                rec["eta"] = bf.scenes.SpectrumId(spectrum(packet, c["parameters"][:4]))
                rec["k"] = bf.scenes.SpectrumId(spectrum(packet, c["parameters"][4:8]))
            closures.append(rec)
        emission = sp["emission"]
        mindex[m["id"]] = s.add_material(
            m["id"], closures,
            emission=bf.scenes.SpectrumId(spectrum(packet, emission)) if any(F32(v) != 0 for v in emission) else None,
            frame_mode=sp["frame_mode"], tangent_rotation=float(F32(sp["tangent_rotation"])))
    for inst, (l2w, w2l) in zip(canon["instances"], canon["world_transforms"]):
        rec = abi.InstanceDesc()
        rec.id, rec.object_id = inst["id"], inst["object"]
        rec.geometry_index, rec.material_index = gindex[inst["geometry"]], mindex[inst["material"]]
        rec.visibility_mask = inst["visibility_mask"]
        for e in range(12):
            rec.local_to_world[e] = float(l2w[e])
            rec.world_to_local[e] = float(w2l[e])
        s.instances.append(rec)
    kinds = {"point": abi.PUNCTUAL_POINT, "directional": abi.PUNCTUAL_DIRECTIONAL, "spot": abi.PUNCTUAL_SPOT}
    for light in canon["punctual_lights"]:
        rec = abi.PunctualDesc()
        rec.kind = kinds[light["kind"]]
        rec.spectrum = int(spectrum(packet, light["spectrum"]))
        for a in range(3):
            if light["kind"] != "directional":
                rec.position[a] = light["position"][a]
                rec.position_error[a] = light["position_error"][a]
            if light["kind"] != "point":
                rec.direction[a] = light["direction"][a]
        if light["kind"] == "spot":
            rec.inner_half_angle, rec.outer_half_angle = light["inner"], light["outer"]
        s.punctuals.append(rec)
    if canon["environment"] is not None:
        s.environment_spectrum = int(spectrum(packet, canon["environment"]["radiance"]))
    # normal / bump maps over the image textures, which keep their canonical (id) order
    image_index = {}
    for t in canon["image_textures"]:
        levels = [np.asarray(l["pixels"], np.float32).reshape(l["height"], l["width"], l["channels"]) for l in t["levels"]]
        image_index[t["id"]] = s.add_image_texture(levels)
    for m in canon["materials"]:
        sp = m["spectral"]
        nm, bm = sp.get("normal_map"), sp.get("bump_map")
        if nm is None and bm is None:
            continue
        normal = None if nm is None else {
            "texture": image_index[nm["texture"]], "channels": (nm["red_channel"], nm["green_channel"], nm["blue_channel"]),
            "u_wrap": nm["u_wrap"], "v_wrap": nm["v_wrap"], "max_anisotropy": nm["maximum_anisotropy"],
            "max_texel_visits": nm["maximum_texel_visits"], "y_convention": nm["y_convention"]}
        bump = None if bm is None else {
            "texture": image_index[bm["texture"]], "channel": bm["channel"], "scale": float(F32(bm["scale"])),
            "u_wrap": bm["u_wrap"], "v_wrap": bm["v_wrap"], "max_anisotropy": bm["maximum_anisotropy"],
            "max_texel_visits": bm["maximum_texel_visits"]}
        s.set_surface_maps(mindex[m["id"]], normal=normal, bump=bump)
    return s.finalize()
