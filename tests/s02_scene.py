"""The reference's Cornell reference-image scene (tests/golden/S02_CornellDiffuse_64x64.json, the
fixture Tests/Renderer/CornellDiffuseReferenceTest.cpp renders) as a bf_scene_desc through the
package's Scene builder, with its camera and transport options: five quads with their winding's
normal, Lambert reflectance and one-sided emission resolved at the fixed four-wavelength packet
(point-sampled spectra), a pinhole camera, diffuse depth 3 and no other scattering, no roulette."""
import json
import math
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
SCENE = GOLDEN / "S02_CornellDiffuse_64x64.json"
IMAGE = GOLDEN / "S02_CornellDiffuse_64x64_scalar_ref_4096spp.exr"


def s02(bf):
    d = json.loads(SCENE.read_text())
    abi = bf.abi
    packet = [float(np.float32(w)) for w in d["sampling"]["wavelength_packet"]["nanometers"]]
    s = bf.scenes.Scene()
    for k, surface in enumerate(d["surfaces"]):
        v = np.asarray(surface["vertices"], dtype=np.float64)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n /= np.linalg.norm(n)
        g = s.add_geometry(k + 1, v, np.tile(n, (len(v), 1)), np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float64),
                           np.asarray(surface["triangulation"]))
        radiance = surface["emission"]["radiance"]
        m = s.add_material(100 + k, [bf.scenes.lambert(bf.scenes.SpectrumId(
            s.sampled_spectrum(packet, surface["reflection"]["reflectance"])))],
            emission=bf.scenes.SpectrumId(s.sampled_spectrum(packet, radiance)) if any(radiance) else None)
        s.add_instance(200 + k, 200 + k, g, m, mask=surface["visibility_mask"])
    s.finalize()
    c = d["camera"]
    origin = np.asarray(c["origin"], dtype=np.float64)
    cam = bf.scenes.look_at_camera(origin, origin + np.asarray(c["forward"]), d["film"]["width"], d["film"]["height"],
                                   math.radians(c["vertical_field_of_view_degrees"]), time=c["time"])
    limits = d["integrator"]["depth_limits"]
    options = bf.scenes.transport_options(max_depth=limits["diffuse"], mis=abi.MIS_BALANCE,
                                          wavelength_mode=abi.WAVELENGTHS_FRAME_PACKET, rr_first_depth=0, packet=packet)
    options.glossy_depth_limit = limits["glossy"]
    options.specular_depth_limit = limits["specular"]
    options.transmission_depth_limit = limits["transmission"]
    return s, cam, options, d
