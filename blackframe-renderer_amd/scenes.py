"""Seeded scene authoring for the wavefront workloads (host side, numpy).

A scene is authored in float64 and rounded once to the float32 arrays of bf_scene_desc, so every
consumer (the HIP library and the test oracle) sees identical bytes. Matrices are 3x4 row-major
affine transforms; world_to_local is the float64 inverse rounded once.

Workloads:
  cornell(...)          Cornell box at 64x64/16 spp: 5 axis-aligned walls, 2 rotated boxes
                        (36 triangles), one emissive quad with constant radiance 17, Lambert walls
                        0.73 / left 0.12 / right 0.45. Wall, emitter and camera coordinates follow
                        Tests/Backends/CPU/Embree/CornellWavefrontScene.hpp:192-413.
  icosphere_grid(...)   32x32 grid of seeded displaced icospheres (subdivision 3, 1280 triangles
                        each, 1.31 M triangles), per-object Lambert / GGX conductor / GGX dielectric,
                        64 emissive triangles and a constant environment.
  textured_cornell(...) the Cornell box with a seeded RGB albedo texture (U[0.1, 0.9]) on the back
                        wall and a GGX conductor box whose widths come from a roughness texture
                        (alpha 0.3): the differentiable workload of the adjoint.
"""

import ctypes as C
import math

import numpy as np

from . import abi

SPLITMIX_GAMMA = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the seeded generator behind every synthetic scene."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + SPLITMIX_GAMMA) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def _affine_inverse(m):
    full = np.eye(4)
    full[:3, :] = m
    return np.linalg.inv(full)[:3, :]


def rotation_y(angle_radians):
    c, s = math.cos(angle_radians), math.sin(angle_radians)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def affine(linear=None, translation=(0.0, 0.0, 0.0)):
    m = np.zeros((3, 4))
    m[:, :3] = np.eye(3) if linear is None else linear
    m[:, 3] = translation
    return m


class SpectrumId(int):
    """An already-built spectrum, accepted wherever Scene.add_material takes spectrum values."""


class Scene:
    """Owns the float32 / uint32 arrays behind one bf_scene_desc."""

    def __init__(self):
        self._positions, self._normals, self._uvs, self._indices = [], [], [], []
        self.geometries, self.instances, self.materials, self.closures = [], [], [], []
        self.spectra, self.spectrum_values = [], []
        self.textures = []
        self.punctuals = []
        self.image_textures = []  # (width, height, channels, mip count, float32 texels of every level)
        self.surface_maps = {}    # material index -> abi.SurfaceMapsDesc
        self.vertex_count = 0
        self.triangle_count = 0
        self.environment_spectrum = abi.INVALID
        self._constant_cache = {}

    # -- spectra ------------------------------------------------------------------------------
    def spectrum(self, values):
        if isinstance(values, SpectrumId):
            return int(values)
        values = [float(v) for v in np.atleast_1d(values)]
        if len(values) == 1 and values[0] in self._constant_cache:
            return self._constant_cache[values[0]]
        sid = len(self.spectra)
        self.spectra.append((len(self.spectrum_values), len(values)))
        self.spectrum_values.extend(values)
        if len(values) == 1:
            self._constant_cache[values[0]] = sid
        return sid

    def sampled_spectrum(self, wavelengths, values):
        """A point-sampled spectrum: four wavelengths (nm) and the value at each; the scene then
        renders in frame-packet mode at exactly those wavelengths."""
        sid = len(self.spectra)
        self.spectra.append((len(self.spectrum_values), abi.SPECTRUM_SAMPLED | 4))
        self.spectrum_values.extend(float(np.float32(v)) for v in list(wavelengths) + list(values))
        return SpectrumId(sid)

    # -- textures -----------------------------------------------------------------------------
    def add_texture(self, texels=None, device=None):
        """A float32 image, (height, width, channels) host texels (copied into the scene) or
        device=(pointer, width, height, channels) read in place by every render. Returns the slot."""
        if texels is not None:
            texels = np.ascontiguousarray(texels, dtype=np.float32)
            if texels.ndim == 2:
                texels = texels[:, :, None]
            height, width, channels = texels.shape
            self.textures.append((width, height, channels, texels, None))
        else:
            pointer, width, height, channels = device
            self.textures.append((width, height, channels, None, int(pointer)))
        return len(self.textures) - 1

    # -- geometry -----------------------------------------------------------------------------
    def add_geometry(self, gid, positions, normals, uvs, indices):
        positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        normals = np.asarray(normals, dtype=np.float64).reshape(-1, 3)
        uvs = np.asarray(uvs, dtype=np.float64).reshape(-1, 2)
        indices = np.asarray(indices, dtype=np.uint32).reshape(-1, 3)
        index = len(self.geometries)
        self.geometries.append((gid, self.vertex_count, len(positions), self.triangle_count, len(indices)))
        self._positions.append(positions.astype(np.float32))
        self._normals.append(normals.astype(np.float32))
        self._uvs.append(uvs.astype(np.float32))
        self._indices.append(indices)
        self.vertex_count += len(positions)
        self.triangle_count += len(indices)
        return index

    # -- materials ----------------------------------------------------------------------------
    def add_material(self, mid, closures, emission=None, frame_mode=0, tangent_rotation=0.0):
        """closures: list of dicts {kind, lobes, weight, eta, k, params, probability}."""
        offset = len(self.closures)
        probabilities = []
        for c in closures:
            rec = abi.ClosureDesc()
            rec.kind = c["kind"]
            rec.lobes = c["lobes"]
            rec.weight_spectrum = self.spectrum(c["weight"]) if c.get("weight") is not None else abi.INVALID
            rec.eta_spectrum = self.spectrum(c["eta"]) if c.get("eta") is not None else abi.INVALID
            rec.k_spectrum = self.spectrum(c["k"]) if c.get("k") is not None else abi.INVALID
            for i, p in enumerate(c.get("params", [])):
                rec.parameters[i] = p
            rec.weight_texture = c["weight_texture"] + 1 if c.get("weight_texture") is not None else 0
            rec.roughness_texture = c["roughness_texture"] + 1 if c.get("roughness_texture") is not None else 0
            self.closures.append(rec)
            probabilities.append(float(c.get("probability", 1.0)))
        m = abi.MaterialDesc()
        m.id = mid
        m.closure_offset = offset
        m.closure_count = len(closures)
        m.emission_spectrum = self.spectrum(emission) if emission is not None else abi.INVALID
        m.frame_mode = frame_mode
        m.tangent_rotation = tangent_rotation
        # host-projected sampler grid (TransportLobesDevice.cuh:76-82): probabilities are whole
        # multiples of 2^-24, so every CDF step equals its probability exactly in binary32
        for i, (p, c) in enumerate(zip(project_probabilities(probabilities), np.cumsum(project_probabilities(probabilities)))):
            m.probabilities[i] = p
            m.cdf[i + 1] = c
        m.cdf[0] = 0.0
        index = len(self.materials)
        self.materials.append(m)
        return index

    # -- instances ----------------------------------------------------------------------------
    def add_instance(self, iid, object_id, geometry, material, local_to_world=None, mask=0xFFFFFFFF):
        m = affine() if local_to_world is None else np.asarray(local_to_world, dtype=np.float64)
        inv = _affine_inverse(m)
        rec = abi.InstanceDesc()
        rec.id = iid
        rec.object_id = object_id
        rec.geometry_index = geometry
        rec.material_index = material
        rec.visibility_mask = mask
        for i, v in enumerate(m.astype(np.float32).reshape(-1)):
            rec.local_to_world[i] = v
        for i, v in enumerate(inv.astype(np.float32).reshape(-1)):
            rec.world_to_local[i] = v
        self.instances.append(rec)
        return len(self.instances) - 1

    def add_image_texture(self, levels):
        """An image texture for the surface maps from its mip chain: a list of (height, width,
        channels) arrays, each level max(1, floor(previous / 2)) in each axis. Returns its index."""
        levels = [np.asarray(l, dtype=np.float32).reshape(np.shape(l)[0], np.shape(l)[1], -1) for l in levels]
        h, w, c = levels[0].shape
        for k in range(1, len(levels)):
            ph, pw, _ = levels[k - 1].shape
            if levels[k].shape != (max(1, ph // 2), max(1, pw // 2), c):
                raise ValueError(f"mip level {k} has shape {levels[k].shape}")
        texels = np.ascontiguousarray(np.concatenate([l.reshape(-1) for l in levels]))
        self.image_textures.append((w, h, c, len(levels), texels))
        return len(self.image_textures) - 1

    def set_surface_maps(self, material, normal=None, bump=None):
        """A material's normal map, dict(texture, channels=(0, 1, 2), u_wrap, v_wrap, max_anisotropy,
        max_texel_visits, y_convention), and / or bump map, dict(texture, channel, scale, u_wrap,
        v_wrap, max_anisotropy, max_texel_visits)."""
        m = abi.SurfaceMapsDesc()
        m.normal_texture = m.bump_texture = abi.INVALID
        if normal is not None:
            m.normal_texture = normal["texture"]
            for k, ch in enumerate(normal.get("channels", (0, 1, 2))):
                m.normal_channels[k] = ch
            m.normal_u_wrap, m.normal_v_wrap = normal.get("u_wrap", 0), normal.get("v_wrap", 0)
            m.normal_max_anisotropy = normal.get("max_anisotropy", 8)
            m.normal_max_texel_visits = normal.get("max_texel_visits", 4096)
            m.normal_y_convention = normal.get("y_convention", 0)
        if bump is not None:
            m.bump_texture, m.bump_channel, m.bump_scale = bump["texture"], bump.get("channel", 0), bump["scale"]
            m.bump_u_wrap, m.bump_v_wrap = bump.get("u_wrap", 0), bump.get("v_wrap", 0)
            m.bump_max_anisotropy = bump.get("max_anisotropy", 8)
            m.bump_max_texel_visits = bump.get("max_texel_visits", 4096)
        self.surface_maps[material] = m

    def set_environment(self, radiance):
        self.environment_spectrum = self.spectrum(radiance)

    # -- punctual lights ------------------------------------------------------------------------
    def add_point_light(self, position, intensity, position_error=(0.0, 0.0, 0.0)):
        return self._punctual(abi.PUNCTUAL_POINT, intensity, position, position_error, (0.0, 0.0, 0.0), 0.0, 0.0)

    def add_spot_light(self, position, axis, inner_half_angle, outer_half_angle, intensity,
                       position_error=(0.0, 0.0, 0.0)):
        return self._punctual(abi.PUNCTUAL_SPOT, intensity, position, position_error, axis, inner_half_angle,
                              outer_half_angle)

    def add_directional_light(self, propagation, irradiance):
        return self._punctual(abi.PUNCTUAL_DIRECTIONAL, irradiance, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), propagation, 0.0, 0.0)

    def _punctual(self, kind, spectrum, position, error, direction, inner, outer):
        rec = abi.PunctualDesc()
        rec.kind = kind
        rec.spectrum = self.spectrum(spectrum)
        for a in range(3):
            rec.position[a], rec.position_error[a], rec.direction[a] = position[a], error[a], direction[a]
        rec.inner_half_angle, rec.outer_half_angle = inner, outer
        self.punctuals.append(rec)
        return len(self.punctuals) - 1

    # -- packing ------------------------------------------------------------------------------
    def finalize(self):
        self.positions = np.ascontiguousarray(np.concatenate(self._positions)).astype(np.float32)
        self.normals = np.ascontiguousarray(np.concatenate(self._normals)).astype(np.float32)
        self.uvs = np.ascontiguousarray(np.concatenate(self._uvs)).astype(np.float32)
        self.indices = np.ascontiguousarray(np.concatenate(self._indices)).astype(np.uint32)
        self._geometry_array = (abi.GeometryDesc * len(self.geometries))()
        for i, (gid, vo, vc, to, tc) in enumerate(self.geometries):
            g = self._geometry_array[i]
            g.id, g.vertex_offset, g.vertex_count, g.triangle_offset, g.triangle_count = gid, vo, vc, to, tc
        self._instance_array = (abi.InstanceDesc * len(self.instances))(*self.instances)
        self._material_array = (abi.MaterialDesc * len(self.materials))(*self.materials)
        self._closure_array = (abi.ClosureDesc * max(1, len(self.closures)))(*self.closures)
        self._spectrum_array = (abi.SpectrumDesc * len(self.spectra))(*[abi.SpectrumDesc(o, c) for o, c in self.spectra])
        self._spectrum_values = np.asarray(self.spectrum_values, dtype=np.float32)
        d = abi.SceneDesc()
        d.geometry_count = len(self.geometries)
        d.instance_count = len(self.instances)
        d.material_count = len(self.materials)
        d.closure_count = len(self.closures)
        d.spectrum_count = len(self.spectra)
        d.has_environment = 0 if self.environment_spectrum == abi.INVALID else 1
        d.vertex_count = self.vertex_count
        d.triangle_count = self.triangle_count
        d.spectrum_value_count = len(self._spectrum_values)
        d.positions = self.positions.ctypes.data
        d.normals = self.normals.ctypes.data
        d.uvs = self.uvs.ctypes.data
        d.indices = self.indices.ctypes.data
        d.geometries = C.addressof(self._geometry_array)
        d.instances = C.addressof(self._instance_array)
        d.materials = C.addressof(self._material_array)
        d.closures = C.addressof(self._closure_array)
        d.spectra = C.addressof(self._spectrum_array)
        d.spectrum_values = self._spectrum_values.ctypes.data
        d.environment_spectrum = self.environment_spectrum
        self._texture_array = (abi.TextureDesc * max(1, len(self.textures)))()
        for i, (width, height, channels, texels, pointer) in enumerate(self.textures):
            t = self._texture_array[i]
            t.width, t.height, t.channels = width, height, channels
            t.data = texels.ctypes.data if texels is not None else None
            t.device_data = pointer
        d.texture_count = len(self.textures)
        d.textures = C.addressof(self._texture_array) if self.textures else None
        self._punctual_array = (abi.PunctualDesc * max(1, len(self.punctuals)))(*self.punctuals)
        d.punctual_count = len(self.punctuals)
        d.punctuals = C.addressof(self._punctual_array) if self.punctuals else None
        self._image_array = (abi.ImageTextureDesc * max(1, len(self.image_textures)))()
        for i, (w, h, c, mips, texels) in enumerate(self.image_textures):
            t = self._image_array[i]
            t.width, t.height, t.channel_count, t.mip_count = w, h, c, mips
            t.texels, t.texel_count = texels.ctypes.data, texels.size
        d.image_texture_count = len(self.image_textures)
        d.image_textures = C.addressof(self._image_array) if self.image_textures else None
        self._maps_array = None
        if self.surface_maps:
            self._maps_array = (abi.SurfaceMapsDesc * len(self.materials))()
            for i in range(len(self.materials)):
                m = self.surface_maps.get(i)
                if m is None:
                    m = abi.SurfaceMapsDesc()
                    m.normal_texture = m.bump_texture = abi.INVALID
                self._maps_array[i] = m
            d.surface_maps = C.addressof(self._maps_array)
        self.desc = d
        return self


def mip_chain(image):
    """Levels of an (height, width, channels) image down to 1x1: each level is the 2x2 box average
    of the previous (an odd last row or column folds into its neighbour)."""
    levels = [np.asarray(image, dtype=np.float64)]
    while levels[-1].shape[0] > 1 or levels[-1].shape[1] > 1:
        prev = levels[-1]
        h, w = max(1, prev.shape[0] // 2), max(1, prev.shape[1] // 2)
        ys = [slice(2 * y, prev.shape[0] if y == h - 1 else 2 * y + 2) for y in range(h)]
        xs = [slice(2 * x, prev.shape[1] if x == w - 1 else 2 * x + 2) for x in range(w)]
        levels.append(np.array([[prev[yy, xx].mean(axis=(0, 1)) for xx in xs] for yy in ys]))
    return [l.astype(np.float32) for l in levels]


def project_probabilities(weights):
    """Closure selection probabilities on the 2^-24 grid: floor of each share, then the missing
    units to the largest remainders (one each, lowest index first on ties), as the reference's
    sampler-grid projection does (TransportLobesDevice.cuh:2643-2718)."""
    weights = np.asarray(weights, dtype=np.float64)
    if len(weights) == 0:
        return np.zeros(0)
    scale = 1 << 24
    scaled = weights / weights.sum() * scale
    masses = np.floor(scaled).astype(np.int64)
    remainders = scaled - masses
    for i in sorted(range(len(weights)), key=lambda k: (-remainders[k], k))[: scale - int(masses.sum())]:
        masses[i] += 1
    return masses.astype(np.float64) / scale


# ---- camera ------------------------------------------------------------------------------------
def look_at_camera(origin, target, width, height, vertical_fov, time=0.5, jitter=abi.JITTER_UNIFORM):
    """Pinhole camera frame as OrthonormalFrame::from_normal_and_tangent(-forward, +x) builds it
    (Tests/Backends/CPU/Embree/CornellWavefrontScene.hpp:391-401, PinholeCamera.hpp:57-90),
    evaluated in float64 and rounded once."""
    origin = np.asarray(origin, dtype=np.float64)
    forward = np.asarray(target, dtype=np.float64) - origin
    forward /= np.linalg.norm(forward)
    normal = -forward
    tangent = np.array([1.0, 0.0, 0.0])
    tangent = tangent - normal * np.dot(tangent, normal)
    tangent /= np.linalg.norm(tangent)
    bitangent = np.cross(normal, tangent)
    bitangent /= np.linalg.norm(bitangent)
    tangent = np.cross(bitangent, normal)
    half_height = math.tan(np.float32(vertical_fov) / 2.0)
    half_width = half_height * (width / height)
    cam = abi.CameraDesc()
    for i in range(3):
        cam.origin[i] = origin[i]
        cam.tangent[i] = tangent[i]
        cam.bitangent[i] = bitangent[i]
        cam.normal[i] = normal[i]
    cam.half_width = half_width
    cam.half_height = half_height
    cam.t_min = 0.0
    cam.t_max = float("inf")
    cam.time = time
    cam.width = width
    cam.height = height
    cam.visibility_mask = 0xFFFFFFFF
    cam.jitter_mode = jitter
    return cam


def frame_desc(camera, spp, seed=0, first_sample_index=0):
    f = abi.FrameDesc()
    f.camera = camera
    f.seed = seed
    f.samples_per_pixel = spp
    f.first_sample_index = first_sample_index
    return f


def transport_options(max_depth=8, mis=abi.MIS_BALANCE, sampler=abi.SAMPLER_SOBOL_OWEN,
                      wavelength_mode=abi.WAVELENGTHS_PER_PATH, rr_first_depth=3, rr_min=0.1, rr_max=0.9,
                      packet=(360.0 + 47.0, 360.0 + 164.5, 360.0 + 282.0, 360.0 + 399.5)):
    o = abi.TransportOptions()
    o.sampler = sampler
    o.wavelength_mode = wavelength_mode
    o.mis_heuristic = mis
    o.diffuse_depth_limit = max_depth
    o.glossy_depth_limit = max_depth
    o.specular_depth_limit = max_depth
    o.transmission_depth_limit = max_depth
    o.volume_depth_limit = 0
    o.russian_roulette_enabled = 1 if rr_first_depth else 0
    o.russian_roulette_first_depth = rr_first_depth
    o.russian_roulette_minimum = rr_min if rr_first_depth else 0.0
    o.russian_roulette_maximum = rr_max if rr_first_depth else 0.0
    for i in range(4):
        o.packet_wavelengths[i] = packet[i]
    return o


# ---- shapes --------------------------------------------------------------------------------------
def quad(corners, normal):
    return (np.asarray(corners, dtype=np.float64), np.tile(np.asarray(normal, dtype=np.float64), (4, 1)),
            np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64), np.array([[0, 1, 2], [0, 2, 3]]))


def unit_box():
    """Axis-aligned cube [-0.5, 0.5]^3 with outward flat normals, 24 vertices, 12 triangles."""
    faces = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sign
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            v = np.cross(n, u)
            c = n * 0.5
            corners = [c - 0.5 * u - 0.5 * v, c + 0.5 * u - 0.5 * v, c + 0.5 * u + 0.5 * v, c - 0.5 * u + 0.5 * v]
            faces.append((corners, n))
    positions, normals, uvs, indices = [], [], [], []
    for f, (corners, n) in enumerate(faces):
        positions.extend(corners)
        normals.extend([n] * 4)
        uvs.extend([[0, 0], [1, 0], [1, 1], [0, 1]])
        b = 4 * f
        indices.extend([[b, b + 1, b + 2], [b, b + 2, b + 3]])
    return np.array(positions), np.array(normals), np.array(uvs, dtype=np.float64), np.array(indices)


def lambert(reflectance, probability=1.0):
    return {"kind": abi.CLOSURE_LAMBERTIAN, "lobes": abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, "weight": reflectance,
            "probability": probability}


def specular_reflection(reflectance, probability=1.0):
    return {"kind": abi.CLOSURE_SPECULAR_REFLECTION, "lobes": abi.LOBE_SPECULAR | abi.LOBE_REFLECTION,
            "weight": reflectance, "probability": probability}


def rough_diffuse(reflectance, roughness, probability=1.0):
    """Energy-preserving Oren–Nayar; parameters[0] = roughness in [0, 1]."""
    return {"kind": abi.CLOSURE_ROUGH_DIFFUSE, "lobes": abi.LOBE_DIFFUSE | abi.LOBE_REFLECTION, "weight": reflectance,
            "params": [roughness], "probability": probability}


def rough_conductor(weight, eta, k, alpha_x, alpha_y=None, probability=1.0):
    """GGX conductor; eta / k are spectra (scalars or samples over [360, 830] nm) resolved into
    parameters[0..3] / [4..7] at the path's wavelengths; parameters[8..9] are the GGX widths."""
    alpha_y = alpha_x if alpha_y is None else alpha_y
    return {"kind": abi.CLOSURE_ROUGH_CONDUCTOR, "lobes": abi.LOBE_GLOSSY | abi.LOBE_REFLECTION, "weight": weight,
            "eta": eta, "k": k, "params": [0.0] * 8 + [alpha_x, alpha_y], "probability": probability}


def rough_dielectric(weight, exterior_eta, interior_eta, alpha_x, alpha_y=None, probability=1.0):
    """GGX dielectric; parameters = [exterior eta, interior eta, alpha_x, alpha_y]."""
    alpha_y = alpha_x if alpha_y is None else alpha_y
    return {"kind": abi.CLOSURE_ROUGH_DIELECTRIC,
            "lobes": abi.LOBE_GLOSSY | abi.LOBE_REFLECTION | abi.LOBE_TRANSMISSION, "weight": weight,
            "params": [exterior_eta, interior_eta, alpha_x, alpha_y], "probability": probability}


def cornell(width=64, height=64, materials="lambert"):
    """Cornell box (see module docstring). materials='mixed' makes the tall box a gold GGX
    conductor (alpha 0.3), the short box a GGX dielectric (eta 1.5, alpha 0.2), the left wall an
    energy-preserving Oren–Nayar surface (roughness 0.5) and the back wall a Lambert + GGX
    conductor mixture."""
    s = Scene()
    white = s.add_material(21, [lambert(0.73)])
    left = s.add_material(22, [lambert(0.12)])
    right = s.add_material(23, [lambert(0.45)])
    emitter = s.add_material(24, [lambert(0.0)], emission=17.0)
    back = white
    if materials == "mixed":
        left = s.add_material(27, [rough_diffuse(0.12, 0.5)])
        back = s.add_material(28, [lambert(0.5, probability=0.7),
                                   rough_conductor(0.8, COPPER_ETA, COPPER_K, 0.15, 0.4, probability=0.3)])
    walls = [
        (11, quad([[-1, -1, 0], [-1, -1, 2], [1, -1, 2], [1, -1, 0]], [0, 1, 0]), white),
        (12, quad([[-1, 1, 0], [1, 1, 0], [1, 1, 2], [-1, 1, 2]], [0, -1, 0]), white),
        (13, quad([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [0, 0, 1]), back),
        (14, quad([[-1, -1, 0], [-1, 1, 0], [-1, 1, 2], [-1, -1, 2]], [1, 0, 0]), left),
        (15, quad([[1, -1, 0], [1, -1, 2], [1, 1, 2], [1, 1, 0]], [-1, 0, 0]), right),
        (16, quad([[-0.32, 0.985, 0.62], [0.32, 0.985, 0.62], [0.32, 0.985, 1.18], [-0.32, 0.985, 1.18]],
                  [0, -1, 0]), emitter),
    ]
    for k, (gid, shape, material) in enumerate(walls):
        g = s.add_geometry(gid, *shape)
        s.add_instance(31 + k, 1 + k, g, material)
    box = s.add_geometry(17, *unit_box())
    short = affine(rotation_y(math.radians(15.0)) @ np.diag([0.55, 0.55, 0.55]), (0.36, -0.725, 1.30))
    tall = affine(rotation_y(math.radians(-18.0)) @ np.diag([0.55, 1.15, 0.55]), (-0.38, -0.425, 0.72))
    short_material = tall_material = white
    if materials == "mixed":
        short_material = s.add_material(25, [rough_dielectric(1.0, 1.0, 1.5, 0.2)])
        tall_material = s.add_material(26, [rough_conductor(1.0, GOLD_ETA, GOLD_K, 0.3)])
    s.add_instance(37, 7, box, short_material, short)
    s.add_instance(38, 8, box, tall_material, tall)
    s.finalize()
    camera = look_at_camera((0.0, 0.0, 4.0), (0.0, -0.05, 0.92), width, height, 0.70)
    return s, camera


def textured_cornell(width=256, height=256, texture_size=512, seed=0xC3_7E57, albedo=None, roughness=None,
                     roughness_size=1, mixed=False, specular=None, tint=None):
    """The differentiable Cornell box: the back wall's reflectance is an RGB texture (seeded
    U[0.1, 0.9] texels unless `albedo` is given, host (h, w, 3) array or device tuple) and the tall
    box is a gold GGX conductor whose widths come from a 1-channel roughness texture (0.3 unless
    `roughness` is given). mixed=True adds every other binding: the back wall becomes a textured
    Lambert + copper GGX mixture, the left wall an Oren-Nayar surface with a grey texture (slot 2)
    and the short box a GGX dielectric with a 2-channel (alpha_x, alpha_y) texture (slot 3).
    specular="mirror" makes the back wall a mirror whose tint is the albedo texture;
    specular="mixture" makes it a Lambert (albedo texture, probability 0.6) + mirror (probability
    0.4) mixture whose mirror tint is a 4x4 RGB texture (`tint`, slot 2).
    Returns (scene, camera, textures) with textures the host arrays used."""
    rng = SplitMix64(seed)
    s = Scene()
    if albedo is None:
        values = np.array([rng.uniform() for _ in range(texture_size * texture_size * 3)])
        albedo = (0.1 + 0.8 * values).reshape(texture_size, texture_size, 3).astype(np.float32)
    if roughness is None:
        roughness = np.full((roughness_size, roughness_size, 1), 0.3, dtype=np.float32)
    albedo_slot = s.add_texture(albedo) if isinstance(albedo, np.ndarray) else s.add_texture(device=albedo)
    rough_slot = s.add_texture(roughness) if isinstance(roughness, np.ndarray) else s.add_texture(device=roughness)
    white = s.add_material(21, [lambert(0.73)])
    left = s.add_material(22, [lambert(0.12)])
    right = s.add_material(23, [lambert(0.45)])
    emitter = s.add_material(24, [lambert(0.0)], emission=17.0)
    back = s.add_material(28, [dict(lambert(1.0), weight_texture=albedo_slot)])
    short_material = white
    textures = {"albedo": albedo, "roughness": roughness}
    if specular == "mirror":
        back = s.add_material(28, [dict(specular_reflection(1.0), weight_texture=albedo_slot)])
    elif specular == "mixture":
        if tint is None:
            tint = (0.3 + 0.6 * np.array([rng.uniform() for _ in range(48)])).reshape(4, 4, 3).astype(np.float32)
        textures["tint"] = tint
        tint_slot = s.add_texture(tint)
        back = s.add_material(28, [dict(lambert(1.0, probability=0.6), weight_texture=albedo_slot),
                                   dict(specular_reflection(1.0, probability=0.4), weight_texture=tint_slot)])
    elif specular is not None:
        raise ValueError(f"unknown specular variant {specular!r}")
    if mixed:
        grey = (0.2 + 0.6 * np.array([rng.uniform() for _ in range(16)])).reshape(4, 4, 1).astype(np.float32)
        widths = (0.15 + 0.3 * np.array([rng.uniform() for _ in range(32)])).reshape(4, 4, 2).astype(np.float32)
        textures.update(grey=grey, widths=widths)
        grey_slot = s.add_texture(grey)
        widths_slot = s.add_texture(widths)
        back = s.add_material(28, [dict(lambert(1.0, probability=0.7), weight_texture=albedo_slot),
                                   rough_conductor(0.8, COPPER_ETA, COPPER_K, 0.15, 0.4, probability=0.3)])
        left = s.add_material(27, [dict(rough_diffuse(0.9, 0.5), weight_texture=grey_slot)])
        short_material = s.add_material(25, [dict(rough_dielectric(1.0, 1.0, 1.5, 0.2), roughness_texture=widths_slot)])
    walls = [
        (11, quad([[-1, -1, 0], [-1, -1, 2], [1, -1, 2], [1, -1, 0]], [0, 1, 0]), white),
        (12, quad([[-1, 1, 0], [1, 1, 0], [1, 1, 2], [-1, 1, 2]], [0, -1, 0]), white),
        (13, quad([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [0, 0, 1]), back),
        (14, quad([[-1, -1, 0], [-1, 1, 0], [-1, 1, 2], [-1, -1, 2]], [1, 0, 0]), left),
        (15, quad([[1, -1, 0], [1, -1, 2], [1, 1, 2], [1, 1, 0]], [-1, 0, 0]), right),
        (16, quad([[-0.32, 0.985, 0.62], [0.32, 0.985, 0.62], [0.32, 0.985, 1.18], [-0.32, 0.985, 1.18]],
                  [0, -1, 0]), emitter),
    ]
    for k, (gid, shape, material) in enumerate(walls):
        g = s.add_geometry(gid, *shape)
        s.add_instance(31 + k, 1 + k, g, material)
    box = s.add_geometry(17, *unit_box())
    short = affine(rotation_y(math.radians(15.0)) @ np.diag([0.55, 0.55, 0.55]), (0.36, -0.725, 1.30))
    tall = affine(rotation_y(math.radians(-18.0)) @ np.diag([0.55, 1.15, 0.55]), (-0.38, -0.425, 0.72))
    tall_material = s.add_material(26, [dict(rough_conductor(1.0, GOLD_ETA, GOLD_K, 0.3), roughness_texture=rough_slot)])
    s.add_instance(37, 7, box, short_material, short)
    s.add_instance(38, 8, box, tall_material, tall)
    s.finalize()
    camera = look_at_camera((0.0, 0.0, 4.0), (0.0, -0.05, 0.92), width, height, 0.70)
    return s, camera, textures


def _icosphere(subdivisions):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
             [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
             [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
             [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(subdivisions):
        cache = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        new_faces = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            new_faces += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = new_faces
    return np.array(verts), np.array(faces, dtype=np.uint32)


# Tabulated complex IOR of gold and copper on 8 samples over [360, 830] nm (smooth fits of
# published optical constants; spectral shape, not a colorimetric reference).
GOLD_ETA = [1.66, 1.48, 0.62, 0.27, 0.18, 0.16, 0.17, 0.20]
GOLD_K = [1.95, 1.89, 2.11, 2.78, 3.35, 3.86, 4.33, 4.78]
COPPER_ETA = [1.34, 1.18, 1.05, 0.64, 0.25, 0.22, 0.24, 0.27]
COPPER_K = [1.94, 2.33, 2.58, 2.58, 3.28, 3.90, 4.42, 4.95]


def icosphere_grid(width=1920, height=1080, grid=32, subdivisions=3, seed=0x5EED_1C05, emitter_count=64,
                   materials="mixed", albedo=None, roughness=None):
    """32x32 displaced icospheres (see module docstring). materials='lambert' keeps every object
    Lambertian (used where a closure kind is not yet available). materials='textured' gives every
    object one shared material bound to two textures: a 50/50 mixture of a Lambert closure and an
    anisotropic gold GGX conductor, both weighted by the RGB `albedo` texture (slot 0), with the
    conductor's widths (alpha_x, alpha_y) read from the 2-channel `roughness` texture (slot 1);
    each texture is a host (h, w, c) array or a device tuple (pointer, width, height, channels).
    The geometry does not depend on `materials`."""
    rng = SplitMix64(seed)
    s = Scene()
    base_vertices, faces = _icosphere(subdivisions)
    spacing = 2.6
    emitter_material = s.add_material(900, [lambert(0.0)], emission=40.0)
    shared = None
    if materials == "textured":
        slots = [s.add_texture(t) if isinstance(t, np.ndarray) else s.add_texture(device=t) for t in (albedo, roughness)]
        shared = s.add_material(1999, [dict(lambert(1.0, probability=0.5), weight_texture=slots[0]),
                                       dict(rough_conductor(1.0, GOLD_ETA, GOLD_K, 0.3, 0.3, probability=0.5),
                                            weight_texture=slots[0], roughness_texture=slots[1])])
    for gy in range(grid):
        for gx in range(grid):
            obj = gy * grid + gx
            # seeded value noise: 8 random lobes per sphere
            lobes = [(np.array([rng.uniform() * 2 - 1, rng.uniform() * 2 - 1, rng.uniform() * 2 - 1]),
                      0.08 * rng.uniform()) for _ in range(8)]
            verts = base_vertices.copy()
            scale = np.ones(len(verts))
            for direction, amplitude in lobes:
                direction /= max(np.linalg.norm(direction), 1e-9)
                scale += amplitude * np.maximum(verts @ direction, 0.0) ** 2
            verts = verts * scale[:, None]
            normals = verts / np.linalg.norm(verts, axis=1, keepdims=True)
            uvs = np.stack([0.5 + np.arctan2(normals[:, 2], normals[:, 0]) / (2 * math.pi),
                            0.5 - np.arcsin(np.clip(normals[:, 1], -1, 1)) / math.pi], axis=1)
            g = s.add_geometry(1000 + obj, verts, normals, uvs, faces)
            kind = 0 if materials == "lambert" else int(rng.uniform() * 3)
            roughness = 0.05 + 0.55 * rng.uniform()
            albedo = 0.1 + 0.8 * rng.uniform()
            if kind == 0:
                closures = [lambert(albedo)]
            elif kind == 1:
                metal = rng.uniform() < 0.5
                closures = [rough_conductor(1.0, GOLD_ETA if metal else COPPER_ETA, GOLD_K if metal else COPPER_K,
                                            roughness * roughness)]
            else:
                closures = [rough_dielectric(1.0, 1.0, 1.5, roughness * roughness)]
            m = shared if shared is not None else s.add_material(2000 + obj, closures)
            x = (gx - (grid - 1) / 2.0) * spacing
            z = (gy - (grid - 1) / 2.0) * spacing
            s.add_instance(3000 + obj, 3000 + obj, g, m, affine(translation=(x, 0.0, z)))
    # 64 emissive triangles: one shared triangle geometry, one light instance each
    tri = s.add_geometry(5000, [[-1.5, 0, -1.5], [1.5, 0, -1.5], [0, 0, 1.8]], [[0, -1, 0]] * 3,
                         [[0, 0], [1, 0], [0, 1]], [[0, 1, 2]])
    side = int(math.sqrt(emitter_count))
    for k in range(emitter_count):
        ex = ((k % side) - (side - 1) / 2.0) * spacing * grid / side
        ez = ((k // side) - (side - 1) / 2.0) * spacing * grid / side
        s.add_instance(10000 + k, 10000 + k, tri, emitter_material, affine(translation=(ex, 14.0, ez)))
    s.set_environment(0.05)
    s.finalize()
    extent = spacing * grid / 2.0
    camera = look_at_camera((0.0, extent * 1.3, extent * 0.75), (0.0, -2.0, -extent * 0.1), width, height, 0.9)
    return s, camera


def batch_scenes(count=32, seed=0xC2_5EED, min_triangles=1.0e4, max_triangles=5.0e5):
    """Seeded batch of uneven scenes: triangle counts log-uniform in [min, max] (icosphere grids of
    subdivision 3, 1 280 triangles per object, same material and light generator as
    icosphere_grid), square resolutions 2^U{7..10} and spp 2^U{2..8}. Returns a list of
    (Scene, camera, spp)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        triangles = math.exp(math.log(min_triangles) + rng.uniform() * (math.log(max_triangles) - math.log(min_triangles)))
        grid = max(1, int(round(math.sqrt(triangles / 1280.0))))
        resolution = 1 << (7 + min(3, int(rng.uniform() * 4)))
        spp = 1 << (2 + min(6, int(rng.uniform() * 7)))
        scene, camera = icosphere_grid(resolution, resolution, grid=grid, subdivisions=3,
                                       seed=rng.next_u64() & 0xFFFFFFFF)
        out.append((scene, camera, spp))
    return out


def fit_specs(count=16, seed=0xC4_5EED, min_triangles=5.0e4, max_triangles=2.0e5, min_resolution=256,
              max_resolution=512):
    """Seeded texture-fitting batch: per scene a triangle count uniform in [min, max] (icosphere
    grids of subdivision 3, 1 280 triangles per object), a square resolution uniform in
    {min_resolution..max_resolution} and a scene seed. Returns a list of dicts (grid, resolution,
    seed); build each with icosphere_grid(..., materials="textured")."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        triangles = min_triangles + rng.uniform() * (max_triangles - min_triangles)
        grid = max(1, int(round(math.sqrt(triangles / 1280.0))))
        resolution = min_resolution + min(max_resolution - min_resolution, int(rng.uniform() * (max_resolution - min_resolution + 1)))
        out.append(dict(grid=grid, resolution=resolution, seed=rng.next_u64() & 0xFFFFFFFF))
    return out


def fit_textures(seed, size=1024):
    """Seeded smooth target textures of a texture fit: RGB albedo in [0.1, 0.9] and GGX widths
    (alpha_x, alpha_y) in [0.05, 0.6], each channel a product of two seeded sinusoids over (u, v).
    Returns float32 arrays (size, size, 3) and (size, size, 2)."""
    rng = SplitMix64(seed)
    u = (np.arange(size, dtype=np.float64) + 0.5) / size

    def channel(lo, hi):
        fu, fv = 1 + int(rng.uniform() * 6), 1 + int(rng.uniform() * 6)
        pu, pv = rng.uniform(), rng.uniform()
        pattern = np.outer(np.cos(2 * math.pi * (fv * u + pv)), np.sin(2 * math.pi * (fu * u + pu)))
        return lo + (hi - lo) * (0.5 + 0.5 * pattern)

    albedo = np.stack([channel(0.1, 0.9) for _ in range(3)], axis=2).astype(np.float32)
    widths = np.stack([channel(0.05, 0.6) for _ in range(2)], axis=2).astype(np.float32)
    return albedo, widths
