"""Specular delta closures in the oracle (TransportLobesDevice.cuh:1689-1960, delta atoms of
mixtures :2394-2480), pinned by closed forms: the mirror direction, Snell's law with the radiance
scaling (eta_i/eta_t)^2 and total internal reflection at the critical angle, and white-furnace
renders whose per-path values are known exactly."""
import ctypes as C
import math

import numpy as np
import pytest

from test_oracle_transport_pins import faceted_sphere

SPEC_R, SPEC_T = 5, 6


def record(abi, kind, weight, params=()):
    rec = abi.ClosureRecord()
    rec.kind = kind
    rec.lobes = abi.LOBE_SPECULAR | (abi.LOBE_REFLECTION if kind == SPEC_R else abi.LOBE_TRANSMISSION)
    for k in range(4):
        rec.weight[k] = weight
    for i, p in enumerate(params):
        rec.parameters[i] = p
    return rec


def sample(oracle, rec, wo):
    f3 = C.c_float * 3
    out = (C.c_float * 9)()
    lobes = C.c_uint32()
    st = oracle.lib.bfo_closure_sample(C.byref(rec), 0x18, f3(*wo), 0.0, 0.0, 0.0, out, C.byref(lobes))
    return st, np.array(out[:4], dtype=np.float64), np.array(out[4:7], dtype=np.float64), out[7], out[8], lobes.value


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def test_mirror_direction_and_value(abi, oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        wo = unit(rng.normal(size=3))
        st, value, wi, pdf, eta, lobes = sample(oracle, record(abi, SPEC_R, 0.7), wo)
        assert st == 0 and pdf == 1.0 and eta == 1.0 and lobes == abi.LOBE_SPECULAR | abi.LOBE_REFLECTION
        assert tuple(wi) == (-wo[0], -wo[1], wo[2])
        assert np.allclose(value, 0.7 / abs(float(wo[2])), rtol=2e-7)


@pytest.mark.parametrize("eta_i,eta_t", [(1.0, 1.5), (1.5, 1.0), (1.33, 1.7)])
def test_snell_law_scaling_and_total_internal_reflection(abi, oracle, eta_i, eta_t):
    rng = np.random.default_rng(7)
    tir = refracted = 0
    for _ in range(400):
        wo = unit(rng.normal(size=3))
        wo[2] = abs(wo[2])  # entering from the exterior side (eta_i = exterior)
        st, value, wi, pdf, eta, lobes = sample(oracle, record(abi, SPEC_T, 1.0, (eta_i, eta_t)), wo)
        sin_i = math.hypot(float(wo[0]), float(wo[1]))
        if sin_i * eta_i / eta_t >= 1.0 - 1e-6:
            if sin_i * eta_i / eta_t > 1.0 + 1e-6:
                assert st == 3  # outside support: total internal reflection
                tir += 1
            continue
        assert st == 0 and pdf == 1.0 and lobes == abi.LOBE_SPECULAR | abi.LOBE_TRANSMISSION
        refracted += 1
        sin_t = math.hypot(wi[0], wi[1])
        assert abs(sin_t - sin_i * eta_i / eta_t) < 4e-6
        assert wi[2] < 0 and abs(np.linalg.norm(wi) - 1.0) < 1e-6
        # tangential components are anti-parallel to the outgoing ones
        assert np.dot(wi[:2], wo[:2]) <= 0.0
        cos_t = abs(wi[2])
        assert np.allclose(value, (eta_i / eta_t) ** 2 / cos_t, rtol=4e-6)
        assert abs(eta - (eta_t / eta_i) ** 2) <= 4e-7 * (eta_t / eta_i) ** 2
    assert refracted > 60 and (eta_i <= eta_t or tir > 10)


def test_equal_indices_pass_straight_through(abi, oracle):
    wo = unit([0.3, -0.2, 0.9])
    st, value, wi, pdf, eta, lobes = sample(oracle, record(abi, SPEC_T, 0.8, (1.2, 1.2)), wo)
    assert st == 0 and tuple(wi) == (-wo[0], -wo[1], -wo[2]) and eta == 1.0
    assert np.allclose(value, 0.8 / abs(float(wo[2])), rtol=2e-7)


def furnace(bf, abi, oracle, closures):
    s = bf.scenes.Scene()
    m = s.add_material(1, closures)
    g = faceted_sphere(bf, s, 2, 1.0)
    s.add_instance(2, 2, g, m)
    s.set_environment(1.0)
    s.finalize()
    cam = bf.scenes.look_at_camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 24, 24, 0.9)
    result = oracle.scene(s).render(bf.scenes.frame_desc(cam, 4), bf.scenes.transport_options(max_depth=8, rr_first_depth=0),
                                    threads=4)
    assert result["report"].failed_paths == 0
    radiance = np.array([list(p.radiance) for p in result["paths"]], dtype=np.float64)
    depth = np.array([p.depth for p in result["paths"]])
    assert np.all(radiance[depth == 0] == 1.0)
    return radiance[depth > 0]


def spec(abi, kind, weight, params=()):
    lobes = abi.LOBE_SPECULAR | (abi.LOBE_REFLECTION if kind == SPEC_R else abi.LOBE_TRANSMISSION)
    return {"kind": kind, "lobes": lobes, "weight": weight, "params": list(params)}


def test_mirror_sphere_furnace(bf, abi, oracle):
    hit = furnace(bf, abi, oracle, [spec(abi, SPEC_R, 0.6)])
    assert len(hit) > 100
    assert np.abs(hit - np.float32(0.6)).max() <= 4 * np.spacing(np.float32(0.6))


def test_glass_sphere_furnace(bf, abi, oracle):
    """Transmission only: a path refracts in and out (the eta factors cancel) or is lost to total
    internal reflection at a facet."""
    hit = furnace(bf, abi, oracle, [spec(abi, SPEC_T, 1.0, (1.0, 1.5))])
    through = hit[hit > 0.5]
    assert len(through) > 100
    assert np.abs(through - 1.0).max() <= 8 * np.spacing(np.float32(1.0))
    assert np.all((hit == 0.0) | (hit > 0.5))


def test_lambert_mirror_mixture_furnace(bf, abi, oracle):
    """Half Lambert (rho 0.3), half mirror (0.8): each path is 2*0.3 (diffuse pick) or 2*0.8 (the
    mirror's delta atom, probability 0.5)."""
    closures = [bf.scenes.lambert(0.3, probability=0.5), dict(spec(abi, SPEC_R, 0.8), probability=0.5)]
    hit = furnace(bf, abi, oracle, closures)
    diffuse = np.abs(hit - 0.6) <= 8e-7
    mirror = np.abs(hit - 1.6) <= 2e-6
    assert np.all(diffuse | mirror) and diffuse.sum() > 50 and mirror.sum() > 50
