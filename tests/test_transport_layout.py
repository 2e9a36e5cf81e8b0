"""The transport records' layout manifest (blackframe_hip_query_transport_abi_layout) against
the reference's frozen layout: the 73 values, in LayoutValue order, that the static_asserts of
Blackframe/XPU/Shared/Include/Blackframe/XPU/Shared/TransportAbi.hpp:387-477 fix (record sizes,
alignments and field offsets of TransportSpectrum, TransportRay, SampleStreamIndex, PathSlot,
QueueHeader, SurfaceIdentifiers, ClosestHit and TransportPathStateLane)."""
import ctypes as C

import pytest

# (name, value) in the reference's LayoutValue order (TransportAbi.hpp:136-210)
EXPECTED = [
    ("spectrum_size", 16), ("spectrum_alignment", 16), ("spectrum_values_offset", 0),
    ("ray_size", 48), ("ray_alignment", 16), ("ray_origin_x_offset", 0), ("ray_origin_y_offset", 4),
    ("ray_origin_z_offset", 8), ("ray_t_min_offset", 12), ("ray_direction_x_offset", 16),
    ("ray_direction_y_offset", 20), ("ray_direction_z_offset", 24), ("ray_t_max_offset", 28), ("ray_time_offset", 32),
    ("ray_visibility_mask_offset", 36), ("ray_current_medium_offset", 40), ("ray_reserved_offset", 44),
    ("sample_stream_index_size", 24), ("sample_stream_index_alignment", 8), ("sample_stream_index_pixel_x_offset", 0),
    ("sample_stream_index_pixel_y_offset", 4), ("sample_stream_index_sample_index_offset", 8),
    ("sample_stream_index_seed_offset", 16),
    ("path_slot_size", 4), ("path_slot_alignment", 4), ("path_slot_value_offset", 0),
    ("queue_header_size", 32), ("queue_header_alignment", 16), ("queue_header_abi_major_offset", 0),
    ("queue_header_abi_minor_offset", 2), ("queue_header_struct_size_offset", 4), ("queue_header_kind_offset", 8),
    ("queue_header_capacity_offset", 12), ("queue_header_size_offset", 16), ("queue_header_overflow_count_offset", 20),
    ("queue_header_rejected_count_offset", 24), ("queue_header_reserved_offset", 28),
    ("surface_identifiers_size", 32), ("surface_identifiers_alignment", 16), ("surface_identifiers_object_offset", 0),
    ("surface_identifiers_instance_offset", 4), ("surface_identifiers_geometry_offset", 8),
    ("surface_identifiers_primitive_offset", 12), ("surface_identifiers_material_offset", 16),
    ("surface_identifiers_reserved_offset", 20),
    ("closest_hit_size", 64), ("closest_hit_alignment", 16), ("closest_hit_parameter_offset", 0),
    ("closest_hit_barycentric_vertex0_offset", 4), ("closest_hit_barycentric_vertex1_offset", 8),
    ("closest_hit_barycentric_vertex2_offset", 12), ("closest_hit_geometric_normal_x_offset", 16),
    ("closest_hit_geometric_normal_y_offset", 20), ("closest_hit_geometric_normal_z_offset", 24),
    ("closest_hit_reserved_offset", 28), ("closest_hit_identifiers_offset", 32),
    ("path_state_size", 128), ("path_state_alignment", 16), ("path_state_beta_offset", 0),
    ("path_state_accumulated_radiance_offset", 16), ("path_state_wavelength_nanometers_offset", 32),
    ("path_state_wavelength_pdf_values_offset", 48), ("path_state_diffuse_depth_offset", 64),
    ("path_state_glossy_depth_offset", 68), ("path_state_specular_depth_offset", 72),
    ("path_state_transmission_depth_offset", 76), ("path_state_volume_depth_offset", 80),
    ("path_state_depth_offset", 84), ("path_state_eta_scale_offset", 88), ("path_state_current_medium_offset", 92),
    ("path_state_delta_flags_offset", 96), ("path_state_wavelength_pdf_measures_offset", 100),
    ("path_state_reserved_offset", 104),
]


def test_expected_table_is_the_reference_manifest(abi):
    assert len(EXPECTED) == abi.LAYOUT_VALUE_COUNT == 73
    assert len({n for n, _ in EXPECTED}) == 73
    assert C.sizeof(abi.LayoutManifest) == 320 and abi.LayoutManifest.values.offset == 16
    assert abi.LayoutManifest.reserved_tail.offset == 308


@pytest.mark.gpu
def test_device_manifest_matches_the_reference_layout(hip):
    manifest, devices = hip.transport_abi_layout()
    assert devices >= 1
    assert (manifest.abi_major, manifest.abi_minor, manifest.value_count) == (1, 0, 73)
    assert manifest.reserved_header == 0 and list(manifest.reserved_tail) == [0, 0, 0]
    got = list(manifest.values)
    wrong = [(name, want, got[k]) for k, (name, want) in enumerate(EXPECTED) if got[k] != want]
    assert not wrong, wrong
