"""Path inputs in the reference's CudaWavefrontPathInput form (sample stream, primary ray, ray
cone, initial path state), generated on the host from a frame exactly as the device camera stage
generates them: the pinhole ray of the path's pixel sample, four stratified wavelengths of the
path's wavelength dimension (pdf 1/470 per nm, wavelength measure) and PathState::create_initial's
unit throughput."""
import ctypes as C

import numpy as np


def frame_inputs(abi, oracle, frame, options, cone=(0.0, 0.0)):
    cam = frame.camera
    pixels = cam.width * cam.height
    n = pixels * frame.samples_per_pixel
    samples = (abi.SampleStreamIndex * n)()
    rays = (abi.TransportRay * n)()
    cones = (abi.RayCone * n)()
    states = (abi.PathStateLane * n)()
    pdf = float(np.float32(1.0) / np.float32(470.0))
    wl = (C.c_float * 4)()
    for p in range(n):
        s, rem = divmod(p, pixels)
        samples[p] = abi.SampleStreamIndex(rem % cam.width, rem // cam.width, frame.first_sample_index + s, frame.seed)
        assert oracle.lib.bfo_camera_ray(C.byref(cam), options.sampler, C.byref(samples[p]), C.byref(rays[p])) == 1
        cones[p] = abi.RayCone(*cone)
        st = states[p]
        if options.wavelength_mode == abi.WAVELENGTHS_PER_PATH:
            oracle.lib.bfo_sample_wavelengths(oracle.lib.bfo_sample_1d(options.sampler, C.byref(samples[p]), 3), wl)
            values = list(wl)
        else:
            values = list(options.packet_wavelengths)
        for k in range(4):
            st.beta[k] = 1.0
            st.wavelength_nanometers[k] = values[k]
            st.wavelength_pdf_values[k] = pdf
            st.wavelength_pdf_measures[k] = abi.WAVELENGTH_MEASURE
        st.eta_scale = 1.0
    return samples, rays, cones, states


def as_u32(records):
    return np.frombuffer(bytes(memoryview(records).cast("B")), dtype=np.uint32).reshape(len(records), -1)
