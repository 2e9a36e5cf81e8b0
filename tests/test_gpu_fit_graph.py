"""One texture-fit step recorded as one graph (blackframe_hip_fit_graph_*, TextureFit(graph=True))
against the direct calls (TextureFit): the same two textured frames and starting textures, four
steps each, the direct fit's eager work and synchronisations between the graph's replays.

Bars:
  - the first replay's forward render (films and tapes) bit-identical to the direct step's (no
    gradient has touched the textures yet), its losses within 1e-12 relative (their binary64 sums
    over workgroups run in any order), and its forward report's lane counts equal;
  - later steps: losses within 1e-4 relative and textures within 1e-4 on all but 0.1 % of texels
    (the adjoint accumulates texel gradients with float atomics, whose order differs between
    runs; a texel whose gradient sums to about zero may take its Adam step with the other sign);
  - the device step counter advances once per replay (the Adam bias correction of step k);
  - a replay beyond max_steps, an empty Adam schedule and stage events (which need host reads)
    are refused, the last while the step is being recorded, after which a new graph records.
"""
import numpy as np
import pytest

from test_gpu_fit import fit_frames

pytestmark = pytest.mark.gpu


def make_fit(bf, hip, graph, max_steps=4096):
    from blackframe_renderer_amd.autodiff import TextureFit

    options, frames, textures = fit_frames(bf, hip)
    params = [(t, 0.05, 0.0, 1.0) if t.shape[-1] == 3 else (t, 0.02, 0.01, 1.0) for t in textures]
    return TextureFit(frames, options, params, graph=graph, max_graph_steps=max_steps), textures


def test_graph_step_equals_direct_calls(bf, hip):
    import torch

    direct, direct_textures = make_fit(bf, hip, graph=False)
    graphed, graph_textures = make_fit(bf, hip, graph=True)
    for a, b in zip(direct_textures, graph_textures):
        assert torch.equal(a, b)
    for step in range(4):
        want = direct.step()
        torch.cuda.synchronize()
        got = graphed.step()
        torch.cuda.synchronize()
        want, got = want.cpu().numpy(), got.cpu().numpy()
        if step == 0:
            assert np.allclose(got, want, rtol=1e-12, atol=0.0), (got, want)  # binary64 sums over blocks, any order
            for fd, fg in zip(direct.frames, graphed.frames):
                assert torch.equal(fd["film"].float_sum, fg["film"].float_sum)
                assert torch.equal(fd["film"].counts, fg["film"].counts)
                assert torch.equal(fd["tape"].radiance, fg["tape"].radiance)
            forward, adjoint = graphed.graph_reports()
            ref = direct.last_forward_report
            assert forward.failed_paths == 0 and adjoint.failed_paths == 0
            assert (forward.closest_hit_queries, forward.shadow_queries, forward.shade_lanes) == \
                (ref.closest_hit_queries, ref.shadow_queries, ref.shade_lanes)
            assert forward.closure_samples == ref.closure_samples and forward.light_samples == ref.light_samples
            assert adjoint.closest_hit_queries == direct.last_adjoint_report.closest_hit_queries
        else:
            assert np.allclose(got, want, rtol=1e-4, atol=0.0), (step, got, want)
        for a, b in zip(direct_textures, graph_textures):
            diff = (a - b).abs()
            assert float((diff > 1e-4).float().mean()) <= 1e-3, (step, float(diff.max()))
            assert float(diff.max()) <= 0.2, (step, float(diff.max()))
    _, _, steps = graphed.graph.report()
    assert steps == 4 and graphed.step_count == 4


def test_graph_step_limits(bf, hip):
    import torch

    fit, _ = make_fit(bf, hip, graph=True, max_steps=2)
    first = float(fit.step().sum())
    second = float(fit.step().sum())
    assert np.isfinite(first) and np.isfinite(second)
    with pytest.raises(hip.XpuError):
        fit.step()
    torch.cuda.synchronize()
    frames = [(f["scene"], f["frame"], f["film"], f["tape"], f["target"], f["film_gradient"], f["gradients"], 1.0)
              for f in fit.frames]
    with pytest.raises(hip.XpuError):  # no replay fits an empty Adam schedule
        hip.FitGraph(fit.workspace, frames, [], fit.losses, fit.options, max_steps=0)
    # stage events need host reads: refused while the step is being recorded, and the capture ends cleanly
    options = type(fit.options).from_buffer_copy(fit.options)
    options.instrumentation = 1
    with pytest.raises(hip.XpuError):
        hip.FitGraph(fit.workspace, frames, [], fit.losses, options)
    fit.graph.close()
    again, _ = make_fit(bf, hip, graph=True, max_steps=1)
    assert np.isfinite(float(again.step().sum()))
