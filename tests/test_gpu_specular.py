"""Specular delta closures on the GPU, bit-exact against the oracle: the microfacet Cornell box
with a mirror tall box, a glass short box (specular transmission, eta 1.5) and a Lambert + mirror
mixture on the back wall (delta atoms of a mixture)."""
import pytest

from test_gpu_transport import check_parity, run_both

pytestmark = pytest.mark.gpu

SPEC_R, SPEC_T = 5, 6


def specular_cornell(bf):
    abi = bf.abi

    def factory(w, h):
        s, cam = bf.scenes.cornell(w, h, materials="mixed")
        ids = {m.id: m for m in s.materials}

        def make(c, kind, params):
            c.kind = kind
            c.lobes = abi.LOBE_SPECULAR | (abi.LOBE_REFLECTION if kind == SPEC_R else abi.LOBE_TRANSMISSION)
            c.eta_spectrum = c.k_spectrum = abi.INVALID
            for i in range(10):
                c.parameters[i] = params[i] if i < len(params) else 0.0
        make(s.closures[ids[26].closure_offset], SPEC_R, ())            # tall box: mirror (gold tint)
        make(s.closures[ids[25].closure_offset], SPEC_T, (1.0, 1.5))    # short box: glass
        make(s.closures[ids[28].closure_offset + 1], SPEC_R, ())        # back wall: Lambert + mirror
        s.finalize()
        return s, cam
    return factory


@pytest.mark.parametrize("opts", [dict(), dict(mis=1, max_depth=12, sampler=0)])
def test_cornell_with_specular_closures_bit_exact(bf, oracle, hip, opts):
    print(f"specular Cornell {opts}", flush=True)
    check_parity(*run_both(bf, oracle, hip, 32, 32, 4, scene_factory=specular_cornell(bf), **dict(opts)))
