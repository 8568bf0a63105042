"""Gradients with respect to 16-bin spectral parameters (binned.gradient_check): the adjoint
chained through the sRGB projection agrees with central differences of the detached estimator for
every bin, roughness, metallic and specular of three surfaces and every emission bin of two lights
of a small bush."""
import pytest

pytestmark = pytest.mark.gpu


def test_binned_gradients_match_finite_differences(cuda):
    from lsystemeditor_amd import abi, binned, diff, inverse
    spec = inverse.c4_specs(count=1, res=32, spp=8, iterations=(3, 3))[0]
    sc = abi.Scene(spec.triangles, spec.materials, bvh="sah")
    r = diff.DiffRenderer(sc, spec.materials, spec.camera, spec.params, spec.width, spec.height, spec.spp)
    bp = binned.BinnedParams(r.theta0(), lights=[3, 4])
    target = r.forward(bp.theta(*bp.perturbed()).detach())
    res = binned.gradient_check(r, bp, target, h=1e-2)
    assert res["params"] == 89
    assert res["median_rel_err"] < 1e-3, res["median_rel_err"]
    assert res["max_rel_err"] < 5e-2, res["max_rel_err"]
