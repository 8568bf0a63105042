"""Batched inverse rendering at the C4 scene size (256x256, 16 spp, depth 8; two scenes, 50 Adam
steps at lr 1e-2), with 16-bin base-colour spectra and with RGB base colours: the base colours
(for spectra, their projected colours), which dominate the images, move toward their targets; the
finite-difference spot checks agree with the adjoint along the trajectory; the loss of 256-spp
images falls most of the way to its value at the target parameters. The step losses are not
asserted: each step renders fresh samples, so they sit on the noise floor of two 16-spp images.

Roughness and metallic are nearly degenerate with the base colours in these images (metallic
starts within 0.01 of its target), so their errors are reported by the optimiser but not
asserted."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("binned_colour", [True, False])
def test_inverse_rendering_recovers_base_colours(cuda, binned_colour):
    from lsystemeditor_amd import inverse
    specs = inverse.c4_specs(count=2, res=256, spp=16, iterations=(3, 4))
    prob = inverse.InverseProblem(specs, binned_colour=binned_colour)
    res = inverse.optimise(prob, steps=50, lr=1e-2, check_every=25)
    e0, e1 = res["param_mae_initial"], res["param_mae_final"]
    print(res)
    assert e1["base_color"] < 0.6 * e0["base_color"], res
    img = res["image_loss"]
    assert img["final"] < img["initial"], img
    assert img["final_excess_fraction"] < 0.3, img
    assert res["target_spp"] == 256
    assert len(res["fd_checks"]) == 2
    for c in res["fd_checks"]:
        assert c["median_rel_err"] < 0.05, c
    assert res["grad_samples_per_s"] > 0.0
    assert prob.parameter_count() == 2 * (54 if binned_colour else 15)
