"""16-bin spectral parameters (lsystemeditor_amd/binned.py) on the CPU: the projection sends a
constant spectrum to white and the rgb2spec spectrum of a colour back to that colour (the 1-nm
spectrum exactly up to the fit, the binned one within the binning error); the parameter map keeps
the scalar parameters and is linear in the bins with the projection as its matrix."""
import numpy as np
import torch

from lsystemeditor_amd import binned, rgb2spec


def test_projection_of_a_constant_spectrum_is_white():
    m = binned.projection()
    assert m.shape == (3, 16)
    assert np.allclose(m.sum(1), 1.0, atol=1e-12)
    assert np.allclose(m @ np.full(16, 0.37), 0.37, atol=1e-12)


def test_projection_bins_partition_the_response():
    """Bin b covers [360 + 29.375 b, 360 + 29.375 (b + 1)) nm: two bins merged equal the projection
    of the same spectrum on 8 bins."""
    m16 = binned.projection(16)
    m8 = binned.projection(8)
    assert np.allclose(m16[:, 0::2] + m16[:, 1::2], m8, atol=1e-12)


def test_uplift_projects_back_to_the_colour():
    m = binned.projection()
    for rgb in [(0.4, 0.3, 0.2), (0.35, 0.35, 0.3), (1.0, 0.95, 0.9), (0.9, 0.95, 1.0), (0.5, 0.5, 0.5)]:
        s = binned.uplift(rgb)
        assert np.all(s > 0.0) and np.all(s < 1.0 + 1e-9)
        assert np.abs(m @ s - np.array(rgb)).max() < 0.01, rgb
    assert np.allclose(binned.uplift((1.0, 1.0, 1.0)), 1.0, atol=1e-3)


def test_unbinned_uplift_is_the_fitted_colour():
    """The same sigmoid spectrum sampled every nanometre (no binning) reproduces the colour to
    the fit's residual: the binned projection's deviation is the binning alone."""
    fit = rgb2spec._Fit()
    rgb = np.array([[0.2, 0.5, 0.1]])
    c, err = fit.solve(np.zeros((1, 3)), rgb, 200)
    assert float(err[0]) < 1e-4


def test_theta_map_is_linear_in_the_bins():
    th0 = torch.tensor(np.random.default_rng(0).uniform(0.1, 0.9, (5, 7)), dtype=torch.float32)
    bp = binned.BinnedParams(th0, lights=[3, 4])
    assert bp.count == 3 * (16 + 3) + 2 * 16
    th = bp.theta()
    m = torch.from_numpy(binned.projection()).float()
    assert torch.allclose(th[bp.surfaces, 0:3], bp.refl @ m.T)
    assert torch.allclose(th[bp.lights, 0:3], bp.emis @ m.T)
    assert torch.equal(th[bp.surfaces, 3:6], th0[bp.surfaces, 3:6])
    assert torch.equal(th[:, 6], th0[:, 6])
    assert torch.equal(th[bp.lights, 3:6], th0[bp.lights, 3:6])
    # chain rule: d(sum w * theta)/d refl = w_rgb @ M
    w = torch.randn(5, 7)
    g = torch.autograd.grad((th * w).sum(), bp.parameters())
    assert torch.allclose(g[0], w[bp.surfaces, 0:3] @ m, atol=1e-6)
    assert torch.allclose(g[1], w[bp.surfaces, 3:6])
    assert torch.allclose(g[2], w[bp.lights, 0:3] @ m, atol=1e-6)
