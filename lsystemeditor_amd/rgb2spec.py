"""rgb2spec coefficient tables for hero-wavelength rendering.

The reference upsamples RGB to spectra with the rgb2spec model of Jakob and Hanika, "A
Low-Dimensional Function Space for Efficient Spectral Upsampling" (Eurographics 2019): a colour's
spectrum is sigmoid(c0 l^2 + c1 l + c2), and (c0, c1, c2) are interpolated from a table indexed
by the dominant channel, a scale axis z (the dominant channel's value) and the two other channels
divided by z (Spectral/Rgb2Spec.h:65-136). Its table file (srgb.coeff, written by the rgb2spec
project's optimiser) is missing from the reference, so this module computes one with the
published method and writes/reads the reference's file format (Spectral/SpectralState.cpp:48-103:
"SPEC", uint32 res, float32 scale[res], float32 data[3][res][res][res][3]).

Optimisation (the paper's, with a linear-RGB residual): for every table cell the target colour is
rgb[l] = z, rgb[l+1] = x z, rgb[l+2] = y z (x, y on a uniform grid, z on the smoothstep-of-
smoothstep scale); the coefficients minimise |rgb(c) - target|^2, where rgb(c) is the spectrum's
linear-sRGB response (CIE 1931 x D65, 1-nm sums over 360..830 nm, the renderer's own response)
divided by that of the constant spectrum 1. Levenberg-Marquardt in the normalised wavelength
t = (lambda - 360) / 470, started at the grey cell and continued slice by slice along z from the
previous slice's solution, then mapped back to coefficients in nanometres.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

LAMBDA_MIN = 360.0
LAMBDA_MAX = 830.0
LAMBDA_RANGE = LAMBDA_MAX - LAMBDA_MIN

_CIE_CACHE: dict[str, np.ndarray] = {}


def _cie_tables() -> dict[str, np.ndarray]:
    """CIE 1931 x, y, z and D65 tables (5 nm from 360 nm), read from the library's data header."""
    if not _CIE_CACHE:
        import re
        src = (Path(__file__).resolve().parent / "csrc" / "pt_cie.h").read_text()
        for name, body in re.findall(r"static constexpr float (\w+)\[kCieSamples\] = \{(.*?)\};", src, re.S):
            vals = [float(v.strip().rstrip("f")) for v in body.replace("\n", " ").split(",") if v.strip()]
            _CIE_CACHE[name] = np.array(vals, np.float64)
    return _CIE_CACHE


def rgb_response(lam: np.ndarray) -> np.ndarray:
    """Linear-sRGB response per nm (CIE matching functions x D65 -> XYZ -> sRGB), shape (n, 3);
    the interpolation of SpectralCore.h:41-48 and the matrix of :125-130, in float64."""
    t = _cie_tables()
    x = (np.asarray(lam, np.float64) - LAMBDA_MIN) * (94.0 / LAMBDA_RANGE)
    o = np.clip(x.astype(np.int64), 0, 93)
    w = x - o

    def interp(tab):
        return tab[o] * (1.0 - w) + tab[o + 1] * w

    ill = interp(t["kCieD65"])
    X, Y, Z = interp(t["kCieXBar"]) * ill, interp(t["kCieYBar"]) * ill, interp(t["kCieZBar"]) * ill
    m = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                  [0.055648, -0.204043, 1.057311]])
    return np.stack([X, Y, Z], -1) @ m.T


def scale_axis(res: int) -> np.ndarray:
    """z values of the table slices: smoothstep(smoothstep(j / (res - 1)))."""
    def smoothstep(v):
        return v * v * (3.0 - 2.0 * v)
    return smoothstep(smoothstep(np.linspace(0.0, 1.0, res)))


def _sigmoid(x):
    return 0.5 + 0.5 * x / np.sqrt(1.0 + x * x)


def _dsigmoid(x):
    return 0.5 / (1.0 + x * x) ** 1.5


class _Fit:
    def __init__(self):
        lam = np.arange(LAMBDA_MIN, LAMBDA_MAX + 1.0, 1.0)
        resp = rgb_response(lam)
        self.R = resp / resp.sum(0)          # constant spectrum 1 -> (1, 1, 1)
        t = (lam - LAMBDA_MIN) / LAMBDA_RANGE
        self.B = np.stack([t * t, t, np.ones_like(t)], -1)  # (n, 3)

    def eval(self, c):
        x = c @ self.B.T                     # (N, n)
        rgb = _sigmoid(x) @ self.R           # (N, 3)
        ds = _dsigmoid(x)
        # d rgb_k / d c_a = sum_l ds_l B_la R_lk
        J = np.stack([(ds * self.B[:, a]) @ self.R for a in range(3)], -1)
        return rgb, J

    def solve(self, c, target, iters: int):
        mu = np.full(len(c), 1e-3)
        rgb, J = self.eval(c)
        r = rgb - target
        err = (r * r).sum(1)
        for _ in range(iters):
            if float(err.max()) < 1e-16:
                break
            JtJ = np.einsum("nka,nkb->nab", J, J)
            g = np.einsum("nka,nk->na", J, r)
            A = JtJ + mu[:, None, None] * (np.einsum("naa->na", JtJ)[:, :, None] * np.eye(3) + 1e-12 * np.eye(3))
            step = np.linalg.solve(A, -g[..., None])[..., 0]
            cn = c + step
            rgbn, Jn = self.eval(cn)
            rn = rgbn - target
            errn = (rn * rn).sum(1)
            ok = errn < err
            c = np.where(ok[:, None], cn, c)
            r = np.where(ok[:, None], rn, r)
            J = np.where(ok[:, None, None], Jn, J)
            err = np.where(ok, errn, err)
            mu = np.where(ok, mu * 0.3, mu * 10.0)
            mu = np.clip(mu, 1e-9, 1e9)
        return c, np.sqrt(err)


def build_table(res: int = 32, iters: int = 40) -> tuple[np.ndarray, np.ndarray, float]:
    """(scale float32 [res], data float32 [3, res, res, res, 3], worst residual)."""
    if res < 2:
        raise ValueError("res must be at least 2")
    fit = _Fit()
    scale = scale_axis(res)
    grid = np.arange(res) / (res - 1)
    data = np.zeros((3, res, res, res, 3), np.float64)
    worst = 0.0
    for l in range(3):
        yy, xx = np.meshgrid(grid, grid, indexing="ij")   # cell (y, x)
        xs, ys = xx.ravel(), yy.ravel()
        start = res // 5
        order = list(range(start, res)) + list(range(start - 1, -1, -1))
        coeffs = {}
        for j in order:
            z = scale[j]
            target = np.zeros((len(xs), 3))
            target[:, l] = z
            target[:, (l + 1) % 3] = xs * z
            target[:, (l + 2) % 3] = ys * z
            if j == start:
                c0 = np.zeros((len(xs), 3))
            else:
                c0 = coeffs[j - 1 if j > start else j + 1]
            c, e = fit.solve(c0, target, iters if j != start else 4 * iters)
            coeffs[j] = c
            # the black slice (z = 0) cannot be fitted and is never fetched with weight > 0
            if z > 1e-3:
                worst = max(worst, float(e.max()))
            # map t = (lambda - 360) / 470 back to nanometres
            a, b, cc = c[:, 0], c[:, 1], c[:, 2]
            A = a / LAMBDA_RANGE ** 2
            Bc = b / LAMBDA_RANGE - 2.0 * a * LAMBDA_MIN / LAMBDA_RANGE ** 2
            Cc = cc - b * LAMBDA_MIN / LAMBDA_RANGE + a * LAMBDA_MIN ** 2 / LAMBDA_RANGE ** 2
            data[l, j] = np.stack([A, Bc, Cc], -1).reshape(res, res, 3)
    return scale.astype(np.float32), data.astype(np.float32), worst


def write_coeff_file(path, scale: np.ndarray, data: np.ndarray) -> None:
    res = len(scale)
    with open(path, "wb") as f:
        f.write(b"SPEC")
        f.write(struct.pack("<I", res))
        f.write(np.ascontiguousarray(scale, np.float32).tobytes())
        f.write(np.ascontiguousarray(data, np.float32).tobytes())


def read_coeff_file(path) -> tuple[np.ndarray, np.ndarray]:
    raw = Path(path).read_bytes()
    if raw[:4] != b"SPEC":
        raise ValueError("invalid rgb2spec coefficient header")
    res = struct.unpack("<I", raw[4:8])[0]
    scale = np.frombuffer(raw, np.float32, res, 8).copy()
    data = np.frombuffer(raw, np.float32, 9 * res ** 3, 8 + 4 * res).copy()
    return scale, data.reshape(3, res, res, res, 3)


_TABLES: dict[int, tuple[np.ndarray, np.ndarray, float]] = {}


def table(res: int = 32) -> tuple[np.ndarray, np.ndarray]:
    """Cached build_table(res) for this process."""
    if res not in _TABLES:
        _TABLES[res] = build_table(res)
    s, d, _ = _TABLES[res]
    return s, d
