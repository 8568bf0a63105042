"""Benchmark: plain adaptive forward rendering of the large L-system bush on one MI355X.

A step is one complete adaptive render of the frame: clear the accumulator, then sample every
pixel until it converges (relative error 0.01, 16..256 spp, max depth 12), with the reference's
per-pixel semantics (pass p uses seed p + 1, convergence tested every 4 samples), executed by the
wavefront scheduler. Samples are counted from the per-pixel counts buffer. Inputs (scene, BVH,
camera, params) are resident in HBM before timing. The BVH is the binned-SAH build by default
(closest hits equal the reference's median-split tree's; --bvh median uses that tree, and the
median-split throughput is reported beside the headline). After the headline measurement the
differentiable C2 workload (adjoint cost, gradient checks) and the CPU baseline are measured.

Prints one JSON line (see the contract in the repository's documentation).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s and grad-samples/s on 1 GPU vs plain fwd at equal RMSE; HBM GB/s %"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of the traversal kernels (the dominant kernels), per unit of work:
NODE_BYTES = 64        # binary internal node fetched: both child boxes + child references
WIDE_NODE_BYTES = 112  # 4-wide node fetched: four child boxes + four child references
TRI_BYTES = 48         # leaf triangle tested: v0, two edges, original index
EXT_RAY_BYTES = 28 + 16  # extension ray: origin, direction, tMin read; t, u, v, triangle written
SHADOW_RAY_BYTES = 32 + 1  # shadow ray record read; occlusion byte written


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["c1", "c0"], default="c1")
    ap.add_argument("--max-paths", type=int, default=0,
                    help="paths in flight per wavefront group of the C1 render (0 = the library default)")
    ap.add_argument("--bvh", choices=["median", "sah", "sah-binary"], default="sah",
                    help="BVH build: the reference's median split, or binned SAH (same results, faster traversal)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every host core this process may use")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-loop", action="store_true", help="skip the per-pass (reference cadence) timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-render (C2) measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 64-scene (C3) measurement")
    ap.add_argument("--no-spectral", action="store_true", help="skip the hero-wavelength (spectral) measurement")
    ap.add_argument("--no-mis", action="store_true", help="skip the light sampling (MIS vs reference) measurement")
    ap.add_argument("--no-inverse", action="store_true", help="skip the batched inverse-rendering (C4) measurement")
    ap.add_argument("--inverse-steps", type=int, default=200)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB device-to-device tensor copy (scale reference for PMC byte counters)")
    return ap.parse_args()


def read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary (key: value lines), if present."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    vals = {}
    for line in p.read_text().splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cores() -> int:
    """Cores this process may run on: its CPU affinity, capped by a cgroup CPU quota if one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if quota != "max":
            n = min(n, max(1, int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(spec, threads: int, runs: int = 5):
    """The oracle (CPU restatement, C with pthreads) on a bounded sample of the same workload: a
    centred 256x256 window of the full-resolution frame for its first 32 passes, on every host core
    this process may use. Two builds of the same sources: the parity build (FP contraction off, the
    order of operations the GPU matches bit for bit) and a contraction build (FMA, AVX2). Each is
    timed as the median of `runs` runs after one warm-up run; the BVH build is timed separately."""
    import ctypes as C
    from lsystemeditor_amd import abi
    from oracle import oracle as O

    t0 = time.perf_counter()
    sc = O.OracleScene(abi.host_to_gpu_triangles(spec.triangles), spec.materials)
    build_s = time.perf_counter() - t0
    w, h = spec.width, spec.height
    win, passes = 256, 32
    x0, y0 = (w - win) // 2, (h - win) // 2
    idx = ((np.arange(y0, y0 + win)[:, None] * w) + np.arange(x0, x0 + win)[None, :]).astype(np.int32).ravel()
    env = O.Env()
    builds = {}
    for variant in ("parity", "contract"):
        L = O.lib(variant)
        times, samples = [], 0
        for run in range(runs + 1):
            acc = O.Accumulator(w, h)
            acc.clear()
            t0 = time.perf_counter()
            for p in range(passes):
                L.orc_adaptive_sample(O.ptr(acc.acc), O.ptr(acc.counts), O.ptr(acc.mean), O.ptr(acc.m2),
                                      O.ptr(acc.conv), O.ptr(idx), idx.size, w, h, O.ptr(spec.camera),
                                      C.byref(sc.scene), C.byref(env), O.ptr(spec.params), p + 1, threads)
            dt = time.perf_counter() - t0
            samples = int(acc.counts[idx].sum())
            if run > 0:
                times.append(dt)
        med = float(np.median(times))
        builds[variant] = {"msamples_s": samples / med / 1e6, "median_s": round(med, 3),
                           "runs_s": [round(t, 3) for t in times], "samples": samples}
    best = max(builds, key=lambda k: builds[k]["msamples_s"])
    return {"value": builds[best]["msamples_s"], "unit": "Msamples/s", "cores": threads, "kind": "port",
            "build": best, "builds": builds,
            "sample": f"oracle (C, pthreads, {threads} threads) on a centred {win}x{win} window of the {w}x{h} frame, "
                      f"passes 1-{passes} ({builds[best]['samples']} samples); median of {runs} runs after a warm-up; "
                      f"parity build (-O3 -ffp-contract=off) {builds['parity']['msamples_s']:.3f}, contraction build "
                      f"(-O3 -march=x86-64-v3 -ffp-contract=fast) {builds['contract']['msamples_s']:.3f} Msamples/s; "
                      f"BVH build {build_s:.1f} s excluded"}


def adjoint_bench(c1, bvh: str, cpu_threads: int) -> dict:
    """C2: differentiable render of the C1 bush at 512x512, 32 spp, depth 8, five material groups
    (bark, leaf, ground, two lights). Times the fixed-spp forward and the adjoint on the same
    samples, checks the adjoint against central differences (h = 1e-3) of the detached estimator
    on the GPU for every live parameter, and against the oracle's FP64 adjoint on a 32x32 window."""
    import ctypes as C
    import torch
    from lsystemeditor_amd import abi, diff, render, scenes
    from oracle import oracle as O

    spec = scenes.config2(c1=c1)
    w, h, spp = spec.width, spec.height, spec.spp
    scene = abi.Scene(spec.triangles, spec.materials, bvh=bvh)
    r = diff.DiffRenderer(scene, spec.materials, spec.camera, spec.params, w, h, spp)
    theta0 = r.theta0()
    target = r.forward(torch.from_numpy(scenes.perturbed_target(theta0.numpy())).cuda())
    img = r.forward(theta0.cuda())
    grad_img = (2.0 * (img - target)).contiguous()
    r.adjoint(grad_img)
    torch.cuda.synchronize()
    reps = 3
    th = theta0.cuda()
    t0 = time.perf_counter()
    for _ in range(reps):
        r.forward(th)   # includes the per-sample radiance capture the adjoint reuses
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        g = r.adjoint(grad_img)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / reps
    samples = w * h * spp
    print(f"[bench] C2 forward {t_fwd * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms ({samples / t_adj / 1e6:.1f} "
          f"grad-Msamples/s)", file=sys.stderr, flush=True)
    # central differences of the FP32 detached estimator on the GPU: the float rounding of the two
    # renders a difference divides by 2h dominates small gradients at h = 1e-3; h = 1e-2 trades
    # it for truncation error. The FP64 differences on a window below carry neither.
    res = diff.gradient_check(r, theta0.cuda(), target, h=1e-3)
    res2 = diff.gradient_check(r, theta0.cuda(), target, h=1e-2)
    print(f"[bench] C2 gradient vs central differences: h=1e-3 max rel err {res['max_rel_err']:.2e} (median "
          f"{res['median_rel_err']:.2e}), h=1e-2 max {res2['max_rel_err']:.2e} (median {res2['median_rel_err']:.2e}) "
          f"over {int(res['live'].sum())} parameters", file=sys.stderr, flush=True)
    # FP64 CPU reference on a 32x32 window (the loss restricted to it)
    win = torch.zeros_like(grad_img).view(h, w, 3)
    y0, x0 = h // 2 - 16, w // 2 - 16
    win[y0:y0 + 32, x0:x0 + 32] = grad_img.view(h, w, 3)[y0:y0 + 32, x0:x0 + 32]
    win = win.reshape(-1, 3).contiguous()
    g_gpu = r.adjoint(win).cpu().numpy()
    ref = O.OracleScene(abi.host_to_gpu_triangles(spec.triangles), spec.materials)
    gi = np.ascontiguousarray(win.cpu().numpy().astype(np.float64))
    g_cpu = np.zeros(7 * spec.materials.shape[0])
    O.lib().orc_render_adjoint(O.ptr(gi), w, h, O.ptr(spec.camera), C.byref(ref.scene), C.byref(O.Env()),
                               O.ptr(r.params_host), spp, O.ptr(g_cpu), cpu_threads)
    g_cpu = g_cpu.reshape(-1, 7)
    live = res["live"] & (g_cpu != 0.0)
    rel_cpu = np.abs(g_gpu - g_cpu)[live] / np.maximum(np.abs(g_cpu[live]), 1e-3 * np.abs(g_cpu).max())
    # central differences (h = 1e-4) of the detached estimator evaluated in double on the same
    # window, for the same linear loss: the derivative without float rounding in the difference
    t0 = time.perf_counter()
    g_fd = np.zeros_like(g_cpu)
    base_mats = np.ascontiguousarray(r.base.copy())
    words = list(diff.PARAM_WORDS)
    win_np = np.ascontiguousarray(gi)

    def window_loss(mats):
        im = np.zeros((w * h, 3))
        O.lib().orc_render_detached_f64_seeded(O.ptr(im), w, h, O.ptr(spec.camera), C.byref(ref.scene),
                                               O.ptr(mats), C.byref(O.Env()), O.ptr(r.params_host), spp, 1,
                                               O.ptr(win_np), cpu_threads)
        return float((win_np * im).sum())

    for i, k in zip(*np.nonzero(res["live"])):
        up, dn = base_mats.copy(), base_mats.copy()
        up[i, words[k]] += np.float32(1e-4)
        dn[i, words[k]] -= np.float32(1e-4)
        g_fd[i, k] = (window_loss(up) - window_loss(dn)) / (float(up[i, words[k]]) - float(dn[i, words[k]]))
    t_fd64 = time.perf_counter() - t0
    live64 = res["live"] & (g_fd != 0.0)
    fd64_floor = 1e-3 * np.abs(g_fd).max()
    rel_fd64 = np.zeros_like(g_fd)
    rel_fd64[live64] = np.abs(g_gpu - g_fd)[live64] / np.maximum(np.abs(g_fd[live64]), fd64_floor)
    print(f"[bench] C2 GPU adjoint vs FP64 central differences of the detached estimator (32x32 window, h=1e-4): "
          f"max rel err {rel_fd64[live64].max():.2e}, median {np.median(rel_fd64[live64]):.2e} over "
          f"{int(live64.sum())} parameters ({t_fd64:.1f} s)", file=sys.stderr, flush=True)
    rel_cpu_all = np.zeros_like(g_cpu)
    rel_cpu_all[live] = rel_cpu
    groups = ["bark", "leaf", "ground", "light 1", "light 2"]
    per_param = []
    for i, k in zip(*np.nonzero(res["live"])):
        per_param.append({"group": groups[i] if i < len(groups) else f"material {i}", "param": diff.PARAM_NAMES[k],
                          "adjoint": float(res["adjoint"][i, k]), "fd_h1e-3": float(res["fd"][i, k]),
                          "fd_rel_err": float(res["rel_err"][i, k]), "fd_h1e-2": float(res2["fd"][i, k]),
                          "fd_h1e-2_rel_err": float(res2["rel_err"][i, k]),
                          "fp64_window_rel_err": float(rel_cpu_all[i, k]) if live[i, k] else None,
                          "fp64_fd_window": float(g_fd[i, k]), "fp64_window_adjoint": float(g_gpu[i, k]),
                          "fp64_fd_window_rel_err": float(rel_fd64[i, k]) if live64[i, k] else None})
    print(f"[bench] C2 GPU adjoint vs the FP64 oracle adjoint (same window): max rel err "
          f"{float(rel_cpu.max()) if rel_cpu.size else float('nan'):.2e}", file=sys.stderr, flush=True)
    # 16-bin spectral parameters: base colour (surfaces) and emission (lights) as binned spectra
    # projected to linear sRGB; gradients through the projection against central differences
    from lsystemeditor_amd import binned
    bp = binned.BinnedParams(theta0, lights=[3, 4])
    target_b = r.forward(bp.theta(*bp.perturbed()).detach())
    t0 = time.perf_counter()
    resb = binned.gradient_check(r, bp, target_b, h=1e-2)
    tb = time.perf_counter() - t0
    proj_dev = float((bp.theta().detach()[:, 0:3] - theta0[:, 0:3]).abs().max())
    print(f"[bench] C2 16-bin spectral parameters ({resb['params']}): gradient vs central differences (h=1e-2) max "
          f"rel err {resb['max_rel_err']:.2e}, median {resb['median_rel_err']:.2e} ({tb:.1f} s)", file=sys.stderr,
          flush=True)
    binned_out = {"params": resb["params"], "bins": binned.BINS,
                  "layout": "3 surfaces x (16 base-colour bins, roughness, metallic, specular) + 2 lights x 16 "
                            "emission bins; bins projected to linear sRGB (CIE 1931 x D65), rendered in RGB",
                  "fd_step": 1e-2, "fd_max_rel_err": resb["max_rel_err"], "fd_median_rel_err": resb["median_rel_err"],
                  "projection_max_abs_dev_from_rgb": proj_dev}
    return {"workload": f"C2: C1 bush, {w}x{h}, {spp} spp fixed, max depth 8, {spec.materials.shape[0]} material "
                        "groups x (base rgb, roughness, metallic, specular, emission); L2 loss against the render "
                        "at parameters perturbed uniformly by +-20% (seeded)",
            "grad_samples_per_s": samples / t_adj, "forward_samples_per_s": samples / t_fwd,
            "cost_ratio_adjoint_vs_forward": t_adj / t_fwd,
            "cost_ratio_forward_plus_adjoint_vs_forward": (t_fwd + t_adj) / t_fwd,
            "equal_rmse": "same samples in both",
            "params_checked": int(res["live"].sum()), "fd_step": 1e-3,
            "fd_max_rel_err": res["max_rel_err"], "fd_median_rel_err": res["median_rel_err"],
            "fd_step_1e-2_max_rel_err": res2["max_rel_err"], "fd_step_1e-2_median_rel_err": res2["median_rel_err"],
            "fp64_cpu_window": "32x32", "fp64_cpu_max_rel_err": float(rel_cpu.max()) if rel_cpu.size else None,
            "fp64_fd_window_max_rel_err": float(rel_fd64[live64].max()),
            "fp64_fd_window_median_rel_err": float(np.median(rel_fd64[live64])),
            "fp64_fd_window_note": "GPU adjoint of the 32x32-window loss against central differences (h = 1e-4) of "
                                   "the detached estimator evaluated in double (orc_render_detached_f64_seeded)",
            "per_parameter": per_param,
            "per_parameter_note": "adjoint, fd_h1e-3, fd_h1e-2: full-frame L2 loss, FP32 detached estimator on the "
                                  "GPU; fp64_*: the 32x32-window linear loss, GPU adjoint vs FP64 oracle adjoint and "
                                  "vs FP64 central differences (h = 1e-4); rel errors floored at 1e-3 of the largest",
            "binned_spectral": binned_out,
            "hero_wavelength": spectral_adjoint_bench(spec, bvh, cpu_threads)}


def spectral_adjoint_bench(spec, bvh: str, cpu_threads: int) -> dict:
    """C2 in hero-wavelength mode (four wavelengths in 380..720 nm): every material's base colour
    and the lights' emission as 16-bin spectra over that range (the uplift of their colours),
    rendered spectrally; pathTracerAdjointRenderSpectral gives the gradient of an L2 loss with
    respect to the bins, roughness, metallic, specular and the emission scales. Times the forward
    and the adjoint, and checks the gradient on windows against the oracle's FP64 spectral adjoint
    and against FP64 central differences of the oracle's spectral detached estimator."""
    import ctypes as C
    import torch
    from lsystemeditor_amd import abi, binned, diff, rgb2spec
    from oracle import oracle as O

    w, h, spp = spec.width, spec.height, spec.spp
    lights = [i for i in range(spec.materials.shape[0]) if spec.materials[i, 9] > 0.0]
    sp = binned.hero_spectra(spec.materials, lights, 380.0, 720.0)
    table = rgb2spec.table(16)
    hero = abi.Spectral(*table)
    hero.enable(380.0, 720.0)
    try:
        scene = abi.Scene(spec.triangles, spec.materials, bvh=bvh)
        r = diff.SpectralDiffRenderer(scene, spec.materials, sp, spec.camera, spec.params, w, h, spp)
        th0 = r.theta0()
        g = np.random.default_rng(5)
        target = r.forward((th0 * torch.from_numpy(g.uniform(0.8, 1.2, tuple(th0.shape))).float()).cuda())
        img = r.forward(th0.cuda())
        grad_img = (2.0 * (img - target)).contiguous()
        r.adjoint(grad_img)
        torch.cuda.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            r.forward(th0.cuda())
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            r.adjoint(grad_img)
        torch.cuda.synchronize()
        t_adj = (time.perf_counter() - t0) / reps
        samples = w * h * spp
        live = r.live_mask()
        print(f"[bench] C2 hero-wavelength: forward {t_fwd * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms "
              f"({samples / t_adj / 1e6:.1f} grad-Msamples/s), {int(live.sum())} live parameters", file=sys.stderr,
              flush=True)
        ref = O.OracleScene(abi.host_to_gpu_triangles(spec.triangles), spec.materials)
        ref.set_spectra(sp)
        ospec = O.OracleSpectral(*table, lambda_min=380.0, lambda_max=720.0)

        def window(size):
            win = torch.zeros_like(grad_img).view(h, w, 3)
            y0, x0 = h // 2 - size // 2, w // 2 - size // 2
            win[y0:y0 + size, x0:x0 + size] = grad_img.view(h, w, 3)[y0:y0 + size, x0:x0 + size]
            win = win.reshape(-1, 3).contiguous()
            return win, np.ascontiguousarray(win.cpu().numpy().astype(np.float64))

        # the 32x32 window: the GPU gradient against the oracle's FP64 spectral adjoint
        win32, gi32 = window(32)
        g32 = r.adjoint(win32).cpu().numpy()
        want = np.zeros(diff.SPECTRAL_PARAMS * spec.materials.shape[0])
        O.lib().orc_render_adjoint_spectral(O.ptr(gi32), w, h, O.ptr(spec.camera), C.byref(ref.scene),
                                            C.byref(O.Env()), O.ptr(r.params_host), C.byref(ospec.model), spp, 1,
                                            O.ptr(want), cpu_threads)
        want = want.reshape(-1, diff.SPECTRAL_PARAMS)
        nz = want != 0.0
        rel_o = np.abs(g32 - want)[nz] / np.maximum(np.abs(want[nz]), 1e-3 * np.abs(want).max())
        # the 16x16 window: central differences (h = 1e-4) of the FP64 spectral detached estimator
        win16, gi16 = window(16)
        g16 = r.adjoint(win16).cpu().numpy()
        t0 = time.perf_counter()
        mats0 = np.ascontiguousarray(r.base.copy())
        sp0 = np.ascontiguousarray(sp.copy())

        def loss(mats, spectra):
            im = np.zeros((w * h, 3))
            O.lib().orc_render_detached_spectral(None, O.ptr(im), O.ptr(gi16), w, h, O.ptr(spec.camera),
                                                 C.byref(ref.scene), O.ptr(mats), O.ptr(spectra), C.byref(O.Env()),
                                                 O.ptr(r.params_host), C.byref(ospec.model), spp, 1, cpu_threads)
            return float((gi16 * im).sum())

        fd = np.zeros_like(g16)
        for i, k in zip(*np.nonzero(live)):
            if k < abi.SPECTRUM_WORDS:
                up, dn = sp0.copy(), sp0.copy()
                up[i, k] += np.float32(1e-4)
                dn[i, k] -= np.float32(1e-4)
                fd[i, k] = (loss(mats0, up) - loss(mats0, dn)) / (float(up[i, k]) - float(dn[i, k]))
            else:
                wd = diff.SPECTRAL_SCALAR_WORDS[k - abi.SPECTRUM_WORDS]
                up, dn = mats0.copy(), mats0.copy()
                up[i, wd] += np.float32(1e-4)
                dn[i, wd] -= np.float32(1e-4)
                fd[i, k] = (loss(up, sp0) - loss(dn, sp0)) / (float(up[i, wd]) - float(dn[i, wd]))
        t_fd = time.perf_counter() - t0
        nzf = live & (fd != 0.0)
        rel_f = np.abs(g16 - fd)[nzf] / np.maximum(np.abs(fd[nzf]), 1e-3 * np.abs(fd).max())
        print(f"[bench] C2 hero-wavelength gradient: vs the FP64 oracle adjoint (32x32 window) max rel err "
              f"{rel_o.max():.2e} over {int(nz.sum())} parameters; vs FP64 central differences (16x16 window, "
              f"h=1e-4) max {rel_f.max():.2e}, median {np.median(rel_f):.2e} over {int(nzf.sum())} ({t_fd:.1f} s)",
              file=sys.stderr, flush=True)
        return {"workload": f"C2 in hero-wavelength mode: four wavelengths in 380..720 nm, {w}x{h}, {spp} spp, "
                            "16-bin reflectance spectra for every material and 16-bin emission for the two lights "
                            "(uplift of their colours over 380..720 nm); L2 loss against the render at every "
                            "parameter scaled by U(0.8, 1.2) (seeded)",
                "parameters_per_material": "16 reflectance bins, 16 emission bins, roughness, metallic, specular, "
                                           "emission scale",
                "live_parameters": int(live.sum()),
                "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3, "grad_samples_per_s": samples / t_adj,
                "cost_ratio_adjoint_vs_forward": t_adj / t_fwd,
                "cost_ratio_forward_plus_adjoint_vs_forward": (t_fwd + t_adj) / t_fwd,
                "fp64_oracle_window": "32x32", "fp64_oracle_params": int(nz.sum()),
                "fp64_oracle_max_rel_err": float(rel_o.max()),
                "fp64_fd_window": "16x16", "fp64_fd_params": int(nzf.sum()), "fp64_fd_step": 1e-4,
                "fp64_fd_max_rel_err": float(rel_f.max()), "fp64_fd_median_rel_err": float(np.median(rel_f))}
    finally:
        abi.Spectral.disable()
        hero.close()


def facade_bench(spec) -> dict:
    """The C++ renderer facade (class PathTracer) rendering the C1 frame to completion on its
    worker thread, SAH build with its 4-wide tree: the reference's cadence (one pass per worker
    iteration with a frame-ready callback each) and 256 passes per launch (one wavefront call)."""
    from lsystemeditor_amd.tracer import Tracer
    p = spec.params
    pi = p.view(np.int32)
    out = {"workload": "C1 through class PathTracer (configure, camera and settings, start, wait until idle)"}
    for per_launch in (1, 256):
        t = Tracer()
        t.set_bvh_build_mode(1)
        t.set_passes_per_launch(per_launch)
        t0 = time.perf_counter()
        if not t.configure(spec.width, spec.height, spec.triangles, spec.materials, None):
            raise RuntimeError("facade configure failed")
        setup_s = time.perf_counter() - t0
        t.set_camera(spec.camera)
        t.set_clear_color(float(p[0]), float(p[1]), float(p[2]))
        t.set_environment_intensity(float(p[3]))
        t.set_max_path_depth(int(pi[4]))
        t.set_russian_roulette_min_depth(int(pi[5]))
        t.set_min_samples(int(pi[6]))
        t.set_max_samples_per_pixel(int(pi[7]))
        t.set_relative_error_threshold(float(p[10]))
        t1 = time.perf_counter()
        t.start()
        if not t.wait_until_idle(600000):
            raise RuntimeError("facade render did not finish")
        dt = time.perf_counter() - t1
        t.stop()
        samples = int(np.asarray(t.read_accumulator()["counts"], np.int64).sum())
        t.close()
        key = "reference_cadence" if per_launch == 1 else f"passes_per_launch_{per_launch}"
        out[key] = {"msamples_s": samples / dt / 1e6, "samples": samples, "seconds": round(dt, 3),
                    "configure_s": round(setup_s, 2)}
    print(f"[bench] facade (class PathTracer): reference cadence {out['reference_cadence']['msamples_s']:.1f}, "
          f"256 passes per launch {out['passes_per_launch_256']['msamples_s']:.1f} Msamples/s", file=sys.stderr,
          flush=True)
    return out


def batch_bench(bvh: str) -> dict:
    """C3: 64 bushes (seeds 0..63, iterations 3..7, 128^2..1024^2, 64 spp, depth 8) rendered by one
    batch wavefront replayed from a graph, against the per-scene graph (every scene's own render
    back to back) and the same 64 renders issued one by one through the adaptive entry point;
    then every scene's tree rebuild and the batch wavefront in one graph, against the sequential
    renders plus the eager builds. Paths-in-flight bounds are autotuned (per-scene: on the
    largest scene; batch: on the batch)."""
    import torch
    from lsystemeditor_amd import batch

    t0 = time.perf_counter()
    specs = batch.batch_specs()
    gen_s = time.perf_counter() - t0
    largest = max(specs, key=lambda s: s.width * s.height * s.triangles.shape[0])
    best, tune = batch.autotune_paths(largest, bvh=bvh)
    br = batch.BatchRenderer(specs, bvh=bvh, max_paths=best)
    samples = br.samples
    br.render_sequential_adaptive_api()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br.render_sequential_adaptive_api()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    br.capture()
    br.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br.replay()
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    # one batch wavefront over all 64 scenes, its paths-in-flight bound autotuned, graph-captured;
    # the per-scene graph goes first (its captures pin the workspace that autotuning regrows)
    br.release_graphs()
    bbest, btune = batch.autotune_batch_paths(br)
    br.max_paths = bbest
    br.capture_batched()
    br.replay_batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    br.replay_batched()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # every scene's tree rebuild and the batch wavefront in one graph: the rebuilds of unchanged
    # geometry enqueue the levels each first build needed; the cold variant enqueues all 49 levels,
    # as a build of new geometry must
    full = None
    if bvh in ("sah", "sah-binary"):
        batch_only = br._batch().acc.acc.clone()
        timings = {}
        # the rebuilds as 4 parallel branches of the graph (the default), and in one branch
        for cold, branches in ((False, 4), (True, 4), (False, 1)):
            br.capture_build_and_render(cold=cold, branches=branches)
            br.replay_build_and_render()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            br.replay_build_and_render()
            torch.cuda.synchronize()
            timings[cold, branches] = time.perf_counter() - t0
            br.check_rebuild()                # every scene's own rebuild status
            if not torch.equal(br._batch().acc.acc, batch_only):
                raise RuntimeError("C3: the build-and-render graph's image differs from the batch-only render")
        t_full, t_cold, t_one = timings[False, 4], timings[True, 4], timings[False, 1]
        t_seq_build = t_seq + br.build_ms / 1e3
        full = {"build_and_render_graph_msamples_s": samples / t_full / 1e6,
                "build_and_render_graph_ms": round(t_full * 1e3, 2),
                "builds_in_graph_ms": round((t_full - t_batch) * 1e3, 2),
                "builds_in_graph_note": "rebuild of unchanged geometry: each rebuild enqueues the levels its "
                                        "scene's first build needed; the 64 rebuilds run as 4 parallel branches "
                                        "of the graph (own stream and build workspace each)",
                "one_branch_build_and_render_graph_msamples_s": samples / t_one / 1e6,
                "one_branch_builds_in_graph_ms": round((t_one - t_batch) * 1e3, 2),
                "cold_build_and_render_graph_msamples_s": samples / t_cold / 1e6,
                "cold_builds_in_graph_ms": round((t_cold - t_batch) * 1e3, 2),
                "build_and_render_image_equals_batch_only": True,
                "sequential_with_builds_msamples_s": samples / t_seq_build / 1e6,
                "build_and_render_speedup_vs_sequential_with_builds": t_seq_build / t_full}
    br.release_graphs()
    tri = sum(s.triangles.shape[0] for s in specs)
    print(f"[bench] C3 {len(specs)} scenes, {tri} triangles, {samples / 1e6:.0f} Msamples: batch wavefront "
          f"{samples / t_batch / 1e6:.1f} Msamples/s (paths in flight {bbest}), per-scene graph "
          f"{samples / t_graph / 1e6:.1f}, sequential {samples / t_seq / 1e6:.1f} Msamples/s",
          file=sys.stderr, flush=True)
    if full is not None:
        print(f"[bench] C3 with the 64 tree rebuilds of unchanged geometry: one graph (rebuilds + batch wavefront) "
              f"{full['build_and_render_graph_msamples_s']:.1f} Msamples/s ({full['builds_in_graph_ms']:.1f} ms of "
              f"rebuilds; cold, all 49 levels: {full['cold_builds_in_graph_ms']:.1f} ms, "
              f"{full['cold_build_and_render_graph_msamples_s']:.1f} Msamples/s; rebuilds in one branch "
              f"{full['one_branch_builds_in_graph_ms']:.1f} ms), sequential renders after eager "
              f"builds {full['sequential_with_builds_msamples_s']:.1f}", file=sys.stderr, flush=True)
    return {"workload": "C3: 64 bushes, seeds 0-63, iterations 3-7, resolutions 128^2-1024^2, 64 spp, depth 8",
            "scenes": len(specs), "triangles": int(tri), "samples": samples,
            "batch_wavefront_msamples_s": samples / t_batch / 1e6,
            "graph_msamples_s": samples / t_graph / 1e6, "sequential_msamples_s": samples / t_seq / 1e6,
            "speedup_vs_sequential": t_seq / t_batch, "per_scene_graph_speedup_vs_sequential": t_seq / t_graph,
            "autotuned_paths_in_flight": best, "autotune_ms": {str(k): round(v, 2) for k, v in tune.items()},
            "batch_autotuned_paths_in_flight": bbest,
            "batch_autotune_ms": {str(k): round(v, 2) for k, v in btune.items()},
            "bvh_build_ms_total": round(br.build_ms, 1), "scene_generation_s": round(gen_s, 1),
            **(full or {})}


def light_sampling_bench(scene, spec) -> dict:
    """The two light sampling modes on the C1 frame, at equal samples and at equal error: mode 0 is
    the reference's (its emissive shadow ray is always occluded, so the area lights are reached by
    BSDF sampling alone), mode 1 adds multiple importance sampling on the area lights. A reference
    image takes 1024 spp per pixel in mode 1 (pass seeds 100001..); each mode then renders 64 spp
    per pixel (seeds 1..64). Reported per mode: Msamples/s, the RMSE of luminance against the
    reference (raw, and with the reference's own expected error removed), and the time the mode
    needs to reach mode 1's RMSE (RMSE falls as 1/sqrt(spp)). Then the adaptive C1 render (relative
    error 0.01, 16-256 spp) in each mode: samples, time and RMSE against the same reference."""
    import math

    import torch

    from lsystemeditor_amd import abi, render
    w, h = spec.width, spec.height
    lum_w = torch.tensor([0.2126, 0.7152, 0.0722], device="cuda", dtype=torch.float64)

    def fixed_inputs(spp: int):
        p = spec.params.copy()
        ints = p[4:10].view(np.int32).copy()
        ints[2] = spp  # minSamples
        ints[3] = spp  # maxSamplesPerPixel
        p[4:10] = ints.view(np.float32)
        return render.DeviceInputs(spec.camera, p)

    def lum(a):
        return a.acc[:, :3].double() @ lum_w

    acc = render.Accumulator.allocate(w, h)
    ref_spp, spp = 1024, 64
    t0 = time.perf_counter()
    with abi.LightSampling(1):
        render.render_wavefront(scene, acc, fixed_inputs(ref_spp), first_iter=100000)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    ref = lum(acc).clone()
    ref_noise = float((acc.m2.double() / (ref_spp - 1)).mean()) / ref_spp  # expected MSE the reference adds
    out = {"workload": "C1 frame, light sampling mode 0 (reference) vs 1 (MIS on area lights)",
           "reference": f"{ref_spp} spp per pixel in mode 1, pass seeds 100001-{100000 + ref_spp} "
                        f"({ref_s:.1f} s)", "fixed_spp": spp}
    modes = {}
    for mode in (0, 1):
        inp = fixed_inputs(spp)
        with abi.LightSampling(mode):
            render.render_wavefront(scene, acc, inp)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            render.render_wavefront(scene, acc, inp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
        mse = float(((lum(acc) - ref) ** 2).mean())
        modes[mode] = {"msamples_s": w * h * spp / dt / 1e6, "seconds": dt, "rmse": math.sqrt(mse),
                       "rmse_without_reference_noise": math.sqrt(max(mse - ref_noise, 0.0))}
    for mode in (0, 1):
        m = modes[mode]
        # seconds to reach mode 1's error at 64 spp: spp scales with (own RMSE / target RMSE)^2
        m["seconds_to_mode1_rmse"] = m["seconds"] * (m["rmse_without_reference_noise"] /
                                                     modes[1]["rmse_without_reference_noise"]) ** 2
    out["mode0"], out["mode1"] = modes[0], modes[1]
    out["equal_rmse_speedup_mode1"] = modes[0]["seconds_to_mode1_rmse"] / modes[1]["seconds_to_mode1_rmse"]
    inputs = render.DeviceInputs(spec.camera, spec.params)
    for mode in (0, 1):
        with abi.LightSampling(mode):
            render.render_wavefront(scene, acc, inputs)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            render.render_wavefront(scene, acc, inputs)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t2
        n = int(acc.counts.to(torch.int64).sum().item())
        out[f"adaptive_mode{mode}"] = {"samples": n, "seconds": dt, "msamples_s": n / dt / 1e6,
                                       "rmse": math.sqrt(float(((lum(acc) - ref) ** 2).mean()))}
    print(f"[bench] light sampling at {spp} spp: mode 0 {modes[0]['msamples_s']:.0f} Msamples/s RMSE "
          f"{modes[0]['rmse_without_reference_noise']:.4g}, mode 1 {modes[1]['msamples_s']:.0f} Msamples/s RMSE "
          f"{modes[1]['rmse_without_reference_noise']:.4g}; time to equal RMSE: mode 1 "
          f"{out['equal_rmse_speedup_mode1']:.2f}x faster; adaptive C1: mode 0 {out['adaptive_mode0']['samples']} "
          f"samples RMSE {out['adaptive_mode0']['rmse']:.4g}, mode 1 {out['adaptive_mode1']['samples']} samples "
          f"RMSE {out['adaptive_mode1']['rmse']:.4g}", file=sys.stderr, flush=True)
    return out


def spectral_bench(scene, spec, inputs, acc, passes: int, bvh: str) -> dict:
    """Hero-wavelength rendering: the C1 frame and the C0 workload (four wavelengths in 380..720 nm)
    with the project's rgb2spec table; Msamples/s beside the RGB figures."""
    import torch

    from lsystemeditor_amd import abi, render, rgb2spec, scenes
    res = 16
    t0 = time.perf_counter()
    table = abi.Spectral(*rgb2spec.table(res))
    table_s = time.perf_counter() - t0
    out = {"hero_wavelengths": 4, "rgb2spec_table_res": res, "rgb2spec_table_build_s": round(table_s, 2)}
    try:
        table.enable(360.0, 830.0)
        render.render_wavefront(scene, acc, inputs, passes=passes)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        render.render_wavefront(scene, acc, inputs, passes=passes)
        torch.cuda.synchronize()
        n = int(acc.counts.to(torch.int64).sum().item())
        out["c1_msamples_s"] = n / (time.perf_counter() - t1) / 1e6
        out["c1_samples"] = n
        c0 = scenes.config0()
        sc0 = abi.Scene(c0.triangles, c0.materials, bvh=bvh)
        in0 = render.DeviceInputs(c0.camera, c0.params)
        acc0 = render.Accumulator.allocate(c0.width, c0.height)
        table.enable(380.0, 720.0)
        render.render_wavefront(sc0, acc0, in0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            render.render_wavefront(sc0, acc0, in0)
        torch.cuda.synchronize()
        n0 = int(acc0.counts.to(torch.int64).sum().item())
        out["c0_workload"] = "C0: 256x256, 16 spp, depth 8, 4 hero wavelengths in 380-720 nm"
        out["c0_msamples_s"] = n0 * reps / (time.perf_counter() - t2) / 1e6
        abi.Spectral.disable()
        render.render_wavefront(sc0, acc0, in0)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(reps):
            render.render_wavefront(sc0, acc0, in0)
        torch.cuda.synchronize()
        out["c0_rgb_msamples_s"] = n0 * reps / (time.perf_counter() - t3) / 1e6
    finally:
        abi.Spectral.disable()
        table.close()
    print(f"[bench] spectral: C1 {out['c1_msamples_s']:.1f} Msamples/s, C0 {out['c0_msamples_s']:.1f} "
          f"(RGB {out['c0_rgb_msamples_s']:.1f})", file=sys.stderr, flush=True)
    return out


def hero_fd_check(problem, h: float = 1e-4, window: int = 16) -> dict:
    """C4's FD spot check at hero wavelengths: scene 0's GPU spectral adjoint of its L2 loss
    restricted to a centred window against central differences (step h) of the oracle's spectral
    detached estimator evaluated in double, for every optimised parameter."""
    import ctypes as C
    import torch
    from lsystemeditor_amd import abi, diff, rgb2spec
    from lsystemeditor_amd.inverse import HERO_RANGE
    from oracle import oracle as O
    r, p, spec = problem.renderers[0], problem.params[0], problem.specs[0]
    w, hh, spp = r.width, r.height, r.spp
    th = p.theta().detach()
    img = r.forward(th.cuda())
    win = torch.zeros_like(img).view(hh, w, 3)
    y0, x0 = hh // 2 - window // 2, w // 2 - window // 2
    win[y0:y0 + window, x0:x0 + window] = (2.0 * (img - problem.targets[0])).view(hh, w, 3)[y0:y0 + window,
                                                                                               x0:x0 + window]
    win = win.reshape(-1, 3).contiguous()
    g = r.adjoint(win).cpu().numpy()
    if getattr(problem, "_oracle0", None) is None:
        ref = O.OracleScene(abi.host_to_gpu_triangles(spec.triangles), spec.materials)
        problem._oracle0 = (ref, O.OracleSpectral(*rgb2spec.table(16), lambda_min=HERO_RANGE[0],
                                                  lambda_max=HERO_RANGE[1]))
    ref, ospec = problem._oracle0
    mats, sp = r._tables(th)
    # the oracle samples with the current parameters too (the adjoint's detached sampling)
    np.copyto(ref.mats, mats.reshape(ref.mats.shape))
    ref.set_spectra(sp)
    gi = np.ascontiguousarray(win.cpu().numpy().astype(np.float64))

    def loss(m, s_):
        im = np.zeros((w * hh, 3))
        O.lib().orc_render_detached_spectral(None, O.ptr(im), O.ptr(gi), w, hh, O.ptr(spec.camera),
                                             C.byref(ref.scene), O.ptr(m), O.ptr(s_), C.byref(O.Env()),
                                             O.ptr(r.params_host), C.byref(ospec.model), spp, 1, 16)
        return float((gi * im).sum())

    rel = []
    fds = {}
    for i, k in zip(*np.nonzero(p.mask.numpy())):
        if k < abi.SPECTRUM_WORDS:
            up, dn = sp.copy(), sp.copy()
            up[i, k] += np.float32(h)
            dn[i, k] -= np.float32(h)
            fds[i, k] = (loss(mats, up) - loss(mats, dn)) / (float(up[i, k]) - float(dn[i, k]))
        else:
            wd = diff.SPECTRAL_SCALAR_WORDS[k - abi.SPECTRUM_WORDS]
            up, dn = mats.copy(), mats.copy()
            up[i, wd] += np.float32(h)
            dn[i, wd] -= np.float32(h)
            fds[i, k] = (loss(up, sp) - loss(dn, sp)) / (float(up[i, wd]) - float(dn[i, wd]))
    top = max(abs(v) for v in fds.values()) or 1.0
    for (i, k), v in fds.items():
        rel.append(abs(g[i, k] - v) / max(abs(v), 1e-3 * top))
    rel = np.array(rel)
    return {"params": int(rel.size), "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel)),
            "reference": f"FP64 central differences (h = {h:g}) of the oracle's spectral detached estimator, "
                         f"{window}x{window} window of scene 0"}



def inverse_bench(bvh: str, steps: int) -> dict:
    """C4: 32 bushes (iterations 3-6, 256x256, 16 spp, depth 8), 200 Adam steps (lr 1e-2) through
    the adjoint on 16-bin base-colour spectra, roughness and metallic of bark, leaf and ground; FD
    spot checks every 50 steps."""
    from lsystemeditor_amd import inverse
    t0 = time.perf_counter()
    specs = inverse.c4_specs()
    gen_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    prob = inverse.InverseProblem(specs, bvh=bvh)
    setup_s = time.perf_counter() - t1
    res = inverse.optimise(prob, steps=steps, lr=1e-2, check_every=50,
                           log=lambda m: print(f"[bench] C4 {m}", file=sys.stderr, flush=True))
    out = {"workload": "C4: 32 bushes, seeds 0-31, iterations 3-6, 256x256, 16 spp, depth 8; Adam lr 1e-2 on "
                       "16-bin base-colour spectra, roughness, metallic of bark, leaf, ground (targets: seeded "
                       "+-20% perturbation, rendered at 256 spp)",
           "scenes": len(specs), "params": prob.parameter_count(),
           "triangles": int(sum(s.triangles.shape[0] for s in specs)),
           "scene_generation_s": round(gen_s, 2), "setup_s": round(setup_s, 2), **res}
    e0, e1 = res["param_mae_initial"], res["param_mae_final"]
    print(f"[bench] C4 {steps} steps in {res['wall_s']:.1f} s ({res['grad_samples_per_s'] / 1e6:.1f} grad-Msamples/s), "
          f"parameter error {e0['all']:.4f} -> {e1['all']:.4f} (projected base colour {e0['base_color']:.4f} -> "
          f"{e1['base_color']:.4f}), step loss {res['loss_first_steps']:.4g} -> {res['loss_last_steps']:.4g}; "
          f"256-spp image loss {res['image_loss']['initial']:.5g} -> {res['image_loss']['final']:.5g} "
          f"(at the target parameters {res['image_loss']['target_params']:.5g}: excess "
          f"{100 * res['image_loss']['final_excess_fraction']:.1f} % of the initial)", file=sys.stderr, flush=True)
    del prob
    out["hero_wavelength"] = inverse_hero_bench(specs, bvh, steps)
    return out


def inverse_hero_bench(specs, bvh: str, steps: int) -> dict:
    """C4 rendered and differentiated at four hero wavelengths in 380..720 nm: the base colours as
    16-bin reflectance spectra over that range (uplift of the C2 colours), optimised with
    roughness and metallic through the batched spectral adjoint; FD spot checks against the
    oracle's FP64 spectral detached estimator."""
    from lsystemeditor_amd import abi, inverse, rgb2spec
    hero = abi.Spectral(*rgb2spec.table(16))
    hero.enable(*inverse.HERO_RANGE)
    try:
        t1 = time.perf_counter()
        prob = inverse.InverseProblem(specs, bvh=bvh, transport="hero", fd_checker=hero_fd_check)
        setup_s = time.perf_counter() - t1
        res = inverse.optimise(prob, steps=steps, lr=1e-2, check_every=50,
                               log=lambda m: print(f"[bench] C4 hero-wavelength {m}", file=sys.stderr, flush=True))
        e0, e1 = res["param_mae_initial"], res["param_mae_final"]
        print(f"[bench] C4 hero-wavelength {steps} steps in {res['wall_s']:.1f} s "
              f"({res['grad_samples_per_s'] / 1e6:.1f} grad-Msamples/s), parameter error {e0['all']:.4f} -> "
              f"{e1['all']:.4f} (bins {e0['bins']:.4f} -> {e1['bins']:.4f}, projected colour {e0['base_color']:.4f} -> "
              f"{e1['base_color']:.4f}, roughness {e0['roughness']:.4f} -> {e1['roughness']:.4f}, metallic "
              f"{e0['metallic']:.4f} -> {e1['metallic']:.4f}); 256-spp image loss {res['image_loss']['initial']:.5g} -> "
              f"{res['image_loss']['final']:.5g} (floor {res['image_loss']['target_params']:.5g})", file=sys.stderr,
              flush=True)
        out = {"workload": "C4 at four hero wavelengths in 380..720 nm: 16-bin reflectance spectra over that range "
                           "(uplift of the colours), roughness, metallic of bark, leaf, ground through the batched "
                           "spectral adjoint; lights' emission spectra fixed; targets: every optimised value scaled "
                           "by a seeded factor in [0.8, 1.2], rendered at 256 spp",
               "params": prob.parameter_count(), "setup_s": round(setup_s, 2), **res}
        del prob
        return out
    finally:
        abi.Spectral.disable()
        hero.close()


def profile_read():
    import ctypes as C
    from lsystemeditor_amd import abi
    out = (C.c_double * 15)()
    abi.check(abi.lib().pathTracerWavefrontProfileRead(C.addressof(out), 15), "pathTracerWavefrontProfileRead")
    return list(out)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the renderer runs on one GPU only; sharding across GPUs is described, "
                                   "not implemented", "n_gpus": args.gpus}))
        return 2
    import torch

    from lsystemeditor_amd import abi, render, scenes

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        return 1
    spec = scenes.config1() if args.config == "c1" else scenes.config0()
    print(f"[bench] scene {spec.name}: {spec.triangles.shape[0]} triangles, {spec.width}x{spec.height}",
          file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    scene = abi.Scene(spec.triangles, spec.materials, stream=int(stream.cuda_stream), bvh=args.bvh)
    inputs = render.DeviceInputs(spec.camera, spec.params)
    acc = render.Accumulator.allocate(spec.width, spec.height)
    passes = spec.spp if spec.spp else inputs.max_spp
    print(f"[bench] BVH build ({args.bvh}) {scene.build_ms:.2f} ms (GPU)", file=sys.stderr, flush=True)
    L = abi.lib()

    def step():
        render.render_wavefront(scene, acc, inputs, passes=passes, max_paths=args.max_paths)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # traversal counters of one step (a counted run of the same work, outside the timed region)
    abi.check(L.pathTracerWavefrontProfile(0, 1), "pathTracerWavefrontProfile")
    step()
    torch.cuda.synchronize()
    prof = profile_read()
    ext_nodes, ext_tris, ext_rays, sh_nodes, sh_tris, sh_rays = (int(v) for v in prof[8:14])
    node_bytes = WIDE_NODE_BYTES if args.bvh == "sah" else NODE_BYTES  # "sah" traces the 4-wide tree
    trace_bytes = (ext_nodes + sh_nodes) * node_bytes + (ext_tris + sh_tris) * TRI_BYTES \
        + ext_rays * EXT_RAY_BYTES + sh_rays * SHADOW_RAY_BYTES

    # timed region: K steps, kernel durations from hipEvents on the render stream
    abi.check(L.pathTracerWavefrontProfile(1, 0), "pathTracerWavefrontProfile")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = profile_read()
    abi.check(L.pathTracerWavefrontProfile(0, 0), "pathTracerWavefrontProfile")
    samples_per_step = int(acc.counts.to(torch.int64).sum().item())
    samples = samples_per_step * args.steps
    ms_ext, ms_sh, ms_shade, ms_other = prof[0:4]
    n_ext, n_sh, n_shade, _ = (int(v) for v in prof[4:8])
    started = int(prof[14])
    trace_ms_per_step = (ms_ext + ms_sh) / args.steps
    achieved = trace_bytes / (trace_ms_per_step / 1e3) / 1e9
    print(f"[bench] wavefront: {args.steps} steps, {samples_per_step} samples/step, wall {wall:.3f} s; "
          f"trace ext {ms_ext / args.steps:.1f} ms/step ({n_ext // args.steps} launches), shadow "
          f"{ms_sh / args.steps:.1f} ms/step, shade {ms_shade / args.steps:.1f} ms/step, gen+accumulate "
          f"{ms_other / args.steps:.1f} ms/step; {started / args.steps:.0f} samples started/step", file=sys.stderr,
          flush=True)

    extra = {}
    if not args.no_loop:
        torch.cuda.synchronize()
        tp = time.perf_counter()
        render.render_persistent(scene, acc, inputs, passes=passes)
        torch.cuda.synchronize()
        extra["persistent_megakernel_msamples_s"] = int(acc.counts.to(torch.int64).sum().item()) / \
            (time.perf_counter() - tp) / 1e6
        tl0 = time.perf_counter()
        render.render_loop(scene, acc, inputs)
        torch.cuda.synchronize()
        extra["reference_cadence_msamples_s"] = int(acc.counts.to(torch.int64).sum().item()) / \
            (time.perf_counter() - tl0) / 1e6
        if args.bvh == "sah":
            # the same wavefront render over the reference's median-split tree
            med = abi.Scene(spec.triangles, spec.materials, stream=int(stream.cuda_stream), bvh="median")
            render.render_wavefront(med, acc, inputs, passes=passes)
            torch.cuda.synchronize()
            tm = time.perf_counter()
            render.render_wavefront(med, acc, inputs, passes=passes)
            torch.cuda.synchronize()
            extra["median_split_bvh_msamples_s"] = int(acc.counts.to(torch.int64).sum().item()) / \
                (time.perf_counter() - tm) / 1e6
            extra["median_split_bvh_build_ms"] = round(med.build_ms, 3)
            del med
        print(f"[bench] persistent megakernel {extra['persistent_megakernel_msamples_s']:.1f}, reference cadence "
              f"(one launch per pass) {extra['reference_cadence_msamples_s']:.1f} Msamples/s", file=sys.stderr,
              flush=True)
        if args.config == "c1":
            extra["facade"] = facade_bench(spec)

    if not args.no_mis and args.config == "c1":
        extra["light_sampling"] = light_sampling_bench(scene, spec)

    if not args.no_spectral and args.config == "c1":
        extra["spectral"] = spectral_bench(scene, spec, inputs, acc, passes, args.bvh)

    if not args.no_batch and args.config == "c1":
        extra["batch"] = batch_bench(args.bvh)

    if not args.no_adjoint and args.config == "c1":
        extra["adjoint"] = adjoint_bench(spec, args.bvh, args.cpu_threads or host_cores())

    if not args.no_inverse and args.config == "c1":
        extra["inverse"] = inverse_bench(args.bvh, args.inverse_steps)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or host_cores()
        cpu = cpu_baseline(spec, threads)
        print(f"[bench] cpu baseline {cpu['value']:.3f} Msamples/s on {threads} threads ({cpu['sample']})",
              file=sys.stderr, flush=True)

    if args.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] copy calibration: 536870912 bytes read, 536870912 bytes written", file=sys.stderr, flush=True)

    adaptive = spec.spp is None
    out = {
        "metric": METRIC,
        "value": samples / wall / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic L-system scene generator, seeded)",
        "config": {"workload": f"{spec.name}: plain forward render, {spec.width}x{spec.height}, "
                               f"{spec.triangles.shape[0]} triangles, "
                               + ("adaptive 16-256 spp at relative error 0.01, max depth 12, 2 area lights + "
                                  "uniform environment" if adaptive else f"{spec.spp} spp, max depth 8"),
                   "samples_per_step": samples_per_step, "bvh": args.bvh, "bvh_build_ms": round(scene.build_ms, 3),
                   "scheduler": "wavefront (first group: the samples no pixel can skip; then 4 samples per "
                                "active pixel between convergence tests, widened speculatively when few remain)",
                   "samples_started_per_step": started // args.steps, **extra},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(f"wf_trace_{spec.name.lower()}_traffic.txt"),
                     "traffic_unit": "HBM bytes per trace launch (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by a "
                                     "512 MiB copy; profiles/wf_trace_c1_traffic.txt)",
                     "algorithmic_bytes_per_launch": trace_bytes / max(1, (n_ext + n_sh) // args.steps),
                     "kernel": ("wf_trace_kernel<false,false,false,W> (extension rays) + "
                                "wf_trace_kernel<true,false,false,W> (shadow rays), W = 4-wide tree"
                                if args.bvh == "sah" else
                                "wf_trace_kernel<false,false,false,false> + wf_trace_kernel<true,false,false,false>"),
                     "kernel_ms_per_step": trace_ms_per_step,
                     "launches_per_step": (n_ext + n_sh) / args.steps,
                     "bytes_per_step": trace_bytes,
                     "bytes_per_step_unit": "algorithmic bytes of all trace launches of one step (model below)",
                     "model": {"ext_nodes": ext_nodes, "ext_triangles": ext_tris, "ext_rays": ext_rays,
                               "shadow_nodes": sh_nodes, "shadow_triangles": sh_tris, "shadow_rays": sh_rays,
                               "node_bytes": node_bytes, "tri_bytes": TRI_BYTES, "ext_ray_bytes": EXT_RAY_BYTES,
                               "shadow_ray_bytes": SHADOW_RAY_BYTES}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
