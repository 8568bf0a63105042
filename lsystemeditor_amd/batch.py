"""Batched rendering of many L-system scenes of uneven size. Two schedules:

- batch wavefront (render_batched / capture_batched): the scenes form one batch
  (ptBatchCreate) and every bounce's trace and shade launches serve the paths of all scenes, so
  the latency tail of a launch (its slowest rays) is paid once per bounce for the whole batch;
- per-scene sequence (render_all / capture): one HIP graph replays every scene's own fixed-spp
  wavefront render back to back.

Both need no host synchronisation; the paths-in-flight bound is picked by a timing autotuner.
With a binned-SAH tree, capture_build_and_render records every scene's tree rebuild
(ptSceneRebuild, one shared workspace) and then the batch wavefront in one graph, so one replay
builds all 64 trees from the scenes' device triangles and renders them.

Workload C3: 64 stochastic bushes, seeds 0..63, iterations uniform in 3..7 (3 k .. 2 M triangles),
resolutions drawn from 128^2 .. 1024^2, 64 spp, max depth 8.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import abi, render, scenes

RESOLUTIONS = (128, 256, 512, 1024)


def batch_specs(count: int = 64, spp: int = 64, depth: int = 8, first_seed: int = 0,
                iterations=(3, 7)) -> list[scenes.SceneSpec]:
    """Scene s: iterations and resolution drawn with SplitMix64 from the seed; the bush itself is
    the C1 generator with that seed."""
    out = []
    lo, hi = iterations
    for s in range(first_seed, first_seed + count):
        g = scenes.SplitMix64(0xC3000000 + s)
        it = lo + g.next() % (hi - lo + 1)
        res = RESOLUTIONS[g.next() % len(RESOLUTIONS)]
        spec = scenes.config1(width=res, height=res, iterations=int(it), seed=s, min_spp=spp, max_spp=spp,
                              threshold=0.01)
        spec.params[4:5] = np.array([depth], np.int32).view(np.float32)
        spec.spp = spp
        spec.info = dict(spec.info, resolution=res)
        out.append(spec)
    return out


class BatchRenderer:
    """Scenes, accumulators and device inputs of a batch, plus its captured graph."""

    def __init__(self, specs: list[scenes.SceneSpec], bvh: str = "sah", max_paths: int = 0):
        self.specs = specs
        self.stream = torch.cuda.Stream()
        h = int(self.stream.cuda_stream)
        self.scenes = [abi.Scene(s.triangles, s.materials, stream=h, bvh=bvh) for s in specs]
        self.build_ms = float(sum(sc.build_ms for sc in self.scenes))
        self.inputs = [render.DeviceInputs(s.camera, s.params) for s in specs]
        self.accs = [render.Accumulator.allocate(s.width, s.height) for s in specs]
        self.max_paths = max_paths
        self.graph = None
        self.batch = None
        self.batch_graph = None
        self.build_graph = None
        self.workspaces = None  # build workspaces of the rebuild branches
        self.branch_streams = []
        self._pins = 0          # workspace pins taken by this renderer's captures
        torch.cuda.synchronize()

    @property
    def samples(self) -> int:
        return int(sum(s.width * s.height * s.spp for s in self.specs))

    def _reserve(self):
        depth = max(render.resolved_depth(i.params_host) for i in self.inputs)
        pixels = max(s.width * s.height for s in self.specs)
        paths = self.max_paths if self.max_paths > 0 else (1 << 26)
        paths = max(paths, pixels)
        abi.check(abi.lib().pathTracerWavefrontReserve(paths, pixels, depth), "pathTracerWavefrontReserve")

    def render_all(self):
        """Every scene's fixed-spp render, back to back on the batch stream (no graph)."""
        self._reserve()
        with torch.cuda.stream(self.stream):
            for sc, acc, inp, spec in zip(self.scenes, self.accs, self.inputs, self.specs):
                render.render_fixed(sc, acc, inp, spec.spp, max_paths=self.max_paths, stream=self.stream)

    def capture(self):
        """Records render_all into one graph."""
        self._reserve()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        p0 = render.workspace_pins()
        with torch.cuda.graph(g, stream=self.stream):
            for sc, acc, inp, spec in zip(self.scenes, self.accs, self.inputs, self.specs):
                render.render_fixed(sc, acc, inp, spec.spp, max_paths=self.max_paths,
                                    stream=torch.cuda.current_stream())
        self._pins += render.workspace_pins() - p0
        self.graph = g

    def replay(self):
        self.graph.replay()

    def _batch(self) -> render.SceneBatch:
        if self.batch is None:
            self.batch = render.SceneBatch(self.scenes, self.inputs, [(s.width, s.height) for s in self.specs])
        return self.batch

    def render_batched(self, max_paths: int | None = None):
        """Every scene's fixed-spp render in one batch wavefront (spp of the first scene; all
        scenes of a batch take the same count)."""
        b = self._batch()
        mp = self.max_paths if max_paths is None else max_paths
        b.reserve(mp)
        with torch.cuda.stream(self.stream):
            b.render_fixed(self.specs[0].spp, max_paths=mp, stream=self.stream)

    def capture_batched(self):
        """Records render_batched into one graph."""
        b = self._batch()
        b.reserve(self.max_paths)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        p0 = render.workspace_pins()
        with torch.cuda.graph(g, stream=self.stream):
            b.render_fixed(self.specs[0].spp, max_paths=self.max_paths, stream=torch.cuda.current_stream())
        self._pins += render.workspace_pins() - p0
        self.batch_graph = g

    def replay_batched(self):
        self.batch_graph.replay()

    def capture_build_and_render(self, cold: bool = False, branches: int = 4):
        """Records every scene's tree rebuild and then render_batched into one graph (binned-SAH
        build modes only). The rebuilds enqueue the levels each scene's first build needed (a
        rebuild of unchanged geometry), or with cold every level the depth cap allows (what a build
        of new geometry enqueues). They form `branches` parallel branches of the graph, each on a
        stream of its own with a build workspace of its own (scenes go largest first to the branch
        with the fewest triangles), so that one branch's small-scene level kernels overlap the
        others'; the batch render waits for all of them."""
        if any(sc.bvh not in ("sah", "sah-binary") for sc in self.scenes):
            raise ValueError("tree rebuilds need a binned-SAH build mode")
        b = self._batch()
        b.reserve(self.max_paths)
        k = max(1, min(int(branches), len(self.scenes)))
        groups = [[] for _ in range(k)]
        loads = [0] * k
        for i in sorted(range(len(self.scenes)), key=lambda i: -self.scenes[i].triangle_count):
            j = loads.index(min(loads))
            groups[j].append(i)
            loads[j] += self.scenes[i].triangle_count
        need = [max(self.scenes[i].triangle_count for i in g) for g in groups]
        if self.workspaces is None or len(self.workspaces) < k or \
                any(ws.max_triangles < n for ws, n in zip(self.workspaces, need)):
            self.workspaces = None
            torch.cuda.synchronize()
            self.workspaces = [abi.BuildWorkspace(n) for n in need]
        if len(self.branch_streams) < k:
            self.branch_streams += [torch.cuda.Stream() for _ in range(k - len(self.branch_streams))]
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        p0 = render.workspace_pins()
        with torch.cuda.graph(g, stream=self.stream):
            s = torch.cuda.current_stream()
            for grp, ws, bs in zip(groups, self.workspaces, self.branch_streams):
                bs.wait_stream(s)
                with torch.cuda.stream(bs):
                    for i in grp:
                        self.scenes[i].rebuild(ws, bs.cuda_stream, cold=cold)
            for bs in self.branch_streams[:k]:
                s.wait_stream(bs)
            b.render_fixed(self.specs[0].spp, max_paths=self.max_paths, stream=s)
        self._pins += render.workspace_pins() - p0
        self.build_graph = g

    def release_graphs(self):
        """Drops this renderer's captured graphs and the workspace pins they took, so later calls
        may grow the shared workspace again."""
        torch.cuda.synchronize()
        self.graph = self.batch_graph = self.build_graph = None
        if self._pins:
            render.release_workspace_pins(self._pins)
            self._pins = 0

    def replay_build_and_render(self):
        self.build_graph.replay()

    def check_rebuild(self):
        """Waits for the batch stream; raises unless every scene's last rebuild reproduced its tree."""
        for sc in self.scenes:
            sc.rebuild_check(self.workspaces[0], int(self.stream.cuda_stream))

    def batched_image(self, i: int) -> torch.Tensor:
        """Scene i's (pixels, 4) accumulator rows of the last batch render."""
        return self._batch().acc.acc[self._batch().pixels(i)]

    def render_sequential_adaptive_api(self):
        """The reference-style sequence: one plain render per scene through the adaptive
        wavefront entry point, with its host synchronisation per group."""
        with torch.cuda.stream(self.stream):
            for sc, acc, inp, spec in zip(self.scenes, self.accs, self.inputs, self.specs):
                render.render_wavefront(sc, acc, inp, passes=spec.spp, stream=self.stream)


def autotune_paths(spec: scenes.SceneSpec, candidates=(1 << 22, 1 << 24, 1 << 25, 1 << 26), bvh: str = "sah",
                   reps: int = 2) -> tuple[int, dict]:
    """Times the fixed-spp render of one scene for each paths-in-flight candidate; returns the
    fastest and all timings (ms)."""
    sc = abi.Scene(spec.triangles, spec.materials, bvh=bvh)
    inp = render.DeviceInputs(spec.camera, spec.params)
    acc = render.Accumulator.allocate(spec.width, spec.height)
    times = {}
    for c in candidates:
        abi.check(abi.lib().pathTracerWavefrontReserve(max(c, spec.width * spec.height), spec.width * spec.height,
                                                        render.resolved_depth(inp.params_host)),
                  "pathTracerWavefrontReserve")
        render.render_fixed(sc, acc, inp, spec.spp, max_paths=c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            render.render_fixed(sc, acc, inp, spec.spp, max_paths=c)
        torch.cuda.synchronize()
        times[c] = (time.perf_counter() - t0) / reps * 1e3
    best = min(times, key=times.get)
    return best, times


def autotune_batch_paths(br: BatchRenderer, candidates=(1 << 24, 1 << 25, 1 << 26, 1 << 27)) -> tuple[int, dict]:
    """Times render_batched for each paths-in-flight candidate (at least the batch's pixel
    count); returns the fastest and all timings (ms)."""
    times = {}
    for c in candidates:
        br.render_batched(max_paths=c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        br.render_batched(max_paths=c)
        torch.cuda.synchronize()
        times[c] = (time.perf_counter() - t0) * 1e3
    best = min(times, key=times.get)
    return best, times
