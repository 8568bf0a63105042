"""A CPU restatement of the library's binned-SAH build (lsystemeditor_amd/csrc/pt_sah.hip), used by
the tests to pin the GPU tree's topology: the same float32 arithmetic, orders and tie rules,
written sequentially.

1. Centroids ((v0 + v1) + v2) * (1/3); 30-bit Morton codes over the centroid bounds (10 bits per
   axis, u = clamp((c - lo) / ext * 1024, 0, 1023)); positions in stable code order.
2. Levels of segments, top down. A segment of one triangle is a leaf (slot = its position). Other
   segments: centroid extent per axis, the longest axis (ties keep the lower axis); unless the
   extent is 0 or the depth cap (level + ceil(log2 count) >= 48) applies, 32 bins of
   b = clamp(int((c - k0) * (32 / ext)), 0, 31); the first strict minimum over boundaries of
   area(left) * nLeft + area(right) * nRight (half areas dx dy + dy dz + dz dx of the bins' triangle
   bounds) with both sides non-empty; otherwise the middle position (count // 2).
3. A stable partition moves each segment's left side to its front; the level's internal segments
   get consecutive node numbers in segment order, and their children form the next level in that
   order (left, right).
Output: nodes (n - 1, 14) float32 words (left box, right box, left ref, right ref as int bits; refs
< 0 are leaf slots ~slot), slot -> triangle, the root box and reference, and the level count.
"""
from __future__ import annotations

import numpy as np

F = np.float32
BINS = 32
MAX_DEPTH = 48


def _expand_bits(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64)
    v = (v * 0x00010001) & 0xFF0000FF
    v = (v * 0x00000101) & 0x0F00F00F
    v = (v * 0x00000011) & 0xC30C30C3
    v = (v * 0x00000005) & 0x49249249
    return v.astype(np.uint32)


def _ceil_log2(c: int) -> int:
    d = 0
    while (1 << d) < c:
        d += 1
    return d


def _half_area(lo, hi) -> np.float32:
    dx, dy, dz = F(hi[0] - lo[0]), F(hi[1] - lo[1]), F(hi[2] - lo[2])
    return F(F(F(dx * dy) + F(dy * dz)) + F(dz * dx))


def build(tris: np.ndarray):
    n = tris.shape[0]
    v = np.ascontiguousarray(tris[:, 0:9], np.float32).reshape(n, 3, 3)
    third = F(1.0) / F(3.0)
    cent = ((v[:, 0] + v[:, 1]) + v[:, 2]) * third
    tlo = v.min(1)
    thi = v.max(1)
    glo, ghi = cent.min(0), cent.max(0)
    code = np.zeros(n, np.uint32)
    for k in range(3):
        ext = F(ghi[k] - glo[k])
        u = ((cent[:, k] - glo[k]) / ext).astype(np.float32) if ext > 0 else np.zeros(n, np.float32)
        u = np.minimum(np.maximum(u * F(1024.0), F(0.0)), F(1023.0))
        code |= _expand_bits(u.astype(np.uint32)) << np.uint32(2 - k)
    perm = np.argsort(code, kind="stable").astype(np.int64)

    nodes = np.zeros((max(n - 1, 0), 14), np.float32)
    refs = nodes[:, 12:14].view(np.int32)
    root_box = np.zeros(6, np.float32)
    root_ref = 0
    segs = [(0, n, -1, 0)]            # begin, count, parent, side
    level = 0
    level_base = 0
    levels_used = 0
    while segs:
        internal = []                 # (segment, leftCount, flags)
        for begin, count, parent, side in segs:
            tri = perm[begin:begin + count]
            box = np.concatenate([tlo[tri].min(0), thi[tri].max(0)]).astype(np.float32)
            if parent < 0:
                root_box[:] = box
            else:
                nodes[parent, 6 * side:6 * side + 6] = box
            if count == 1:
                ref = ~begin
                if parent < 0:
                    root_ref = ref
                else:
                    refs[parent, side] = ref
                continue
            c = cent[tri]
            cmin, cmax = c.min(0), c.max(0)
            ext = (cmax - cmin).astype(np.float32)
            axis = 0
            if ext[1] > ext[axis]:
                axis = 1
            if ext[2] > ext[axis]:
                axis = 2
            left = None
            if ext[axis] > 0 and level + _ceil_log2(count) < MAX_DEPTH:
                k0 = cmin[axis]
                scale = F(BINS) / ext[axis]
                b = np.clip(((c[:, axis] - k0) * scale).astype(np.float32).astype(np.int64), 0, BINS - 1)
                cnt = np.bincount(b, minlength=BINS)
                blo = np.full((BINS, 3), np.inf, np.float32)
                bhi = np.full((BINS, 3), -np.inf, np.float32)
                np.minimum.at(blo, b, tlo[tri])
                np.maximum.at(bhi, b, thi[tri])
                suffix = np.zeros(BINS, np.float32)
                suffix_n = np.zeros(BINS, np.int64)
                lo3 = np.full(3, np.inf, np.float32)
                hi3 = np.full(3, -np.inf, np.float32)
                acc = 0
                for k in range(BINS - 1, -1, -1):
                    if cnt[k]:
                        lo3 = np.minimum(lo3, blo[k])
                        hi3 = np.maximum(hi3, bhi[k])
                        acc += int(cnt[k])
                    suffix_n[k] = acc
                    suffix[k] = _half_area(lo3, hi3) if acc > 0 else F(0.0)
                lo3 = np.full(3, np.inf, np.float32)
                hi3 = np.full(3, -np.inf, np.float32)
                acc = 0
                best, best_bin, best_left = F(np.inf), -1, 0
                for k in range(BINS - 1):
                    if cnt[k]:
                        lo3 = np.minimum(lo3, blo[k])
                        hi3 = np.maximum(hi3, bhi[k])
                        acc += int(cnt[k])
                    right = int(suffix_n[k + 1])
                    if acc == 0 or right == 0:
                        continue
                    cost = F(F(_half_area(lo3, hi3) * F(acc)) + F(suffix[k + 1] * F(right)))
                    if cost < best:
                        best, best_bin, best_left = cost, k, acc
                if best_bin >= 0:
                    left = (b <= best_bin, best_left)
            if left is None:
                flags = np.arange(count) < count // 2
                left = (flags, count // 2)
            internal.append(((begin, count, parent, side), left))
        if not internal:
            break
        levels_used = level + 1
        nxt = []
        for rank, ((begin, count, parent, side), (flags, lc)) in enumerate(internal):
            idx = level_base + rank
            if parent < 0:
                root_ref = idx
            else:
                refs[parent, side] = idx
            seg = perm[begin:begin + count]
            perm[begin:begin + count] = np.concatenate([seg[flags], seg[~flags]])
            nxt.append((begin, lc, idx, 0))
            nxt.append((begin + lc, count - lc, idx, 1))
        level_base += len(internal)
        level += 1
        segs = nxt
    return nodes, perm.astype(np.int32), root_box, int(root_ref), levels_used + 1
