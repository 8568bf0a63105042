"""pathTracerCompactActiveListNoAlloc: the allocation-free compaction against
pathTracerCompactActiveList (same kernels, library-allocated tile counts) on the same inputs.

Bar: bit-exact (index work). Two calls in a row, the second with a smaller input, so tile counts
left in the fixed device array by the first call would show if the second read them; then a
region-filtered call; then an input beyond the array's maximum, which must be refused on the host.
"""
import numpy as np
import pytest

from lsystemeditor_amd import abi

pytestmark = pytest.mark.gpu

MAX_ENTRIES = 65536 * 1024


def _state(torch, w, h, seed, max_spp):
    rng = np.random.default_rng(seed)
    n = w * h
    conv = (rng.random(n) < 0.45).astype(np.uint8)
    counts = rng.integers(0, max_spp + 8, n).astype(np.uint32)
    dev = lambda a: torch.from_numpy(a).cuda()
    return dev(conv), dev(counts.view(np.int32)), conv, counts


def _compact(torch, fn, active, conv, counts, max_spp, n, w, h, region):
    out = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")
    cnt = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    r = region or (0, 0, 0, 0)
    ok = fn(active.data_ptr(), out.data_ptr(), cnt.data_ptr(), conv.data_ptr(), counts.data_ptr(), max_spp, n, w, h,
            region is not None, r[0], r[1], r[2], r[3], None)
    assert ok, abi.last_error()
    torch.cuda.synchronize()
    k = int(cnt.item())
    return k, out[:k].cpu().numpy()


def _want(active, conv, counts, max_spp, w, region):
    keep = []
    for p in active:
        if p < 0 or conv[p] or (max_spp > 0 and counts[p] >= max_spp):
            continue
        if region is not None:
            x, y = p % w, p // w
            if x < region[0] or x > region[2] or y < region[1] or y > region[3]:
                continue
        keep.append(p)
    return np.array(keep, dtype=np.int32)


@pytest.mark.parametrize("region", [None, (37, 11, 900, 700)])
def test_no_alloc_compaction_matches_allocating_entry_point(cuda, region):
    torch = cuda
    L = abi.lib()
    max_spp = 64
    # first call: a 1024 x 1024 frame's list (1024 tiles); second call: a smaller list whose
    # tiles the first call's tile counts cover, in a different frame
    for (w, h, frac, seed) in ((1024, 1024, 1.0, 1), (1000, 777, 0.3, 2)):
        print(f"no-alloc compaction: {w}x{h}, fraction {frac}, region {region}", flush=True)
        conv_d, counts_d, conv, counts = _state(torch, w, h, seed, max_spp)
        rng = np.random.default_rng(seed + 100)
        n = int(frac * w * h)
        active = rng.permutation(w * h)[:n].astype(np.int32)
        active[::97] = -1   # entries below zero are dropped
        active_d = torch.from_numpy(active).cuda()
        k_ref, got_ref = _compact(torch, L.pathTracerCompactActiveList, active_d, conv_d, counts_d, max_spp, n, w, h,
                                  region)
        k_na, got_na = _compact(torch, L.pathTracerCompactActiveListNoAlloc, active_d, conv_d, counts_d, max_spp, n,
                                w, h, region)
        assert k_na == k_ref
        assert np.array_equal(got_na, got_ref)
        want = _want(active, conv, counts, max_spp, w, region)
        assert k_ref == want.size and np.array_equal(got_ref, want)


def test_no_alloc_compaction_identity_and_tiny(cuda):
    """Lists of one entry and of one partial tile, after a large call in the same process."""
    torch = cuda
    L = abi.lib()
    w, h = 33, 3
    conv_d, counts_d, conv, counts = _state(torch, w, h, 5, 16)
    for n in (w * h, 1):
        active = np.arange(n, dtype=np.int32)
        k_na, got = _compact(torch, L.pathTracerCompactActiveListNoAlloc, torch.from_numpy(active).cuda(), conv_d,
                             counts_d, 16, n, w, h, None)
        want = _want(active, conv, counts, 16, w, None)
        assert k_na == want.size and np.array_equal(got, want)


def test_no_alloc_compaction_refuses_oversize(cuda):
    torch = cuda
    L = abi.lib()
    buf = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = buf.data_ptr()
    ok = L.pathTracerCompactActiveListNoAlloc(p, p, p, p, p, 16, MAX_ENTRIES + 1, 8192, 8193, False, 0, 0, 0, 0, None)
    assert ok is False
    assert "exceed the maximum" in abi.last_error()
