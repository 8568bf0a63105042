"""Oracle math and RNG: accuracy bounds against libm, identities, published known answers."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def ulp_err(got: np.ndarray, exact: np.ndarray) -> np.ndarray:
    """Error in units of the float32 spacing at the exact value."""
    e32 = exact.astype(np.float32)
    spacing = np.spacing(np.abs(e32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / spacing


@pytest.fixture(scope="module")
def L():
    return O.lib()


def sample_inputs(lo, hi, n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, n).astype(np.float32)


@pytest.mark.parametrize("name,ref,lo,hi", [
    ("orc_sinf", np.sin, -8 * math.pi, 8 * math.pi),
    ("orc_cosf", np.cos, -8 * math.pi, 8 * math.pi),
    ("orc_tanf", np.tan, -1.5, 1.5),
    ("orc_expf", np.exp, -87.0, 88.0),
    ("orc_logf", np.log, 1e-30, 1e30),
    ("orc_acosf", np.arccos, -1.0, 1.0),
])
def test_unary_accuracy(L, name, ref, lo, hi):
    fn = getattr(L, name)
    x = sample_inputs(lo, hi, 20000)
    if name == "orc_logf":
        x = np.exp(np.random.default_rng(1).uniform(-80, 80, 20000)).astype(np.float32)
    got = np.array([fn(float(v)) for v in x], dtype=np.float32)
    exact = ref(x.astype(np.float64))
    err = ulp_err(got, exact)
    # evaluated in double and rounded once: within 0.5 ulp up to the double evaluation error
    assert err.max() <= 0.5001, f"{name}: max error {err.max():.4f} ulp at x={x[err.argmax()]!r}"


def test_atan2_accuracy(L):
    rng = np.random.default_rng(2)
    y = rng.normal(size=5000).astype(np.float32)
    x = rng.normal(size=5000).astype(np.float32)
    got = np.array([L.orc_atan2f(float(a), float(b)) for a, b in zip(y, x)], dtype=np.float32)
    err = ulp_err(got, np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert err.max() <= 0.5001


def test_special_values(L):
    assert L.orc_sinf(0.0) == 0.0 and math.copysign(1, L.orc_sinf(-0.0)) == -1
    assert L.orc_cosf(0.0) == 1.0
    assert math.isnan(L.orc_sinf(float("inf"))) and math.isnan(L.orc_cosf(float("nan")))
    assert L.orc_expf(0.0) == 1.0 and L.orc_expf(-200.0) == 0.0 and math.isinf(L.orc_expf(100.5))
    assert L.orc_logf(1.0) == 0.0 and L.orc_logf(0.0) == float("-inf") and math.isnan(L.orc_logf(-1.0))
    assert L.orc_atan2f(0.0, -1.0) == np.float32(math.pi)
    assert L.orc_acosf(1.0) == 0.0 and L.orc_acosf(-1.0) == np.float32(math.pi)
    # odd/even symmetry is exact (the reduction is symmetric)
    for v in sample_inputs(-20, 20, 200, 5):
        assert L.orc_sinf(-float(v)) == -L.orc_sinf(float(v))
        assert L.orc_cosf(-float(v)) == L.orc_cosf(float(v))


def test_expf_subnormal_and_ldexp_path(L):
    for v in (-100.0, -103.0, -87.5, 88.7):
        x = float(np.float32(v))
        assert L.orc_expf(x) == np.float32(math.exp(x))


# Random123 known-answer vectors for philox4x32_10 (Salmon et al., SC'11); the same three
# vectors are produced by rocRAND's philox4x32_10 engine (/opt/rocm/include/rocrand).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(L, ctr, key, want):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    L.orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    assert tuple(int(v) for v in out) == want


def test_seed_hash_matches_reference_formula(L):
    # Sampling/RandCore.h:9-10, u32 wrap-around
    for idx, s, g in [(0, 0, 1), (12345, 7, 3), (1 << 20, 255, 1 << 31), (2**31 - 1, 2**31 - 1, 2**32 - 1)]:
        want = (idx * 747796405 + s + g * 2891336453) % 2**32
        assert L.orc_seed_hash(idx, s, g) == want


def test_uniform_conversion_range(L):
    assert L.orc_u32_to_unit(0) == np.float32(2.0 ** -24)
    assert L.orc_u32_to_unit(0xFFFFFFFF) == 1.0
    assert L.orc_u32_to_unit(0xFF) == np.float32(2.0 ** -24)   # low byte discarded
    assert L.orc_u32_to_unit(1 << 31) == np.float32(0.5 + 2.0 ** -24)


def test_uniform_moments(L):
    c = np.zeros(4, dtype=np.uint32)
    k = np.array([7, 9], dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    vals = []
    for i in range(5000):
        c[0] = i
        L.orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
        vals.extend(L.orc_u32_to_unit(int(v)) for v in out)
    v = np.array(vals)
    assert abs(v.mean() - 0.5) < 0.01 and abs(v.var() - 1 / 12) < 0.005
    assert v.min() > 0.0 and v.max() <= 1.0


def test_normalize_matches_reference_core():
    ref = O.ref_policy()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = O.lib()
    rng = np.random.default_rng(4)
    for v in list(rng.normal(size=(500, 3)).astype(np.float32)) + [np.zeros(3, np.float32)]:
        out = np.zeros(3, dtype=np.float32)
        ref.ref_normalize(np.ascontiguousarray(v).ctypes.data, out.ctypes.data)
        got = L.orc_vec_normalize(O.v3(v)).np()
        assert np.array_equal(got.view(np.uint32), out.view(np.uint32))
