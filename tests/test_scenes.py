"""Scene generator: counts, determinism (triangle-byte SHA-256) and generator known answers."""
import numpy as np

from lsystemeditor_amd import scenes

C0_SHA256 = "b948139ea4274a06"  # prefix of the triangle-byte digest, recorded when the generator was written


def test_splitmix64_known_answers():
    # SplitMix64 reference outputs for seed 1234567 (Vigna's splitmix64.c test sequence)
    g = scenes.SplitMix64(1234567)
    assert [g.next() for _ in range(3)] == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_lsystem_expansion():
    ls = scenes.LSystem("F", {"F": [(1.0, "F[+F]F[-F]F")]}, 2)
    s = ls.expand()
    assert s.count("F") == 25 and s.count("[") == s.count("]") == 12  # 2 + 5 * 2


def test_config0_geometry():
    s = scenes.config0()
    assert s.info["segments"] == 625
    assert s.triangles.shape == (625 * 16 + 4, 25)
    assert s.sha256().startswith(C0_SHA256)
    mats = s.triangles[:, 24].view(np.uint32)
    assert np.bincount(mats).tolist() == [10000, 2, 2]
    assert s.params[4:8].view(np.int32).tolist() == [8, 3, 16, 16]


def test_config0_is_deterministic():
    assert scenes.config0().sha256() == scenes.config0().sha256()


def test_cylinder_vertices_on_radius():
    t = scenes.run_turtle("F", scenes.TurtleParams(radius=0.05))
    tris = scenes.cylinders(t, 8, 0)
    v = tris[:, 0:9].reshape(-1, 3).astype(np.float64)
    r = np.linalg.norm(v[:, [0, 2]], axis=1)
    assert np.allclose(r, 0.05, atol=1e-7) and set(np.round(v[:, 1], 6)) == {0.0, 1.0}
