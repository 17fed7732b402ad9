"""Seeded generators: known answers and invariants."""
import numpy as np

from cookiedough_amd import workloads as wl


def test_splitmix64_known_answer():
    # published first outputs of splitmix64 seeded with 0
    out = wl.splitmix64(0, 3)
    assert [hex(int(v)) for v in out] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4",
                                          "0x6c45d188009454f"]


def test_streams_are_deterministic_and_distinct():
    a = wl.Rng(5, 1).uniform(8)
    b = wl.Rng(5, 1).uniform(8)
    c = wl.Rng(5, 2).uniform(8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert ((a >= 0) & (a < 1)).all()


def test_scene_distributions_and_prefix_property():
    s8, s32 = wl.sdf_scene(0, 8), wl.sdf_scene(0, 32)
    for i in range(32):
        cx, cy, cz, r = s32.sphere[i]
        assert -4 <= cx <= 4 and -4 <= cy <= 4 and -4 <= cz <= 4 and 0.3 <= r <= 1.2
    for i in range(8):
        assert list(s8.sphere[i]) == list(s32.sphere[i])


def test_landscape_maps_shapes_and_palette():
    m = wl.landscape_maps(0)
    assert m.height.shape == (1024, 1024) and m.height.dtype == np.uint8
    assert m.height.min() == 0 and m.height.max() == 255
    assert m.color.shape == (1024, 1024) and (m.color >> 24 == 255).all()
    assert m.fog.shape == (256,)
    # the noise tiles (maps wrap with mask 1023)
    assert abs(int(m.height[0, 0]) - int(m.height[0, -1])) < 24


def test_tunnel_grid_shape_and_range():
    g = wl.tunnel_grid(1280, 720, 5.0)
    assert g.shape == (91, 161, 2) and g.dtype == np.int32
    assert np.abs(g).max() < 2 ** 25  # the 8x8 interpolation multiplies by 64


def test_batch_mix_and_resolutions():
    items = wl.batch_requests(0, 512)
    eff = np.array([i.effect for i in items])
    frac = [(eff == k).mean() for k in range(3)]
    assert abs(frac[0] - 0.4) < 0.07 and abs(frac[1] - 0.3) < 0.07 and abs(frac[2] - 0.3) < 0.07
    assert {(i.W, i.H) for i in items} == set(wl.RESOLUTIONS)
    assert all(0 <= i.t <= 60 and 1 <= i.radius <= 32 for i in items)
