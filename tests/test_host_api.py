"""Host C++ API, CPU side: the scene JSON reader against Python's json
module, the wall and lattice builders against the Python scene generator
and closed forms, the fine-particle cloud against the standard library's
minstd_rand0 + generate_canonical arithmetic, and error behaviour.  No GPU
calls."""
import json
from pathlib import Path

import numpy as np
import pytest

from granular_gpu_amd import host, scenes

SCENES = sorted((Path(__file__).parent / "golden" / "scenes").glob("*.json"))
f32 = np.float32


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_scene_json_matches_python(path):
    s = host.load_scene(path)
    j = json.loads(path.read_text())
    v3 = lambda k: [f32(j[k][a]) for a in "xyz"]
    assert list(s.space_size) == v3("space_size")
    assert list(s.gravity) == v3("gravity")
    assert list(s.box_size) == v3("box_size")
    assert list(s.boundary_translation) == v3("boundary_translation")
    assert list(s.particle_translation) == v3("particle_translation")
    for k, a in (("dt", "dt"), ("sph_spacing", "sph_spacing"), ("init_spacing", "init_spacing"),
                 ("smoothing_radius", "smoothing_radius"), ("cell_length", "cell_length"),
                 ("density", "density")):
        assert f32(getattr(s, a)) == f32(j[k]), k
    assert s.scene == j["scene"] and s.is_adaptive == j["is_adaptive"]
    assert s.obj_file.decode() == j["obj_file"] and s.save_file.decode() == j["save_file"]
    L = f32(j["cell_length"])
    want = [int(np.ceil(f32(j["space_size"][a]) / L)) for a in "xyz"]
    assert list(s.cell_size) == want


def test_scene_json_errors():
    with pytest.raises(host.HostError, match="Unable to open"):
        host.load_scene("/nonexistent/scene.json")
    good = json.loads(SCENES[0].read_text())
    del good["dt"]
    with pytest.raises(host.HostError, match="missing key dt"):
        host.parse_scene(json.dumps(good))
    with pytest.raises(host.HostError):
        host.parse_scene('{"space_size": {"x": 1,')


def test_box_scene_counts():
    # the reference box scene: 50 x 50 x 40 lattice, 51,088 wall particles
    s = host.load_scene(SCENES[[p.stem for p in SCENES].index("box")])
    assert host.scene_lattice(s).shape == (100000, 3)
    assert host.scene_walls(s).shape == (51088, 3)


@pytest.mark.parametrize("name,recipe", [("box", "box"), ("piling", "piling"),
                                         ("corner", "corner"), ("box_opening_up", "box_opening")])
def test_walls_match_python_recipe(name, recipe):
    s = host.load_scene(Path(__file__).parent / "golden" / "scenes" / f"{name}.json")
    w = host.scene_walls(s)
    pw = scenes.box_walls(tuple(s.space_size), s.cell_length, recipe,
                          tuple(s.boundary_translation))
    assert np.array_equal(w, pw)
    # every wall particle lies on a face of the scaled box
    S = np.asarray(s.space_size, np.float32)
    lo = f32(0.005) * S + np.asarray(s.boundary_translation, np.float32)
    hi = lo + f32(0.99) * S
    on_face = np.isclose(w, lo, atol=1e-6) | np.isclose(w, hi, atol=1e-6)
    assert on_face.any(axis=1).all()


@pytest.mark.parametrize("name", ["excavator", "funnel"])
def test_open_top_walls(name):
    # box without its top wall, never translated (main.cu:889-925, 985-1021)
    s = host.load_scene(Path(__file__).parent / "golden" / "scenes" / f"{name}.json")
    w = host.scene_walls(s)
    assert np.array_equal(w, scenes.box_walls(tuple(s.space_size), s.cell_length, name))
    nx, ny, nz = (2 * c for c in scenes.cell_counts(tuple(s.space_size), s.cell_length))
    assert w.shape[0] == 2 * nx * ny + nx * (nz - 2) + 2 * (ny - 2) * (nz - 2)
    S = np.asarray(s.space_size, np.float32)
    lo, hi = f32(0.005) * S, f32(0.99) * S + f32(0.005) * S
    # at the top height only the upper edges of the front and back walls
    at_top = np.isclose(w[:, 1], hi[1], atol=1e-6)
    assert at_top.sum() == 2 * nx
    assert (np.isclose(w[at_top, 2], lo[2], atol=1e-6) |
            np.isclose(w[at_top, 2], hi[2], atol=1e-6)).all()
    assert np.isclose(w[:, 1], lo[1], atol=1e-6).sum() == nx * nz
    # without a point-cloud file the boundary is the walls alone
    pos, animated, anim = host.scene_boundary(s)
    assert np.array_equal(pos, w) and not animated.any() and anim.is_animating == 0


def _scene_with_cloud(name, path):
    s = host.load_scene(Path(__file__).parent / "golden" / "scenes" / f"{name}.json")
    s.obj_file = str(path).encode()
    return s


def _cloud(n=257, seed=3):
    rng = np.random.default_rng(seed)
    return (rng.random((n, 3)) * [0.6, 0.7, 0.4] - [0.3, 0.0, 0.2]).astype(np.float32)


def _float_centroid(c):
    """Insertion-order float32 sum, then division by float(n)."""
    acc = np.zeros(3, np.float32)
    for p in c:
        acc = acc + p
    return acc / f32(len(c))


def test_point_cloud_reader(tmp_path):
    # whitespace-separated triples; reading stops at the first incomplete one
    f = tmp_path / "c.txt"
    f.write_text("0.5 0.25 -1\n  2e-3\t4 5\n\n6 7 8 9 10\n")
    s = _scene_with_cloud("excavator", f)
    pos, animated, anim = host.scene_boundary(s)
    nw = host.scene_walls(s).shape[0]
    assert pos.shape[0] == nw + 3 and anim.num_particles == 3
    want = np.float32([[0.5, 0.25, -1], [2e-3, 4, 5], [6, 7, 8]])
    assert np.array_equal(pos[nw:], want + np.float32([2.5, 0.0, 0.5]))


def test_excavator_bucket(tmp_path):
    c = _cloud()
    f = tmp_path / "bucket.txt"
    f.write_text("".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in c.tolist()))
    s = _scene_with_cloud("excavator", f)
    pos, animated, anim = host.scene_boundary(s)
    w = host.scene_walls(s)
    nw = w.shape[0]
    assert np.array_equal(pos[:nw], w)
    # cloud shifted by (2.5, 0, 0.5) in float32, marked animated
    assert np.array_equal(pos[nw:], c + np.float32([2.5, 0.0, 0.5]))
    assert not animated[:nw].any() and (animated[nw:] == 1).all()
    assert (anim.is_animating, anim.start_index, anim.num_particles) == (1, nw, len(c))
    assert anim.translation_complete == 0 and anim.animation_time == 0
    # rotation centre: centroid of the cloud as read, before the shift
    g = np.array(anim.rotation_center, np.float32)
    assert np.array_equal(g, _float_centroid(c))
    assert np.allclose(g, c.astype(np.float64).mean(0), atol=1e-6)


def test_funnel_rotation(tmp_path):
    c = _cloud(401, seed=5)
    f = tmp_path / "funnel.txt"
    f.write_text("".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in c.tolist()))
    s = _scene_with_cloud("funnel", f)
    pos, animated, anim = host.scene_boundary(s)
    nw = host.scene_walls(s).shape[0]
    got = pos[nw:].astype(np.float64)
    g = _float_centroid(c).astype(np.float64)
    d = c.astype(np.float64) - g
    # +90 degrees about x through the centroid: (y, z) -> (-z, y), then
    # + (1.25, 0.3, 1.5) (main.cu:1043-1070); float32 rounding only
    want = np.stack([d[:, 0], -d[:, 2], d[:, 1]], 1) + g + [1.25, 0.3, 1.5]
    assert np.abs(got - want).max() < 4e-7
    # cos of the float-rounded right angle is not zero: its trace is visible
    # where |y - g.y| is large against |z - g.z|
    assert f32(np.cos(np.float64(f32(np.pi / 2)))) == f32(-4.371139e-08)
    assert (animated[nw:] == 1).all() and anim.num_particles == len(c)
    # the funnel keeps the default rotation centre (main.cu:1036-1041)
    assert list(anim.rotation_center) == [0.0, 0.0, 0.0]
    assert (anim.is_animating, anim.start_index) == (1, nw)


REF_OBJ = Path("/root/reference/obj")


@pytest.mark.skipif(not REF_OBJ.is_dir(), reason="reference point clouds not present")
@pytest.mark.parametrize("name,count", [("excavator", 13045), ("funnel", 83581)])
def test_reference_point_clouds(name, count):
    fname = {"excavator": "bucket.txt", "funnel": "funnel.txt"}[name]
    s = _scene_with_cloud(name, REF_OBJ / fname)
    pos, animated, anim = host.scene_boundary(s)
    nw = host.scene_walls(s).shape[0]
    assert anim.num_particles == count and pos.shape[0] == nw + count
    raw = np.loadtxt(REF_OBJ / fname, dtype=np.float32)
    if name == "excavator":
        assert np.array_equal(pos[nw:], raw + np.float32([2.5, 0.0, 0.5]))
    else:
        ext = pos[nw:].max(0) - pos[nw:].min(0)
        rext = raw.max(0) - raw.min(0)
        assert np.allclose(ext, [rext[0], rext[2], rext[1]], rtol=1e-5)


def test_lattice_formula():
    s = host.load_scene(SCENES[[p.stem for p in SCENES].index("box")])
    L = host.scene_lattice(s)
    sp, t = f32(s.init_spacing), np.asarray(s.particle_translation, np.float32)
    nx, ny, nz = (int(v) for v in s.box_size)
    i, j, k = 3, 7, 11
    p = L[(i * ny + j) * nz + k]
    want = np.array([f32(0.27) + sp * f32(j), f32(0.13) + sp * f32(i),
                     f32(0.17) + sp * f32(k)], np.float32) + t
    assert np.array_equal(p, want)


def _minstd_canonical(count):
    """minstd_rand0 (x <- 16807 x mod 2^31-1, seed 1) mapped to [0,1) as
    float(x - 1) / float(2^31 - 2) in float arithmetic."""
    x, out = 1, []
    for _ in range(count):
        x = (16807 * x) % 2147483647
        out.append(f32(x - 1) / f32(2147483646.0))
    return np.array(out, np.float32)


def test_upsampled_cloud_generator():
    c = np.float32([[0.5, 0.6, 0.7], [0.1, 0.2, 0.3]])
    r, per = 0.01, 4
    got = host.upsampled_cloud(c, r, per)
    u = _minstd_canonical(2 * per * 3).reshape(2, per, 3)
    off = u * (f32(r) - f32(-r)) + f32(-r)
    want = (c[:, None, :] + off).reshape(-1, 3)
    assert np.array_equal(got, want)
    assert np.all(np.abs(got.reshape(2, per, 3) - c[:, None]) <= r)


def test_vtk_writer_format(tmp_path):
    # the reference's point files (main.cu:188-253): NaN points dropped,
    # "x z y" with the stream's default float formatting, one vertex cell each
    pts = np.float32([[0.1, 0.2, 0.3], [np.nan, 0.0, 0.0], [1.5, -2.25, 3e-7]])
    assert host.write_vtk(tmp_path / "a.vtk", pts) == 2
    lines = (tmp_path / "a.vtk").read_text().splitlines()
    assert lines[:5] == ["# vtk DataFile Version 3.0", "Upsampled Granular Particles", "ASCII",
                         "DATASET UNSTRUCTURED_GRID", "POINTS 2 float"]
    assert lines[5] == "0.1 0.3 0.2"
    assert lines[6] == "1.5 3e-07 -2.25"
    assert lines[7:] == ["CELLS 2 4", "1 0", "1 1", "CELL_TYPES 2", "1", "1"]
    # every point NaN: the file is created and left empty, as in the reference
    assert host.write_vtk(tmp_path / "b.vtk", np.full((2, 3), np.nan, np.float32)) == 0
    assert (tmp_path / "b.vtk").read_text() == ""
    with pytest.raises(host.HostError):
        host.write_vtk(tmp_path / "no_such_dir" / "c.vtk", pts)


def test_checkpoint_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n, nb, nf = 1000, 300, 8000
    arrs = dict(pos=rng.normal(size=(n, 3)).astype(np.float32),
                vel=rng.normal(size=(n, 3)).astype(np.float32),
                mass=rng.uniform(0.5, 4, n).astype(np.float32),
                adaptive_last_step=rng.integers(0, 2, n).astype(np.int32),
                boundary=rng.normal(size=(nb, 3)).astype(np.float32),
                boundary_animated=rng.integers(0, 2, nb).astype(np.int32),
                fine_pos=rng.normal(size=(nf, 3)).astype(np.float32),
                fine_vel=rng.normal(size=(nf, 3)).astype(np.float32))
    arrs["pos"][5] = np.nan  # bit patterns travel as they are
    path = tmp_path / "state.ggck"
    host.write_checkpoint(path, **arrs)
    got = host.read_checkpoint(path)
    pairs = {("granular", "pos"): "pos", ("granular", "vel"): "vel",
             ("granular", "mass"): "mass", ("granular", "adaptive_last_step"):
             "adaptive_last_step", ("boundary", "pos"): "boundary",
             ("boundary", "is_animated"): "boundary_animated", ("fine", "pos"): "fine_pos",
             ("fine", "vel"): "fine_vel"}
    for key, name in pairs.items():
        assert got[key].tobytes() == arrs[name].tobytes(), key
    raw = path.read_bytes()
    assert raw[:4] == b"GGCK" and len(raw) == 4 + 4 + 3 * 8 + n * 32 + nb * 16 + nf * 24
    (tmp_path / "short.ggck").write_bytes(raw[:-1])
    (tmp_path / "magic.ggck").write_bytes(b"XXCK" + raw[4:])
    (tmp_path / "long.ggck").write_bytes(raw + b"\0")
    # a count far beyond the file (would size a vector of 2^33 float3) and
    # one above INT32_MAX are refused before anything is allocated
    huge = bytearray(raw)
    huge[8:16] = (1 << 33).to_bytes(8, "little")
    (tmp_path / "huge.ggck").write_bytes(bytes(huge))
    wide = bytearray(raw)
    wide[8:16] = (1 << 31).to_bytes(8, "little")
    (tmp_path / "wide.ggck").write_bytes(bytes(wide))
    for bad in ("short", "magic", "long", "huge", "wide"):
        with pytest.raises(host.HostError):
            host.read_checkpoint(tmp_path / f"{bad}.ggck")
