"""GPU parity of the surface reconstruction (SRTool.dual_contouring,
SPH/surf_recon/srtool.py:1264-1304) through the C-ABI against the oracle, bit
for bit: the field before and after smoothing, G and x_bar, the smoothing's sd,
the mesh's vertices, triangles and normals.  Cases: the reference's torus
example (SPH/surf_recon/example.py on SPH/data/models/torus.ply, kept as a
fixture), the isotropic field alone, a ball, a jittered fluid block with two
smoothing passes and a non-zero isolevel, a shifted origin, positions already
on the device, object reuse, the SRTool mirror with its PLY export, and the
error cases."""
import os

import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import ply, scenes
from nonnewton_amd.surface import SRTool, Surface
from oracle import surface as O

from test_oracle_surface import ball, check_closed_oriented, in_own_cube_in_grid_order, torus

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_both(x, **kw):
    s = Surface(**kw)
    info = s.run(x)
    o = O.reconstruct(x, **kw)
    return s, info, o


def assert_same(s, info, o, normals=False):
    assert info["lo"] == o["lo"] and info["dims"] == o["dims"]
    assert info["active_blocks"] == o["n_active"]
    assert np.array_equal(s.field(0), o["field0"])
    assert np.array_equal(s.field(1), o["field"], equal_nan=True)
    assert info["sd"] == o["sd"] or (np.isnan(info["sd"]) and np.isnan(o["sd"]))
    v, f, n = s.mesh()
    assert np.array_equal(v, o["vertices"])
    assert np.array_equal(f, o["faces"])
    if normals:
        assert np.array_equal(n, o["normals"])
    return v, f


def test_torus_example_bit_exact():
    kw = dict(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.06, isolevel=0.0, normals=True)
    x = torus()
    s, info, o = run_both(x, **kw)
    G, xb = s.aniso()
    assert np.array_equal(G, o["G"]) and np.array_equal(xb, o["xbar"])
    v, f = assert_same(s, info, o, normals=True)
    assert check_closed_oriented(v.astype(np.float64), f) == 0
    s.close()


def test_isotropic_field_without_smoothing():
    s, info, o = run_both(torus(), voxel_size=0.01, particle_radius=0.02, smooth_radius=0.03, isolevel=-0.0,
                          anisotropic=False, smooth_iterations=0)
    assert_same(s, info, o)
    s.close()


def test_ball_bit_exact_and_closed():
    s, info, o = run_both(ball(), voxel_size=0.01, particle_radius=0.02, smooth_radius=0.06, normals=True)
    v, f = assert_same(s, info, o, normals=True)
    assert check_closed_oriented(v.astype(np.float64), f) == 2
    in_own_cube_in_grid_order(v, 0.01)
    s.close()


def fluid_block(n_side=(16, 10, 14), sp=0.02, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    g = [np.arange(k) * sp for k in n_side]
    x = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3) + 0.13
    x += rng.uniform(-0.3 * sp, 0.3 * sp, size=x.shape)
    return x.astype(np.float32)


def test_jittered_block_two_passes_nonzero_isolevel():
    x = fluid_block()
    s, info, o = run_both(x, voxel_size=0.01, particle_radius=0.02, smooth_radius=0.05, isolevel=-0.3,
                          smooth_iterations=2, normals=True)
    G, xb = s.aniso()
    assert np.array_equal(G, o["G"]) and np.array_equal(xb, o["xbar"])
    v, f = assert_same(s, info, o, normals=True)
    assert len(f) > 1000
    s.close()


def test_shifted_origin_and_voxel():
    x = torus() + np.float32([0.07, -0.31, 0.02])
    s, info, o = run_both(x, origin=(0.05, -0.4, -0.01), voxel_size=0.008, particle_radius=0.02,
                          smooth_radius=0.05)
    assert_same(s, info, o)
    s.close()


def test_device_positions_equal_host_positions():
    """nn_surface_run_device on positions copied into device memory with the HIP
    runtime the library itself links (the same copy, already loaded)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    x = fluid_block(seed=11)
    s = Surface(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.05)
    s.run(x)
    v0, f0, _ = s.mesh()
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(x.nbytes)) == 0
    try:
        assert hip.hipMemcpy(d, x.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(x.nbytes), 1) == 0  # host to device
        s.run_device(d.value, len(x))
        v1, f1, _ = s.mesh()
    finally:
        hip.hipFree(d)
    assert np.array_equal(v0, v1) and np.array_equal(f0, f1)
    s.close()


def test_reuse_with_a_smaller_input_equals_a_fresh_object():
    kw = dict(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.05)
    s = Surface(**kw)
    s.run(fluid_block(seed=3))
    small = torus()
    info = s.run(small)
    o = O.reconstruct(small, **kw)
    assert_same(s, info, o)
    s.close()


def test_srtool_mirror_and_ply_export(tmp_path):
    x = torus()
    sr = SRTool([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], voxel_size=0.01, particle_radius=0.02, record_normals=True,
                max_num_vertices=20000, max_num_indices=600000)
    sr.dual_contouring(isolevel=0, smooth_radius=0.06, num_particles=len(x), pos=x)
    o = O.reconstruct(x, voxel_size=0.01, particle_radius=0.02, smooth_radius=0.06, normals=True)
    assert sr.num_vertices[None] == len(o["vertices"]) and sr.num_indices[None] == 3 * len(o["faces"])
    assert np.array_equal(sr.mesh_vertex, o["vertices"])
    assert np.array_equal(sr.mesh_index.reshape(-1, 3), o["faces"])
    path = str(tmp_path / "torus_mesh.ply")
    sr.export_mesh(path, output_normal=True)
    back = ply.read_vertices(path)
    assert np.array_equal(np.stack([back["x"], back["y"], back["z"]], 1), o["vertices"])
    assert np.array_equal(np.stack([back["nx"], back["ny"], back["nz"]], 1), o["normals"])
    with pytest.raises(ValueError):
        small = SRTool([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], voxel_size=0.01, particle_radius=0.02, max_num_vertices=100)
        small.dual_contouring(isolevel=0, smooth_radius=0.06, num_particles=len(x), pos=x)


@pytest.mark.parametrize("case", ["below_origin", "nan", "crowded_voxel"])
def test_errors_match_the_oracle(case):
    if case == "below_origin":
        x, code = np.array([[-0.02, 0.1, 0.1], [0.2, 0.2, 0.2]], np.float32), L.NN_E_OUT_OF_GRID
    elif case == "nan":
        x, code = np.array([[0.1, np.nan, 0.1]], np.float32), L.NN_E_OUT_OF_GRID
    else:
        x, code = np.full((65, 3), 0.105, np.float32), L.NN_E_CELL_OVERFLOW
    s = Surface(voxel_size=0.01, particle_radius=0.01, smooth_radius=0.03)
    with pytest.raises(L.NNError) as e:
        s.run(x)
    assert e.value.code == code
    with pytest.raises(O.SurfaceError):
        O.reconstruct(x, voxel_size=0.01, particle_radius=0.01, smooth_radius=0.03)
    # the object still works after an error
    info = s.run(torus())
    o = O.reconstruct(torus(), voxel_size=0.01, particle_radius=0.01, smooth_radius=0.03)
    assert_same(s, info, o)
    s.close()


def test_velocity_channels_bit_exact_and_exported(tmp_path):
    x = fluid_block(seed=5)
    rng = np.random.Generator(np.random.PCG64(4))
    u = rng.normal(size=x.shape).astype(np.float32)
    kw = dict(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.05)
    s = Surface(velocity=True, normals=True, **kw)
    info = s.run(x, u)
    o = O.reconstruct(x, normals=True, velocity=u, **kw)
    assert_same(s, info, o, normals=True)
    assert np.array_equal(s.mesh_velocity(), o["mesh_velocity"])
    s.close()
    sr = SRTool([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], voxel_size=0.01, particle_radius=0.02, record_velocity=True)
    sr.dual_contouring(isolevel=0, smooth_radius=0.05, num_particles=len(x), pos=x, velocity=u)
    assert np.array_equal(sr.mesh_velocity, o["mesh_velocity"])
    path = str(tmp_path / "vel.ply")
    sr.export_mesh(path, output_velocity=True)
    back = ply.read_vertices(path)
    assert np.array_equal(np.stack([back["vx"], back["vy"], back["vz"]], 1), o["mesh_velocity"])


def test_budget_caps_the_grid():
    s = Surface(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.06)
    s.set_budget(1 << 20)
    with pytest.raises(L.NNError) as e:
        s.run(torus())
    assert e.value.code == L.NN_E_OOM
    s.set_budget(0)
    info = s.run(torus())
    assert info["num_triangles"] > 0
    s.close()


def test_mesh_of_a_context_equals_mesh_of_its_downloaded_fluid():
    """nn_surface_run_context meshes the fluid where the step left it in device
    memory; the mesh equals the one of the downloaded fluid positions (and the
    oracle's), bit for bit."""
    from nonnewton_amd.params import KIND_FLUID

    sc = scenes.dam_break(fluid_counts=(16, 12, 16))
    ctx = L.Context(sc.params, sc.n)
    ctx.upload(sc.x, sc.kind, sc.v)
    ctx.init()
    for _ in range(5):
        ctx.step()
    x = ctx.download(L.FIELD_X).reshape(-1, 3)
    fluid = x[sc.kind == KIND_FLUID]
    assert np.all(np.flatnonzero(sc.kind == KIND_FLUID) == np.arange(len(fluid)))  # fluid numbered first
    kw = dict(voxel_size=0.01, particle_radius=0.02, smooth_radius=0.05)
    s = Surface(**kw)
    info = s.run_context(ctx)
    assert info["particles"] == len(fluid)
    v1, f1, _ = s.mesh()
    info2 = s.run(fluid)
    v2, f2, _ = s.mesh()
    assert np.array_equal(v1, v2) and np.array_equal(f1, f2) and len(f1) > 0
    o = O.reconstruct(fluid, **kw)
    assert np.array_equal(v1, o["vertices"]) and np.array_equal(f1, o["faces"])
    s.close()
    # with the velocity channels: the fluid's velocities where the step left them
    u = ctx.download(L.FIELD_V).reshape(-1, 3)[sc.kind == KIND_FLUID]
    sv = Surface(velocity=True, **kw)
    sv.run_context(ctx)
    ov = O.reconstruct(fluid, velocity=u, **kw)
    assert np.array_equal(sv.mesh_velocity(), ov["mesh_velocity"]) and np.any(ov["mesh_velocity"] != 0)
    sv.close()
    ctx.close()
