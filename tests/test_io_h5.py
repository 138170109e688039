"""The reference's file formats through libzfv_io.so (HDF5, host only):
.msh.h5 meshes (load_grid_gmsh_h5, src/zisa/grid/grid.cpp:889-901), the
renumbering tool (src/renumber_grid.cpp:46-135) and AllVariables snapshots
(src/zisa/model/models.cpp:7-27, src/zisa/model/all_variables.cpp:74-113).

The tool's exact cell order comes from the un-vendored ZisaSFC library's
Hilbert curve, which is not in the reference's tree: the order is pinned here
by what any Hilbert renumbering must satisfy (a permutation of the cells with
the same vertices, idempotent, local), not cell for cell."""
import os
import subprocess

import numpy as np
import pytest

import zisafvm_amd as z
from zisafvm_amd import io as zio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _centres(mesh):
    return mesh.corners.mean(axis=1)


def test_mesh_round_trip(tmp_path):
    m = z.lattice_mesh_2d(12, jitter=0.2, seed=3, periodic=(False, False), ghost_layers=2)
    p = str(tmp_path / "lattice.msh.h5")
    zio.mesh_to_msh_h5(p, m)
    r = zio.read_msh_h5(p)
    assert r.n_dims == 2 and r.n_cells == m.n_cells
    np.testing.assert_array_equal(r.vertex_ids, m.vertex_ids)
    np.testing.assert_array_equal(r.corners, m.corners)
    assert not r.ghost.any()
    m3 = z.ball_mesh_3d(8, radius=1.0, ghost_radius=0.8, jitter=0.1, seed=2)
    p3 = str(tmp_path / "ball.msh.h5")
    zio.mesh_to_msh_h5(p3, m3)
    r3 = zio.read_msh_h5(p3)
    np.testing.assert_array_equal(r3.vertex_ids, m3.vertex_ids)
    np.testing.assert_array_equal(r3.corners, m3.corners)


def test_read_rejects_malformed_files(tmp_path):
    p = str(tmp_path / "bad.msh.h5")
    with pytest.raises(z.ZfvError):
        zio.read_msh_h5(str(tmp_path / "missing.msh.h5"))
    zio.write_snapshot(p, np.ones((4, 5)), 0.0, 0)  # a snapshot is not a mesh
    with pytest.raises(z.ZfvError, match="n_dims"):
        zio.read_msh_h5(p)
    with pytest.raises(z.ZfvError, match="out of range"):
        zio.write_msh_h5(p, 2, np.zeros((3, 3)), np.array([[0, 1, 3]]))


@pytest.mark.parametrize("dims", [2, 3])
def test_renumbering_is_a_local_idempotent_permutation(tmp_path, dims):
    m = (z.lattice_mesh_2d(24, jitter=0.2, seed=5, periodic=(False, False), ghost_layers=2) if dims == 2
         else z.ball_mesh_3d(10, radius=1.0, ghost_radius=0.8, jitter=0.1, seed=5))
    # scramble the cells first, so that the file's order carries no locality
    rng = np.random.default_rng(1)
    shuffle = rng.permutation(m.n_cells)
    scrambled = z.Mesh(dims, m.vertex_ids[shuffle], m.corners[shuffle], m.ghost[shuffle])
    p = str(tmp_path / "m.msh.h5")
    zio.mesh_to_msh_h5(p, scrambled)
    order = zio.renumber_msh_h5(p)
    r = zio.read_msh_h5(p)
    # a permutation of the cells, vertices untouched
    assert sorted(order.tolist()) == list(range(m.n_cells))
    np.testing.assert_array_equal(r.vertex_ids, scrambled.vertex_ids[order])
    np.testing.assert_array_equal(r.corners, scrambled.corners[order])
    # local: consecutive cells are neighbours in space, unlike the scrambled order
    step = lambda mesh: np.linalg.norm(np.diff(_centres(mesh), axis=0), axis=1).mean()
    assert step(r) < 0.2 * step(scrambled)
    # idempotent: a Hilbert-ordered file stays as it is
    order2 = zio.renumber_msh_h5(p)
    np.testing.assert_array_equal(order2, np.arange(m.n_cells))


def test_renumbering_tool_and_volume_check(tmp_path):
    tool = os.path.join(ROOT, "zisafvm_amd", "_lib", "zfv_renumber_grid")
    m = z.ball_mesh_3d(6, radius=1.0, ghost_radius=0.8, jitter=0.1, seed=2)
    p = str(tmp_path / "ball.msh.h5")
    zio.mesh_to_msh_h5(p, m)
    r = subprocess.run([tool, "--grid", p], capture_output=True, text=True)
    assert r.returncode == 0 and "renumbered" in r.stdout
    # a tetrahedron far smaller than the rest: the tool's sanity check refuses
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1e-3, 1e-3, 1e-3], [1, 1, 1.0]])
    cells = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [1, 2, 3, 5]])
    zio.write_msh_h5(p, 3, verts, cells)
    with pytest.raises(z.ZfvError, match="Faulty cell"):
        zio.renumber_msh_h5(p)


def test_snapshot_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(3)
    u = rng.normal(size=(1000, 5))
    p = str(tmp_path / "snap.h5")
    zio.write_snapshot(p, u, 0.125, 37)
    v, t, k = zio.read_snapshot(p)
    np.testing.assert_array_equal(v, u)
    assert t == 0.125 and k == 37
    s = str(tmp_path / "steady.h5")
    zio.write_steady_state(s, u)
    np.testing.assert_array_equal(zio.read_snapshot(s, with_time=False), u)
    with pytest.raises(z.ZfvError, match="time"):
        zio.read_snapshot(s)


def test_solver_input_from_a_mesh_file_equals_the_generated_mesh(tmp_path):
    # the instance built from the file reconstructs exactly like the one
    # built from the generator's arrays (host setup and kernel math through
    # the CPU emulation)
    import emu_binding as E
    from zisafvm_amd import workloads as W

    m = z.lattice_mesh_2d(14, jitter=0.2, seed=9, periodic=(False, False), ghost_layers=3)
    p = str(tmp_path / "iso.msh.h5")
    zio.mesh_to_msh_h5(p, m)
    r = zio.read_msh_h5(p)
    r.ghost = m.ghost.copy()  # ghosts are per experiment, not in the file
    u = W.cell_average(m, lambda x: W.isentropic_vortex(x, beta=3.0))
    params = z.make_params(n_dims=2, order=3, flux_bc=1)
    np.testing.assert_array_equal(E.Emu(r, params).rate(u), E.Emu(m, params).rate(u))
