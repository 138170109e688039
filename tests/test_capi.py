"""The C-ABI library loads and exports every function include/zfv.h declares;
host-only entry points (mesh generation) work without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import zisafvm_amd as z

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zfv_[a-z0-9_]+)\s*\(", text)))


# each header and the library that implements it
LIBRARIES = {"zfv.h": z.LIB_PATH, "zfv_io.h": os.path.join(ROOT, "zisafvm_amd", "_lib", "libzfv_io.so")}


def test_every_header_has_its_library():
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == sorted(LIBRARIES)


@pytest.mark.parametrize("header", sorted(LIBRARIES))
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIBRARIES[header])
    names = declared_functions(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_a_clean_error():
    if os.path.exists("/dev/kfd"):
        return  # covered by the GPU tests
    wl_mesh = z.lattice_mesh_2d(6)
    p = z.make_params()
    try:
        z.Solver(wl_mesh, p)
    except z.ZfvError as e:
        assert "device" in str(e)
    else:
        raise AssertionError("zfv_create succeeded without a device")


def test_invalid_configuration_is_rejected_before_device_work():
    p = z.make_params(order=3)
    p.steps_per_recompute = 0
    mesh = z.lattice_mesh_2d(6)
    try:
        z.Solver(mesh, p)
    except z.ZfvError as e:
        assert "steps_per_recompute" in str(e) or "device" in str(e)
    else:
        raise AssertionError("accepted steps_per_recompute = 0")


def test_lattice_mesh_counts_and_volumes():
    m = z.lattice_mesh_2d(16, jitter=0.2, seed=4, periodic=(True, True))
    assert m.n_cells == 2 * 16 * 16
    assert abs(m.volumes().sum() - 1.0) < 1e-12
    assert (m.ghost == 0).all()
    m2 = z.lattice_mesh_2d(16, jitter=0.2, seed=4, periodic=(True, False), ghost_layers=3)
    assert m2.n_cells == 2 * 16 * 22
    interior = m2.ghost == 0
    assert abs(m2.volumes()[interior].sum() - 1.0) < 1e-12


def test_lattice_mesh_is_delaunay():
    # every interior edge is locally Delaunay: opposite angles sum to <= pi
    m = z.lattice_mesh_2d(12, jitter=0.25, seed=6, periodic=(False, False), ghost_layers=0)
    edges = {}
    for i, tri in enumerate(m.vertex_ids):
        for k in range(3):
            key = tuple(sorted((tri[k], tri[(k + 1) % 3])))
            edges.setdefault(key, []).append((i, (k + 2) % 3))
    worst = 0.0
    for key, sides in edges.items():
        if len(sides) != 2:
            continue
        ang = 0.0
        for i, opp in sides:
            c = m.corners[i]
            p, a, b = c[opp], c[(opp + 1) % 3], c[(opp + 2) % 3]
            u, v = a - p, b - p
            ang += np.arccos(np.clip(u @ v / np.linalg.norm(u) / np.linalg.norm(v), -1, 1))
        worst = max(worst, ang)
    assert worst <= np.pi + 1e-12


def test_ball_mesh():
    m = z.ball_mesh_3d(10, radius=1.0, ghost_radius=0.8, jitter=0.1, seed=2)
    assert m.n_cells > 0
    assert (m.volumes() > 0).all()
    r = np.linalg.norm(m.centroids(), axis=1)
    assert r.max() <= 1.0
    assert ((r > 0.8) == (m.ghost == 1)).all()
    # total volume close to the ball's (jagged boundary)
    assert abs(m.volumes().sum() / (4.0 / 3.0 * np.pi) - 1.0) < 0.1


def test_mesh_generation_is_deterministic():
    a = z.lattice_mesh_2d(8, seed=42)
    b = z.lattice_mesh_2d(8, seed=42)
    np.testing.assert_array_equal(a.corners, b.corners)
    c = z.lattice_mesh_2d(8, seed=43)
    assert not np.array_equal(a.corners, c.corners)


def _estimate(mesh, p):
    from zisafvm_amd import serving as S

    return S.estimate_device_bytes(mesh, p)


def test_enumerated_fields_are_range_checked():
    # an unknown mode is an error (the reference's factories LOG_ERR on
    # unknown modes), never another physics; shape_for runs before any device
    # work, so zfv_estimate_device_bytes reports it on the CPU
    mesh = z.lattice_mesh_2d(6)
    assert _estimate(mesh, z.make_params()) > 0
    bad = [("flux_bc", 3), ("flux_bc", -1), ("flux", 2), ("reconstruction", 2), ("equilibrium", 3), ("gravity", 4),
           ("gamma", 1.0), ("steps_per_recompute", 0), ("scaling", 2), ("scaling", -1)]
    for field, value in bad:
        p = z.make_params()
        setattr(p, field, value)
        try:
            _estimate(mesh, p)
        except z.ZfvError as e:
            assert field.split("_")[0] in str(e), (field, str(e))
        else:
            raise AssertionError(f"accepted {field} = {value}")
    # the isentropic flux BC rides on the isentropic well-balancing
    p = z.make_params(flux_bc=z.ZFV_FLUX_BC_ISENTROPIC)
    try:
        _estimate(mesh, p)
    except z.ZfvError as e:
        assert "isentropic" in str(e)
    else:
        raise AssertionError("accepted the isentropic flux BC without isentropic well-balancing")
    p = z.make_params(flux_bc=z.ZFV_FLUX_BC_ISENTROPIC, equilibrium=z.ZFV_EQ_ISENTROPIC, gravity=z.ZFV_GRAVITY_AXIAL,
                      gravity_g=1.0)
    assert _estimate(mesh, p) > 0
    # kept equilibria with the isentropic flux BC (which solves the current
    # equilibrium every call) are not implemented
    p = z.make_params(flux_bc=z.ZFV_FLUX_BC_ISENTROPIC, equilibrium=z.ZFV_EQ_ISENTROPIC, gravity=z.ZFV_GRAVITY_AXIAL,
                      gravity_g=1.0, steps_per_recompute=2)
    try:
        _estimate(mesh, p)
    except z.ZfvError as e:
        assert "steps_per_recompute" in str(e)
    else:
        raise AssertionError("accepted steps_per_recompute 2 with the isentropic flux BC")
    # both of the reference's scalings (EulerScaling, UnityScaling)
    p = z.make_params()
    p.scaling = 1
    assert _estimate(mesh, p) == _estimate(mesh, z.make_params())
    # kept equilibria: 8 doubles and a counter per cell on top
    assert _estimate(mesh, z.make_params(steps_per_recompute=3)) == _estimate(mesh, z.make_params()) + 68 * mesh.n_cells


def test_batches_above_the_gpu_run_sizes_are_refused():
    # zfv_batch_create checks its capacity before any device work
    p = z.make_params()
    for cells, members in ((65537, 2), (1024, 9), (4031374, 64)):
        try:
            z.Batch(p, cell_capacity=cells, face_capacity=2 * cells, max_instances=members)
        except z.ZfvError as e:
            assert "refused" in str(e), str(e)
        else:
            raise AssertionError(f"accepted a batch of {cells} cells, {members} members")
    try:
        z.Batch(p, cell_capacity=65536, face_capacity=131072, max_instances=8)
    except z.ZfvError as e:
        assert "refused" not in str(e) and "device" in str(e), str(e)
