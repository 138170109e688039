import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must run on a GPU box; fail loudly without one."""
    if not _has_gpu():
        pytest.fail("no GPU visible for a test marked gpu")
    import zisafvm_amd

    return zisafvm_amd.load_library()


def rel_err(a, b, mask=None):
    """Per-variable max |a-b| / max |b| (absolute where b vanishes)."""
    import numpy as np

    if mask is not None:
        a, b = a[mask], b[mask]
    out = []
    for v in range(b.shape[1]):
        den = np.abs(b[:, v]).max()
        diff = np.abs(a[:, v] - b[:, v]).max()
        out.append(diff / den if den > 0 else diff)
    return out


def pressure_scale(mesh, u, gamma):
    """max_i sum_faces p_i |face| / vol_i: the size of the pressure terms that
    cancel in a hydrostatic equilibrium (the normalisation of well-balanced
    rate-of-change errors)."""
    import numpy as np

    kin = 0.5 * (u[:, 1] ** 2 + u[:, 2] ** 2 + u[:, 3] ** 2) / u[:, 0]
    p = (u[:, 4] - kin) * (gamma - 1.0)
    c = mesh.corners
    if mesh.n_dims == 2:
        perim = sum(np.linalg.norm(c[:, (a + 1) % 3] - c[:, a], axis=1) for a in range(3))
    else:
        faces = [(0, 1, 3), (0, 2, 1), (0, 3, 2), (1, 2, 3)]
        perim = sum(0.5 * np.linalg.norm(np.cross(c[:, b] - c[:, a], c[:, d] - c[:, a]), axis=1) for a, b, d in faces)
    return float((np.abs(p) * perim / mesh.volumes()).max())


def wb_err(a, b, mesh, u, gamma, mask=None):
    """Per-variable max |a-b| / max(max |b|, pressure scale)."""
    import numpy as np

    if mask is None:
        mask = mesh.ghost == 0
    s = pressure_scale(mesh, u, gamma)
    out = []
    for v in range(b.shape[1]):
        den = max(np.abs(b[mask, v]).max(), s)
        out.append(np.abs(a[mask, v] - b[mask, v]).max() / den)
    return out


def shape_cases():
    """One (mesh workload, params) pair per reconstruction kernel shape the
    library compiles (stencil sizes x quadrature points): 2D orders 2, 3, 4
    with both edge-rule sizes, 3D order 3 with both face rules and with the
    reference's default of three one-sided stencils."""
    import zisafvm_amd as z
    from zisafvm_amd import workloads as W

    def p3(**kw):
        p = W.polytrope_3d(n=10, amplitude=1e-3).params
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    v2 = lambda: W.vortex_2d(n=14)
    b3 = lambda: W.polytrope_3d(n=10, amplitude=1e-3)
    return {
        "2d_order3_edge_deg2": (v2, lambda: z.make_params(n_dims=2, order=3, edge_deg=2)),
        "2d_order3_edge_deg4": (v2, lambda: z.make_params(n_dims=2, order=3, edge_deg=4)),
        "2d_order4_edge_deg2": (v2, lambda: z.make_params(n_dims=2, order=4, edge_deg=2)),
        "2d_order4_edge_deg4": (v2, lambda: z.make_params(n_dims=2, order=4, edge_deg=4)),
        "2d_order2_edge_deg1": (v2, lambda: z.make_params(n_dims=2, order=2, edge_deg=1)),
        "2d_order2_edge_deg2": (v2, lambda: z.make_params(n_dims=2, order=2, edge_deg=2)),
        "3d_order3_face_deg2": (b3, lambda: p3()),
        "3d_order3_face_deg3": (b3, lambda: p3(edge_deg=3)),
        "3d_order3_three_one_sided": (b3, lambda: p3(n_stencils=4)),
    }


def with_recompute(params, steps, threshold):
    """A copy of params with steps_per_recompute and recompute_threshold set."""
    q = type(params).from_buffer_copy(params)
    q.steps_per_recompute = steps
    q.recompute_threshold = threshold
    return q


def recompute_states(u, n_calls=7, rel=1e-4, seed=0):
    """States for successive rate evaluations of kept-equilibrium tests: the
    second repeats the first; each later one scales every cell's conserved
    variables (velocity kept, rho and E_int scaled alike) by 1 + rel N(0, 1)
    against the first, so that under a threshold of about rel some cells
    recompute and others keep their equilibrium."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = [u, u]
    for _ in range(n_calls - 2):
        f = 1.0 + rel * rng.standard_normal(u.shape[0])
        out.append(u * f[:, None])
    return out
