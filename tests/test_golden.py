"""Golden fixtures (tests/golden/*.json, written by make_golden.py): the
oracle must reproduce them (regression pin); the GPU path must match them to
the parity bar."""
import glob
import json
import os

import numpy as np
import pytest

import zisafvm_amd as z
from conftest import rel_err, wb_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# single-state cases; vortex_convergence.json (time runs) has its own tests
FILES = sorted(f for f in glob.glob(os.path.join(HERE, "*.json")) if not f.endswith("_convergence.json"))


def load_case(path):
    d = json.load(open(path))
    mesh = z.Mesh(d["n_dims"], np.array(d["vertex_ids"], dtype=np.int32), np.array(d["corners"]),
                  np.array(d["ghost"], dtype=np.uint8))
    p = z.Params()
    for name, _ in p._fields_:
        v = d["params"][name]
        if isinstance(v, list):
            arr = getattr(p, name)
            for k, x in enumerate(v):
                arr[k] = x
        else:
            setattr(p, name, v)
    u0 = np.array(d["u0"])
    steady = None if d["steady"] is None else np.array(d["steady"])
    return d, mesh, p, u0, steady


def errors(d, mesh, p, u0, a, b):
    if p.equilibrium != 0:
        return wb_err(a, b, mesh, u0, p.gamma)
    return rel_err(a, b, mask=mesh.ghost == 0)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-5] for f in FILES])
def test_oracle_reproduces_golden(path):
    from oracle.zfv_ref import Oracle

    d, mesh, p, u0, steady = load_case(path)
    o = Oracle(mesh, p)
    if steady is not None:
        o.set_steady_state(steady)
    assert max(errors(d, mesh, p, u0, o.rate(u0), np.array(d["rate"]))) <= 1e-14
    u1 = o.step(u0, 0.0, d["dt"])
    du_ref = np.array(d["step"]) - u0
    assert max(errors(d, mesh, p, u0, (u1 - u0) / d["dt"], du_ref / d["dt"])) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-5] for f in FILES])
def test_gpu_matches_golden(gpu, path):
    d, mesh, p, u0, steady = load_case(path)
    s = z.Solver(mesh, p)
    errs = errors(d, mesh, p, u0, s.rate(u0), np.array(d["rate"]))
    print(os.path.basename(path), "rate error per variable", errs)
    assert max(errs) <= 1e-12
    s.set_state(u0)
    if steady is not None:
        s.set_steady_state(steady)
    s.step(0.0, d["dt"])
    u1 = s.get_state()
    du_ref = np.array(d["step"]) - u0
    errs = errors(d, mesh, p, u0, (u1 - u0) / d["dt"], du_ref / d["dt"])
    print(os.path.basename(path), "step error per variable", errs)
    assert max(errs) <= 1e-12
