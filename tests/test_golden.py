"""Committed regression fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py from
the oracle): the oracle must reproduce them bit for bit (CPU), and so must the HIP path (gpu) --
the residual of the stored initial state, the state after the stored number of steps and the
per-step dt."""
import numpy as np
import pytest

from tests.golden import make_fixtures as MF

NAMES = sorted(MF.cases())


def _load(name):
    with np.load(MF.HERE / f"{name}.npz") as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    U0, R, U, dts = MF.compute(name)
    f = _load(name)
    assert np.array_equal(f["dts"], dts)
    for n in U0:
        assert np.array_equal(f[f"U0_{n}"], U0[n])
        assert np.array_equal(f[f"R_{n}"], R[n])
        assert np.array_equal(f[f"U_{n}"], U[n])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_path_reproduces_fixture(name):
    from pyhype_amd import Euler2D, PrimitiveState
    from pyhype_amd import workloads as W
    from pyhype_amd.initial_conditions import ExplosionInitialCondition

    mesh, nx, ny, kw, steps = MF.cases()[name]
    cfg = W._config(nx, ny, ic=ExplosionInitialCondition(), flux=kw["flux"], integrator=kw["integrator"],
                    cfl=kw["CFL"])
    cfg.fvm_slope_limiter_type = kw.get("limiter", "Venkatakrishnan")
    if kw.get("reconstruction") == "Primitive":
        cfg.reconstruction_type = PrimitiveState
    cfg.fvm_num_quadrature_points = kw.get("quadrature_points", 1)
    f = _load(name)
    s = Euler2D(cfg, mesh, rows_per_item=2)
    s.set_state([{n: f[f"U0_{n}"] for n in mesh}])
    R = s.plan.get_states(s.plan.residual())[0]
    dts = []
    for _ in range(steps):
        s.run_steps(1)
        dts.append(float(s.plan.status()[3][0]))
    U = s.states()[0]
    for n in mesh:
        np.testing.assert_array_equal(R[n], f[f"R_{n}"], err_msg=f"residual, block {n}")
        np.testing.assert_array_equal(U[n], f[f"U_{n}"], err_msg=f"state, block {n}")
    assert np.array_equal(np.array(dts), f["dts"])
