"""Parity of the HIP backend with the oracle (gpu).  Every comparison is bit-for-bit unless noted;
the parity tolerance of the path (max relative error <= 1e-9 per variable) is asserted as well."""
import numpy as np
import pytest

from oracle import pyhype_oracle as O
from pyhype_amd import Euler2D, Euler2DBatch, RealizabilityError
from pyhype_amd import workloads as W
from tests.helpers import bitwise_equal_blocks, max_rel_err_blocks, oracle_for, skewed_mesh

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _solver(cfg, mesh, **kw):
    s = Euler2D(cfg, mesh, **kw)
    s.apply_initial_condition()
    return s


def _perturbed(states, seed=0, amp=1e-3):
    rng = np.random.default_rng(seed)
    out = {}
    for n, a in states.items():
        b = a.copy()
        b *= 1 + amp * rng.standard_normal(b.shape)
        b[..., 1:3] += amp * rng.standard_normal(b[..., 1:3].shape) * np.abs(a[..., 0:1]).max()
        out[n] = b
    return out


@pytest.mark.parametrize("case", ["single", "multiblock", "oblique"])
def test_residual_matches_oracle(case):
    if case == "single":
        cfg, mesh = W.explosion_single(24)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(12, 2, 4)
    else:
        cfg, mesh = W.oblique_shock(16, 3)
    s = _solver(cfg, mesh)
    st = _perturbed(s.states()[0], seed=3)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        ref = O.residual(b, sim.flux_fn)
        np.testing.assert_array_equal(R[n], ref, err_msg=f"block {n}")


def test_ghost_fill_matches_oracle():
    cfg, mesh = W.explosion_multiblock(10, 2, 3)
    s = _solver(cfg, mesh)
    st = _perturbed(s.states()[0], seed=4)
    s.set_state([st])
    G = s.plan.ghosts()
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        for d in "EWNS":
            np.testing.assert_array_equal(G[(0, n)][d], b.ghost[d], err_msg=f"block {n} side {d}")


@pytest.mark.parametrize("case", ["single", "multiblock"])
def test_dt_matches_oracle(case):
    cfg, mesh = W.explosion_single(30) if case == "single" else W.explosion_multiblock(10, 2, 4)
    s = _solver(cfg, mesh)
    st = _perturbed(s.states()[0], seed=5)
    s.set_state([st])
    dt = float(s.plan.dt().cpu().numpy()[0])
    assert dt == oracle_for(cfg, mesh, st).get_dt()


@pytest.mark.parametrize("use_graph", [False, True])
def test_explosion_single_steps(use_graph):
    cfg, mesh = W.explosion_single(40)
    s = _solver(cfg, mesh)
    sim = oracle_for(cfg, mesh, s.states()[0])
    s.run_steps(12, use_graph=use_graph)
    sim.run(12)
    g, r = s.states()[0], sim.states()
    assert max_rel_err_blocks(g, r) <= TOL
    assert bitwise_equal_blocks(g, r)
    assert s.t[0] == sim.t


def test_explosion_multiblock_steps():
    cfg, mesh = W.explosion_multiblock(16, 2, 4)
    s = _solver(cfg, mesh)
    sim = oracle_for(cfg, mesh, s.states()[0])
    s.run_steps(8)
    sim.run(8)
    g, r = s.states()[0], sim.states()
    assert max_rel_err_blocks(g, r) <= TOL
    assert bitwise_equal_blocks(g, r)


def test_oblique_shock_hlll_ssprk2_steps():
    cfg, mesh = W.oblique_shock(20, 3)
    s = _solver(cfg, mesh)
    sim = oracle_for(cfg, mesh, s.states()[0])
    s.run_steps(10)
    sim.run(10)
    g, r = s.states()[0], sim.states()
    assert max_rel_err_blocks(g, r) <= TOL
    assert bitwise_equal_blocks(g, r)


def test_batched_uneven_cases_match_independent_oracles():
    """Uneven independent cases advanced in the same launches, each with its own dt; case 1 is one
    the reference scheme cannot advance (pressure ratio 7.4 at CFL 0.8): it fails at step 1 in
    both, is frozen on the device, and the other cases are unaffected."""
    cfgs = W.explosion_batch(seed=5, ncases=2, sizes=[36, 52], density_scaled=True)[0]
    cfgs += W.explosion_batch(seed=0, ncases=1, sizes=[28], density_scaled=True)[0]
    cfgs += W.explosion_batch(seed=3, ncases=1, sizes=[44])[0]
    meshes = [W.single_block_mesh() for _ in cfgs]
    b = Euler2DBatch(cfgs, meshes)
    b.apply_initial_condition()
    st0 = b.states()
    b.run_steps(6)
    got = b.states()
    assert b.failed == {2: 1}
    for c in (0, 1, 3):
        sim = oracle_for(cfgs[c], meshes[c], st0[c])
        sim.run(6)
        assert bitwise_equal_blocks(got[c], sim.states()), f"case {c}"
        assert b.t[c] == sim.t
    sim = oracle_for(cfgs[2], meshes[2], st0[2])
    with pytest.raises(FloatingPointError, match="after step 1"):
        sim.run(1)


def test_unstable_case_fails_like_the_oracle():
    """A drawn case the reference scheme cannot advance (24x24, pressure ratio 7.4, CFL 0.8): the
    oracle and the device both report an unrealizable state after step 1."""
    cfgs, meshes, _ = W.explosion_batch(seed=0, ncases=1, sizes=[24], density_scaled=True)
    s = Euler2D(cfgs[0], meshes[0])
    s.apply_initial_condition()
    sim = oracle_for(cfgs[0], meshes[0], s.states()[0])
    with pytest.raises(FloatingPointError, match="after step 1"):
        sim.run(1)
    with pytest.raises(RealizabilityError, match="after step 1"):
        s.run_steps(1)


def test_explosion_single_full_size():
    """The single-block workload at its full size and step count (100x100, 50 RK4 steps)."""
    cfg, mesh = W.explosion_single(100)
    s = _solver(cfg, mesh)
    sim = oracle_for(cfg, mesh, s.states()[0])
    s.run_steps(50)
    sim.run(50)
    g, r = s.states()[0], sim.states()
    assert max_rel_err_blocks(g, r) <= TOL
    assert bitwise_equal_blocks(g, r)


def test_unrealizable_state_raises():
    cfg, mesh = W.explosion_single(16)
    s = _solver(cfg, mesh)
    st = s.states()[0]
    st[0][3, 3, 0] = -1.0
    s.set_state([st])
    with pytest.raises(RealizabilityError):
        s.run_steps(1)


def test_solve_writes_reference_files_and_stops_at_t_final(tmp_path):
    """Euler2D.solve with write_solution: mesh files, solution directories k = 0, n, 2n, ... holding
    the state after step k+1 (Solver.write_solution runs before the step counter is incremented),
    and the step count at which the reference loop stops (t >= t_final)."""
    cfg, mesh = W.explosion_single(20)
    a = cfg.fluid.far_field.a
    cfg.t_final = 0.004            # seconds; 0.004*343 non-dimensional
    cfg.write_solution = True
    cfg.write_solution_base = str(tmp_path)
    cfg.write_solution_name = "expl"
    cfg.write_every_n_timesteps = 7
    s = Euler2D(cfg, mesh)
    s.solve(chunk=5)
    # oracle: the reference loop with its write points
    h = Euler2D(W.explosion_single(20)[0], mesh)
    h.apply_initial_condition()
    sim = oracle_for(cfg, mesh, h.states()[0])
    written = {}
    k = 0
    while sim.t < cfg.t_final * a:
        sim.step()
        if k % 7 == 0:
            written[k] = sim.states()
        k += 1
    assert s.num_time_step == k
    root = tmp_path / "expl"
    np.testing.assert_array_equal(np.load(root / "mesh" / "mesh_x_blk_0.npy"), s.case_blocks[0][0].mesh.x)
    dirs = sorted(int(p.name) for p in root.iterdir() if p.name != "mesh")
    assert dirs == sorted(written)
    for step, st in written.items():
        np.testing.assert_array_equal(np.load(root / str(step) / "expl_blk_0.npy"), st[0], err_msg=f"dir {step}")


@pytest.mark.parametrize("variant", [0, 1, 2, 4, 5, 6, 7, 8])
def test_double_mach_reflection_example_matches_oracle(variant):
    """The reference's double-Mach-reflection example (examples/dmr: 4 x 1 transfinite blocks over a
    30 degree ramp, outlets and slip walls, HLLL, RK2, PrimitiveState reconstruction) at 12 x 12
    cells per block: one Cartesian block ahead of the ramp and three rotated-frame blocks on it, in
    one plan.  Residual and 8 steps bit for bit on every stage variant."""
    cfg, mesh = W.double_mach_reflection(12)
    s = _solver(cfg, mesh, rows_per_item=2)
    assert {e.cartesian for e in s.plan.blocks} == {True, False}
    if variant:
        s.plan.set_launch(2, 0, 3 if variant in (1, 5, 6, 7, 8) else 2, variant=variant, rows_b=2)
    st = s.states()[0]
    sim = oracle_for(cfg, mesh, st)
    R = s.plan.get_states(s.plan.residual())[0]
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(8)
    sim.run(8)
    assert bitwise_equal_blocks(s.states()[0], sim.states())
    assert s.t[0] == sim.t


@pytest.mark.parametrize("example", ["implosion", "shockbox"])
def test_reference_examples_match_oracle(example):
    """The reference's implosion and shock-box examples (one walled block, Roe, RK2, CFL 0.4) at
    24 x 24: 20 steps bit for bit, time included."""
    cfg, mesh = getattr(W, example)(24)
    s = _solver(cfg, mesh)
    sim = oracle_for(cfg, mesh, s.states()[0])
    s.run_steps(20)
    sim.run(20)
    assert bitwise_equal_blocks(s.states()[0], sim.states())
    assert s.t[0] == sim.t


def test_skewed_multiblock_residual_and_steps_match_oracle():
    cfg, _ = W.explosion_single(18)
    mesh = skewed_mesh()
    s = _solver(cfg, mesh)
    assert not any(e.cartesian for e in s.plan.blocks)
    st = _perturbed(s.states()[0], seed=8)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(6)
    sim.run(6)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("case,rows,rows_b,waves", [
    ("single", 1, 4, 3), ("single", 3, 32, 4), ("multiblock", 2, 8, 3), ("oblique", 4, 5, 4),
    ("skewed", 1, 3, 3), ("wide", 2, 7, 3), ("wide_multiblock", 3, 16, 4), ("narrow", 1, 1, 3),
    ("narrow", 2, 64, 4)])
@pytest.mark.parametrize("variant", [1, 5, 6, 7, 8])
def test_split_variant_matches_oracle(case, rows, rows_b, waves, variant):
    """The split stage (reconstruction + x-faces, then y-faces, then the update; variant 5 with the
    y-faces and the update in one kernel, variant 6 with one y-face per wave, variant 7 that with
    the row-major reconstruction kernel, variant 8 with the y-faces over flat cell ranges) gives the oracle's results bit for bit: residual, steps
    (RK4 / SSP-RK2 with the CFL reduction) and the ghost semantics."""
    if case == "single":
        cfg, mesh = W.explosion_single(26)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(14, 2, 4)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(18, 3)
    elif case == "wide":            # three column strips per row (interior strip edges)
        cfg, mesh = W.explosion_single(130)
    elif case == "wide_multiblock":
        cfg, mesh = W.explosion_multiblock(66, 2, 2)
    elif case == "narrow":          # 3 x 40 cells: one strip, most lanes outside the block
        from pyhype_amd.initial_conditions import ExplosionInitialCondition
        cfg, mesh = W._config(3, 40, ic=ExplosionInitialCondition()), W.single_block_mesh()
    else:
        cfg, mesh = W.explosion_single(20)[0], skewed_mesh()
    s = _solver(cfg, mesh, rows_per_item=rows)
    s.plan.set_launch(rows, 0, waves, variant=variant, rows_b=rows_b)
    st = _perturbed(s.states()[0], seed=11)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("chunk", [1, 4])
@pytest.mark.parametrize("variant,case", [(3, "oblique"), (6, "multiblock"), (7, "oblique"), (7, "skewed"),
                                          (8, "wide_multiblock"), (1, "wide")])
def test_split_work_lists_dealt_to_the_xcds_match_oracle(variant, case, chunk):
    """phx_set_xcd_order: the split and element-wise stages with their work lists dealt to the XCDs
    in chunks of 1 and 4 items give the oracle's results bit for bit (residual and 5 steps through
    the step graphs), as the contiguous order does."""
    if case == "multiblock":
        cfg, mesh = W.explosion_multiblock(30, 2, 4)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(40, 3)
    elif case == "wide":
        cfg, mesh = W.explosion_single(130)
    elif case == "wide_multiblock":
        cfg, mesh = W.explosion_multiblock(66, 2, 2)
    else:
        cfg, mesh = W.explosion_single(40)[0], skewed_mesh()
    s = _solver(cfg, mesh, rows_per_item=1)
    s.plan.set_launch(1, 0, 3, variant=variant, rows_b=3, xcd_chunk=chunk)
    assert s.plan.launch_info()["xcd_chunk"] == chunk
    st = _perturbed(s.states()[0], seed=13)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("case", ["single", "multiblock", "oblique", "skewed", "wide", "wide_multiblock", "tall"])
def test_element_wise_variant_matches_oracle(case):
    """The element-wise stage (variant 3: one lane per cell and variable for the reconstruction, one
    per face for the fluxes, one per cell for the update) gives the oracle's results bit for bit:
    residual, steps with the CFL reduction, ghosts of every boundary kind; tiles that end inside a
    row and partial last tiles (block sizes not multiples of 64 or 256)."""
    if case == "single":
        cfg, mesh = W.explosion_single(26)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(14, 2, 4)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(18, 3)
    elif case == "wide":
        cfg, mesh = W.explosion_single(130)
    elif case == "wide_multiblock":
        cfg, mesh = W.explosion_multiblock(66, 2, 2)
    elif case == "tall":            # 3 x 40 cells: a tile spans many rows
        from pyhype_amd.initial_conditions import ExplosionInitialCondition
        cfg, mesh = W._config(3, 40, ic=ExplosionInitialCondition()), W.single_block_mesh()
    else:
        cfg, mesh = W.explosion_single(20)[0], skewed_mesh()
    s = _solver(cfg, mesh)
    s.plan.set_launch(4, 0, 2, variant=3)
    assert s.plan.launch["variant"] == 3
    st = _perturbed(s.states()[0], seed=19)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


def test_element_wise_variant_batched_cases_with_failures():
    """Batched uneven cases on the element-wise stage, one of which becomes unrealizable at step 1
    and is frozen (its tiles are skipped): states, times and failures equal the fused stage's."""
    cfgs = W.explosion_batch(seed=5, ncases=2, sizes=[36, 52], density_scaled=True)[0]
    cfgs += W.explosion_batch(seed=0, ncases=1, sizes=[28], density_scaled=True)[0]
    meshes = [W.single_block_mesh() for _ in cfgs]
    a = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b.plan.set_launch(4, 0, 2, variant=3)
    for s in (a, b):
        s.apply_initial_condition()
        s.run_steps(7)
    for x, y in zip(a.states(), b.states()):
        assert bitwise_equal_blocks(x, y)
    assert a.failed == b.failed == {2: 1}
    assert list(a.t) == list(b.t)


@pytest.mark.parametrize("case,rows,waves", [("single", 1, 2), ("single", 3, 3), ("multiblock", 2, 2),
                                             ("oblique", 2, 3), ("skewed", 1, 2), ("wide", 2, 2),
                                             ("wide_multiblock", 4, 3), ("narrow", 1, 2), ("narrow", 7, 3)])
def test_paired_variant_matches_oracle(case, rows, waves):
    """Two waves per work item (variant 4: each wave reconstructs two variables, one evaluates the
    x-face flux and the other the y-face flux, halves exchanged through LDS): residual and steps
    bit for bit, the CFL reduction included."""
    if case == "single":
        cfg, mesh = W.explosion_single(26)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(14, 2, 4)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(18, 3)
    elif case == "wide":
        cfg, mesh = W.explosion_single(130)
    elif case == "wide_multiblock":
        cfg, mesh = W.explosion_multiblock(66, 2, 2)
    elif case == "narrow":
        from pyhype_amd.initial_conditions import ExplosionInitialCondition
        cfg, mesh = W._config(3, 40, ic=ExplosionInitialCondition()), W.single_block_mesh()
    else:
        cfg, mesh = W.explosion_single(20)[0], skewed_mesh()
    s = _solver(cfg, mesh)
    s.plan.set_launch(rows, 0, waves, variant=4)
    st = _perturbed(s.states()[0], seed=23)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


def test_paired_variant_batched_cases_with_failures():
    """Batched uneven cases on the paired stage, one of which fails at step 1 and is frozen."""
    cfgs = W.explosion_batch(seed=5, ncases=2, sizes=[36, 52], density_scaled=True)[0]
    cfgs += W.explosion_batch(seed=0, ncases=1, sizes=[28], density_scaled=True)[0]
    meshes = [W.single_block_mesh() for _ in cfgs]
    a = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b.plan.set_launch(3, 0, 2, variant=4)
    for s in (a, b):
        s.apply_initial_condition()
        s.run_steps(7)
    for x, y in zip(a.states(), b.states()):
        assert bitwise_equal_blocks(x, y)
    assert a.failed == b.failed == {2: 1}
    assert list(a.t) == list(b.t)


@pytest.mark.parametrize("variant", [1, 5, 6, 7, 8])
def test_split_variant_batched_cases(variant):
    cfgs, meshes, _ = W.explosion_batch(seed=3, ncases=4, sizes=[40, 64, 33, 70])
    a = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b.plan.set_launch(2, 0, 3, variant=variant, rows_b=8)
    for s in (a, b):
        s.apply_initial_condition()
        s.run_steps(6)
    for x, y in zip(a.states(), b.states()):
        assert bitwise_equal_blocks(x, y)
    assert a.failed == b.failed


@pytest.mark.parametrize("limiter", ["BarthJespersen", "VanAlbada", "VanLeer"])
@pytest.mark.parametrize("variant", [0, 1, 3, 4, 5, 6, 7, 8])
def test_other_limiters_match_oracle(limiter, variant):
    """SlopeLimiterFactory's other limiters (limiters/limiters.py:25-62) on the general-geometry
    multiblock mesh: residual and RK4 steps bit for bit, fused, split, element-wise and paired
    stage."""
    cfg, mesh = W.explosion_multiblock(12, 2, 2)
    cfg.fvm_slope_limiter_type = limiter
    s = _solver(cfg, mesh, rows_per_item=3)
    if variant:
        s.plan.set_launch(2, 0, 3 if variant == 1 else 2, variant=variant, rows_b=5)
    st = _perturbed(s.states()[0], seed=21)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(4)
    sim.run(4)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("case,rows", [("single", 1), ("single", 3), ("multiblock", 1), ("multiblock", 2),
                                       ("oblique", 2), ("skewed", 1), ("wide", 2), ("wide_multiblock", 4)])
def test_shared_halo_variant_matches_oracle(case, rows):
    """Workgroups of 4 waves sharing halo rows through LDS (variant 2): residual and steps bit for
    bit, including items whose upper waves have no rows (block heights not a multiple of 4*rows)."""
    if case == "single":
        cfg, mesh = W.explosion_single(26)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(14, 2, 4)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(18, 3)
    elif case == "wide":
        cfg, mesh = W.explosion_single(130)
    elif case == "wide_multiblock":
        cfg, mesh = W.explosion_multiblock(66, 2, 2)
    else:
        cfg, mesh = W.explosion_single(20)[0], skewed_mesh()
    s = _solver(cfg, mesh)
    s.plan.set_launch(rows, 0, 2, variant=2)
    st = _perturbed(s.states()[0], seed=13)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


def test_shared_halo_variant_batched_cases_and_limiter():
    cfgs, meshes, _ = W.explosion_batch(seed=3, ncases=4, sizes=[40, 64, 33, 70])
    for c in cfgs:
        c.fvm_slope_limiter_type = "VanAlbada"
    a = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b.plan.set_launch(1, 0, 2, variant=2)
    for s in (a, b):
        s.apply_initial_condition()
        s.run_steps(6)
    for x, y in zip(a.states(), b.states()):
        assert bitwise_equal_blocks(x, y)
    assert a.failed == b.failed


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("case", ["multiblock", "oblique"])
def test_primitive_reconstruction_matches_oracle(case, variant):
    """reconstruction_type = PrimitiveState (fvm/base.py:286-289, blocks/quad_block.py:225-240):
    residual and steps bit for bit on every stage variant."""
    from pyhype_amd import PrimitiveState

    if case == "multiblock":
        cfg, mesh = W.explosion_multiblock(14, 2, 4)
    else:
        cfg, mesh = W.oblique_shock(18, 3)
    cfg.reconstruction_type = PrimitiveState
    s = _solver(cfg, mesh, rows_per_item=2)
    if variant:
        s.plan.set_launch(2, 0, 3 if variant == 1 else 2, variant=variant, rows_b=4)
    st = _perturbed(s.states()[0], seed=17)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(5)
    sim.run(5)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("variant", [1, 5, 6])
@pytest.mark.parametrize("nq", [2, 3])
@pytest.mark.parametrize("case,rows,rows_b,waves", [
    ("single", 2, 8, 3), ("multiblock", 1, 4, 3), ("oblique", 3, 5, 4), ("skewed", 2, 3, 3),
    ("wide_multiblock", 2, 16, 4)])
def test_gauss_faces_match_oracle(case, rows, rows_b, waves, nq, variant):
    """fvm_num_quadrature_points = 2 and 3 (mesh/quadratures.py:33-95, fvm/base.py:168-190): one
    flux per face and Gauss point, limiter over every point; the plan runs the split stage (and the
    split stage with fused y-faces and update).  Residual and steps bit for bit, Cartesian and
    rotated-frame blocks, walls, outlets and neighbours, interior strip edges."""
    if case == "single":
        cfg, mesh = W.explosion_single(22)
    elif case == "multiblock":
        cfg, mesh = W.explosion_multiblock(12, 2, 3)
    elif case == "oblique":
        cfg, mesh = W.oblique_shock(16, 3)
    elif case == "wide_multiblock":   # two column strips per block
        cfg, mesh = W.explosion_multiblock(66, 2, 1)
    else:
        cfg, mesh = W.explosion_single(16)[0], skewed_mesh()
    cfg.fvm_num_quadrature_points = nq
    s = _solver(cfg, mesh, rows_per_item=rows)
    assert s.plan.launch_info()["variant"] == 6   # the default of Gauss-face plans
    s.plan.set_launch(rows, 0, waves, variant=variant, rows_b=rows_b)
    st = _perturbed(s.states()[0], seed=21)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(4)
    sim.run(4)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("variant", [1, 5, 6])
def test_gauss_faces_with_primitive_reconstruction_and_van_leer(variant):
    from pyhype_amd import PrimitiveState

    cfg, mesh = W.explosion_multiblock(12, 2, 2)
    cfg.fvm_num_quadrature_points = 2
    cfg.fvm_slope_limiter_type = "VanLeer"
    cfg.reconstruction_type = PrimitiveState
    s = _solver(cfg, mesh, rows_per_item=2)
    s.plan.set_launch(2, 0, 2, variant=variant, rows_b=3)
    st = _perturbed(s.states()[0], seed=23)
    s.set_state([st])
    R = s.plan.get_states(s.plan.residual())[0]
    sim = oracle_for(cfg, mesh, st)
    for n, b in sim.blocks.items():
        np.testing.assert_array_equal(R[n], O.residual(b, sim.flux_fn), err_msg=f"block {n}")
    s.run_steps(4)
    sim.run(4)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


def test_flat_y_face_ranges_refuse_blocks_wider_than_their_lds():
    """Variant 8 evaluates the nx + n y-faces of a range of n >= 64 cells in at most 2048 LDS
    slots, so a block with nx + 64 > 2048 is refused; the plan keeps its stage and still runs."""
    from pyhype_amd.initial_conditions import ExplosionInitialCondition
    cfg, mesh = W._config(2000, 2, ic=ExplosionInitialCondition()), W.single_block_mesh()
    s = _solver(cfg, mesh, rows_per_item=2)
    with pytest.raises(ValueError, match="too wide"):
        s.plan.set_launch(1, 0, 3, variant=8, rows_b=7)
    assert s.plan.launch_info().get("variant", 0) != 8
    st = s.states()[0]
    sim = oracle_for(cfg, mesh, st)
    s.run_steps(2)
    sim.run(2)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


def test_gauss_faces_refuse_fused_variants():
    cfg, mesh = W.explosion_single(12)
    cfg.fvm_num_quadrature_points = 3
    s = _solver(cfg, mesh, rows_per_item=2)
    for v in (0, 2, 3, 4, 7, 8):
        with pytest.raises(ValueError, match="Gauss point"):
            s.plan.set_launch(2, 0, 2, variant=v)
    s.run_steps(2)   # the plan still runs after the refusals


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("integrator", ["ExplicitEuler1", "RK2", "Ralston2", "Generic2", "RK3", "RK3SSP",
                                        "Ralston3", "Ralston4", "DormandPrince5", "Generic3"])
def test_every_butcher_table_matches_oracle(integrator, variant):
    """Every table of TimeIntegratorFactory (time_marching/explicit_runge_kutta.py:91-274) in
    accumulator form on every stage variant: up to 6 outputs per stage (Dormand-Prince), negative
    coefficients (RK3), zero entries skipped; 3 steps bit for bit."""
    from pyhype_amd.time_marching import butcher_table

    cfg, mesh = W.explosion_single(16)
    cfg.time_integrator = integrator
    cfg.CFL = 0.4
    if integrator.startswith("Generic"):
        cfg.alpha = 0.7 if integrator == "Generic2" else 0.4
    s = _solver(cfg, mesh, rows_per_item=2)
    if variant:
        s.plan.set_launch(2, 0, 3 if variant == 1 else 2, variant=variant, rows_b=3)
    st = _perturbed(s.states()[0], seed=31)
    s.set_state([st])
    sim = oracle_for(cfg, mesh, st)
    assert sim.a == butcher_table(integrator, cfg)
    s.run_steps(3)
    sim.run(3)
    assert bitwise_equal_blocks(s.states()[0], sim.states())


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5, 6, 7, 8])
def test_batch_mixing_cartesian_and_rotated_frame_blocks(variant):
    """One plan whose cases mix Cartesian blocks (one block), rotated-frame blocks with neighbours
    (a 2x2 rectangular mesh, off-by-ulp interior corners) and a skewed two-block mesh: both kernel
    kinds launch in every stage; each case keeps its own dt and matches its own oracle."""
    c1, m1 = W.explosion_single(18)
    c2, m2 = W.explosion_multiblock(12, 2, 2)
    c3, m3 = W.explosion_single(14)[0], skewed_mesh()
    cfgs, meshes = [c1, c2, c3], [m1, m2, m3]
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    kinds = {e.cartesian for e in b.plan.blocks}
    assert kinds == {True, False}
    if variant:
        b.plan.set_launch(2, 0, 3 if variant == 1 else 2, variant=variant, rows_b=4)
    b.apply_initial_condition()
    st0 = b.states()
    b.run_steps(5)
    got = b.states()
    assert not b.failed
    for c in range(3):
        sim = oracle_for(cfgs[c], meshes[c], st0[c])
        sim.run(5)
        assert bitwise_equal_blocks(got[c], sim.states()), f"case {c}"
        assert b.t[c] == sim.t


def test_multi_step_graph_with_failing_case_matches_oracles():
    """19 graph steps (two replays of the 8-step graph, then three of the one-step graph) on uneven
    cases, one of which fails at step 1 and is frozen: every other case matches its own oracle bit
    for bit, time and the last dt included."""
    cfgs = W.explosion_batch(seed=5, ncases=2, sizes=[36, 52], density_scaled=True)[0]
    cfgs += W.explosion_batch(seed=0, ncases=1, sizes=[28], density_scaled=True)[0]
    meshes = [W.single_block_mesh() for _ in cfgs]
    b = Euler2DBatch(cfgs, meshes, rows_per_item=2)
    b.apply_initial_condition()
    st0 = b.states()
    b.run_steps(19)
    got = b.states()
    assert b.failed == {2: 1}
    _, step, t, dt = b.plan.status()
    assert step == 19
    for c in (0, 1):
        sim = oracle_for(cfgs[c], meshes[c], st0[c])
        sim.run(19)
        assert bitwise_equal_blocks(got[c], sim.states()), f"case {c}"
        assert t[c] == sim.t and dt[c] == sim.dts[-1]
