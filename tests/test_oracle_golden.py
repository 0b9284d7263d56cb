"""The FP64 oracle against the reference's own solver core.

Pins: every committed fixture under tests/golden/ was produced by the
reference's CPU solver (src/lbm_solver.tpp, src/qtensor_solver.tpp,
src/analysis_fields.h) compiled from its sources (oracle/build_ref.sh,
tests/golden/make_golden.py). The oracle must reproduce each one bit for bit
(rho, u, Q, F, populations), the free energy and the SimIO::Log sums
exactly (single-threaded summation order).
"""
import os

import numpy as np
import pytest

from conftest import golden_config, golden_names, load_golden, run_golden_on

NAMES = golden_names()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference_bitwise(name, oracle_lib):
    g = load_golden(name)
    o = oracle_lib.Oracle(golden_config(g), threads=1)
    run_golden_on(o, g)
    st = o.state()
    np.testing.assert_array_equal(st["q"], g["q"])
    np.testing.assert_array_equal(o.force(), g["F"])
    if "rho" in g:
        np.testing.assert_array_equal(st["rho"], g["rho"])
        np.testing.assert_array_equal(st["u"], g["u"])
    if "f" in g:
        np.testing.assert_array_equal(o.populations(), g["f"])
    assert o.free_energy() == float(g["free_energy"])
    # SimIO::Log sums (src/sim_io.cc:78-165), single-threaded order
    np.testing.assert_array_equal(np.array(list(o.log_sums().values())), g["log"])


def test_fixture_set_is_complete():
    # Every boundary tag, both advection schemes, all three stepping modes and
    # both body-force models appear in the committed fixtures.
    gs = [load_golden(n) for n in NAMES]
    assert {str(g["mode"]) for g in gs} == {"full", "lbm", "q"}
    assert {str(g["advection"]) for g in gs} == {"centred", "upwind"}
    assert {int(g["force_model"]) for g in gs} == {0, 1}
    assert {"Channel", "FullyPeriodic", "SlitFreeSlip", "AnchoredZ", "MovingLidZ", "SpecularZ"} <= {
        str(g["bc"]) for g in gs}


@pytest.mark.skipif(not os.path.exists("/root/reference/src/lbm_solver.tpp"), reason="reference sources absent")
def test_oracle_matches_reference_live(oracle_lib):
    """Fresh seed, not a committed fixture: run the reference binary now."""
    from lbman3d_amd import abi
    if not os.path.exists(oracle_lib.ref_binary("actup")):
        pytest.skip("oracle/_ref not built")
    dims = (12, 10, 8)
    q = abi.seed_director_field(424242, *dims, 0.5, 0.01)
    r = oracle_lib.run_ref("actup", "SlitFreeSlip", "full", 0, 25, dims, q)
    cfg = abi.make_config(*dims, abi.SlitConfig, abi.Params(LAMBDA=0.7, ALPHA=0.046), advection=abi.ADV_UPWIND)
    o = oracle_lib.Oracle(cfg)
    o.set_state(q=q)
    o.step(25)
    st = o.state()
    np.testing.assert_array_equal(st["q"], r["q"])
    np.testing.assert_array_equal(st["u"], r["u"])
    np.testing.assert_array_equal(o.populations(), r["f"])


def test_oracle_thread_count_independent(oracle_lib):
    """Every population slot has one writer per step, so the state is
    independent of the OpenMP thread count (src/lbm_solver.tpp:125)."""
    g = load_golden("act_channel")
    a = oracle_lib.Oracle(golden_config(g), threads=1)
    b = oracle_lib.Oracle(golden_config(g), threads=4)
    for s in (a, b):
        run_golden_on(s, g)
    np.testing.assert_array_equal(a.state()["q"], b.state()["q"])
    np.testing.assert_array_equal(a.populations(), b.populations())


def test_order_director_closed_form(oracle_lib):
    """Q = S (n n - I/3) has largest eigenvalue 2S/3 with eigenvector n. The
    reference reports that eigenvalue as the order (QtensorToOrderDirector,
    src/analysis_fields.cc:59-93; its relaxation test expects -B/(3C) =
    2 S_eq / 3, tests/integration/test_qtensor_relaxation.cc:154-157) and
    +-n as the director."""
    from lbman3d_amd import abi
    dims = (6, 5, 4)
    V = 6 * 5 * 4
    rng = np.random.default_rng(12)
    n = rng.normal(size=(V, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    S = rng.uniform(0.1, 0.9, V)
    q = np.stack([S * (n[:, 0] ** 2 - 1 / 3), S * n[:, 0] * n[:, 1], S * n[:, 0] * n[:, 2],
                  S * (n[:, 1] ** 2 - 1 / 3), S * n[:, 1] * n[:, 2]])
    o = oracle_lib.Oracle(abi.make_config(*dims))
    o.set_state(q=q)
    order, director = o.order_director()
    np.testing.assert_allclose(order, 2 * S / 3, rtol=1e-10)
    np.testing.assert_allclose(np.abs(np.sum(director * n, axis=1)), 1.0, rtol=1e-10)


@pytest.mark.skipif(not os.path.exists("/root/reference/src/lbm_solver.tpp"), reason="reference sources absent")
def test_oracle_matches_reference_over_the_32cube_1000_step_run(oracle_lib):
    """The long-trajectory pin: the whole 32^3 FullyPeriodic centred run
    (1000 steps, seed 12345) through the reference's ActiveNematicSim step
    order, run now by the reference binary (oracle/_ref/ref_c0), against the
    oracle. Rounding differences would have 1000 steps to grow; the states
    must still be equal bit for bit."""
    import concurrent.futures as cf
    from lbman3d_amd import abi
    if not os.path.exists(oracle_lib.ref_binary("c0")):
        pytest.skip("oracle/_ref not built")
    dims = (32, 32, 32)
    q = abi.seed_director_field(12345, *dims, 0.5, 0.01)
    with cf.ThreadPoolExecutor(1) as ex:
        fut = ex.submit(oracle_lib.run_ref, "c0", "FullyPeriodic", "full", 0, 1000, dims, q)
        o = oracle_lib.Oracle(abi.make_config(*dims), threads=4)
        o.set_state(q=q)
        o.step(1000)
        r = fut.result()
    st = o.state()
    np.testing.assert_array_equal(st["q"], r["q"])
    np.testing.assert_array_equal(st["u"], r["u"])
    np.testing.assert_array_equal(st["rho"], r["rho"])
    np.testing.assert_array_equal(o.populations(), r["f"])
    o.set_threads(1)        # the reference binary sums the free energy on one thread
    assert o.free_energy() == float(r["free_energy"])
    np.testing.assert_array_equal(np.array(list(o.log_sums().values())), r["log"])
