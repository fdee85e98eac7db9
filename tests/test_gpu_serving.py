"""Serving: batched multi-system PCG equals the single-system solve of every request,
continuous batching admits requests while others iterate, every result is the same alone
and in the mix (1e-8 relative), and the HBM budget bounds the admitted set."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(reqs, precision="fp64"):
    from yasps_amd.serving import SolveServer, SolveRequest
    out = {}
    for r in reqs:
        srv = SolveServer(max_active=1, keep_solutions=True, precision=precision)
        q = SolveRequest(id=r.id, n=r.n, young=r.young, seed=r.seed, arrival=0.0)
        (d,) = srv.serve([q])
        out[r.id] = d
    return out


def test_batched_pcg_matches_single_system_solver():
    from yasps_amd import device as dev
    from yasps_amd.scenes import cube_scene, build_yasps_scene, lame_parameters
    from yasps_amd.serving import SolveServer, SolveRequest
    print("batched vs single", flush=True)
    reqs = [SolveRequest(id=i, n=n, young=y, seed=10 + i)
            for i, (n, y) in enumerate([(4, 1e5), (7, 3e4), (5, 8e5), (6, 2e5)])]
    # sequential admission: the four requests arriving together share every chunk
    srv = SolveServer(max_active=8, keep_solutions=True, overlap_loads=False)
    done = {r.id: r for r in srv.serve(reqs)}
    assert max(srv.batch_sizes) == 4
    for r in reqs:
        cs = cube_scene(r.n, young=r.young, seed=0)
        rng = np.random.default_rng(r.seed)
        nf = cs.n_fixed
        cs.position[nf:] = cs.rest[nf:] + rng.normal(0.0, 0.02 / r.n, size=(cs.rest.shape[0] - nf, 3))
        s0, _ = build_yasps_scene(cs, threshold_mode="relative_residual")
        x = dev.to_host(s0.minimizeEnergy(tolerance=1e-6)[0])
        it = s0.minimizer.linearSolver.statistics["iterations"]
        d = done[r.id]
        assert d.status >= 0 and abs(d.iterations - it) <= 1
        assert np.linalg.norm(d.solution - x) <= 1e-8 * np.linalg.norm(x)


def test_continuous_batching_same_alone_and_in_mix():
    from yasps_amd.serving import SolveServer, serving_stream, poisson_arrivals, summary
    print("continuous batching", flush=True)
    reqs = serving_stream(count=20, n_min=4, n_max=12, seed=3)
    srv = SolveServer(max_active=16, keep_solutions=True, passes=4)
    srv.warm([r.n for r in reqs])
    poisson_arrivals(reqs, rate=400.0, seed=2)
    done = srv.serve(reqs)
    assert len(done) == 20 and all(r.status >= 0 for r in done)
    # requests joined while others were iterating
    assert max(srv.batch_sizes) > 1
    first_finish = min(r.finished for r in done)
    assert any(r.admitted > min(q.admitted for q in done) and r.admitted < first_finish + 1e9
               for r in done)
    alone = _alone(done)
    for r in done:
        a = alone[r.id]
        assert abs(a.iterations - r.iterations) <= 1
        assert np.linalg.norm(a.solution - r.solution) <= 1e-8 * np.linalg.norm(a.solution)
    st = summary(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_hbm_budget_limits_admission():
    from yasps_amd.serving import SolveServer, SolveRequest
    reqs = [SolveRequest(id=i, n=6, young=1e5, seed=i) for i in range(4)]
    srv = SolveServer(max_active=8)
    srv.warm([6])
    srv.budget = srv.pool[6][0].footprint          # room for exactly one request
    done = srv.serve(reqs)
    assert len(done) == 4 and max(srv.batch_sizes) == 1


def test_payload_positions_match_seeded_positions():
    """Requests carrying their positions (pinned payload, asynchronous upload) give the
    same solutions, bit for bit, as requests whose positions the server derives, and the
    loader-thread server gives the same bits as the sequential one."""
    from yasps_amd.serving import SolveServer, SolveRequest, attach_payloads
    print("payload positions", flush=True)
    spec = [(5, 2e5), (7, 6e4), (9, 4e5)]
    plain = [SolveRequest(id=i, n=n, young=y, seed=40 + i) for i, (n, y) in enumerate(spec)]
    carried = attach_payloads([SolveRequest(id=i, n=n, young=y, seed=40 + i)
                               for i, (n, y) in enumerate(spec)])
    a = {r.id: r for r in SolveServer(max_active=4, keep_solutions=True).serve(plain)}
    b = {r.id: r for r in SolveServer(max_active=4, keep_solutions=True).serve(carried)}
    seq = [SolveRequest(id=r.id, n=r.n, young=r.young, seed=r.seed, positions=r.positions)
           for r in carried]
    c = {r.id: r for r in SolveServer(max_active=4, keep_solutions=True,
                                      overlap_loads=False).serve(seq)}
    for i in a:
        assert a[i].status >= 0 and a[i].iterations == b[i].iterations == c[i].iterations
        np.testing.assert_array_equal(a[i].solution, b[i].solution)
        np.testing.assert_array_equal(b[i].solution, c[i].solution)


def test_request_larger_than_the_budget_is_rejected():
    from yasps_amd.serving import SolveServer, SolveRequest
    print("budget rejection", flush=True)
    for overlap in (True, False):
        srv = SolveServer(max_active=2, overlap_loads=overlap)
        srv.warm([5])
        srv.budget = srv.pool[5][0].footprint - 1
        with pytest.raises(ValueError, match="does not fit"):
            srv.serve([SolveRequest(id=0, n=5, young=1e5, seed=0)])


def _oracle_system(r):
    """FP64 operator and right-hand side of a request from the oracle (same positions as
    the server derives from the request's seed)."""
    from yasps_amd.scenes import cube_scene
    from yasps_amd.serving import request_positions
    from oracle import newton
    cs = cube_scene(r.n, young=r.young, seed=0)
    nf = cs.n_fixed
    cs.position[nf:] = request_positions(cs, r)
    st = newton.fast_structure(cs)
    asm = newton.evaluate(cs, st)
    return newton.bsr_matrix(st, asm.values), asm.gradient


def test_batched_mas_matches_single_system_mas():
    """Every system of a batch carries its own MAS: the batched restriction, local solves
    and prolongation give each request the single-system MAS solve."""
    from yasps_amd import device as dev
    from yasps_amd.scenes import cube_scene, build_yasps_scene
    from yasps_amd.serving import SolveServer, SolveRequest, request_positions
    print("batched MAS vs single-system MAS", flush=True)
    reqs = [SolveRequest(id=i, n=n, young=y, seed=60 + i)
            for i, (n, y) in enumerate([(5, 1e5), (8, 3e4), (6, 8e5), (7, 2e5)])]
    srv = SolveServer(max_active=8, keep_solutions=True, overlap_loads=False, solver="mas")
    done = {r.id: r for r in srv.serve(reqs)}
    assert max(srv.batch_sizes) == 4
    for r in reqs:
        cs = cube_scene(r.n, young=r.young, seed=0)
        nf = cs.n_fixed
        cs.position[nf:] = request_positions(cs, r)
        s0, _ = build_yasps_scene(cs, solver_name="mas", threshold_mode="relative_residual",
                                  max_levels=4, coarsest_level_weight=1.0)
        x = dev.to_host(s0.minimizeEnergy(tolerance=1e-6)[0])
        it = s0.minimizer.linearSolver.statistics["iterations"]
        d = done[r.id]
        assert d.status >= 0 and abs(d.iterations - it) <= 1, (d.iterations, it)
        assert np.linalg.norm(d.solution - x) <= 1e-8 * np.linalg.norm(x)


@pytest.mark.parametrize("solver", ["mas", "jacobian"])
def test_mixed_serving_reaches_fp64_residual_same_alone_and_in_mix(solver):
    """Mixed precision serving (FP32 values and preconditioner data in the batched Krylov
    loop, FP64 iterative refinement): every request reaches ||g - H x|| <= 1e-6 ||g|| on
    the oracle's FP64 operator, requests join while others iterate, and each result is
    the same alone and in the mix (1e-8 relative: the per-system arithmetic does not
    depend on the batch)."""
    from yasps_amd.serving import SolveServer, serving_stream, poisson_arrivals
    print(f"mixed serving {solver}", flush=True)
    reqs = serving_stream(count=12, n_min=4, n_max=10, seed=5)
    srv = SolveServer(max_active=16, keep_solutions=True, passes=4, precision="mixed",
                      solver=solver)
    srv.warm([r.n for r in reqs])
    poisson_arrivals(reqs, rate=300.0, seed=4)
    done = srv.serve(reqs)
    assert len(done) == 12 and all(r.status >= 0 for r in done)
    assert max(srv.batch_sizes) > 1
    for r in done:
        assert r.rel_residual <= 1e-6 and r.outer_iterations >= 1
        A, g = _oracle_system(r)
        assert np.linalg.norm(g - A @ r.solution) <= 1.01e-6 * np.linalg.norm(g)
    from yasps_amd.serving import SolveRequest
    for r in done[:6]:
        (a,) = SolveServer(max_active=1, keep_solutions=True, precision="mixed",
                           solver=solver).serve(
            [SolveRequest(id=r.id, n=r.n, young=r.young, seed=r.seed)])
        assert a.status >= 0 and abs(a.iterations - r.iterations) <= 1
        assert np.linalg.norm(a.solution - r.solution) <= 1e-8 * np.linalg.norm(a.solution)
