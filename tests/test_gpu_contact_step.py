"""A contact time loop run end to end on the GPU as the reference's example drives it
(examples/one_bunny/one_bunny.py:173-295): a soft tet cube falls onto a fixed floor under
gravity; each Newton direction is bounded by the additive-CCD step of its sweep
(compute_largest_step_size(0.8)), and every trial of the backtracking search refreshes the
pp / pe / pt / ee barrier sets from the proximity detection before the energy is
evaluated. Checked: every time step's Newton loop converges, no vertex of the cube ever
reaches the floor plane, the cube lands and stays within the barrier's range, and the
contact sets the loop last loaded equal the oracle's proximity pairs at the final
positions (bit-exact multisets)."""
import numpy as np
import pytest

from oracle import ccd as occd

pytestmark = pytest.mark.gpu


def test_cube_drops_onto_floor_without_penetration():
    from yasps_amd import device as dev
    from yasps_amd.ccd import CCD
    from yasps_amd.scenes import drop_scene, build_drop_scene
    from yasps_amd.stepping import ContactSets, newton_time_step
    ds = drop_scene()
    s0, h = build_drop_scene(ds, threshold_mode="relative_residual")
    nf, nv = ds.n_fixed, ds.rest.shape[0]
    c = CCD(len(ds.surface), nv, mesh_indices=ds.mesh_ids, print_timings=False)
    x0 = dev.to_device(ds.position.reshape(-1), "float64")
    c.init_faces(x0, ds.faces.reshape(-1), ds.surface, len(ds.faces))
    c.init_edges(x0, x0, ds.edges.reshape(-1), len(ds.edges))
    prims = {k: (h[k + "_primitive"], h[k + "_connectivity"]) for k in ("pp", "pe", "pt", "ee")}
    contact = ContactSets(c, ds.position[:nf], prims, ds.dhat)
    lowest = []
    # gradient entries start near m g dt^2 ~ 1e-4; at 1e-7 a Newton step still lowers the
    # energy far above the rounding of its sum (~1e-17), so the backtracking test resolves it
    for step in range(24):
        print(f"drop step {step}", flush=True)
        out = newton_time_step(s0, h["moving"], ds.dt, tolerance=1e-8, gradient_tol=1e-7,
                               max_newton=40, step_bound=contact.step_bound,
                               on_positions=contact.refresh)
        print(f"  newton {out['newton_iterations']} steps {['%.3g' % t for t in out['steps']]} "
              f"contacts {contact.counts[-1]}", flush=True)
        assert out["converged"], out
        x = dev.to_host(h["moving"]["position"].value).reshape(-1, 3)
        lowest.append(float(x[:, 1].min()))
        assert lowest[-1] > 0.0
    # the cube fell from 0.02 onto a corner and stays in contact within the barrier's range
    assert lowest[0] < 0.02 and 0.0 < lowest[-1] < np.sqrt(ds.dhat)
    assert max(sum(cnt) for cnt in contact.counts) > 0
    # the sets loaded last are the oracle's proximity pairs at the final positions
    xa = np.concatenate([ds.position[:nf], x])
    ref = occd.proximity(xa, ds.faces, ds.surface, ds.edges, ds.dhat, ds.mesh_ids)
    got = c.separated_counts
    assert got == [len(ref[k]) for k in ("pp", "pe", "pt", "ee")] and sum(got) > 0
    for k, w in (("pp", 2), ("pe", 3), ("pt", 4), ("ee", 4)):
        a = dev.to_host(getattr(c, k)).astype(np.int64).reshape(-1, w)
        assert sorted(map(tuple, a)) == sorted(ref[k])
