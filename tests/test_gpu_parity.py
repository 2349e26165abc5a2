# SPDX-License-Identifier: MIT
"""GPU parity: the HIP path against the oracle on identical inputs.

The GPU sums each body's contact rows as L interleaved partial blocks
(L = 32 lanes per body by default; 8, 16 and 64 are world options)
combined by an xor butterfly; the oracle is run with the same order
(lanes=L), so integer/index results must be equal and float results equal
up to the transcendental functions (atan2f on large rotations), for which
the stated tolerance is |a-b| <= 1e-5 + 1e-5|b| per component after one
step.  The oracle in reference order (lanes=0) is then within the
single-step tolerance |a-b| <= 1e-4 + 1e-4|b| (positions, velocities).
"""
import sys

import numpy as np
import pytest

import _oracle as O
import pmamd
import scenarios as S
from pmamd import layouts as L

pytestmark = pytest.mark.gpu
LANES = 32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_world(scene, lanes=LANES, **kw):
    return scene.build(pmamd.World(solve_lanes=lanes, **kw))


def oracle_world(scene, lanes=LANES, **kw):
    return scene.build(O.OracleWorld(lanes=lanes, **kw))


def exact_fraction(a, b):
    a = np.asarray(a).view(np.uint32)
    b = np.asarray(b).view(np.uint32)
    return float(np.mean(a == b))


def assert_states_close(g, o, rtol=1e-5, atol=1e-5):
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        np.testing.assert_allclose(g[f], o[f], rtol=rtol, atol=atol, err_msg=f)


def test_device_is_gfx950():
    name, cus = pmamd.device_info()
    say(f"device {name} with {cus} CUs")
    assert name.startswith("gfx950")


def test_seam_realtime_fixture():
    say("seam: rigid_realtime fixture")
    p, prev, cur, m, pairs = S.realtime_fixture()
    start = cur.copy()
    g_cur, g_m = cur.copy(), m.copy()
    sched = pmamd.prepare_solve(p, prev, g_cur, g_m, pairs, gravity=(0, 0, 0), base=64, requested=64)
    S.check_realtime_fixture(g_cur, start, sched)
    o_cur, o_m = cur.copy(), m.copy()
    O.prepare_solve(p, prev, o_cur, o_m, pairs, gravity=(0, 0, 0), base=64, requested=64, lanes=LANES)
    say(f"bit-exact state fraction {exact_fraction(g_cur.view(np.float32), o_cur.view(np.float32)):.4f}")
    assert_states_close(g_cur, o_cur)
    np.testing.assert_allclose(g_m["contacts"]["avbd_dual"], o_m["contacts"]["avbd_dual"], rtol=1e-5, atol=1e-5)


def test_seam_with_joints_matches_the_oracle():
    say("seam with joints: joint rig inputs, bodies named by shifted ids")
    inp = S.joint_seam_inputs()
    runs = []
    for side in ("gpu", "oracle"):
        st, man, bod, jo = inp["states"].copy(), inp["manifolds"].copy(), inp["bodies"].copy(), inp["joints"].copy()
        args = (inp["params"], inp["previous"], st, man, inp["pairs"])
        kw = dict(bodies=bod, base=10, profile=0, joints=jo, ids=inp["ids"])
        sched = (pmamd.prepare_solve(*args, **kw) if side == "gpu" else O.prepare_solve(*args, lanes=LANES, **kw))
        runs.append((st, jo, bod, sched))
    (gs, gj, gb, gc), (os_, oj, ob, oc) = runs
    say(f"bit-exact state fraction {exact_fraction(gs.view(np.float32), os_.view(np.float32)):.4f}")
    assert_states_close(gs, os_)
    for f in ("avbd_count", "avbd_next_a", "avbd_next_b", "dense_a", "dense_b", "enabled", "broken"):
        np.testing.assert_array_equal(gj[f], oj[f], err_msg=f)
    valid = np.arange(18)[None, :] < oj["avbd_count"][:, None]
    for f in ("lambda", "penalty", "error", "a", "b"):
        np.testing.assert_allclose(gj["rows"][f][valid], oj["rows"][f][valid], rtol=1e-4, atol=1e-4, err_msg=f)
    assert gj["options"].tobytes() == inp["joints"]["options"].tobytes()
    for f in ("joint_head", "color", "parent", "impact"):
        np.testing.assert_array_equal(gb[f], ob[f], err_msg=f)
    assert int(gc["budget"]) == int(oc["budget"])


def test_seam_stack_window():
    say("seam: two-body stack window, 720 ticks")

    def run(p, prev, cur, m, pairs, bodies):
        return pmamd.prepare_solve(p, prev, cur, m, pairs, bodies=bodies)
    out = S.stack_window(run)
    say(str(out))
    S.check_stack_window(out)


@pytest.mark.parametrize("profile", [1, 0])
def test_pyramid_single_step(profile):
    say(f"pyramid single step, profile {profile}")
    scene = S.pyramid()
    g = gpu_world(scene, profile=profile, base_iterations=10)
    o = oracle_world(scene, profile=profile, base_iterations=10)
    g.step(1)
    o.step(1)
    np.testing.assert_array_equal(g.pairs(), o.pairs())
    assert_manifolds_equal(g.manifolds(), o.manifolds())
    gb, ob = g.bodies(), o.bodies()
    np.testing.assert_array_equal(gb["color"], ob["color"])
    np.testing.assert_array_equal(gb["parent"], ob["parent"])
    np.testing.assert_array_equal(gb["impact"], ob["impact"])
    gs, os_ = g.states(), o.states()
    say(f"bit-exact state fraction {exact_fraction(gs.view(np.float32), os_.view(np.float32)):.4f}")
    assert_states_close(gs, os_)
    assert g.stats()["colors"] == o.stats()["colors"]
    for f in ("unsupported_pairs", "triangle_pairs", "fallback_pairs"):
        assert g.stats()[f] == o.stats()[f], f
    # Reference accumulation order stays within the single-step tolerance.
    r = oracle_world(scene, lanes=0, profile=profile, base_iterations=10)
    r.step(1)
    assert_states_close(gs, r.states(), rtol=1e-4, atol=1e-4)


def assert_manifolds_equal(gm, om):
    np.testing.assert_array_equal(gm["count"], om["count"])
    # Only slots below `count` are contacts: face-patch pruning leaves the
    # tail of the fixed 8-slot record stale in the reference layout.
    valid = np.arange(8)[None, :] < gm["count"][:, None]
    for f in ("normal", "point", "penetration", "impact_fraction", "avbd_dual", "avbd_anchor_a",
              "avbd_anchor_b"):
        np.testing.assert_array_equal(gm["contacts"][f][valid], om["contacts"][f][valid], err_msg=f)
    for h in ("body_hinge", "collider_hinge"):
        for f in gm["contacts"][h].dtype.names:
            np.testing.assert_array_equal(gm["contacts"][h][f][valid], om["contacts"][h][f][valid],
                                          err_msg=f"{h}.{f}")
    np.testing.assert_array_equal(gm["contacts"]["persistent"][valid], om["contacts"]["persistent"][valid],
                                  err_msg="persistent")


def test_joints_that_shape_contacts_follow_the_oracle():
    say("hinge contact frames: guided slider and piston into walls, hinged lid, welded pair on a prism; 40 steps")
    scene = S.hinge_contacts()
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    guided_steps = 0
    for step in range(40):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs())
        gm, om = g.manifolds(), o.manifolds()
        assert_manifolds_equal(gm, om)
        valid = np.arange(8)[None, :] < gm["count"][:, None]
        guided_steps += int((gm["contacts"]["body_hinge"]["axial"][valid] != 0).any())
        assert_states_close(g.states(), o.states())
    assert guided_steps > 0                   # the guided pairs were in contact
    assert g.stats()["triangle_pairs"] == o.stats()["triangle_pairs"]


def test_crossing_edges_take_the_triangle_path():
    say("crossing box edges: triangle-pair path")
    scene = S.crossing_edges(far=2)   # steps of worlds under four bodies are refused
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    g.step(1)
    o.step(1)
    assert g.stats()["triangle_pairs"] == o.stats()["triangle_pairs"] == 1
    assert_manifolds_equal(g.manifolds(), o.manifolds())
    assert_states_close(g.states(), o.states())


def test_concave_drop_trajectory():
    say("boxes thrown onto a concave prism, 90 steps")
    scene = S.concave_drop()
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    worst = 0.0
    for k in range(90):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k}")
        gs, os_ = g.stats(), o.stats()
        assert gs["triangle_pairs"] == os_["triangle_pairs"], k
        if k == 0:
            assert_manifolds_equal(g.manifolds(), o.manifolds())
        worst = max(worst, float(np.abs(g.states()["position"] - o.states()["position"]).max()))
    say(f"worst position difference {worst:.3e}")
    assert worst < 1e-3


def test_triangle_path_fallback_matches_oracle():
    say("mesh pile side 8: triangle pairs beyond the candidate slots take the BVH traversal")
    scene = S.pile(side=8, seed=3, analytic_spheres=False)
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    fallback = 0
    for k in range(40):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k}")
        gs, os_ = g.stats(), o.stats()
        assert (gs["triangle_pairs"], gs["fallback_pairs"]) == (os_["triangle_pairs"], os_["fallback_pairs"]), k
        fallback += gs["fallback_pairs"]
        if gs["fallback_pairs"] and fallback == gs["fallback_pairs"]:
            assert_manifolds_equal(g.manifolds(), o.manifolds())
    say(f"fallback pairs over 40 steps: {fallback}")
    assert fallback > 0
    assert_states_close(g.states(), o.states(), rtol=1e-3, atol=1e-3)


def test_inputs_and_kinematic_targets_follow_the_oracle():
    say("pusher: kinematic targets and rigid inputs, 40 steps")
    scene = S.pusher()
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    k = scene.n - 1
    x0 = float(scene.states["position"][k][0])
    for step in range(40):
        t = S.states(1, np.array([[x0 + 0.04 * (step + 1), 0.5, 0.0]], np.float32))
        t["orientation"][0] = S.quat((0, 1, 0), 0.01 * step)
        g.set_kinematic_targets(k, t)
        o.set_kinematic_targets(k, t)
        if step % 10 == 0:
            inp = np.zeros(2, L.RIGID_INPUT)
            inp["body_index"] = [scene.n - 2, 1]
            inp["linear_impulse"][0] = (0.0, 500.0, 100.0)
            inp["torque"][1] = (0.0, 2000.0, 0.0)
            g.apply_inputs(inp)
            o.apply_inputs(inp)
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {step}")
    gs, os_ = g.states(), o.states()
    say(f"bit-exact state fraction {exact_fraction(gs.view(np.float32), os_.view(np.float32)):.4f}")
    np.testing.assert_array_equal(gs[k]["position"], os_[k]["position"])
    assert_states_close(gs, os_, rtol=1e-4, atol=1e-4)


def assert_joints_match(gj, oj, rtol=1e-5, atol=1e-5):
    for f in ("enabled", "broken", "dense_a", "dense_b", "avbd_count", "avbd_next_a", "avbd_next_b"):
        np.testing.assert_array_equal(gj[f], oj[f], err_msg=f)
    np.testing.assert_allclose(gj["applied_impulse"], oj["applied_impulse"], rtol=rtol, atol=atol)
    valid = np.arange(18)[None, :] < oj["avbd_count"][:, None]
    for f in ("lambda", "penalty", "error", "a", "b"):
        np.testing.assert_allclose(gj["rows"][f][valid], oj["rows"][f][valid], rtol=rtol, atol=atol, err_msg=f)


@pytest.mark.parametrize("profile", [0, 1])
def test_joint_rig_single_step(profile):
    say(f"joint rig single step, profile {profile}")
    scene = S.joint_rig()
    g = gpu_world(scene, profile=profile, base_iterations=10)
    o = oracle_world(scene, profile=profile, base_iterations=10)
    g.step(1)
    o.step(1)
    np.testing.assert_array_equal(g.pairs(), o.pairs())
    gb, ob = g.bodies(), o.bodies()
    for f in ("color", "parent", "impact", "joint_head"):
        np.testing.assert_array_equal(gb[f], ob[f], err_msg=f)
    assert g.schedule()["budget"] == o.schedule()["budget"]
    assert g.stats()["islands"] == o.stats()["islands"]
    assert_states_close(g.states(), o.states())
    assert_joints_match(g.joints(), o.joints())


def test_joint_rig_trajectory():
    say("joint rig, 90 steps, profile 0")
    scene = S.joint_rig()
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    for k in range(90):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k}")
    gj, oj = g.joints(), o.joints()
    np.testing.assert_array_equal(gj["broken"], oj["broken"])
    assert oj["broken"][-1] == 1
    assert_states_close(g.states(), o.states(), rtol=1e-3, atol=1e-3)


def test_weld_group_and_disabled_pairs_are_excluded():
    say("joint collision exclusions")
    scene = S.joint_rig()
    g = gpu_world(scene, profile=0)
    o = oracle_world(scene, profile=0)
    g.step(2)
    o.step(2)
    gp = g.pairs()
    np.testing.assert_array_equal(gp, o.pairs())
    n = scene.n
    codes = set(int(c) for c in gp)
    for j in scene.joints:
        a, b = int(j["body_a"][0]), int(j["body_b"][0])
        assert a * n + b not in codes and b * n + a not in codes


@pytest.mark.parametrize("lanes", [8, 16, 64])
def test_lane_counts_follow_the_oracle(lanes):
    say(f"pyramid 5 steps with {lanes} lanes per body")
    scene = S.pyramid()
    g = gpu_world(scene, lanes=lanes)
    o = oracle_world(scene, lanes=lanes)
    g.step(5)
    o.step(5)
    np.testing.assert_array_equal(g.pairs(), o.pairs())
    gs, os_ = g.states(), o.states()
    say(f"bit-exact state fraction {exact_fraction(gs.view(np.float32), os_.view(np.float32)):.4f}")
    assert_states_close(gs, os_)


def test_invalid_lane_count_is_rejected():
    with pytest.raises(pmamd.PmhError):
        pmamd.World(solve_lanes=12)


def test_pyramid_trajectory():
    say("pyramid 60 steps")
    scene = S.pyramid()
    g = gpu_world(scene)
    o = oracle_world(scene)
    for k in range(60):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k}")
    gs, os_ = g.states(), o.states()
    say(f"bit-exact state fraction {exact_fraction(gs.view(np.float32), os_.view(np.float32)):.4f}")
    assert_states_close(gs, os_, rtol=1e-4, atol=1e-4)
    assert g.stats()["cache_hits"] == o.stats()["cache_hits"]


@pytest.mark.parametrize("analytic", [True, False])
def test_pile_drop_trajectory(analytic):
    side = 6
    say(f"pile side {side} analytic={analytic}, 45 steps through landing")
    scene = S.pile(side=side, seed=3, analytic_spheres=analytic)
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    worst = 0.0
    for k in range(45):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k}")
        gs, os_ = g.states(), o.states()
        worst = max(worst, float(np.abs(gs["position"] - os_["position"]).max()))
    st = g.stats()
    say(f"worst position difference {worst:.3e}; stats {st}")
    assert st["points"] > 0
    assert worst < 1e-3


def test_full_size_broadphase_is_bit_exact():
    say("32k pile: broadphase pairs against the oracle")
    scene = S.pile(side=32, seed=2)
    g = gpu_world(scene, profile=0, base_iterations=10)
    g.step(1)
    o = oracle_world(scene, profile=0, base_iterations=10)
    o.step(1)
    gp, op = g.pairs(), o.pairs()
    say(f"pairs {len(gp)}")
    assert len(gp) > 100000
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(g.proxies(), o.proxies())


def test_full_size_determinism():
    say("32k pile: two runs agree bit for bit")
    scene = S.pile(side=32, seed=2)
    a = gpu_world(scene, profile=0, base_iterations=10)
    b = gpu_world(scene, profile=0, base_iterations=10)
    a.step(3)
    b.step(3)
    np.testing.assert_array_equal(a.states().view(np.uint32), b.states().view(np.uint32))


def first_fit_colours(n, dynamic, edges):
    """Sequential first-fit colouring in index order over dynamic bodies
    (geometry_constraints.slang:1520-1545)."""
    nbr = [[] for _ in range(n)]
    for a, b in edges:
        nbr[a].append(b)
        nbr[b].append(a)
    colour = np.full(n, L.INVALID, dtype=np.uint32)
    for i in range(n):
        if not dynamic[i]:
            continue
        used = {int(colour[j]) for j in nbr[i] if j < i}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return colour


def test_contact_rich_graph_matches_recomputation():
    """Islands (smallest-index roots) and first-fit colours of a settled
    4096-body pile, recomputed on the host from the GPU's own contact
    graph, and two worlds stepped side by side agree bit for bit."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    say("4096-body pile, 150 steps: islands, colours and determinism")
    scene = S.pile(side=16, seed=3)
    g = gpu_world(scene, profile=0, base_iterations=10)
    h = gpu_world(scene, profile=0, base_iterations=10)
    for k in range(0, 150, 30):
        g.step(30)
        h.step(30)
        np.testing.assert_array_equal(g.states().view(np.uint32), h.states().view(np.uint32),
                                      err_msg=f"step {k + 30}")
    n = scene.n
    dynamic = scene.params["motion"] == L.MOTION_DYNAMIC
    codes = g.pairs().astype(np.int64)
    count = g.manifolds()["count"]
    a, b = codes // n, codes % n
    live = (count > 0) & dynamic[a] & dynamic[b]
    say(f"{len(codes)} pairs, {int(live.sum())} dynamic contact edges")
    assert live.sum() > 2 * n
    graph = coo_matrix((np.ones(int(live.sum())), (a[live], b[live])), shape=(n, n))
    _, label = connected_components(graph, directed=False)
    root = np.full(label.max() + 1, n, dtype=np.int64)
    np.minimum.at(root, label, np.arange(n))
    bodies = g.bodies()
    np.testing.assert_array_equal(bodies["parent"][dynamic], root[label][dynamic])
    colour = first_fit_colours(n, dynamic, zip(a[live].tolist(), b[live].tolist()))
    np.testing.assert_array_equal(bodies["color"][dynamic], colour[dynamic])


def assert_residuals_equal(g, o):
    gr, gx, gm = g.residuals(maximum=True)
    orr, ox, om = o.residuals(maximum=True)
    np.testing.assert_array_equal(gx, ox)
    np.testing.assert_array_equal(np.isnan(gr), np.isnan(orr))
    m = ~np.isnan(orr)
    np.testing.assert_array_equal(gr[m].view(np.uint32), orr[m].view(np.uint32))
    np.testing.assert_array_equal(gm[m].view(np.uint32), om[m].view(np.uint32))


@pytest.mark.parametrize("scene_name", ["pyramid", "joint_rig"])
def test_residuals_match_the_oracle(scene_name):
    """The residual is a max over points and rows of values both sides
    compute in the same order, so it is equal bit for bit."""
    say(f"residual log, {scene_name}")
    scene = S.pyramid(rows=6) if scene_name == "pyramid" else S.joint_rig()
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    for k in range(30):
        g.step(1)
        o.step(1)
        assert_residuals_equal(g, o)


@pytest.mark.parametrize("scene_name", ["pyramid", "joint_rig"])
def test_to_residual_mode_matches_the_oracle(scene_name):
    say(f"to-residual mode (1e-4, cap 40), {scene_name}")
    scene = S.pyramid(rows=6) if scene_name == "pyramid" else S.joint_rig()
    g = gpu_world(scene, profile=0, base_iterations=40, residual_tolerance=1e-4)
    o = oracle_world(scene, profile=0, base_iterations=40, residual_tolerance=1e-4)
    ran = []
    for k in range(40):
        g.step(1)
        o.step(1)
        assert_residuals_equal(g, o)
        ran.append(int(g.residuals()[1][0]))
    assert_states_close(g.states(), o.states(), rtol=1e-4, atol=1e-4)
    say(f"pose-phase iterations per step: {ran}")
    assert min(ran) >= 1 and max(ran) <= 40


def joint_gaps(scene, states):
    o = scene.joints
    qa = states["orientation"][o["body_a"][:, 0]].astype(np.float64)
    qb = states["orientation"][o["body_b"][:, 0]].astype(np.float64)

    def rot(q, v):
        u = q[:, :3]
        t = 2.0 * np.cross(u, v)
        return v + q[:, 3:4] * t + np.cross(u, t)
    pa = states["position"][o["body_a"][:, 0]] + rot(qa, o["local_anchor_a"].astype(np.float64))
    pb = states["position"][o["body_b"][:, 0]] + rot(qb, o["local_anchor_b"].astype(np.float64))
    return np.linalg.norm(pa - pb, axis=1)


def test_deep_hinge_schedule_follows_the_oracle():
    """One 80-link chain of hinges only: each hinge shares a body with the
    next, so the anchor projection's schedule has 80 levels and runs as the
    one-block walk; 10 steps against the oracle (serial in joint order)."""
    say("one 80-link hinge chain (80 projection levels)")
    scene = S.chains(count=1, links=80, hinges_only=True)
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    g.step(1)
    o.step(1)
    assert_states_close(g.states(), o.states())
    g.step(9)
    o.step(9)
    assert_states_close(g.states(), o.states(), rtol=1e-4, atol=1e-4)
    assert_joints_match(g.joints(), o.joints(), rtol=1e-3, atol=1e-3)


def test_chains_single_step_and_trajectory():
    """Four 64-link capsule chains (ball and hinge joints in turn, heavy
    links): one step to the single-step tolerance; 20 steps with equal pair
    sets at every step and states within 1e-3; then, at step 60, the
    aggregates the long-run criterion names (joint drift and kinetic
    energy) within 10% -- the chains whip, so ulp differences of the
    rotation logarithms' atan2f/sinf grow into millimetres by then."""
    say("4 chains x 64 capsule links")
    scene = S.chains(count=4, links=64)
    g = gpu_world(scene, profile=0, base_iterations=10)
    o = oracle_world(scene, profile=0, base_iterations=10)
    g.step(1)
    o.step(1)
    gb, ob = g.bodies(), o.bodies()
    for f in ("color", "parent", "impact", "joint_head"):
        np.testing.assert_array_equal(gb[f], ob[f], err_msg=f)
    assert_states_close(g.states(), o.states())
    assert_joints_match(g.joints(), o.joints())
    for k in range(19):
        g.step(1)
        o.step(1)
        np.testing.assert_array_equal(g.pairs(), o.pairs(), err_msg=f"step {k + 2}")
    assert_states_close(g.states(), o.states(), rtol=1e-3, atol=1e-3)
    g.step(40)
    o.step(40)
    gs, os_ = g.states(), o.states()
    m = scene.params["inverse_mass"][2:]
    ke = [float(np.sum(0.5 / m * np.sum(s["linear_velocity"][2:].astype(np.float64) ** 2, axis=1)))
          for s in (gs, os_)]
    dg, do = joint_gaps(scene, gs), joint_gaps(scene, os_)
    say(f"step 60: kinetic energy {ke[0]:.4g} / {ke[1]:.4g} J, mean joint gap {dg.mean():.3g} / {do.mean():.3g} m")
    assert abs(ke[0] - ke[1]) <= 0.1 * ke[1]
    assert abs(dg.mean() - do.mean()) <= 0.1 * do.mean()


def test_batched_scenes_step_as_they_would_alone():
    """Eight scenes of 9-300 bodies in one world step bit for bit as each
    does alone: pairs never cross scenes and the triangle-path rules are
    the scene's own."""
    say("batch of 8 scenes vs each alone, 90 steps")
    b = S.batch(count=8, seed=11, max_bodies=300)
    g = b.build(pmamd.World(profile=0, base_iterations=10))
    assert g.scene_count == 8
    alone = [b.build_one(pmamd.World(profile=0, base_iterations=10), s) for s in range(8)]
    for k in range(90):
        g.step(1)
        for w in alone:
            w.step(1)
    st = g.states()
    off = 0
    for s, w in enumerate(alone):
        n = len(b.scenes[s][0])
        np.testing.assert_array_equal(st[off:off + n].view(np.uint32), w.states().view(np.uint32),
                                      err_msg=f"scene {s}")
        off += n
    assert g.stats()["manifolds"] == sum(w.stats()["manifolds"] for w in alone)


def test_autotuner_picks_a_lane_tile_and_keeps_simulating():
    b = S.batch(count=16, seed=3, max_bodies=64)
    w = b.build(pmamd.World(profile=0, base_iterations=10))
    w.step(5)
    best, times = pmamd.autotune_lanes(w, steps=2)
    say(f"lane tile {best}: {times}")
    assert best in (8, 16, 32, 64) and set(times) == {8, 16, 32, 64}
    w.step(2)
    assert np.all(np.isfinite(w.states()["position"]))

