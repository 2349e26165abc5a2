# SPDX-License-Identifier: MIT
"""The oracle's triangle-pair narrowphase (geometry.slang:949-1117,
1272-1584, 1855-2285; runtime_deformables.cu.inc:140-230;
runtime_domains.cu.inc:1111-1245) against independent checks.

The reference holds no vectors for this path, so each piece is pinned by
what the mathematics fixes:
  * closest points of segments, of a point and a triangle and of two
    triangles against double-precision minimisation by region enumeration
    (a different method from the oracle's Voronoi-region tests);
  * certified shell normals against the outward directions of closed convex,
    closed concave, open and flat meshes;
  * the BVH against its structural invariants (preorder, median split,
    leaf ranges partition the reordered triangles, bounds);
  * contacts against closed forms: crossing box edges (the case the convex
    face path declines) and a box resting on a concave mesh;
  * no tunnelling of a fast box into a concave mesh (swept pass).
The order in which leaf candidates reach the contact merge rule is not
covered by any of these (parity unpinned for that order).
"""
import numpy as np
import pytest

import _oracle as O
import scenarios as S
from pmamd import layouts as L

RNG_SEED = 11


# ---- independent double-precision geometry ---------------------------------
def seg_seg_distance(p0, p1, q0, q1):
    """min |p0 + s d0 - q0 - t d1| over [0,1]^2: the convex quadratic's
    interior stationary point if feasible, else the best of the four edges."""
    p0, p1, q0, q1 = (np.asarray(v, np.float64) for v in (p0, p1, q0, q1))
    d0, d1, r = p1 - p0, q1 - q0, p0 - q0

    def f(s, t):
        return np.linalg.norm(r + s * d0 - t * d1)

    def clamp1(val):
        return min(1.0, max(0.0, val))

    cands = []
    a, b, e = d0 @ d0, d0 @ d1, d1 @ d1
    c, g = d0 @ r, d1 @ r
    det = a * e - b * b
    if det > 1e-18:
        s = (b * g - c * e) / det
        t = (a * g - b * c) / det
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append(f(s, t))
    for s in (0.0, 1.0):
        t = clamp1((g + s * b) / e) if e > 0 else 0.0
        cands.append(f(s, t))
    for t in (0.0, 1.0):
        s = clamp1((t * b - c) / a) if a > 0 else 0.0
        cands.append(f(s, t))
    return min(cands)


def point_triangle_distance(p, tri):
    """Plane projection when inside, else the nearest of the three edges."""
    p = np.asarray(p, np.float64)
    a, b, c = (np.asarray(v, np.float64) for v in tri)
    n = np.cross(b - a, c - a)
    q = p - n * ((p - a) @ n) / (n @ n)
    inside = all(np.cross(v1 - v0, q - v0) @ n >= 0 for v0, v1 in ((a, b), (b, c), (c, a)))
    if inside:
        return np.linalg.norm(p - q)
    best = np.inf
    for v0, v1 in ((a, b), (b, c), (c, a)):
        d = v1 - v0
        t = min(1.0, max(0.0, (p - v0) @ d / (d @ d)))
        best = min(best, np.linalg.norm(p - (v0 + t * d)))
    return best


def triangles_intersect(A, B):
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)

    def seg_hits(p, q, T):
        n = np.cross(T[1] - T[0], T[2] - T[0])
        den = n @ (q - p)
        if abs(den) < 1e-15:
            return False
        t = n @ (T[0] - p) / den
        if t < 0 or t > 1:
            return False
        x = p + t * (q - p)
        return all(np.cross(T[(k + 1) % 3] - T[k], x - T[k]) @ n >= -1e-12 for k in range(3))

    return any(seg_hits(A[k], A[(k + 1) % 3], B) for k in range(3)) or \
        any(seg_hits(B[k], B[(k + 1) % 3], A) for k in range(3))


def triangle_distance(A, B):
    if triangles_intersect(A, B):
        return 0.0
    best = min(point_triangle_distance(v, B) for v in A)
    best = min(best, min(point_triangle_distance(v, A) for v in B))
    for i in range(3):
        for j in range(3):
            best = min(best, seg_seg_distance(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3]))
    return best


# ---- kernels -------------------------------------------------------------------
def test_closest_segments_against_region_enumeration():
    rng = np.random.default_rng(RNG_SEED)
    for k in range(400):
        p0, p1, q0, q1 = rng.uniform(-1, 1, size=(4, 3)).astype(np.float32)
        if k % 10 == 0:
            q1 = q0 + (p1 - p0) * np.float32(0.5)        # parallel
        if k % 17 == 0:
            p1 = p0.copy()                               # degenerate
        u, w = O.closest_segments(p0, p1, q0, q1)
        exact = seg_seg_distance(p0, p1, q0, q1)
        got = np.linalg.norm(u.astype(np.float64) - w)
        assert abs(got - exact) <= 2e-6, (k, got, exact)
        # The witnesses lie on their segments.
        assert seg_seg_distance(u, u, p0, p1) <= 2e-6
        assert seg_seg_distance(w, w, q0, q1) <= 2e-6


def test_closest_point_on_triangle_against_region_enumeration():
    rng = np.random.default_rng(RNG_SEED + 1)
    for k in range(400):
        tri = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        p = rng.uniform(-1.5, 1.5, size=3).astype(np.float32)
        q = O.closest_on_triangle(p, tri)
        exact = point_triangle_distance(p, tri)
        assert abs(np.linalg.norm(p.astype(np.float64) - q) - exact) <= 3e-6, k
        assert point_triangle_distance(q, tri) <= 3e-6


def test_closest_triangle_pair_against_region_enumeration():
    rng = np.random.default_rng(RNG_SEED + 2)
    hits = 0
    for k in range(300):
        A = rng.uniform(-1, 1, size=(3, 3)).astype(np.float32)
        B = (rng.uniform(-1, 1, size=(3, 3)) + rng.uniform(-0.8, 0.8, size=3)).astype(np.float32)
        a, b = O.closest_triangle_pair(A, B)
        exact = triangle_distance(A, B)
        got = np.linalg.norm(a.astype(np.float64) - b)
        assert abs(got - exact) <= 5e-6, (k, got, exact)
        assert point_triangle_distance(a, A) <= 5e-6
        assert point_triangle_distance(b, B) <= 5e-6
        if exact == 0.0:
            hits += 1
            np.testing.assert_array_equal(a, b)
    assert hits > 20


def test_feature_normal_face_and_edge_cases():
    # Vertex of A above the interior of B's face: the face normal, oriented
    # from B towards A.
    A = np.array([[0.2, 0.3, 0.1], [0.9, 0.8, 0.6], [0.1, 0.9, 0.7]], np.float32)
    B = np.array([[-1, 0, -1], [0, 0, 2], [2, 0, -1]], np.float32)
    a, b = O.closest_triangle_pair(A, B)
    np.testing.assert_allclose(a, [0.2, 0.3, 0.1], atol=1e-7)
    n = O.triangle_feature_normal(A, B, a, b, np.array([0, 0, 1], np.float32))
    np.testing.assert_allclose(n, [0, 1, 0], atol=1e-6)
    # Crossing edges: the normal is the normalised separation, parallel to
    # the cross product of the edge directions.
    A = np.array([[-1, 0.5, 0], [1, 0.5, 0], [0, 2, 0.3]], np.float32)
    B = np.array([[0, 0, -1], [0, 0, 1], [0.2, -2, 0]], np.float32)
    a, b = O.closest_triangle_pair(A, B)
    np.testing.assert_allclose(a, [0, 0.5, 0], atol=1e-6)
    np.testing.assert_allclose(b, [0, 0, 0], atol=1e-6)
    n = O.triangle_feature_normal(A, B, a, b, np.array([1, 0, 0], np.float32))
    np.testing.assert_allclose(n, [0, 1, 0], atol=1e-6)


# ---- meshes ------------------------------------------------------------------
def world_with(meshes):
    w = O.OracleWorld(gravity=(0, 0, 0), profile=0)
    for v, i in meshes:
        w.add_mesh(v, i)
    return w


def reordered_triangles(w, shape, vertices, count):
    idx = w.mesh_indices(shape, count).reshape(-1, 3)
    return vertices[idx]


def test_shell_normals_of_convex_box_are_outward_face_normals():
    v, i = S.box_mesh(0.5, 0.3, 0.2)
    w = world_with([(v, i)])
    shell = w.mesh_shell(0)
    tris = reordered_triangles(w, 0, v, len(i))
    assert shell.shape == (12, 3)
    centroids = tris.mean(axis=1)
    np.testing.assert_allclose(np.linalg.norm(shell, axis=1), 1.0, atol=1e-6)
    for n, c in zip(shell, centroids):
        assert np.abs(n).max() == pytest.approx(1.0)            # an axis
        assert n @ c > 0                                        # outward


def outward_normals_of_l(tris):
    """Geometric outward normal of each triangle of the L prism."""
    out = []
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0]).astype(np.float64)
        n /= np.linalg.norm(n)
        c = t.mean(axis=0)
        probe = c + 1e-3 * n
        inside = (abs(probe[2]) < 1) and (0 < probe[0] < 2) and (0 < probe[1] < 2) and \
            not (probe[0] > 1 and probe[1] > 1)
        out.append(-n if inside else n)
    return np.array(out)


@pytest.mark.parametrize("flip", [False, True])
def test_shell_normals_of_closed_concave_prism(flip):
    v, i = S.l_prism(flip)
    w = world_with([(v, i)])
    assert w.shape(0)[15] == 0                      # not convex: no solid planes
    shell = w.mesh_shell(0)
    tris = reordered_triangles(w, 0, v, len(i))
    np.testing.assert_allclose(np.linalg.norm(shell, axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(shell, outward_normals_of_l(tris), atol=1e-6)


def test_shell_normals_of_open_box_are_support_planes():
    v, i = S.box_mesh(0.5, 0.5, 0.5)
    tri = i.reshape(-1, 3)
    keep = [t for t in tri if not np.all(v[t][:, 1] > 0)]      # drop the top face
    i_open = np.array(keep, np.uint32).reshape(-1)
    w = world_with([(v, i_open)])
    shell = w.mesh_shell(0)
    tris = reordered_triangles(w, 0, v, len(i_open))
    assert shell.shape == (10, 3)
    np.testing.assert_allclose(np.linalg.norm(shell, axis=1), 1.0, atol=1e-6)
    for n, t in zip(shell, tris):
        assert n @ t.mean(axis=0) > 0


def test_flat_sheet_has_no_certified_normals():
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    i = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    w = world_with([(v, i)])
    assert w.mesh_shell(0).shape == (0, 3)


def test_saddle_sheet_marks_interior_triangles_two_sided():
    g = 5
    xs = np.linspace(-1, 1, g, dtype=np.float32)
    v = np.array([[x, x * z, z] for z in xs for x in xs], np.float32)
    idx = []
    for r in range(g - 1):
        for c in range(g - 1):
            a, b, d, e = r * g + c, r * g + c + 1, (r + 1) * g + c, (r + 1) * g + c + 1
            idx += [a, d, b, b, d, e]
    i = np.array(idx, np.uint32)
    w = world_with([(v, i)])
    shell = w.mesh_shell(0)
    lengths = np.linalg.norm(shell, axis=1)
    tris = w.mesh_indices(0, len(i)).reshape(-1, 3)
    edges = {}
    for t in tris:
        for k in range(3):
            e = tuple(sorted((int(t[k]), int(t[(k + 1) % 3]))))
            edges[e] = edges.get(e, 0) + 1
    boundary = np.array([any(edges[tuple(sorted((int(t[k]), int(t[(k + 1) % 3]))))] == 1
                             for k in range(3)) for t in tris])
    assert np.all(np.isclose(lengths[boundary], 0.0) | np.isclose(lengths[boundary], 1.0))
    assert np.all(np.isclose(lengths[~boundary], 0.25) | np.isclose(lengths[~boundary], 1.0))
    assert np.isclose(lengths, 0.25).sum() > 0


def test_bvh_structure():
    g = 11
    xs = np.linspace(0, 3, g, dtype=np.float32)
    rng = np.random.default_rng(RNG_SEED + 3)
    v = np.array([[x, rng.uniform(0, 0.2), z] for z in xs for x in xs], np.float32)
    idx = []
    for r in range(g - 1):
        for c in range(g - 1):
            a, b, d, e = r * g + c, r * g + c + 1, (r + 1) * g + c, (r + 1) * g + c + 1
            idx += [a, d, b, b, d, e]
    i = np.array(idx, np.uint32)
    T = len(i) // 3
    w = world_with([(v, i)])
    nodes, bounds = w.mesh_bvh(0)
    tris = w.mesh_indices(0, len(i)).reshape(-1, 3)
    # The reordered buffer is a permutation of the input triangles.
    assert sorted(map(tuple, tris)) == sorted(map(tuple, i.reshape(-1, 3)))

    def walk(k, begin, end):
        left, right, first, count = nodes[k]
        tri_v = v[tris[begin:end]].reshape(-1, 3)
        np.testing.assert_array_equal(bounds[k, :3], tri_v.min(axis=0))
        np.testing.assert_array_equal(bounds[k, 3:], tri_v.max(axis=0))
        if end - begin <= 4:
            assert (first, count) == (begin, end - begin)
            return k + 1
        assert count == 0 and left == k + 1
        mid = begin + (end - begin) // 2
        nxt = walk(left, begin, mid)
        assert right == nxt
        return walk(right, mid, end)

    assert walk(0, 0, T) == len(nodes)


# ---- contacts ------------------------------------------------------------------
def rot(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis /= np.linalg.norm(axis)
    s = np.sin(angle / 2)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(angle / 2)], np.float32)


def two_body_world(mesh_a, mesh_b, pos_a, q_a, pos_b, q_b, vel_a=(0, 0, 0), margin=0.01):
    w = O.OracleWorld(gravity=(0, 0, 0), profile=0, base_iterations=10)
    w.add_mesh(*mesh_b)
    w.add_mesh(*mesh_a)
    p = S.params(2, motion=np.array([L.MOTION_STATIC, L.MOTION_DYNAMIC], np.uint32),
                 shape=np.array([0, 1]), margin=margin)
    st = S.states(2, np.array([pos_b, pos_a], np.float32))
    st["orientation"][0] = q_b
    st["orientation"][1] = q_a
    st["linear_velocity"][1] = vel_a
    w.add_bodies(p, st)
    return w


def test_crossing_box_edges_take_the_triangle_path():
    h = np.float32(0.5)
    gap = np.float32(5e-4)
    top = h * np.float32(np.sqrt(2.0))
    w = two_body_world(S.box_mesh(h, h, h), S.box_mesh(h, h, h),
                       (0, 2 * top + gap, 0), rot((1, 0, 0), np.pi / 4),
                       (0, 0, 0), rot((0, 0, 1), np.pi / 4))
    w.step(1)
    st = w.stats()
    assert st["triangle_pairs"] == 1 and st["fallback_pairs"] == 0
    m = w.manifolds()
    assert m["count"][0] == 1
    c = m["contacts"][0][0]
    np.testing.assert_allclose(c["normal"], [0, 1, 0], atol=1e-5)
    np.testing.assert_allclose(c["point"], [0, top + gap / 2, 0], atol=1e-5)
    # penetration = rest offset - separation + 1e-5, rest = min(margin, 1e-3)
    assert c["penetration"] == pytest.approx(1e-3 - gap + 1e-5, abs=2e-6)
    assert c["collider_hinge"]["static_body"] == 1 and c["body_hinge"]["static_body"] == 0


def test_box_resting_on_concave_mesh():
    h = np.float32(0.1)
    gap = np.float32(5e-4)
    w = two_body_world(S.box_mesh(h, h, h), S.l_prism(), (0.5, 2 + h + gap, 0.0), rot((0, 1, 0), 0.0),
                       (0, 0, 0), rot((0, 1, 0), 0.0))
    w.step(1)
    assert w.stats()["triangle_pairs"] == 1
    m = w.manifolds()
    # One witness per triangle pair: parallel faces yield the first vertex
    # of each bottom triangle at the minimum distance, never an interior or
    # side point, so the contacts are distinct bottom corners.
    k = int(m["count"][0])
    assert 1 <= k <= 4
    cs = m["contacts"][0][:k]
    np.testing.assert_allclose(cs["normal"], np.tile([0, 1, 0], (k, 1)), atol=1e-6)
    corners = np.array([(x, z) for x in (0.4, 0.6) for z in (-0.1, 0.1)])
    d = np.abs(cs["point"][:, None, [0, 2]] - corners[None]).max(axis=2)
    assert np.all(d.min(axis=1) < 1e-5)
    assert len(set(d.argmin(axis=1))) == k
    np.testing.assert_allclose(cs["point"][:, 1], 2 + gap / 2, atol=1e-5)
    np.testing.assert_allclose(cs["penetration"], 1e-3 - gap + 1e-5, atol=2e-6)


def test_fast_box_does_not_tunnel_into_concave_mesh():
    h = np.float32(0.1)
    w = two_body_world(S.box_mesh(h, h, h), S.l_prism(), (0.5, 2.6, 0.0), rot((0, 1, 0), 0.3),
                       (0, 0, 0), rot((0, 1, 0), 0.0), vel_a=(0, -30.0, 0))
    lowest = np.inf
    for k in range(20):
        w.step(1)
        lowest = min(lowest, float(w.states()["position"][1][1]))
    # Bottom face stays above the top surface (y = 2) up to the rest offset
    # and solver tolerance, although one step moves the box 0.5 m.
    assert lowest > 2 + h - 0.02, lowest


def test_mesh_pile_settles_without_launching_bodies():
    side = 6
    scene = S.pile(side=side, seed=3, analytic_spheres=False)
    w = scene.build(O.OracleWorld(profile=0, base_iterations=10))
    top = float(scene.states["position"][:, 1].max())
    bound = np.sqrt(2 * 9.81 * (top + 1.0)) + 1.0      # free fall from the highest start
    worst, triangle_pairs = 0.0, 0
    for k in range(120):
        w.step(1)
        st = w.states()
        worst = max(worst, float(np.linalg.norm(st["linear_velocity"][5:], axis=1).max()))
        triangle_pairs += w.stats()["triangle_pairs"]
    assert triangle_pairs > 0
    assert worst < bound, (worst, bound)
    assert float(st["position"][5:, 1].min()) > 0.0     # nothing below the floor's top face
