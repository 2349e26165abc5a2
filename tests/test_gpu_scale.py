# SPDX-License-Identifier: MIT
"""A world past the sizes that grids of one thread per body reach when
capped (4096 blocks of 256 threads: 1 048 576 bodies; 4096 blocks of 64:
262 144 bodies): a lattice of 102^3 analytic spheres whose candidate pairs
and contacts have closed forms.

Spheres of radius 0.1 m with no collision margin have proxy boxes of
half-extent h = 0.1 + 0.25 * 0.2 (the correction reach) + 1e-5 ~ 0.15 m. At a
spacing of 0.19 m (h < 0.19 < 2h) the boxes of the 26 lattice neighbours
overlap and those of second neighbours do not; only the 6 face neighbours
(0.19 < 0.2 apart) touch, the diagonal ones (0.19 * sqrt 2 = 0.27 apart) do
not. A body past either limit that is left out changes both counts.
"""
import sys

import numpy as np
import pytest

import pmamd
import scenarios as S


def lattice_pairs(n):
    """Unordered 26-neighbour pairs of an n x n x n lattice: 3 axis, 6 face-
    diagonal and 4 body-diagonal directions."""
    return 3 * n * n * (n - 1) + 6 * n * (n - 1) ** 2 + 4 * (n - 1) ** 3


def lattice_faces(n):
    return 3 * n * n * (n - 1)


def test_closed_forms_on_small_lattices():
    # brute force on 2^3 and 3^3 lattices
    for n in (2, 3):
        p = np.indices((n, n, n)).reshape(3, -1).T
        d = np.abs(p[:, None, :] - p[None, :, :]).max(axis=2)
        s = np.abs(p[:, None, :] - p[None, :, :]).sum(axis=2)
        assert (d == 1).sum() // 2 == lattice_pairs(n)
        assert (s == 1).sum() // 2 == lattice_faces(n)


def dimer_positions(nx, ny, nz, gap=1.0, offset=0.19):
    """Two bodies per site of an nx x ny x nz grid of spacing `gap`: the
    site and the site + (offset, 0, 0)."""
    site = np.indices((nx, ny, nz)).reshape(3, -1).T.astype(np.float32) * np.float32(gap)
    pos = np.repeat(site, 2, axis=0)
    pos[1::2, 0] += np.float32(offset)
    return pos


@pytest.mark.gpu
@pytest.mark.skip(reason="the first step of this 305 252-body world (zero gravity, zero collision margin, analytic "
                         "spheres only) did not return within 120 s; no cause has been shown, so it stays off the GPU")
def test_dimers_past_262144_bodies_find_every_pair():
    """152 626 isolated dimers (305 252 bodies, two-body islands): each
    dimer's spheres (r 0.1 m, 0.19 m apart) touch and overlap as proxies,
    and no two dimers' proxies overlap (0.81 m apart, proxy half-extent
    ~0.15 m), so one step has exactly one candidate pair, manifold and
    contact point per dimer.  A query that skipped the bodies past index
    262 144 loses the dimers those bodies own."""
    pos = dimer_positions(67, 67, 34)
    n = len(pos)
    print(f"{n // 2} dimers ({n} spheres), one step", flush=True)
    sys.stdout.flush()
    w = pmamd.World(gravity=(0.0, 0.0, 0.0), profile=0, base_iterations=2, pairs_per_body=16, points_per_body=16)
    shape = w.add_sphere(0.1)
    w.add_bodies(S.params(n, shape=shape, margin=0.0, mass=1.0, inertia=(0.004, 0.004, 0.004)), S.states(n, pos))
    w.step(1)
    st = w.stats()
    print(f"  pairs {st['pairs']}, manifolds {st['manifolds']}, points {st['points']}, islands {st['islands']}",
          flush=True)
    assert st["status"] == 0
    assert st["pairs"] == st["manifolds"] == st["points"] == n // 2
    out = w.states()["position"]
    assert np.isfinite(out).all()
    # the last dimer (both bodies past the limit) was pushed apart along x
    assert out[n - 1][0] > pos[n - 1][0] and out[n - 2][0] < pos[n - 2][0]
    w.close()


@pytest.mark.gpu
@pytest.mark.skip(reason="the first step of this 1 061 208-body world (default pools: 256 points per body, about "
                         "180 GB of device memory; one island of every body) did not return within 120 s; no cause "
                         "has been shown, so it stays off the GPU")
def test_lattice_beyond_a_million_bodies_finds_every_neighbour():
    side = 102
    n = side ** 3
    print(f"lattice of {n} spheres (r 0.1 m, spacing 0.19 m), one step", flush=True)
    sys.stdout.flush()
    w = pmamd.World(gravity=(0.0, 0.0, 0.0), profile=0, base_iterations=2)
    shape = w.add_sphere(0.1)
    pos = np.indices((side, side, side)).reshape(3, -1).T.astype(np.float32) * np.float32(0.19)
    w.add_bodies(S.params(n, shape=shape, margin=0.0, mass=1.0, inertia=(0.004, 0.004, 0.004)),
                 S.states(n, pos))
    w.step(1)
    st = w.stats()
    print(f"  pairs {st['pairs']}, manifolds {st['manifolds']}, points {st['points']}", flush=True)
    assert st["status"] == 0
    assert st["pairs"] == lattice_pairs(side)
    assert st["manifolds"] == lattice_faces(side)
    assert st["points"] == lattice_faces(side)
    # The last body (a corner, above both limits) was integrated and pushed
    # outward by its three contacts.
    out = w.states()["position"]
    assert np.isfinite(out).all()
    corner = n - 1
    assert (out[corner] > pos[corner]).all()
    w.close()
