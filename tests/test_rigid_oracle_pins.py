"""Pins of the oracle's rigid-body restatement (oracle/rigid_oracle.cpp) and
bitwise agreement of the product's host rigid pipeline with it.

The reference has no numeric fixtures for its rigid step; its tests are
property checks (tests/python/test_collision.py, test_free_body_sim.py).  The
oracle is pinned by those properties, by brute force (sweep and prune against
all pairs) and by closed forms of the solver (a converged single contact, the
supported weight of a resting box, Coulomb deceleration, momentum), each
chosen so that a sign, index or term slip fails it.  Then the product's World
(novaphy_amd, host C++) must equal the oracle bit for bit over long rollouts.
"""
import math

import numpy as np
import pytest

import novaphy_amd as nv
from oracle import oracle as O

F32 = np.float32
G = 9.81


def _pose(p, q=(1, 0, 0, 0)):
    return np.array([*p, *q], F32)


def _box_inertia(m, he):
    b = nv.RigidBody.from_box(m, np.array(he, F32))
    return b.inertia


# ------------------------------------------------------------ broadphase
def test_oracle_sap_equals_brute_force():
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 3, (80, 3)).astype(F32)
    he = rng.uniform(0.05, 0.4, c.shape).astype(F32)
    c[:4] = c[4]
    lo, hi = c - he, c + he
    st = rng.uniform(size=80) < 0.3
    want = [(i, j) for i in range(80) for j in range(i + 1, 80)
            if not (st[i] and st[j]) and np.all(lo[i] <= hi[j]) and np.all(hi[i] >= lo[j])]
    assert O.sap_pairs(lo, hi, st) == want
    assert O.sap_pairs(np.array([[-1, -1, -1], [2, 2, 2]]), np.array([[0, 0, 0], [3, 3, 3]]), [0, 0]) == []
    assert O.sap_pairs(np.array([[-1, -1, -1], [0, 0, 0]]), np.array([[1, 1, 1], [2, 2, 2]]), [0, 0]) == [(0, 1)]
    assert O.sap_pairs(np.array([[-1, -1, -1], [0, 0, 0]]), np.array([[1, 1, 1], [2, 2, 2]]), [1, 1]) == []


# ------------------------------------------------------------ narrowphase
def test_oracle_narrowphase_reference_properties():  # test_collision.py:50-161
    sa, sb = O.rigid_shape(O.SHAPE_R_SPHERE, 0, radius=1.0), O.rigid_shape(O.SHAPE_R_SPHERE, 1, radius=1.0)
    c = O.collide_shapes(sa, _pose((0, 0, 0)), sb, _pose((1.5, 0, 0)))
    assert len(c) == 1 and c[0]["penetration"] > 0
    np.testing.assert_allclose(c[0]["normal"], [1, 0, 0], atol=1e-5)
    s5a, s5b = O.rigid_shape(O.SHAPE_R_SPHERE, 0, radius=0.5), O.rigid_shape(O.SHAPE_R_SPHERE, 1, radius=0.5)
    assert O.collide_shapes(s5a, _pose((0, 0, 0)), s5b, _pose((2, 0, 0))) == []
    plane = O.rigid_shape(O.SHAPE_R_PLANE, -1, normal=(0, 1, 0), offset=0.0, restitution=0.0)
    c = O.collide_shapes(sa, _pose((0, 0.5, 0)), plane, _pose((0, 0, 0)))
    assert len(c) == 1 and c[0]["penetration"] == pytest.approx(0.5, abs=1e-5)
    assert (c[0]["body_a"], c[0]["body_b"]) == (-1, 0)  # the plane is side A
    box = O.rigid_shape(O.SHAPE_R_BOX, 0, half=(0.5, 0.5, 0.5))
    c = O.collide_shapes(box, _pose((0, 0.4, 0)), plane, _pose((0, 0, 0)))
    assert len(c) == 4 and all(x["penetration"] > 0 for x in c)
    bb = O.rigid_shape(O.SHAPE_R_BOX, 1, half=(0.5, 0.5, 0.5))
    c = O.collide_shapes(box, _pose((0, 0.5, 0)), bb, _pose((0, 1.49, 0)))
    assert len(c) == 4 and all(x["penetration"] > 0 and x["normal"][1] > 0.9 for x in c)


def test_oracle_box_box_axes():
    a = O.rigid_shape(O.SHAPE_R_BOX, 0, half=(0.5, 0.5, 0.5))
    b = O.rigid_shape(O.SHAPE_R_BOX, 1, half=(0.5, 0.5, 0.5))
    assert O.collide_shapes(a, _pose((0, 0, 0)), b, _pose((0.2, 0.3, 1.01))) == []
    c = O.collide_shapes(a, _pose((0, 0, 0)), b, _pose((-0.9, 0.1, 0.05)))
    assert c and all(np.allclose(x["normal"], [-1, 0, 0], atol=1e-5) for x in c)
    assert all(x["penetration"] == pytest.approx(0.1, abs=1e-5) for x in c)
    # b turned 45 degrees about z, one vertex 0.03 m into a's top face
    h = math.sin(math.pi / 8)
    q = (math.cos(math.pi / 8), 0.0, 0.0, h)
    c = O.collide_shapes(a, _pose((0, 0, 0)), b, _pose((0.0, 0.5 + 0.5 * math.sqrt(2) - 0.03, 0.0), q))
    assert c
    for x in c:
        np.testing.assert_allclose(x["normal"], [0, 1, 0], atol=1e-4)
        assert x["penetration"] == pytest.approx(0.03, abs=2e-5)
    # edge against edge: b turned 45 degrees about x then placed so that its
    # lower edge (along x) crosses a's upper edge (along z) of a turned copy;
    # the contact is the centres' midpoint with the cross-product axis
    qa = (math.cos(math.pi / 8), 0.0, 0.0, math.sin(math.pi / 8))   # 45 deg about z: an edge on top along z
    qb = (math.cos(math.pi / 8), math.sin(math.pi / 8), 0.0, 0.0)   # 45 deg about x: an edge at the bottom along x
    top = 0.5 * math.sqrt(2)
    pb = (0.0, 2 * top - 0.02, 0.0)
    c = O.collide_shapes(a, _pose((0, 0, 0), qa), b, _pose(pb, qb))
    assert len(c) == 1
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-4)
    assert c[0]["penetration"] == pytest.approx(0.02, abs=1e-4)
    np.testing.assert_allclose(c[0]["position"], np.array(pb) * 0.5, atol=1e-6)


def test_oracle_box_sphere_flip():
    box = O.rigid_shape(O.SHAPE_R_BOX, 0, half=(0.5, 0.25, 0.5))
    s = O.rigid_shape(O.SHAPE_R_SPHERE, 1, radius=0.2)
    c = O.collide_shapes(box, _pose((0, 0, 0)), s, _pose((0.1, 0.4, 0.0)))
    np.testing.assert_allclose(c[0]["normal"], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(c[0]["position"], [0.1, 0.25, 0.0], atol=1e-6)
    assert c[0]["penetration"] == pytest.approx(0.05, abs=1e-6)
    c2 = O.collide_shapes(s, _pose((0.1, 0.4, 0.0)), box, _pose((0, 0, 0)))
    np.testing.assert_allclose(c2[0]["normal"], [0, -1, 0], atol=1e-6)
    assert (c2[0]["body_a"], c2[0]["body_b"]) == (1, 0)


# ------------------------------------------------------------ solver
def test_oracle_head_on_contact_closed_form():
    """Two unit spheres meet head on at 2 m/s each, no gravity, overlap below
    the slop (no Baumgarte bias), restitution 0.3: the converged single
    contact ends with relative normal velocity e * v0 = 0.3 * (-4), as the
    reference's restitution bias (rel_vel_n < -1 -> bias += e rel_vel_n,
    free_body_solver.cpp:183-197) prescribes, with no friction impulse and
    the total momentum unchanged."""
    ma, mb = 1.0, 3.0
    bodies = [O.rigid_body(ma, np.eye(3) * 0.4 * ma * 0.25), O.rigid_body(mb, np.eye(3) * 0.4 * mb * 0.25)]
    shapes = [O.rigid_shape(O.SHAPE_R_SPHERE, 0, radius=0.5), O.rigid_shape(O.SHAPE_R_SPHERE, 1, radius=0.5)]
    w = O.RigidWorld(bodies, shapes, [_pose((0, 1, 0)), _pose((0.998, 1, 0))], gravity=(0, 0, 0))
    w.lin[0] = [2.0, 0, 0]
    w.lin[1] = [-2.0, 0, 0]
    w.step(1.0 / 120.0)
    assert len(w.contacts) == 1
    c = w.contacts[0]
    assert c["acc_t1"] == 0.0 and c["acc_t2"] == 0.0
    vrel = w.lin[1][0] - w.lin[0][0]
    assert vrel == pytest.approx(0.3 * -4.0, rel=1e-5)
    assert ma * w.lin[0][0] + mb * w.lin[1][0] == pytest.approx(2.0 * ma - 2.0 * mb, abs=1e-5)
    # impulse = reduced mass * change of relative velocity
    assert c["acc_n"] == pytest.approx(ma * mb / (ma + mb) * (-1.2 + 4.0), rel=1e-5)


def _resting_box_world(mu_pair=0.5, v0=0.0, steps=0):
    he = (0.5, 0.25, 0.5)
    bodies = [O.rigid_body(2.0, _box_inertia(2.0, he))]
    shapes = [O.rigid_shape(O.SHAPE_R_PLANE, -1, normal=(0, 1, 0), friction=mu_pair, restitution=0.0),
              O.rigid_shape(O.SHAPE_R_BOX, 0, half=he, friction=mu_pair)]
    w = O.RigidWorld(bodies, shapes, [_pose((0.0, 0.25, 0.0))])
    w.lin[0] = [v0, 0, 0]
    for _ in range(steps):
        w.step(1.0 / 120.0)
    return w


def test_oracle_resting_box_carries_its_weight():
    w = _resting_box_world(steps=300)
    w.step(1.0 / 120.0)
    assert len(w.contacts) == 4
    assert sum(c["acc_n"] for c in w.contacts) == pytest.approx(2.0 * G / 120.0, rel=0.02)
    assert abs(w.lin[0][1]) < 0.02 and abs(w.pose7[0][1] - 0.25) < 0.01


def test_oracle_coulomb_deceleration():
    """A box sliding at 3 m/s on the ground loses mu g dt of speed per step
    (mu = sqrt(0.5 * 0.5) = 0.5, combine_friction) while it slides."""
    w = _resting_box_world(v0=0.0, steps=120)  # settle first
    w.lin[0][0] = 3.0
    v = [float(w.lin[0][0])]
    for _ in range(30):
        w.step(1.0 / 120.0)
        v.append(float(w.lin[0][0]))
    dv = -np.diff(v)[5:]
    np.testing.assert_allclose(dv, 0.5 * G / 120.0, rtol=0.05)


def test_oracle_free_fall_and_body_frame_torque():
    he = (0.1, 0.2, 0.3)
    I = _box_inertia(2.0, he)
    w = O.RigidWorld([O.rigid_body(2.0, I)], [O.rigid_shape(O.SHAPE_R_BOX, 0, half=he)], [_pose((0, 10, 0))])
    dt = F32(1.0 / 120.0)
    y, v = F32(10.0), F32(0.0)
    for _ in range(5):
        w.step(float(dt))
        v = F32(v + F32(F32(-G) * dt))
        y = F32(y + F32(v * dt))
    assert w.lin[0][1] == v and w.pose7[0][1] == y
    # torque about z on a body turned 90 degrees about x: the reference applies
    # the BODY-frame inverse inertia (integrator.cpp:19), so w_z uses I_zz
    q = (math.cos(math.pi / 4), math.sin(math.pi / 4), 0.0, 0.0)
    w = O.RigidWorld([O.rigid_body(2.0, I)], [O.rigid_shape(O.SHAPE_R_BOX, 0, half=he)], [_pose((0, 10, 0), q)],
                     gravity=(0, 0, 0))
    w.apply_torque(0, [0.0, 0.0, 1.0])
    w.step(float(dt))
    assert w.ang[0][2] == pytest.approx(0.999 / I[2, 2] * float(dt), rel=1e-5)


# ------------------------------------------------------------ product == oracle
def _scene(seed, n_bodies=10):
    """The same random scene as a product Model and as oracle structs: boxes
    and spheres with rotated local transforms, some on rotated bodies, one
    static box, a ground plane, bodies dropped onto each other."""
    rng = np.random.default_rng(seed)
    b = nv.ModelBuilder()
    ob, os_, poses = [], [], []
    b.add_ground_plane(y=0.0, friction=0.6, restitution=0.1)
    os_.append(O.rigid_shape(O.SHAPE_R_PLANE, -1, normal=(0, 1, 0), offset=0.0, friction=0.6, restitution=0.1))
    for k in range(n_bodies):
        box = k % 2 == 0
        pos = np.array([rng.uniform(-0.6, 0.6), 0.3 + 0.45 * k, rng.uniform(-0.6, 0.6)], F32)
        ax = rng.normal(size=3)
        ang = rng.uniform(0, math.pi)
        q = np.array([math.cos(ang / 2), *(math.sin(ang / 2) * ax / np.linalg.norm(ax))], F32)
        q = (q / np.linalg.norm(q)).astype(F32)
        m = float(rng.uniform(0.5, 3.0)) if k != 3 else 0.0
        fr, re = float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.0, 0.5))
        if box:
            he = rng.uniform(0.1, 0.3, 3).astype(F32)
            body = nv.RigidBody.from_box(m, he) if m > 0 else nv.RigidBody.make_static()
        else:
            r = float(rng.uniform(0.1, 0.25))
            body = nv.RigidBody.from_sphere(m, r) if m > 0 else nv.RigidBody.make_static()
        i = b.add_body(body, nv.Transform(pos, np.array([q[1], q[2], q[3], q[0]], F32)))
        lq = np.array([0.98, 0.1, -0.1, 0.12], F32)
        lq = (lq / np.linalg.norm(lq)).astype(F32)
        lp = np.array([0.01, -0.02, 0.015], F32)
        local = nv.Transform(lp, np.array([lq[1], lq[2], lq[3], lq[0]], F32))
        if box:
            b.add_shape(nv.CollisionShape.make_box(he, i, local, fr, re))
            os_.append(O.rigid_shape(O.SHAPE_R_BOX, i, half=he, local_pos=lp, local_rot=lq, friction=fr,
                                     restitution=re))
        else:
            b.add_shape(nv.CollisionShape.make_sphere(r, i, local, fr, re))
            os_.append(O.rigid_shape(O.SHAPE_R_SPHERE, i, radius=r, local_pos=lp, local_rot=lq, friction=fr,
                                     restitution=re))
        ob.append(O.rigid_body(body.mass, body.inertia))
    model = b.build()
    world = nv.World(model)
    # the product's initial transforms (normalised as its Transform does)
    for t in world.state.transforms:
        r = np.array(t.rotation, F32)  # x, y, z, w
        poses.append(np.array([*t.position, r[3], r[0], r[1], r[2]], F32))
    return world, O.RigidWorld(ob, os_, poses)


def _product_state(world):
    p = np.array([[*t.position, t.rotation[3], *t.rotation[:3]] for t in world.state.transforms], F32)
    lv = np.array(world.state.linear_velocities, F32)
    av = np.array(world.state.angular_velocities, F32)
    return p, lv, av


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_product_world_equals_oracle_bitwise(seed):
    """200 steps of a pile of boxes and spheres (box-box SAT faces and edges,
    box-sphere, sphere-sphere, plane contacts, a static body, external forces
    and torques on some steps): every body's pose and velocities, and every
    step's contacts with their accumulated impulses, equal the oracle's."""
    world, ow = _scene(seed)
    dt = 1.0 / 120.0
    seen = set()
    for k in range(200):
        if k % 7 == 0:
            f, t = np.array([0.5, 0.0, -0.3], F32), np.array([0.0, 0.05, 0.02], F32)
            world.apply_force(1, f)
            world.apply_torque(1, t)
            ow.apply_force(1, f)
            ow.apply_torque(1, t)
        world.step(dt)
        ow.step(dt)
        p, lv, av = _product_state(world)
        np.testing.assert_array_equal(p, ow.pose7, err_msg=f"step {k}")
        np.testing.assert_array_equal(lv, ow.lin, err_msg=f"step {k}")
        np.testing.assert_array_equal(av, ow.ang, err_msg=f"step {k}")
        pc = world.contacts
        assert len(pc) == len(ow.contacts), k
        for a, b in zip(pc, ow.contacts):
            assert (a.body_a, a.body_b) == (b["body_a"], b["body_b"])
            np.testing.assert_array_equal(np.array(a.position, F32), b["position"])
            np.testing.assert_array_equal(np.array(a.normal, F32), b["normal"])
            assert a.accumulated_normal_impulse == b["acc_n"] and a.penetration == b["penetration"]
            assert a.accumulated_tangent_impulse_1 == b["acc_t1"] and a.accumulated_tangent_impulse_2 == b["acc_t2"]
            seen.add((a.body_a < 0, len(pc)))
    assert len(seen) > 1  # plane and body-body contacts both occurred
