"""GPU tests of the rigid-body step (lrh_rigid_*): the reference's own known answers for its GPU
rigid bodies (test_gpu_rigid.c:48-122: free fall, a bounce's peak, momentum through collisions, a
resting stack, a tilted box landing on a face) on the device, each also stepped by the reference's
CPU solver (engine/phys/rigid.c built from its sources: tests/phys_ref.py) where the two must
agree; and a pile of 60 balls and boxes dropped in the room on both, compared by what must hold
whatever the solver's order (test_gpu_rigid.c:246-277: everyone inside the room, no ball sunk into
another, settled). Sleeping is off on both (the solver alone, test_gpu_rigid.c:26-31)."""
import math
import sys

import numpy as np
import pytest

import little_renderer_amd as lr
import phys_ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DT = 1.0 / 240   # the reference's test steps: 1/240 s in 4 substeps (test_gpu_rigid.c:37-42)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "these tests need the GPU"
    c = lr.Context(0)
    yield c
    c.close()


def max_speed(b):
    """the fastest point of any body (centre speed + spin x half-size diagonal; test_phys_util.h)"""
    return max(math.sqrt(sum(x * x for x in v.vel)) + math.sqrt(sum(x * x for x in v.angvel)) *
               math.sqrt(sum(x * x for x in v.half)) for v in b)


def world(ctx, **params):
    r = lr.Rigid(ctx, max_bodies=256, max_contacts=8192)
    if params:
        r.set_params(**params)
    return r


def axis_angle(axis, angle):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    s = math.sin(angle / 2)
    return (a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2))


def test_free_fall_matches_the_closed_form_and_the_reference(ctx):
    r = world(ctx, angular_damping=0.0)
    r.add_sphere((0, 1000, 0), 10, 1)
    r.steps(DT, 4, 240)
    b = r.bodies()[0]
    say(f"free fall 1 s: v {b.vel[1]:.3f}, y {b.pos[1]:.3f}")
    assert abs(b.vel[1] + 981) < 0.5 and abs(b.pos[1] - (1000 - 0.5 * 981)) < 1.5
    if phys_ref.lib():
        w = phys_ref.World(angular_damping=0.0)
        w.sphere((0, 1000, 0), 10, 1)
        w.steps(DT, 4, 240)
        o = w.bodies()[0]
        w.close()
        # the same semi-implicit Euler in the same order: the same numbers to float rounding
        assert abs(o[1] - b.pos[1]) < 1e-3 and abs(o[8] - b.vel[1]) < 1e-3
    r.close()


def test_bounce_peak(ctx):
    r = world(ctx, room=((-500, 0, -500), (500, 1000, 500)), wall_restitution=0.5, angular_damping=0.0)
    i = r.add_sphere((0, 410, 0), 10, 1)
    b = r.bodies()
    b[0].restitution = 0.5
    r.set_bodies(0, [b[0]])
    top, bounced = 0.0, False
    for _ in range(240 * 2):
        r.steps(DT, 4, 1)
        s = r.bodies(i, 1)[0]
        if s.vel[1] > 0:
            bounced = True
        if bounced:
            top = max(top, s.pos[1])
    say(f"bounce: peak {top:.2f} (expected 110)")
    assert bounced and abs(top - (10 + 0.25 * 400)) < 4
    r.close()


def test_momentum_through_collisions(ctx):
    r = world(ctx, gravity=(0, 0, 0), angular_damping=0.0)
    a = r.add_sphere((0, 0, 0), 20, 1)
    r.add_sphere((100, 8, 0), 25, 2)
    r.add_box((0, 120, 3), (20, 15, 25), axis_angle((1, 2, 0.5), 0.7), 1.5)
    r.add_box((110, 125, -6), (18, 22, 15), axis_angle((0.3, 1, 1), 1.2), 0.8)
    b = list(r.bodies())
    for k, vx in enumerate((200, -150, 180, -160)):
        b[k].vel[0] = vx
    r.set_bodies(0, b)

    def momentum(bs):
        return np.sum([np.array(x.vel[:]) / x.inv_mass for x in bs], axis=0)

    p0 = momentum(r.bodies())
    scale = 200 * (1 / b[0].inv_mass) * 4
    r.steps(DT, 4, 240)
    bs = r.bodies()
    p1 = momentum(bs)
    say(f"momentum: change / total {np.linalg.norm(p1 - p0) / scale:.2e}, sphere a vx {bs[a].vel[0]:.1f}")
    assert np.linalg.norm(p1 - p0) / scale < 1e-3
    assert bs[a].vel[0] < 0          # it hit the other sphere and turned around
    r.close()


def test_resting_stack(ctx):
    r = world(ctx, room=((0, 0, 0), (555, 555, 555)))
    for k in range(3):
        r.add_box((278, 30 + 60 * k, 278), (40, 30, 40))
    s = r.add_sphere((278, 180 + 25, 278), 25)
    r.steps(DT, 4, 720)
    b = r.bodies()
    st = r.stats()
    say(f"stack: top box y {b[2].pos[1]:.3f}, ball y {b[s].pos[1]:.3f}, fastest {max_speed(b):.3f}; "
        f"{st.contacts} contacts, {st.colours} colours, {st.ms * 1e3:.1f} us a step")
    assert abs(b[2].pos[1] - 150) < 1.5
    assert abs(b[2].pos[0] - 278) + abs(b[2].pos[2] - 278) < 1
    assert abs(b[s].pos[1] - 205) < 1.5
    assert max_speed(b) < 1
    if phys_ref.lib():
        w = phys_ref.World()
        w.room((0, 0, 0), (555, 555, 555))
        for k in range(3):
            w.box((278, 30 + 60 * k, 278), (40, 30, 40))
        w.sphere((278, 205, 278), 25)
        w.steps(DT, 4, 720)
        o = w.bodies()
        w.close()
        # the same rest: heights within the contact slop (0.5) of each other
        assert np.all(np.abs(o[:, 1] - np.array([x.pos[1] for x in b])) < 0.6)
    r.close()


def test_tilted_box_lands_on_a_face(ctx):
    r = world(ctx, room=((0, 0, 0), (555, 555, 555)))
    i = r.add_box((278, 300, 278), (50, 20, 35), axis_angle((1, 0.3, 0.5), 0.9))
    r.steps(DT, 4, 960)
    b = r.bodies()
    y = b[i].pos[1]
    say(f"tilted box: y {y:.3f}, fastest {max_speed(b):.3f}")
    assert min(abs(y - 50), abs(y - 20), abs(y - 35)) < 1.5
    assert max_speed(b) < 1
    r.close()


def pile_spec(n, side, seed=1234):
    """the reference's pile (test_gpu_rigid.c:224-242): a grid of drop points over the room's
    floor, jittered; every third body a box of random size and turn (density 0.7), the others
    balls of radius 18-32 and density 1-7 (the generator is numpy's: the same shapes, not the
    same draws)"""
    rng = np.random.default_rng(seed)
    per = int(side / 75)
    out = []
    for i in range(n):
        at = np.array([45 + 75 * (i % per), 60 + 75 * (i // (per * per)), 45 + 75 * ((i // per) % per)], np.float64)
        at += [rng.random() * 8 - 4, 0, rng.random() * 8 - 4]
        if i % 3 == 2:
            h = 15 + 15 * rng.random(3)
            ax = rng.random(3) - 0.5
            out.append(("box", tuple(at), tuple(h), axis_angle(ax, 3 * rng.random()), 0.7))
        else:
            out.append(("sphere", tuple(at), 18 + 14 * rng.random(), 1 + 6 * rng.random()))
    return out


def pile_checks(who, pos, half, shape, side, calm):
    outside, worst = 0, 0.0
    for i in range(len(pos)):
        r = half[i][0] if shape[i] == 0 else min(half[i])
        x, y, z = pos[i]
        if x < r - 2 or z < r - 2 or x > side - r + 2 or z > side - r + 2 or y < r - 2:
            outside += 1
        if shape[i] != 0:
            continue
        for j in range(i + 1, len(pos)):
            if shape[j] != 0:
                continue
            worst = max(worst, half[i][0] + half[j][0] - float(np.linalg.norm(np.array(pos[i]) - pos[j])))
    say(f"pile ({who}): outside {outside}, deepest ball-ball overlap {worst:.3f}, calm {calm:.2f} units/s")
    assert outside == 0
    assert worst < 2.5
    assert calm < 10


def test_pile_settles_like_the_reference(ctx):
    n, side = 60, 555.0
    spec = pile_spec(n, side)
    r = lr.Rigid(ctx, max_bodies=256, max_contacts=16384)
    r.set_params(room=((0, 0, 0), (side, 4 * side, side)))
    for s in spec:
        if s[0] == "box":
            r.add_box(s[1], s[2], s[3], s[4])
        else:
            r.add_sphere(s[1], s[2], s[3])
    r.steps(DT, 4, 2400)        # 10 s

    def fastest_vy(bs):
        return max(abs(x.vel[1]) for x in bs)

    calm = fastest_vy(r.bodies())
    for _ in range(2):
        r.steps(DT, 4, 240)
        calm = min(calm, fastest_vy(r.bodies()))
    b = r.bodies()
    st = r.stats()
    say(f"pile on the device: {st.contacts} contacts, {st.pairs} pairs, {st.colours} colours, {st.ms * 1e3:.1f} us a step")
    pile_checks("device", [tuple(x.pos) for x in b], [tuple(x.half) for x in b], [x.shape for x in b], side, calm)
    if phys_ref.lib():
        w = phys_ref.World()
        w.room((0, 0, 0), (side, 4 * side, side))
        for s in spec:
            if s[0] == "box":
                w.box(s[1], s[2], s[3], s[4])
            else:
                w.sphere(s[1], s[2], s[3])
        w.steps(DT, 4, 2400)
        o = w.bodies()
        calm_c = float(np.max(np.abs(o[:, 8])))
        for _ in range(2):
            w.steps(DT, 4, 240)
            o = w.bodies()
            calm_c = min(calm_c, float(np.max(np.abs(o[:, 8]))))
        w.close()
        pile_checks("reference CPU", [tuple(p) for p in o[:, 0:3]], [tuple(x.half) for x in b], [x.shape for x in b], side,
                    calm_c)
        # the same pile: its mean height within a few units of the reference's
        hd, hc = np.mean([x.pos[1] for x in b]), float(np.mean(o[:, 1]))
        say(f"  mean height: device {hd:.2f}, reference {hc:.2f}")
        assert abs(hd - hc) < 10
    r.close()


def test_rigid_step_moves_an_instanced_scene(ctx):
    # the physics-driven instanced workload at a small size: its dynamic instances as bodies
    # dropped among its props, each frame one step (1/60 s), their poses the scene's moved
    # records and the top level rebuilt (lrh_pt_scene_move); after 2 s the bodies lie on the
    # ground inside the room, and the scene equals one made anew with their final poses
    w = lr.instanced_workload(0, 0, n_dynamic=120, n_static=150, n_props=3)
    rigid, idx, scales = lr.instanced_rigid(ctx, w, n_dynamic=120)
    ps = lr.PtScene.instanced(ctx, w["mtris"], w["mesh_first"], w["mats"], w["instances"])
    recs = w["instances"][idx].copy()
    for _ in range(120):
        rigid.steps(1.0 / 60, 4, 1)
        recs["xform"] = lr.rigid_xforms(rigid.bodies(0, len(idx)), scales)
        ps.move(idx, recs, builder=lr.PLOC)   # (the top level as a scene made anew builds it)
    b = rigid.bodies(0, len(idx))
    st = rigid.stats()
    low = min(x.pos[1] - (x.half[0] if x.shape == lr.BODY_SPHERE else 0.0) for x in b)
    say(f"instanced: {len(idx)} moving bodies, lowest surface {low:.2f} cm, {st.contacts} contacts, "
        f"{st.colours} colours, {st.ms * 1e3:.0f} us a step")
    for x in b:
        r = x.half[0] if x.shape == lr.BODY_SPHERE else min(x.half)
        assert x.pos[1] > r - 2.0 and abs(x.pos[0]) < 2000 and abs(x.pos[2]) < 2000
    final = w["instances"].copy()
    final["xform"][idx] = recs["xform"]
    fresh = lr.PtScene.instanced(ctx, w["mtris"], w["mesh_first"], w["mats"], final)
    rng = np.random.default_rng(5)
    n = 20000
    o = rng.uniform(-19, 19, size=(n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(2, 6, n)
    t = rng.uniform(-19, 19, size=(n, 3)).astype(np.float32)
    t[:, 1] = 0.0
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 1e-4, d, 1e30
    ha, hb = ps.trace(rays), fresh.trace(rays)
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    ps.close()
    fresh.close()
    rigid.close()
