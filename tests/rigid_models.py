"""Test helper: the oracle's rigid-world description of a novaphy_amd Model
(bodies, shapes and the current transforms of a World), so that a product
World / FluidWorld and the oracle step the same bodies."""
import numpy as np

from oracle import oracle as O

F32 = np.float32
_KIND = {"Box": O.SHAPE_R_BOX, "Sphere": O.SHAPE_R_SPHERE, "Plane": O.SHAPE_R_PLANE}


def pose7(t):
    r = np.array(t.rotation, F32)  # x, y, z, w
    return np.array([*t.position, r[3], r[0], r[1], r[2]], F32)


def oracle_rigid(model, transforms, solver_settings=None, gravity=(0.0, -9.81, 0.0)):
    bodies = [O.rigid_body(b.mass, np.array(b.inertia, F32)) for b in model.bodies]
    shapes = []
    for s in model.shapes:
        lt = s.local_transform
        lq = np.array(lt.rotation, F32)
        shapes.append(O.rigid_shape(_KIND[s.type.name], s.body_index, half=np.array(s.half_extents, F32),
                                    radius=s.radius, normal=np.array(s.normal, F32), offset=s.offset,
                                    local_pos=np.array(lt.position, F32),
                                    local_rot=(lq[3], lq[0], lq[1], lq[2]), friction=s.friction,
                                    restitution=s.restitution))
    kw = {}
    if solver_settings is not None:
        kw = dict(velocity_iterations=solver_settings.velocity_iterations, baumgarte=solver_settings.baumgarte,
                  slop=solver_settings.slop, warm_starting=solver_settings.warm_starting)
    poses = np.array([pose7(t) for t in transforms], F32).reshape(-1, 7)
    return O.RigidWorld(bodies, shapes, poses, gravity=gravity, **kw)


def world_state(world):
    p = np.array([pose7(t) for t in world.state.transforms], F32).reshape(-1, 7)
    lv = np.array(world.state.linear_velocities, F32).reshape(-1, 3)
    av = np.array(world.state.angular_velocities, F32).reshape(-1, 3)
    return p, lv, av


def boundary_arrays(world):
    bp = world.boundary_particles
    local = np.array([b.local_position for b in bp], F32).reshape(-1, 3)
    body = np.array([b.body_index for b in bp], np.int32)
    psi = np.array([b.volume for b in bp], F32)
    return local, body, psi
