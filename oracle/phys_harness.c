// phys_harness.c - the reference's own CPU rigid bodies (engine/phys/rigid.c, convex.c), built
// from its sources where they lie, behind plain-array entry points for the tests. TEST
// INFRASTRUCTURE ONLY: the rigid-body tests load it as the checker of the device's rigid-body
// step; the product library never does. It calls the reference's functions and restates none.
//
// A body as a flat array of 14 floats: position (3), orientation quaternion x, y, z, w (4),
// velocity (3), angular velocity (3), and 1 if it sleeps.
#include "phys/rigid.h"

#include <stdlib.h>

void *ph_create(void)
{
    RigidWorld *w = calloc(1, sizeof(RigidWorld));
    if (w) rigid_init(w);   // rigid.c:122-132 (gravity 981 down, 10 iterations, warm start)
    return w;
}
void ph_free(void *p)
{
    RigidWorld *w = p;
    if (!w) return;
    rigid_free(w);
    free(w);
}
// sleep: 0 for the solver alone, as the reference's GPU tests run it (test_gpu_rigid.c:26-31)
void ph_params(void *p, float gx, float gy, float gz, float linear_damping, float angular_damping, int iterations, int sleep,
               float wall_restitution, float wall_friction)
{
    RigidWorld *w = p;
    w->gravity = vec3(gx, gy, gz);
    w->linear_damping = linear_damping;
    w->angular_damping = angular_damping;
    w->iterations = iterations;
    w->sleep = sleep;
    w->wall_restitution = wall_restitution;
    w->wall_friction = wall_friction;
}
void ph_room(void *p, const float *lo, const float *hi)
{
    rigid_set_room((RigidWorld *)p, (Aabb){vec3(lo[0], lo[1], lo[2]), vec3(hi[0], hi[1], hi[2])});
}
int ph_sphere(void *p, const float *c, float radius, float density)
{
    return rigid_add_sphere((RigidWorld *)p, vec3(c[0], c[1], c[2]), radius, density);
}
// q: x, y, z, w
int ph_box(void *p, const float *c, const float *half, const float *q, float density)
{
    return rigid_add_box((RigidWorld *)p, vec3(c[0], c[1], c[2]), vec3(half[0], half[1], half[2]), quat(q[3], q[0], q[1], q[2]),
                         density);
}
// a body's velocities, bounce and friction; fixed (inv_mass and inverse inertia 0) if fixed
void ph_set(void *p, int i, const float *vel, const float *angvel, float restitution, float friction, int fixed)
{
    Body *b = &((RigidWorld *)p)->bodies[i];
    b->vel = vec3(vel[0], vel[1], vel[2]);
    b->angvel = vec3(angvel[0], angvel[1], angvel[2]);
    b->restitution = restitution;
    b->friction = friction;
    if (fixed) {
        b->inv_mass = 0;
        b->inv_inertia = vec3(0, 0, 0);
    }
    b->dirty = 1;
}
void ph_steps(void *p, float dt, int substeps, int n)
{
    for (int k = 0; k < n; k++) rigid_steps((RigidWorld *)p, dt, substeps, NULL);
}
void ph_get(void *p, int i, float *o)
{
    const Body *b = &((RigidWorld *)p)->bodies[i];
    o[0] = b->pos.x; o[1] = b->pos.y; o[2] = b->pos.z;
    o[3] = b->rot.x; o[4] = b->rot.y; o[5] = b->rot.z; o[6] = b->rot.w;
    o[7] = b->vel.x; o[8] = b->vel.y; o[9] = b->vel.z;
    o[10] = b->angvel.x; o[11] = b->angvel.y; o[12] = b->angvel.z;
    o[13] = (float)b->sleeping;
}
int ph_contacts(void *p) { return ((RigidWorld *)p)->ncontacts; }
