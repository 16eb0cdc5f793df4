/*
 * cdx_scene.h -- CPU contact generation feeding the device contact buffer.
 *
 * Contact generation stays on the CPU (the reference runs it through Coal in
 * CollisionCoal::detectAll, src/collision/collision_coal.cpp:405-548).  These functions
 * produce contacts with the reference's conventions:
 *   - normal from A to B, contact point at the midpoint of the two witness points, tangents
 *     from tangentFrameFromNormal (tangent_frame.hpp:23-50), body-frame copies through R^T
 *     (makeContact, collision_coal.cpp:112-163);
 *   - friction combined per pair with "min" (combineFrictionMu, :94-106);
 *   - angular Jacobian rows r_body x d_body per side, the input of buildContactRowByDof
 *     (contact_jacobian.hpp:17-23);
 *   - warm start by greedy nearest-point matching per body pair within a maximum distance
 *     (ContactTracker::matchContactsForPair, contact_tracker.cpp:43-72).
 * Output arrays follow the layout of include/cdx/cdx.h.
 */
#ifndef CDX_SCENE_H
#define CDX_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Right-handed orthonormal (t1, t2) from a contact normal (tangent_frame.hpp:23-50). */
void cdx_tangent_frame(const double n[3], double t1[3], double t2[3]);

/* Sphere / static half-space contacts.
 *   q[7*nb]       : x,y,z, qx,qy,qz,qw per sphere (rigid bodies);
 *   radius[nb], mu_body[nb];
 *   walls         : nw half-spaces {x : n.x >= d} (n[3*nw] unit inward normal, d[nw]), mu_wall[nw];
 *   margin        : report pairs with gap <= margin (the collision security margin,
 *                   config.hpp:31);
 * Contacts are emitted per body a (ascending): walls first, then spheres b > a (ascending).
 * Sphere-wall contacts have A = the wall (-1), B = the sphere.
 * Writes at most cap contacts; returns the number found (which may exceed cap) or -1.
 * point[3*c] and depth[c] (penetration >= 0) are written for warm-start matching. */
int cdx_sphere_contacts(int nb, const double* q, const double* radius, const double* mu_body, int nw,
                        const double* wall_n, const double* wall_d, const double* mu_wall, double margin, int cap,
                        int* body_a, int* body_b, double* mu, double* frame, double* angA, double* angB,
                        double* point, double* depth);

/* Spheres and boxes in nscenes independent scenes (bodies of scene s: [body_off[s], body_off[s+1]);
 * only bodies of the same scene collide; every scene sees the same static half-spaces).
 * shape[b]: 0 sphere (dims[3b] = radius), 1 box (dims[3b..3b+2] = half extents).
 * Box-halfspace: every corner within the margin; box-sphere: closest point; box-box:
 * separating-axis test, then the incident face clipped against the reference face (face axes)
 * or one closest-point contact (edge-edge axes).  Contacts of scene s are
 * [contact_off[s], contact_off[s+1]) when contact_off is given.  Return value as
 * cdx_sphere_contacts. */
int cdx_shape_contacts(int nscenes, const int* body_off, const double* q, const int* shape, const double* dims,
                       const double* mu_body, int nw, const double* wall_n, const double* wall_d, const double* mu_wall,
                       double margin, int cap, int* body_a, int* body_b, double* mu, double* frame, double* angA,
                       double* angB, double* point, double* depth, int* contact_off);

/* A static heightfield shared by all scenes: samples h[i + nx*j] at (x0 + i*dx, y0 + j*dy),
 * nx, ny >= 2; each cell is split along its (i+1, j)-(i, j+1) diagonal into two triangles;
 * no terrain outside the sampled rectangle. */
typedef struct {
    int nx, ny;
    double x0, y0, dx, dy;
    const double* h;
    double mu; /* friction of the terrain (combined with the body's by "min") */
} cdx_heightfield;

/* cdx_shape_contacts with capsules (shape 2: dims[3b] = radius, dims[3b+1] = half length of the
 * segment along the body z axis) and an optional heightfield (NULL: none).  Capsule pairs:
 * wall -- the two segment ends; sphere -- nearest segment point; capsule -- closest points of
 * the two segments (the two ends of the overlap when parallel); box -- sphere probes at the
 * ends and at the segment point nearest the box centre.  Heightfield probes (one contact each,
 * terrain = side A): sphere centre, capsule ends and midpoint, box corners.  Pairs listed in
 * (dis_a[k], dis_b[k]), k < ndisabled, never collide (CollisionCoal::disablePair,
 * collision_coal.cpp:558-570; e.g. bodies sharing a joint).  Per body a the order is walls,
 * heightfield, then pairs (a, b > a). */
int cdx_shape_contacts_hf(int nscenes, const int* body_off, const double* q, const int* shape, const double* dims,
                          const double* mu_body, int nw, const double* wall_n, const double* wall_d, const double* mu_wall,
                          const cdx_heightfield* terrain, int ndisabled, const int* dis_a, const int* dis_b, double margin,
                          int cap, int* body_a, int* body_b, double* mu, double* frame, double* angA, double* angB,
                          double* point, double* depth, int* contact_off);

/* Random sequential placement of bounding spheres (radius bound[b]) without overlap inside the
 * box [lo, hi] of each scene, largest first; a scene's column grows upward when the box is too
 * full.  pos[3*nb]. */
int cdx_scene_rsa(int nscenes, const int* body_off, const double* bound, const double* lo, const double* hi, uint64_t seed,
                  double* pos);

/* Warm start: for each current contact, the impulse of the nearest unused previous contact of
 * the same body pair within max_dist, greedily by increasing distance
 * (contact_tracker.cpp:43-72); zero if none. */
int cdx_match_warmstart(int nprev, const int* prev_a, const int* prev_b, const double* prev_point, const double* prev_imp,
                        int ncur, const int* cur_a, const int* cur_b, const double* cur_point, double max_dist,
                        double* warm_out);

/* Jittered-grid sphere drop (1000 spheres in a 1x1x2 box): positions on an nx x ny x nz
 * lattice of pitch `pitch`, first layer centre at z0, uniform jitter +-jitter in x and y;
 * random orientations.  Returns the count placed. */
int cdx_scene_sphere_grid(int n, int nx, int ny, double pitch, double z0, double jitter, uint64_t seed, double* q);

/* Settled polydisperse sphere pile: radii log-uniform in [rmin, rmax], random sequential
 * placement in the box [0,Lx]x[0,Ly]x[0,Lz], then a position-based gravity settle of
 * `settle_steps` steps.  Bodies are returned in Morton (spatial) order.  Returns n or -1. */
int cdx_scene_sphere_pile(int n, double rmin, double rmax, double Lx, double Ly, double Lz, uint64_t seed, int settle_steps,
                          double* q, double* radius);

#ifdef __cplusplus
}
#endif
#endif
