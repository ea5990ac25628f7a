/*
 * beep_host.h -- host-side scene construction (C ABI).
 *
 * The caller side of the BEM mat-vec: reference meshes, rigid mesh instances,
 * node patches with their qualocation / quadrature / galerkin point sets, and
 * the adaptive octree that fixes the explicit (near-field) pair list.  These
 * replace, for the hot path, the reference's host objects:
 *   Mesh / MeshInstance            src/bem/mesh.h, src/bem/mesh_instance.h:23-87
 *   BasicNodePatch point sets      src/bem/node_patch.cpp:55-273
 *   BEEP::calculate_bounding_cube  src/beep/beep.cpp:1652-1693
 *   BEEP::init_fmm_octree          src/beep/beep.cpp:1585-1641 (octree.h)
 * and emit the flat arrays consumed by the device boundary (beep_hip.h).
 *
 * All functions return 0 on success and a negative code on error; the text of
 * the last error on the calling thread is available from beep_host_last_error().
 */
#ifndef BEEP_HOST_H
#define BEEP_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct beep_mesh beep_mesh;
typedef struct beep_scene beep_scene;

const char* beep_host_last_error(void);

/* ---- reference meshes (planar node patches, one per vertex) ---- */
int beep_mesh_icosphere(int subdivisions, double radius, beep_mesh** out);
/* r(theta,phi) = r0 (1 + sum_{l=2..lmax} a_lm Y_lm), a_lm ~ N(0, sigma0/l) */
int beep_mesh_star(int subdivisions, double r0, int lmax, double sigma0, uint64_t seed,
                   beep_mesh** out);
/* charges: xyzq = n x (x, y, z, q) in the mesh frame */
int beep_mesh_set_charges(beep_mesh* m, int n, const double* xyzq);
int beep_mesh_charges_ball(beep_mesh* m, int n, double rmax, double qmin, double qmax,
                           uint64_t seed);
/* random-walk chain inside a star mesh (needs the mesh from beep_mesh_star) */
int beep_mesh_charges_walk(beep_mesh* m, int n, double step, double clearance, double qsigma,
                           uint64_t seed);
/* counts and the reference-frame centre (charge centroid) and radius */
int beep_mesh_info(const beep_mesh* m, int* n_vertices, int* n_triangles, int* n_charges,
                   double* centre3, double* radius);
int beep_mesh_vertices(const beep_mesh* m, double* xyz, int* tri3);
/* Surface files: .off (src/bem/off_utils.h:26-160) or .gts (gts_utils.h:32-140);
 * charges: lines "x y z q [r]" (the xyzqr member of a .mtz, mesh_tarball.h:52-75) */
int beep_mesh_read(const char* path, beep_mesh** out);
int beep_mesh_write_off(const beep_mesh* m, const char* path);
int beep_mesh_read_charges(beep_mesh* m, const char* path);
/* A .mtz mesh tarball (MeshTarball, mesh_tarball.h:27-140 and
 * mesh_tarball.cpp:118-270, as Mesh::init(mtz) loads it, mesh.cpp:88-151):
 * a tar archive, gzip-compressed or plain (bzip2 and compress are refused
 * with a message), whose definition.xml (root <contents>) names the members
 * <mesh> (.off or .gts), <xyzqr> (charges), <centre> (x y z: the mesh centre,
 * kept instead of the charge centroid) and <energies> (mandatory in the
 * definition; the curved-patch energy precalcs, recomputed here as for the
 * reference's planar patches), and optionally <fh> (f h per node patch,
 * beep_mesh_initial_fh). */
int beep_mesh_read_mtz(const char* path, beep_mesh** out);
/* f and h per node patch (n_vertices each) from a .mtz fh member
 * (Mesh::init_fh_vals, mesh.cpp:630-650): returns 0, or 1 if the mesh has
 * none. */
int beep_mesh_initial_fh(const beep_mesh* m, double* f, double* h);
/* Curved node patches (on = 1) or planar ones (0, the default): the
 * reference's PNG1 triangles (PNG1_Triangle, src/bem/png1.h:86-346; chosen
 * by Mesh::init unless force_planar, mesh.cpp:65-72), with vertex normals
 * from the area-weighted planar normals (Vertex::set_normal, vertex.cpp:
 * 86-141) and the planar normals across each edge (triangle.h:737-757).  The
 * node patches' points, normals, weights (surface jacobian), centroids and
 * areas follow the curved surface.  Scenes with a curved mesh need host point
 * sets (the device point generation is planar). */
int beep_mesh_set_curved(beep_mesh* m, int on);
/* Point and unit normal of curved triangle `tri` (the mesh's triangle order)
 * at barycentric (u, v) (w = 1 - u - v weights its first vertex): the
 * surface PNG1_Triangle::get_point / get_normal evaluate (png1.h:135-207). */
int beep_mesh_curved_point(const beep_mesh* m, int tri, double u, double v, double* p3, double* n3);
/* independent copy (surface, patches, charges, star shape) */
int beep_mesh_clone(const beep_mesh* m, beep_mesh** out);
void beep_mesh_destroy(beep_mesh* m);

/* ---- scene: BEEP(Dsolvent, kappa, quad, qual, nbsize) + instances ---- */
int beep_scene_create(double Dsolvent, double kappa, int quad_rule, int qual_rule, int nbsize,
                      beep_scene** out);
/* pose: rotation quaternion (w,x,y,z) about the mesh centre, then the centre
   is placed at offset3 (MeshInstance, mesh_instance.cpp:23-87) */
int beep_scene_add_instance(beep_scene* s, const beep_mesh* m, const double* quat4,
                            const double* offset3, double Dprotein);
/* patches, bounding cube, octree, point sets */
int beep_scene_build(beep_scene* s);

typedef struct {
    int n_patches;
    int n_instances;
    int n_charges;
    int n_qual;        /* total qualocation points */
    int n_quad;        /* total quadrature points */
    int n_galerkin;    /* total galerkin (4-point rule) points */
    int n_leaves;
    int n_near_cols;   /* sum over leaves of the neighbourhood size */
    long long n_near_pairs; /* sum over leaves of |leaf| * |neighbourhood| */
    int tree_depth;    /* bottom level of the octree */
    double cube_centre[3];
    double cube_edge;
    int n_triangles;         /* surface triangles over all instances */
    int n_vertex_triangles;  /* sum over vertices of the triangles around them */
} beep_scene_sizes;
int beep_scene_get_sizes(const beep_scene* s, beep_scene_sizes* out);

/* exports (caller-allocated, sizes from beep_scene_get_sizes) */
/* centroid3, normal3: 3N; area, eps_ratio (Dsolvent/Dprotein), gc: N; inst: N */
int beep_scene_patches(const beep_scene* s, double* centroid3, double* normal3, double* area,
                       double* eps_ratio, double* gc, int* instance);
/* kind 0 = qualocation, 1 = quadrature, 2 = galerkin;
   ptr: N+1, pts: n x 7 (x, y, z, nx, ny, nz, w) */
int beep_scene_points(const beep_scene* s, int kind, int* ptr, double* pts);
/* Leave the point sets to the device (on = 0): build() then skips them and
 * beep_scene_points returns only the pointer arrays; the device makes the
 * points from beep_scene_surface (beep_hip_set_surface).  Default on. */
int beep_scene_set_host_points(beep_scene* s, int on);
/* Leave out the octree (on = 0): build() then makes the patches only, with
 * no leaves (for a scene whose near field comes from elsewhere, as the
 * instances of a pose batch, beep_serve.h).  Default on. */
int beep_scene_set_tree(beep_scene* s, int on);
/* The surface in universe coordinates: placed vertices (3 x n_patches: patch
 * p is vertex p), triangles (3 x n_triangles, vertex indices), and for every
 * vertex the triangles around it (vt_ptr: n_patches + 1, vt_tri and the
 * corner 0/1/2 of the vertex in each: n_vertex_triangles), in the order the
 * patch's points are generated (node_patch.cpp:55-273). */
int beep_scene_surface(const beep_scene* s, double* verts3, int* tris3, int* vt_ptr, int* vt_tri,
                       int* vt_corner);
/* The curved triangles of the surface (meshes set curved, PNG1_Triangle,
 * src/bem/png1.h:86-121): n_curved, then per triangle (n_triangles, the
 * order of beep_scene_surface's tris) its index among the curved ones or -1
 * for a planar triangle, and the 21 placed control points (63 doubles) of
 * each curved one, in the order b300 b030 b003, p1_210 p1_201 p2_120 p2_021
 * p3_102 p3_012, p2_201 p2_102 p3_120 p3_210 p1_012 p1_021, c12_1 c12_2
 * c23_2 c23_3 c31_3 c31_1 (evaluated by beep3_amd/csrc/curved.hpp).  Any
 * pointer may be NULL; the device generates curved point sets from these
 * (beep_hip_surface). */
int beep_scene_surface_curved(const beep_scene* s, int* n_curved, int* ctl_idx, double* ctl);
/* leaf_ptr: L+1, leaf_idx: N, nb_ptr: L+1, nb_idx: n_near_cols */
int beep_scene_tree(const beep_scene* s, int* leaf_ptr, int* leaf_idx, int* nb_ptr, int* nb_idx);
/* posed charges: xyzq n x 4, charge_instance n */
int beep_scene_charges(const beep_scene* s, double* xyzq, int* charge_instance);
/* per-instance: Dprotein, first patch, first charge (n_instances + 1 entries for
   the two offsets) */
int beep_scene_instances(const beep_scene* s, double* Dprotein, int* patch_ptr, int* charge_ptr);
void beep_scene_destroy(beep_scene* s);

#ifdef __cplusplus
}
#endif

#endif /* BEEP_HOST_H */
