/*
 * qpso_scene.h -- synthetic RGB-D scenes and cluster ROIs.
 *
 * No counterpart in the reference's hot path: this stands in for the RGB-D
 * capture plus the host-side GraphCannySeg segmentation (GraphCannySeg.h,
 * GraphCannySeg.cpp), which only supply the per-cluster ROI that the
 * estimate-pose block consumes (main.cu:1269-1322):
 *   rect_aabb_   -> qpso_cluster.rect (cv::Rect; br exclusive)
 *   centroid3D   -> qpso_cluster.centroid_mm (GraphCannySeg.h:2569-2597:
 *                   the cluster's non-zero depths back-projected through the
 *                   float K^-1 and averaged in float)
 *   num_points   -> qpso_cluster.num_points (every cluster pixel, zero depth
 *                   included, GraphCannySeg.h:1824)
 *   clusterDepth -> the cluster-masked depth frame (GraphCannySeg.h:1805-1809)
 * Meshes are produced in millimetres, fp32, in the reference's loadOBJ3
 * struct-of-arrays layout (main.h:620-679): for face f and vertex slot j,
 * x = A[f + jF], y = A[3F + f + jF], z = A[6F + f + jF].
 * All functions are host-only and deterministic for a given seed.
 */
#ifndef QPSO_SCENE_H_
#define QPSO_SCENE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Icosphere with `subdiv` subdivisions (20 * 4^subdiv faces), scaled to the
 * semi-axes (ax, ay, az) mm.  Call with out == NULL to get the face count. */
int qpso_mesh_icosphere(int subdiv, float ax, float ay, float az, float* out_soa);

/* UV ellipsoid (2 * slices * (stacks - 1) faces) plus `n_boxes` axis-aligned
 * boxes (12 faces each) of edge U(box_min, box_max) mm centred on random
 * surface points, seeded by `seed`.  Call with out == NULL for the count. */
int qpso_mesh_uv_ellipsoid(int slices, int stacks, float ax, float ay, float az,
                           int n_boxes, float box_min, float box_max, uint64_t seed,
                           float* out_soa);

/* Uniformly random unit quaternion (w,x,y,z) from 4 x N(0,1), seeded. */
void qpso_random_quat(uint64_t seed, float q[4]);

typedef struct {
    const float* mesh_soa;   /* mm, SoA */
    int n_faces;
    float pose[7];           /* tx,ty,tz (mm), q0..q3 (scalar first) */
} qpso_scene_object;

/* Ray-cast depth sensor: for every pixel centre (x, y) the nearest
 * intersection depth (z, mm) over all objects and a background plane at
 * z = plane_z (plane_z <= 0: no plane), then + N(0, sigma) mm, rounded to
 * uint16, then a fraction `hole_frac` of all pixels set to 0.
 * depth_out: 480*640 uint16; label_out (optional): 480*640 int32, the index of
 * the object seen at each pixel or -1. */
int qpso_scene_render(const qpso_scene_object* objs, int n_objs,
                      double fx, double fy, double cx, double cy,
                      float plane_z, float sigma, float hole_frac, uint64_t seed,
                      uint16_t* depth_out, int32_t* label_out);

typedef struct {
    int rect[4];             /* minx, miny, maxx, maxy (max exclusive) */
    float centroid_mm[3];
    uint32_t num_points;
} qpso_cluster_roi;

/* Cluster of object `obj` (its ground-truth mask from label): ROI fields and
 * the cluster-masked depth frame (480*640 uint16, zero outside the mask).
 * Returns 0, or -1 if the object is not visible. */
int qpso_scene_cluster(const uint16_t* depth, const int32_t* label, int obj,
                       double fx, double fy, double cx, double cy,
                       qpso_cluster_roi* roi, uint16_t* cluster_depth_out);

#ifdef __cplusplus
}
#endif
#endif
