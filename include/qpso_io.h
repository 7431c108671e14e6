/*
 * qpso_io.h -- input/output formats around the hot path (host only).
 *
 *   qpso_read_obj         replaces loadOBJ3 (main.h:620-679, assimp): first
 *                         object of a Wavefront OBJ, polygons fan-triangulated,
 *                         emitted in the SoA layout x(f,j) = A[f + jF],
 *                         y = A[3F + f + jF], z = A[6F + f + jF]
 *   qpso_read_dpt         replaces loadDepth (main.cu:35-64): ACCV .dpt =
 *                         int32 rows, int32 cols, rows*cols uint16 (mm)
 *   qpso_write_dpt        the same format, for label maps and synthetic depth
 *   qpso_write_bestpose   results/<rgb>-<N>-bestpose.csv (main.cu:1752-1757)
 *   qpso_render_cpu_csv   RenderCPU + saveBestParticleRender2MatlabCSV
 *                         (main.h:1810-2043): visualisation render with
 *                         iround = (int)(v + 0.5f) and a float buffer
 *   qpso_clusters_from_labels
 *                         cluster-ROI provider from a label map (label k + 1
 *                         marks cluster k): the SegResults fields the
 *                         estimate-pose block reads (main.cu:1269-1322)
 */
#ifndef QPSO_IO_H_
#define QPSO_IO_H_

#include <stdint.h>

#include "qpso.h"

#ifdef __cplusplus
extern "C" {
#endif

int qpso_read_obj(const char* path, float** soa_out, uint32_t* n_faces);
int qpso_read_dpt(const char* path, uint16_t** data_out, int32_t* rows, int32_t* cols);
int qpso_write_dpt(const char* path, const uint16_t* data, int32_t rows, int32_t cols);
void qpso_free(void* p);

int qpso_write_bestpose(const char* path, float fit, const float pose[7]);
int qpso_render_cpu(const float pose[7], const float* soa, uint32_t n_faces, const qpso_camera* cam,
                    float* buffer /* 480*640, filled by the caller (60000.0f) */);
int qpso_render_cpu_csv(const char* path, const float pose[7], const float* soa, uint32_t n_faces,
                        const qpso_camera* cam);

/* Cluster `k` of a 480x640 label map; depth_out (480*640) receives the
 * cluster-masked depth.  Returns 0, or QPSO_E_NOTFOUND if the label is empty. */
int qpso_clusters_from_labels(const uint16_t* labels, const uint16_t* depth, int k,
                              const qpso_camera* cam, int32_t rect[4], float centroid_mm[3],
                              uint32_t* num_points, uint16_t* depth_out);

#ifdef __cplusplus
}
#endif
#endif
