// qpso_main.cpp -- drop-in CLI for the reference's `main` (main.cu:1198-1780):
//
//   qpso_main <object_model.obj> <rgb_image> <depth_image.dpt> <cluster_idx>
//
// Same positional arguments and the same usage message on a wrong count
// (main.cu:1200-1204).  The mesh is read like loadOBJ3 (metres), the depth
// like loadDepth (ACCV .dpt).  The cluster ROIs come from a label map
// `<depth_image>.labels.dpt` (uint16, label k+1 = cluster k), the place the
// reference runs GraphCannySeg; rgb_image only names the outputs, as the
// segmentation that would read it stays outside this hot path.  Outputs match
// main.cu:1749-1780: results/<rgb>-<N>-bestpose.csv, results/<rgb>-<N>-best.csv
// and the CPU renders of the first 10 particles' final poses, 0.csv .. 9.csv.
// Like the reference's 6x512-thread render grid (main.h:162-163) it renders
// faces 0..3071 only; QPSO_FACE_CAP overrides that (0: every face).
// QPSO_OFFFRAME=1 selects the reference's own off-frame indexing (pixels past
// the left or right edge land in the neighbouring row; qpso_params.offframe).
// Optional environment: QPSO_PARTICLES, QPSO_ITERS, QPSO_FX/FY/CX/CY,
// QPSO_FACE_CAP, QPSO_OFFFRAME.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "qpso.h"
#include "qpso_io.h"

static double env_d(const char* k, double d) {
    const char* v = std::getenv(k);
    return v ? std::atof(v) : d;
}

int main(int argc, char** argv) {
    if (argc != 5) {
        std::printf("Usage: %s path/to/.obj color_image depth_image cluster_idx\n", argv[0]);
        return 0;
    }
    const std::string obj = argv[1], rgb = argv[2], dpt = argv[3];
    const int idx = std::atoi(argv[4]);
    std::string rgb_name = rgb.substr(rgb.find_last_of('/') == std::string::npos ? 0 : rgb.find_last_of('/') + 1);

    float* mesh = nullptr;
    uint32_t F = 0;
    if (int r = qpso_read_obj(obj.c_str(), &mesh, &F)) {
        std::fprintf(stderr, "cannot read mesh %s: %s\n", obj.c_str(), qpso_status_str(r));
        return 1;
    }
    uint16_t *depth = nullptr, *labels = nullptr;
    int32_t rows = 0, cols = 0, lr = 0, lc = 0;
    if (qpso_read_dpt(dpt.c_str(), &depth, &rows, &cols) || rows != QPSO_ROWS || cols != QPSO_COLS) {
        std::fprintf(stderr, "cannot read 480x640 depth %s\n", dpt.c_str());
        return 1;
    }
    const std::string lab = dpt + ".labels.dpt";
    if (qpso_read_dpt(lab.c_str(), &labels, &lr, &lc) || lr != QPSO_ROWS || lc != QPSO_COLS) {
        std::fprintf(stderr, "cannot read cluster label map %s\n", lab.c_str());
        return 1;
    }
    qpso_params p;
    qpso_params_default(&p);
    p.n_particles = (int)env_d("QPSO_PARTICLES", p.n_particles);
    p.n_iters = (int)env_d("QPSO_ITERS", p.n_iters);
    p.cam.fx = env_d("QPSO_FX", p.cam.fx);
    p.cam.fy = env_d("QPSO_FY", p.cam.fy);
    p.cam.cx = env_d("QPSO_CX", p.cam.cx);
    p.cam.cy = env_d("QPSO_CY", p.cam.cy);
    p.face_cap = (int)env_d("QPSO_FACE_CAP", 3072);  // fxptBlocks * fxptThsPerBlock
    p.offframe = (int)env_d("QPSO_OFFFRAME", 0);     // 1: the reference's off-frame indexing

    std::vector<uint16_t> cdepth(QPSO_ROWS * QPSO_COLS);
    qpso_cluster cl{};
    if (qpso_clusters_from_labels(labels, depth, idx, &p.cam, cl.rect, cl.centroid_mm, &cl.num_points,
                                  cdepth.data())) {
        std::fprintf(stderr, "cluster %d not present\n", idx);
        return 1;
    }
    cl.depth = cdepth.data();
    std::printf("Segmented Cluster Points: %u\n", cl.num_points);

    qpso_ctx* ctx = nullptr;
    if (int r = qpso_create(0, &ctx)) {
        std::fprintf(stderr, "qpso_create: %s\n", qpso_status_str(r));
        return 1;
    }
    uint32_t mid = 0;
    qpso_result res{};
    qpso_solve* solve = nullptr;
    std::vector<float> pos(7 * (size_t)std::max(p.n_particles, 0));
    int r = qpso_mesh_upload(ctx, mesh, F, &mid);
    if (!r) r = qpso_solve_create(ctx, mid, &cl, &p, 0, &solve);
    if (!r) r = qpso_solve_run(solve, &res);
    if (!r) {
        qpso_trace tr{};
        tr.pos = pos.data();  // d_pso_pos after the last iteration (main.cu:1764)
        r = qpso_solve_get_trace(solve, &tr);
    }
    if (solve) qpso_solve_destroy(solve);
    if (r) {
        std::fprintf(stderr, "estimate_pose: %s (%s)\n", qpso_status_str(r), qpso_last_error(ctx));
        qpso_destroy(ctx);
        return 1;
    }
    std::printf("Solution Best Fit: %f\nSolution Best Pose: [ %f, %f, %f, %f, %f, %f, %f ]\n", res.fit,
                res.pose[0], res.pose[1], res.pose[2], res.pose[3], res.pose[4], res.pose[5], res.pose[6]);
    const std::string base = "results/" + rgb_name + "-" + std::to_string(idx);
    qpso_write_bestpose((base + "-bestpose.csv").c_str(), res.fit, res.pose);
    qpso_render_cpu_csv((base + "-best.csv").c_str(), res.pose, mesh, F, &p.cam);
    // the first 10 particles' final poses (main.cu:1766-1780)
    for (int k = 0; k < 10 && k < p.n_particles; ++k) {
        float pose[7];
        for (int d = 0; d < 7; ++d) pose[d] = pos[(size_t)d * p.n_particles + k];
        qpso_render_cpu_csv((std::to_string(k) + ".csv").c_str(), pose, mesh, F, &p.cam);
    }
    qpso_destroy(ctx);
    qpso_free(mesh);
    qpso_free(depth);
    qpso_free(labels);
    return 0;
}
