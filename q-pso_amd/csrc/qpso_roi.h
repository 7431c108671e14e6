// qpso_roi.h -- the 3D centroid GraphCannySeg hands to the estimate-pose
// block, host only.  Shared by the label-map ROI provider (qpso_io.cpp) and
// the synthetic scenes (qpso_scene.cpp).
//
// compute3DCentroid (GraphCannySeg.h:2569-2597) back-projects every cluster
// pixel with non-zero depth through K_inv (projectPixel2CameraRF,
// GraphCannySeg.h:2510-2517), sums the points in float, and divides by their
// number in float.  K_inv = K_.inv() of the float camera matrix
// (GraphCannySeg.cpp:406-407).  Two steps are OpenCV's (2.4, not in this
// image): a 3x3 float matrix is inverted from its cofactors in double over
// the float entries, each entry rounded once; a float matrix product sums its
// products in double and rounds each output once.  The pixel order is the
// segmentation's cluster list (GraphCannySeg.h:1716); here it is row-major.
#ifndef QPSO_ROI_H_
#define QPSO_ROI_H_

#include <stdint.h>

namespace qpso_roi {

struct KInv {
    float k00, k02, k11, k12;  // the non-trivial entries of K^-1; row 2 is (0, 0, 1)
};

inline KInv kinv(double fx, double fy, double cx, double cy) {
    const double f0 = (float)fx, f1 = (float)fy, c0 = (float)cx, c1 = (float)cy;
    const double d = 1.0 / (f0 * f1);  // det [[fx, 0, cx], [0, fy, cy], [0, 0, 1]]
    KInv k;
    k.k00 = (float)(f1 * d);                  // cofactor (0,0) = fy
    k.k02 = (float)((0.0 * c1 - c0 * f1) * d);  // cofactor (0,2) = 0 cy - cx fy
    k.k11 = (float)(f0 * d);                  // cofactor (1,1) = fx
    k.k12 = (float)((c0 * 0.0 - f0 * c1) * d);  // cofactor (1,2) = cx 0 - fx cy
    return k;
}

struct Centroid {
    KInv k;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t n = 0;  // pixels with non-zero depth

    explicit Centroid(const KInv& kk) : k(kk) {}

    // one cluster pixel (column u, row v, depth z in mm)
    void add(int u, int v, uint16_t z) {
        if (z == 0) return;  // holes are skipped (GraphCannySeg.h:2576)
        const float p0 = (float)(u * (int)z), p1 = (float)(v * (int)z), p2 = (float)z;
        sx += (float)((double)k.k00 * p0 + 0.0 * (double)p1 + (double)k.k02 * p2);
        sy += (float)(0.0 * (double)p0 + (double)k.k11 * p1 + (double)k.k12 * p2);
        sz += (float)(0.0 * (double)p0 + 0.0 * (double)p1 + 1.0 * (double)p2);
        ++n;
    }

    bool result(float out[3]) const {
        if (n == 0) return false;
        out[0] = sx / (float)n;
        out[1] = sy / (float)n;
        out[2] = sz / (float)n;
        return true;
    }
};

}  // namespace qpso_roi

#endif
