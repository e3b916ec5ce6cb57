// Face-pressure element loads on the host (cheap: one face per loaded
// element). Follows elem.rs:892-950 (C3D8, 2x2 Gauss on the bilinear face)
// and quadratic.rs:266-354 (C3D20, 3x3 Gauss on the serendipity face):
// traction = p * (dr/dxi x dr/deta); the CalculiX face winding makes that
// normal point into the element, so positive pressure compresses. Built with
// -ffp-contract=off so the loads carry the reference's FP64 bits.
#include <cmath>

#include "axg_internal.h"

namespace axg {

static const double G2 = 0.5773502691896257;
static const double G3[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
static const double W3[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};
static const double QUAD_XI[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
static const double QUAD8_XI[8][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0},
                                      {0.0, -1.0},  {1.0, 0.0},  {0.0, 1.0}, {-1.0, 0.0}};
static const int FACE8[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1},
                                {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
static const int FACE20[6][8] = {{0, 1, 2, 3, 8, 9, 10, 11},   {4, 7, 6, 5, 15, 14, 13, 12},
                                 {0, 4, 5, 1, 16, 12, 17, 8},  {1, 5, 6, 2, 17, 13, 18, 9},
                                 {2, 6, 7, 3, 18, 14, 19, 10}, {3, 7, 4, 0, 19, 15, 16, 11}};

static void face_normal(const double *xyz, const int *fn, int nf, const double *dxi,
                        const double *deta, double nrm[3]) {
    double rxi[3] = {0.0, 0.0, 0.0}, reta[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < nf; a++) {
        const double *q = xyz + 3 * fn[a];
        for (int k = 0; k < 3; k++) {
            rxi[k] += dxi[a] * q[k];
            reta[k] += deta[a] * q[k];
        }
    }
    nrm[0] = rxi[1] * reta[2] - rxi[2] * reta[1];
    nrm[1] = rxi[2] * reta[0] - rxi[0] * reta[2];
    nrm[2] = rxi[0] * reta[1] - rxi[1] * reta[0];
}

int face_pressure_host(int nn, const double *xyz, int face, double p, double *fe) {
    if (face < 1 || face > 6) return AXG_E_INVALID;
    const int fi = face - 1;
    for (int i = 0; i < 3 * nn; i++) fe[i] = 0.0;
    if (nn == 8) {
        const double pts[2] = {-G2, G2};
        for (int s = 0; s < 2; s++)
            for (int t = 0; t < 2; t++) {
                double xi = pts[s], eta = pts[t], n[4], dxi[4], deta[4], nrm[3];
                for (int i = 0; i < 4; i++) {
                    double x = QUAD_XI[i][0], e = QUAD_XI[i][1];
                    n[i] = 0.25 * (1.0 + x * xi) * (1.0 + e * eta);
                    dxi[i] = 0.25 * x * (1.0 + e * eta);
                    deta[i] = 0.25 * e * (1.0 + x * xi);
                }
                face_normal(xyz, FACE8[fi], 4, dxi, deta, nrm);
                double tx = p * nrm[0], ty = p * nrm[1], tz = p * nrm[2];
                for (int a = 0; a < 4; a++) {
                    int gd = 3 * FACE8[fi][a];
                    fe[gd] += n[a] * tx;
                    fe[gd + 1] += n[a] * ty;
                    fe[gd + 2] += n[a] * tz;
                }
            }
        return AXG_OK;
    }
    if (nn == 20) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double xi = G3[i], eta = G3[j], w = W3[i] * W3[j];
                double n[8], dxi[8], deta[8], nrm[3];
                for (int k = 0; k < 4; k++) {
                    double x = QUAD8_XI[k][0], e = QUAD8_XI[k][1];
                    double a = 1.0 + x * xi, b = 1.0 + e * eta;
                    n[k] = 0.25 * a * b * (x * xi + e * eta - 1.0);
                    dxi[k] = 0.25 * x * b * (2.0 * x * xi + e * eta);
                    deta[k] = 0.25 * e * a * (x * xi + 2.0 * e * eta);
                }
                for (int k = 4; k < 8; k++) {
                    double x = QUAD8_XI[k][0], e = QUAD8_XI[k][1];
                    if (std::fabs(x) < 0.5) {
                        n[k] = 0.5 * (1.0 - xi * xi) * (1.0 + e * eta);
                        dxi[k] = -xi * (1.0 + e * eta);
                        deta[k] = 0.5 * (1.0 - xi * xi) * e;
                    } else {
                        n[k] = 0.5 * (1.0 - eta * eta) * (1.0 + x * xi);
                        dxi[k] = 0.5 * (1.0 - eta * eta) * x;
                        deta[k] = -eta * (1.0 + x * xi);
                    }
                }
                face_normal(xyz, FACE20[fi], 8, dxi, deta, nrm);
                double tx = p * nrm[0] * w, ty = p * nrm[1] * w, tz = p * nrm[2] * w;
                for (int a = 0; a < 8; a++) {
                    int gd = 3 * FACE20[fi][a];
                    fe[gd] += n[a] * tx;
                    fe[gd + 1] += n[a] * ty;
                    fe[gd + 2] += n[a] * tz;
                }
            }
        return AXG_OK;
    }
    return AXG_E_INVALID;
}

}  // namespace axg
