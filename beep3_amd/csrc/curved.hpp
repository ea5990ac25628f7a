// Evaluation of a curved (PNG1) triangle from its control points, shared by
// the host point sets (host_mesh.cpp) and the device point generation
// (kernels_aux.hip), so both evaluate the same surface with the same
// arithmetic.  The construction of the control points is host_mesh.cpp's
// (CurvedTri, after PNG1_Triangle, src/bem/png1.h:86-121); the blends and
// the surface and its tangents follow png1.h:135-270.
//
// Control points: 21 points, 63 doubles, in this order:
//   0 b300  1 b030  2 b003                          corners (vertices 1, 2, 3)
//   3 p1_210  4 p1_201  5 p2_120  6 p2_021  7 p3_102  8 p3_012
//                                                   corner tangent-plane points
//   9 p2_201  10 p2_102  11 p3_120  12 p3_210  13 p1_012  14 p1_021
//                                                   the same carried to the next corner
//   15 c12_1  16 c12_2  17 c23_2  18 c23_3  19 c31_3  20 c31_1
//                                                   centre variants per edge
// Barycentric (u, v, w = 1 - u - v): w weights vertex 1, u vertex 2, v vertex 3.
// Every output is an affine (point) or linear (tangent) combination of the
// control points, so a rigid motion of the control points moves the surface
// with them.
#ifndef BEEP_CURVED_HPP
#define BEEP_CURVED_HPP

#if defined(__HIPCC__)
#define CURVED_FN __host__ __device__ inline
#else
#define CURVED_FN inline
#endif

namespace beepcurve {

constexpr int CTL_DOUBLES = 63;

// rational blend of control points a and b by a coordinate t: a at t = 0, b at t = 1
CURVED_FN void blend(double t, const double* a, const double* b, double* out) {
    const double s = 1.0 - t, ss = s * s, tt = t * t;
    const double d = tt + ss;
    for (int k = 0; k < 3; ++k) out[k] = (a[k] * ss + b[k] * tt) / d;
}

// the seven blended control points at (u, v): b210 b120 b201 b021 b102 b012 b111
CURVED_FN void controls(const double* c, double u, double v, double* k) {
    const double w = 1.0 - u - v;
    const double* P = c;
    blend(v, P + 3 * 3, P + 3 * 12, k + 0);    // b210: p1_210 -> p3_210
    blend(v, P + 3 * 5, P + 3 * 11, k + 3);    // b120: p2_120 -> p3_120
    blend(u, P + 3 * 4, P + 3 * 9, k + 6);     // b201: p1_201 -> p2_201
    blend(w, P + 3 * 6, P + 3 * 14, k + 9);    // b021: p2_021 -> p1_021
    blend(u, P + 3 * 7, P + 3 * 10, k + 12);   // b102: p3_102 -> p2_102
    blend(w, P + 3 * 8, P + 3 * 13, k + 15);   // b012: p3_012 -> p1_012
    // the centre point: the three edge variants blended by weights that pick
    // the variant of the nearest edge (shifted off the corners, where the
    // weights are 0/0)
    const double eps = 1e-6;
    double a = u + eps, b = v + eps, g = w + eps;
    const double tot = a + b + g;
    a /= tot;
    b /= tot;
    g /= tot;
    const double sum = a * b + a * g + b * g;
    const double f1 = (a * g / sum) / (g + a - 2.0 * a * g);
    const double f2 = (a * b / sum) / (a + b - 2.0 * a * b);
    const double f3 = (b * g / sum) / (g + b - 2.0 * b * g);
    for (int q = 0; q < 3; ++q) {
        const double t1 = P[3 * 15 + q] * ((1.0 - a) * g) + P[3 * 16 + q] * ((1.0 - g) * a);
        const double t2 = P[3 * 17 + q] * ((1.0 - b) * a) + P[3 * 18 + q] * ((1.0 - a) * b);
        const double t3 = P[3 * 19 + q] * ((1.0 - g) * b) + P[3 * 20 + q] * ((1.0 - b) * g);
        k[18 + q] = t1 * f1 + t2 * f2 + t3 * f3;
    }
}

// surface point p and the tangents du = dS/du, dv = dS/dv (with the blended
// control points held fixed, as the construction's jacobian and normal take
// them) at (u, v)
CURVED_FN void eval(const double* c, double u, double v, double* p, double* du, double* dv) {
    const double w = 1.0 - u - v;
    double k[21];
    controls(c, u, v, k);
    const double* b300 = c;
    const double* b030 = c + 3;
    const double* b003 = c + 6;
    const double* b210 = k + 0;
    const double* b120 = k + 3;
    const double* b201 = k + 6;
    const double* b021 = k + 9;
    const double* b102 = k + 12;
    const double* b012 = k + 15;
    const double* b111 = k + 18;
    for (int q = 0; q < 3; ++q) {
        p[q] = b300[q] * (w * w * w) + b030[q] * (u * u * u) + b003[q] * (v * v * v) +
               b210[q] * (3 * w * w * u) + b120[q] * (3 * w * u * u) + b201[q] * (3 * w * w * v) +
               b021[q] * (3 * u * u * v) + b102[q] * (3 * w * v * v) + b012[q] * (3 * u * v * v) +
               b111[q] * (6 * w * u * v);
        du[q] = b030[q] * (3 * u * u) - b300[q] * (3 * w * w) + b120[q] * (6 * u * w - 3 * u * u) +
                b021[q] * (6 * u * v) - b102[q] * (3 * v * v) + b012[q] * (3 * v * v) +
                b210[q] * (3 * w * w - 6 * w * u) - b201[q] * (6 * w * v) + b111[q] * (6 * w * v - 6 * u * v);
        dv[q] = b003[q] * (3 * v * v) - b300[q] * (3 * w * w) - b120[q] * (3 * u * u) + b021[q] * (3 * u * u) +
                b102[q] * (6 * v * w - 3 * v * v) + b012[q] * (6 * v * u) - b210[q] * (6 * w * u) +
                b201[q] * (3 * w * w - 6 * w * v) + b111[q] * (6 * w * u - 6 * u * v);
    }
}

}  // namespace beepcurve

#endif
