// fsgh_rz_geom.h -- quadrilateral geometry of the 2D axisymmetric (RZ) path,
// shared by the host set-up and the device kernels so that both evaluate the
// same operations.  Coordinates are (x, y) = (axial, radius); node order
// counter-clockwise 0-1-2-3.  Expressions keep the reference's operation
// order (the library is compiled with -ffp-contract=off).
#pragma once

#include <cmath>

#ifndef FSGH_HD
#define FSGH_HD __host__ __device__
#endif

namespace fsgh {
namespace rz {

// outward corner normals B[i][d] = d(planar area)/d(x_i, y_i)  (K/geometry.cpp:315-378)
FSGH_HD inline void bmatrix(const double x[4], const double y[4], double B[8])
{
    B[0] = -0.5 * (y[3] - y[1]);
    B[2] = -0.5 * (y[0] - y[2]);
    B[4] = -0.5 * (y[1] - y[3]);
    B[6] = -0.5 * (y[2] - y[0]);
    B[1] = -0.5 * (x[1] - x[3]);
    B[3] = -0.5 * (x[2] - x[0]);
    B[5] = -0.5 * (x[3] - x[1]);
    B[7] = -0.5 * (x[0] - x[2]);
}

// planar area from the diagonals  (K/geometry.cpp:435-468)
FSGH_HD inline double area(const double x[4], const double y[4])
{
    return 0.5 * ((x[0] - x[2]) * (y[1] - y[3]) + (x[3] - x[1]) * (y[0] - y[2]));
}

// volume per radian: two triangles, each area x mean radius  (K/geometry.cpp:381-431)
FSGH_HD inline double volume(const double x[4], const double y[4])
{
    return ((y[2] + y[3] + y[0]) * ((y[2] - y[3]) * (x[0] - x[3]) - (y[0] - y[3]) * (x[2] - x[3])) +
            (y[0] + y[1] + y[2]) * ((y[0] - y[1]) * (x[2] - x[1]) - (y[2] - y[1]) * (x[0] - x[1]))) /
           6.0;
}

// triangle area from its side lengths a, b, c, summed in the reference's
// order  (K/geometry.cpp:472-498)
FSGH_HD inline double heron(double a, double b, double c)
{
    double S = 0.0;
    S += a;
    S += b;
    S += c;
    S *= 0.5;
    return sqrt(S * (S - a) * (S - b) * (S - c));
}

// corner areas by Barlow's partition of the four node-centroid-node
// triangles  (K/geometry.cpp:503-543).  The reference's triangle (node i,
// centroid, node i+1) has sides |centroid - node i|, |node i+1 - centroid|
// and |node i+1 - node i|; each node-centroid length appears in two
// triangles with opposite signs of the differences, whose squares are
// bitwise equal (fl(a - b) = -fl(b - a)), so each is formed once: 12 square
// roots instead of 16, with the same bits.
FSGH_HD inline void corner_areas(const double x[4], const double y[4], double A[4])
{
    double rc = 0.0, zc = 0.0;
    for (int a = 0; a < 4; ++a) {
        rc += 0.25 * y[a];
        zc += 0.25 * x[a];
    }
    double d[4], e[4];
    for (int a = 0; a < 4; ++a) {
        const int b = (a + 1) & 3;
        d[a] = sqrt((x[a] - zc) * (x[a] - zc) + (y[a] - rc) * (y[a] - rc));
        e[a] = sqrt((x[b] - x[a]) * (x[b] - x[a]) + (y[b] - y[a]) * (y[b] - y[a]));
    }
    const double A12 = heron(d[0], d[1], e[0]);
    const double A23 = heron(d[1], d[2], e[1]);
    const double A34 = heron(d[2], d[3], e[2]);
    const double A41 = heron(d[3], d[0], e[3]);
    A[0] = (5. * A41 + 5. * A12 + A23 + A34) / 12.;
    A[1] = (A41 + 5. * A12 + 5. * A23 + A34) / 12.;
    A[2] = (A41 + A12 + 5. * A23 + 5. * A34) / 12.;
    A[3] = (5. * A41 + A12 + A23 + 5. * A34) / 12.;
}

// velocity divergence in RZ  (K/momentum.cpp:298-378)
FSGH_HD inline double divergence(const double B[8], const double u[4], const double v[4],
                                 double elem_area, double vol)
{
    const double inverse_area = 1.0 / elem_area;
    const double mean_radius = vol / elem_area;
    const double elem_vel_r = 0.25 * (v[0] + v[1] + v[2] + v[3]);
    double d = 0.0;
    d += (u[0] * B[0] + u[1] * B[2] + u[2] * B[4] + u[3] * B[6]) * inverse_area;
    d += (v[0] * B[1] + v[1] * B[3] + v[2] * B[5] + v[3] * B[7]) * inverse_area +
         elem_vel_r / mean_radius;
    return d;
}

// shortest of the 6 node-pair distances  (K/time_integration.cpp:193-215);
// sqrt is monotone, so the minimum distance is sqrt of the minimum square
FSGH_HD inline double min_pair_distance(const double x[4], const double y[4])
{
    double m = HUGE_VAL;
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j <= 3; ++j) {
            const double dx = x[i] - x[j], dy = y[i] - y[j];
            m = fmin(m, dx * dx + dy * dy);
        }
    return sqrt(m);
}

}  // namespace rz
}  // namespace fsgh
