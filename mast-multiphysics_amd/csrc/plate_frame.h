// The quad4 plate element's local frame and its surface map, shared by the
// device kernels (stiffness, sensitivity) and the host (pressure load).
//
// An element whose four nodes share node 0's z (compared exactly) is used in
// the global x-y plane.  Any other element is computed in its own frame
// (GeomElem::_init_local_elem_2d, src/mesh/geom_elem.cpp:460-511):
//   v1 = unit(x1 - x0), v3 = unit(v1 x unit(x2 - x0)), v2 = unit(v3 x v1),
//   T = [v1 v2 v3] (columns, row-major T[i*3 + j] = v_j[i]),
//   local coordinates T^T (x - x0);
// the element's vectors and matrices then go to the global system through
// the same T on the translations and on the rotations of every node
// (StructuralElementBase::transform_*_system, src/elasticity/
// structural_element_base.cpp:1568-1660).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace mastg {

__host__ __device__ inline void pf_unit(double v[3]) {
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
}

__host__ __device__ inline void pf_cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// T and the local node coordinates; false (T = I, XL = X) for an element in
// a z = const plane
__host__ __device__ inline bool quad4_frame(const double X[4][3], double T[9], double XL[4][3]) {
    const bool use = X[1][2] != X[0][2] || X[2][2] != X[0][2] || X[3][2] != X[0][2];
    if (!use) {
        for (int i = 0; i < 9; i++) T[i] = (i % 4 == 0) ? 1.0 : 0.0;
        for (int n = 0; n < 4; n++)
            for (int d = 0; d < 3; d++) XL[n][d] = X[n][d];
        return false;
    }
    double v1[3], v2[3], v3[3];
    for (int i = 0; i < 3; i++) {
        v1[i] = X[1][i] - X[0][i];
        v2[i] = X[2][i] - X[0][i];
    }
    pf_unit(v1);
    pf_unit(v2);
    pf_cross(v1, v2, v3);
    pf_unit(v3);
    pf_cross(v3, v1, v2);
    pf_unit(v2);
    for (int i = 0; i < 3; i++) {
        T[i * 3 + 0] = v1[i];
        T[i * 3 + 1] = v2[i];
        T[i * 3 + 2] = v3[i];
    }
    for (int n = 0; n < 4; n++)
        for (int j = 0; j < 3; j++)
            XL[n][j] = T[0 * 3 + j] * (X[n][0] - X[0][0]) + T[1 * 3 + j] * (X[n][1] - X[0][1]) +
                       T[2 * 3 + j] * (X[n][2] - X[0][2]);
    return true;
}

// bilinear shape functions and their reference derivatives at (xi, eta);
// libMesh quad4 node order (-1,-1), (1,-1), (1,1), (-1,1)
__host__ __device__ inline void quad4_shape_all(double xi, double eta, double N[4], double dN[4][2]) {
    for (int n = 0; n < 4; n++) {
        const double sx = (n == 1 || n == 2) ? 1.0 : -1.0;
        const double sy = (n >= 2) ? 1.0 : -1.0;
        const double fx = 0.5 * (1.0 + sx * xi), fy = 0.5 * (1.0 + sy * eta);
        N[n] = fx * fy;
        dN[n][0] = 0.5 * sx * fy;
        dN[n][1] = fx * 0.5 * sy;
    }
}

// the surface map of a 2-D element in 3-D (libMesh FEMap for dim 2 in 3-D
// space): tangents t_r = sum_a x_a dN_a/dxi_r, metric g = t^T t, area
// Jacobian sqrt(det g), physical gradients dN/dx = dN/dxi_r (g^-1 t^T)_r;
// g[a] receives the x and y components (the plate's local in-plane axes)
__host__ __device__ inline double quad4_grad_surface(const double X[4][3], double xi, double eta,
                                                    double g[4][2], double N[4]) {
    double dN[4][2];
    quad4_shape_all(xi, eta, N, dN);
    double t[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int a = 0; a < 4; a++)
        for (int i = 0; i < 3; i++) {
            t[i][0] += X[a][i] * dN[a][0];
            t[i][1] += X[a][i] * dN[a][1];
        }
    const double g11 = t[0][0] * t[0][0] + t[1][0] * t[1][0] + t[2][0] * t[2][0];
    const double g12 = t[0][0] * t[0][1] + t[1][0] * t[1][1] + t[2][0] * t[2][1];
    const double g22 = t[0][1] * t[0][1] + t[1][1] * t[1][1] + t[2][1] * t[2][1];
    const double det = g11 * g22 - g12 * g12;
    const double id = 1.0 / det;
    const double i11 = g22 * id, i12 = -g12 * id, i22 = g11 * id;
    for (int i = 0; i < 2; i++) {
        const double dxi = i11 * t[i][0] + i12 * t[i][1];
        const double deta = i12 * t[i][0] + i22 * t[i][1];
        for (int a = 0; a < 4; a++) g[a][i] = dN[a][0] * dxi + dN[a][1] * deta;
    }
    return sqrt(det);
}

}  // namespace mastg
