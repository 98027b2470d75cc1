// Oracle (test infrastructure only): FE maps, material/section matrices and
// element kernels, restated from the MAST files cited in oracle.h.
#include "oracle.h"
#include <cmath>
#include <cstring>
#include <algorithm>

namespace oracle {

// ----------------------------------------------------------- quadrature ----
// libMesh QGauss for first-order Lagrange: default order 2p+1 = 3 -> 2 points
// per direction; MindlinBendingOperator lowers it by 2 -> 1 point
// (src/mesh/fe_base.cpp:102-107, src/elasticity/mindlin_bending_operator.cpp:37,104).
// Points are tensor products with the first direction fastest.
static const double GP2[2] = {-0.57735026918962576451, 0.57735026918962576451};

// hex8 reference node coordinates (libMesh Hex8 node order)
static const double HEX_NODE[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
static const double QUAD_NODE[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};

static void hex8_shape(double xi, double eta, double zeta, double N[8], double dN[8][3]) {
    for (int a = 0; a < 8; a++) {
        double sx = 0.5 * (1 + HEX_NODE[a][0] * xi), sy = 0.5 * (1 + HEX_NODE[a][1] * eta),
               sz = 0.5 * (1 + HEX_NODE[a][2] * zeta);
        N[a] = sx * sy * sz;
        dN[a][0] = 0.5 * HEX_NODE[a][0] * sy * sz;
        dN[a][1] = sx * 0.5 * HEX_NODE[a][1] * sz;
        dN[a][2] = sx * sy * 0.5 * HEX_NODE[a][2];
    }
}

static void quad4_shape(double xi, double eta, double N[4], double dN[4][2]) {
    for (int a = 0; a < 4; a++) {
        double sx = 0.5 * (1 + QUAD_NODE[a][0] * xi), sy = 0.5 * (1 + QUAD_NODE[a][1] * eta);
        N[a] = sx * sy;
        dN[a][0] = 0.5 * QUAD_NODE[a][0] * sy;
        dN[a][1] = sx * 0.5 * QUAD_NODE[a][1];
    }
}

// libMesh FEMap 3D volume map: JxW and physical gradients dphi[a][0..2]
static double hex8_map(const double X[8][3], const double dN[8][3], double g[8][3]) {
    double dx[3][3] = {{0}};  // dx[i][r] = d x_i / d xi_r
    for (int a = 0; a < 8; a++)
        for (int i = 0; i < 3; i++)
            for (int r = 0; r < 3; r++) dx[i][r] += X[a][i] * dN[a][r];
    const double dxdxi = dx[0][0], dydxi = dx[1][0], dzdxi = dx[2][0];
    const double dxdeta = dx[0][1], dydeta = dx[1][1], dzdeta = dx[2][1];
    const double dxdzeta = dx[0][2], dydzeta = dx[1][2], dzdzeta = dx[2][2];
    const double jac = dxdxi * (dydeta * dzdzeta - dzdeta * dydzeta) +
                       dydxi * (dzdeta * dxdzeta - dxdeta * dzdzeta) +
                       dzdxi * (dxdeta * dydzeta - dydeta * dxdzeta);
    const double ij = 1.0 / jac;
    const double dxidx = (dydeta * dzdzeta - dzdeta * dydzeta) * ij;
    const double dxidy = (dzdeta * dxdzeta - dxdeta * dzdzeta) * ij;
    const double dxidz = (dxdeta * dydzeta - dydeta * dxdzeta) * ij;
    const double detadx = (dzdxi * dydzeta - dydxi * dzdzeta) * ij;
    const double detady = (dxdxi * dzdzeta - dzdxi * dxdzeta) * ij;
    const double detadz = (dydxi * dxdzeta - dxdxi * dydzeta) * ij;
    const double dzetadx = (dydxi * dzdeta - dzdxi * dydeta) * ij;
    const double dzetady = (dzdxi * dxdeta - dxdxi * dzdeta) * ij;
    const double dzetadz = (dxdxi * dydeta - dydxi * dxdeta) * ij;
    for (int a = 0; a < 8; a++) {
        g[a][0] = dN[a][0] * dxidx + dN[a][1] * detadx + dN[a][2] * dzetadx;
        g[a][1] = dN[a][0] * dxidy + dN[a][1] * detady + dN[a][2] * dzetady;
        g[a][2] = dN[a][0] * dxidz + dN[a][1] * detadz + dN[a][2] * dzetadz;
    }
    return jac;
}

// libMesh FEMap 2D-manifold map (element in 3D space): metric-tensor form
static double quad4_map(const double X[4][3], const double dN[4][2], double g[4][3]) {
    double t[3][2] = {{0}};
    for (int a = 0; a < 4; a++)
        for (int i = 0; i < 3; i++)
            for (int r = 0; r < 2; r++) t[i][r] += X[a][i] * dN[a][r];
    const double g11 = t[0][0] * t[0][0] + t[1][0] * t[1][0] + t[2][0] * t[2][0];
    const double g12 = t[0][0] * t[0][1] + t[1][0] * t[1][1] + t[2][0] * t[2][1];
    const double g22 = t[0][1] * t[0][1] + t[1][1] * t[1][1] + t[2][1] * t[2][1];
    const double det = g11 * g22 - g12 * g12;
    const double idet = 1.0 / det;
    const double g11i = g22 * idet, g12i = -g12 * idet, g22i = g11 * idet;
    double dxi[3], deta[3];
    for (int i = 0; i < 3; i++) {
        dxi[i] = g11i * t[i][0] + g12i * t[i][1];
        deta[i] = g12i * t[i][0] + g22i * t[i][1];
    }
    for (int a = 0; a < 4; a++)
        for (int i = 0; i < 3; i++) g[a][i] = dN[a][0] * dxi[i] + dN[a][1] * deta[i];
    return std::sqrt(det);
}

// ------------------------------------------------------------ materials ----
void material_D3(double E, double nu, double D[36]) {
    std::memset(D, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            D[i * 6 + j] = (i == j) ? E * (1. - nu) / (1. - nu - 2. * nu * nu)
                                    : E * nu / (1. - nu - 2. * nu * nu);
        D[(i + 3) * 6 + i + 3] = E / 2. / (1. + nu);
    }
}

void material_dD3(double E, double nu, double dE, double dnu, double m[36]) {
    double dm[36];
    std::memset(m, 0, 36 * sizeof(double));
    std::memset(dm, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            dm[i * 6 + j] = (i == j) ? (1. - nu) / (1. - nu - 2. * nu * nu) : nu / (1. - nu - 2. * nu * nu);
        dm[(i + 3) * 6 + i + 3] = 1. / 2. / (1. + nu);
    }
    for (int k = 0; k < 36; k++) m[k] += dE * dm[k];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            dm[i * 6 + j] = (i == j)
                ? -E / (1. - nu - 2. * nu * nu) + E * (1. - nu) / std::pow(1. - nu - 2. * nu * nu, 2) * (1. + 4. * nu)
                : E / (1. - nu - 2. * nu * nu) + E * nu / std::pow(1. - nu - 2. * nu * nu, 2) * (1. + 4. * nu);
        dm[(i + 3) * 6 + i + 3] = -E / 2. / std::pow(1. + nu, 2);
    }
    for (int k = 0; k < 36; k++) m[k] += dnu * dm[k];
}

void material_D2_plane_stress(double E, double nu, double m[9]) {
    std::memset(m, 0, 9 * sizeof(double));
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) m[i * 3 + j] = (i == j) ? E / (1. - nu * nu) : E * nu / (1. - nu * nu);
    m[8] = E / 2. / (1. + nu);
}

void material_dD2_plane_stress(double E, double nu, double dE, double dnu, double m[9]) {
    double dm[9] = {0};
    std::memset(m, 0, 9 * sizeof(double));
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) dm[i * 3 + j] = (i == j) ? 1. / (1. - nu * nu) : 1. * nu / (1. - nu * nu);
    dm[8] = 1. / 2. / (1. + nu);
    for (int k = 0; k < 9; k++) m[k] += dE * dm[k];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            dm[i * 3 + j] = (i == j) ? E / std::pow(1. - nu * nu, 2) * 2. * nu
                                     : E / (1. - nu * nu) + E * nu / std::pow(1. - nu * nu, 2) * 2. * nu;
    dm[8] = -E / 2. / std::pow(1. + nu, 2);
    for (int k = 0; k < 9; k++) m[k] += dnu * dm[k];
}

void material_G(double E, double nu, double G[4]) {
    G[0] = E / 2. / (1. + nu); G[1] = 0; G[2] = 0; G[3] = G[0];
}

void section_A(const double Dps[9], double h, double A[9]) {
    for (int k = 0; k < 9; k++) A[k] = Dps[k] * h;
}
void section_B(const double Dps[9], double h, double off, double B[9]) {
    for (int k = 0; k < 9; k++) B[k] = Dps[k] * (h * off);
}
void section_Db(const double Dps[9], double h, double off, double Db[9]) {
    for (int k = 0; k < 9; k++) Db[k] = Dps[k] * (std::pow(h, 3) / 12. + h * std::pow(off, 2));
}
void section_S(const double G[4], double h, double kappa, double S[4]) {
    for (int k = 0; k < 4; k++) { S[k] = G[k] * h; S[k] *= kappa; }
}

// ------------------------------------------------------- hex8 elasticity ----
// Dense linear strain operator B (6 x 24), rows exx,eyy,ezz,gxy,gyz,gzx, columns
// var-major (ux0..7, uy0..7, uz0..7): StructuralElement3D::
// initialize_green_lagrange_strain_operator, linear branch (solid_element_3d.cpp:1359-1402).
static void hex8_B(const double g[8][3], double B[6][24]) {
    std::memset(B, 0, sizeof(double) * 6 * 24);
    for (int a = 0; a < 8; a++) {
        B[0][0 * 8 + a] = g[a][0];  // exx = du/dx
        B[3][1 * 8 + a] = g[a][0];  // gxy: dv/dx
        B[5][2 * 8 + a] = g[a][0];  // gzx: dw/dx
        B[1][1 * 8 + a] = g[a][1];  // eyy = dv/dy
        B[3][0 * 8 + a] = g[a][1];  // gxy: du/dy
        B[4][2 * 8 + a] = g[a][1];  // gyz: dw/dy
        B[2][2 * 8 + a] = g[a][2];  // ezz = dw/dz
        B[4][1 * 8 + a] = g[a][2];  // gyz: dv/dz
        B[5][0 * 8 + a] = g[a][2];  // gzx: du/dz
    }
}

template <bool JAC>
static void hex8_elastic_impl(const double X[8][3], const double D[36], const double* u,
                              double* f, double* K, int nd) {
    // nd: element vector size (48 with rotations, 24 without); f, K sized nd
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                double N[8], dN[8][3], g[8][3], B[6][24];
                hex8_shape(GP2[i], GP2[j], GP2[k], N, dN);
                const double JxW = hex8_map(X, dN, g) * 1.0;
                hex8_B(g, B);
                double strain[6] = {0}, stress[6] = {0};
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 24; c++) strain[r] += B[r][c] * u[c];
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 6; c++) stress[r] += D[r * 6 + c] * strain[c];
                for (int c = 0; c < 24; c++) {
                    double v = 0;
                    for (int r = 0; r < 6; r++) v += B[r][c] * stress[r];
                    f[c] += JxW * v;
                }
                if (JAC) {
                    double DB[6][24];
                    for (int r = 0; r < 6; r++)
                        for (int c = 0; c < 24; c++) {
                            double v = 0;
                            for (int s = 0; s < 6; s++) v += B[s][c] * D[r * 6 + s];
                            DB[r][c] = v;
                        }
                    for (int a = 0; a < 24; a++)
                        for (int b = 0; b < 24; b++) {
                            double v = 0;
                            for (int s = 0; s < 6; s++) v += B[s][a] * DB[s][b];
                            K[a * nd + b] += JxW * v;
                        }
                }
            }
    if (JAC && nd == 48) {
        // small diagonal for the rotational dofs (solid_element_3d.cpp:413-417)
        double mx = K[0];
        for (int a = 1; a < 48; a++) mx = std::max(mx, K[a * 48 + a]);
        for (int a = 24; a < 48; a++) K[a * 48 + a] += 1.0e-20 * mx;
    }
}

void hex8_elastic(const double X[8][3], const double D[36], const double* u, double* f, double* K) {
    // full MAST element (48 dofs); callers wanting the active 24 use the top-left block
    double ff[48] = {0};
    static thread_local double KK[48 * 48];
    if (K) std::memset(KK, 0, sizeof(KK));
    if (K) hex8_elastic_impl<true>(X, D, u, ff, KK, 48);
    else hex8_elastic_impl<false>(X, D, u, ff, nullptr, 48);
    for (int a = 0; a < 48; a++) f[a] += ff[a];
    if (K)
        for (int a = 0; a < 48 * 48; a++) K[a] += KK[a];
}

// StructuralElement3D::calculate_stress (solid_element_3d.cpp:1147-1236), linear
// strain: at every quadrature point (libMesh QGauss order, xi fastest) the
// strain B u and the stress D B u; JxW beside it.  u: the 24 translations.
void hex8_stress(const double X[8][3], const double D[36], const double* u, double sigma[8][6],
                 double jxw[8]) {
    int q = 0;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++, q++) {
                double N[8], dN[8][3], g[8][3], B[6][24];
                hex8_shape(GP2[i], GP2[j], GP2[k], N, dN);
                jxw[q] = hex8_map(X, dN, g) * 1.0;
                hex8_B(g, B);
                double strain[6] = {0};
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 24; c++) strain[r] += B[r][c] * u[c];
                for (int r = 0; r < 6; r++) {
                    double v = 0;
                    for (int c = 0; c < 6; c++) v += D[r * 6 + c] * strain[c];
                    sigma[q][r] = v;
                }
            }
}

void hex8_elastic_sens(const double X[8][3], const double dD[36], const double* u, double* df) {
    double ff[48] = {0};
    hex8_elastic_impl<false>(X, dD, u, ff, nullptr, 48);
    for (int a = 0; a < 48; a++) df[a] += ff[a];
}

// ------------------------------------------------------ quad4 Mindlin ----
// Operators over the 24 element dofs (var-major u,v,w,tx,ty,tz x 4 nodes):
//  Bm (membrane, structural_element_2d.cpp:143-253 linear branch),
//  Bb (bending, mindlin_bending_operator.cpp:57-87, z = 1),
//  Bs (transverse shear, mindlin_bending_operator.cpp:312-361).
static void quad4_ops(const double g[4][3], const double N[4], double Bm[3][24],
                      double Bb[3][24], double Bs[2][24]) {
    std::memset(Bm, 0, sizeof(double) * 72);
    std::memset(Bb, 0, sizeof(double) * 72);
    std::memset(Bs, 0, sizeof(double) * 48);
    for (int a = 0; a < 4; a++) {
        const double gx = g[a][0], gy = g[a][1];
        Bm[0][0 * 4 + a] = gx;   // exx = du/dx
        Bm[2][1 * 4 + a] = gx;   // gxy: dv/dx
        Bm[1][1 * 4 + a] = gy;   // eyy = dv/dy
        Bm[2][0 * 4 + a] = gy;   // gxy: du/dy
        Bb[0][4 * 4 + a] = gx;   // kx  =  d(ty)/dx
        Bb[2][3 * 4 + a] = -gx;  // kxy: -d(tx)/dx
        Bb[2][4 * 4 + a] = gy;   // kxy:  d(ty)/dy
        Bb[1][3 * 4 + a] = -gy;  // ky  = -d(tx)/dy
        Bs[0][2 * 4 + a] = gx;   // gxz: dw/dx
        Bs[1][2 * 4 + a] = gy;   // gyz: dw/dy
        Bs[0][4 * 4 + a] = N[a]; // gxz: +ty
        Bs[1][3 * 4 + a] = -N[a];// gyz: -tx
    }
}

template <int R>
static void matvec(const double (*B)[24], const double* u, double* out) {
    for (int r = 0; r < R; r++) {
        double v = 0;
        for (int c = 0; c < 24; c++) v += B[r][c] * u[c];
        out[r] = v;
    }
}
template <int R>
static void add_BtS(const double (*B)[24], const double* s, double w, double* f) {
    for (int c = 0; c < 24; c++) {
        double v = 0;
        for (int r = 0; r < R; r++) v += B[r][c] * s[r];
        f[c] += w * v;
    }
}
// K += w * Bl^T M Br   (M is R x R row-major)
template <int R>
static void add_BtMB(const double (*Bl)[24], const double* M, const double (*Br)[24],
                     double w, double* K) {
    double MB[R][24];
    for (int r = 0; r < R; r++)
        for (int c = 0; c < 24; c++) {
            double v = 0;
            for (int s = 0; s < R; s++) v += Br[s][c] * M[r * R + s];
            MB[r][c] = v;
        }
    for (int a = 0; a < 24; a++)
        for (int b = 0; b < 24; b++) {
            double v = 0;
            for (int s = 0; s < R; s++) v += Bl[s][a] * MB[s][b];
            K[a * 24 + b] += w * v;
        }
}

static void quad4_mindlin_impl(const double X[4][3], const PlateSection& s, const double* u,
                               double* f, double* K) {
    double BT[9];  // B^T of the coupling matrix
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) BT[r * 3 + c] = s.B[c * 3 + r];
    // membrane / coupling / bending on the 2x2 rule (_internal_residual_operation)
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            double N[4], dN[4][2], g[4][3], Bm[3][24], Bb[3][24], Bs[2][24];
            quad4_shape(GP2[i], GP2[j], N, dN);
            const double JxW = quad4_map(X, dN, g) * 1.0;
            quad4_ops(g, N, Bm, Bb, Bs);
            double em[3], kb[3], sig[3], t[3];
            matvec<3>(Bm, u, em);
            matvec<3>(Bb, u, kb);
            for (int r = 0; r < 3; r++) {
                sig[r] = 0;
                for (int c = 0; c < 3; c++) sig[r] += s.A[r * 3 + c] * em[c];
                for (int c = 0; c < 3; c++) sig[r] += s.B[r * 3 + c] * kb[c];
            }
            add_BtS<3>(Bm, sig, JxW, f);                 // membrane force
            for (int r = 0; r < 3; r++) {
                t[r] = 0;
                for (int c = 0; c < 3; c++) t[r] += BT[r * 3 + c] * em[c];
            }
            add_BtS<3>(Bb, t, JxW, f);                   // coupling B^T e_m
            for (int r = 0; r < 3; r++) {
                t[r] = 0;
                for (int c = 0; c < 3; c++) t[r] += s.Db[r * 3 + c] * kb[c];
            }
            add_BtS<3>(Bb, t, JxW, f);                   // bending moment
            if (K) {
                add_BtMB<3>(Bm, s.A, Bm, JxW, K);         // membrane-membrane
                add_BtMB<3>(Bb, BT, Bm, JxW, K);          // bending-membrane
                add_BtMB<3>(Bm, s.B, Bb, JxW, K);         // membrane-bending
                add_BtMB<3>(Bb, s.Db, Bb, JxW, K);        // bending-bending
            }
        }
    // transverse shear on the reduced 1x1 rule (calculate_transverse_shear_residual)
    {
        double N[4], dN[4][2], g[4][3], Bm[3][24], Bb[3][24], Bs[2][24];
        quad4_shape(0.0, 0.0, N, dN);
        const double JxW = quad4_map(X, dN, g) * 4.0;
        quad4_ops(g, N, Bm, Bb, Bs);
        double gam[2], q[2];
        matvec<2>(Bs, u, gam);
        for (int r = 0; r < 2; r++) q[r] = s.S[r * 2] * gam[0] + s.S[r * 2 + 1] * gam[1];
        add_BtS<2>(Bs, q, JxW, f);
        if (K) add_BtMB<2>(Bs, s.S, Bs, JxW, K);
    }
}

void quad4_mindlin(const double X[4][3], const PlateSection& s, const double* u, double* f,
                   double* K, bool add_thetaz_diag) {
    quad4_mindlin_impl(X, s, u, f, K);
    if (K && add_thetaz_diag)
        for (int a = 0; a < 4; a++) K[(20 + a) * 24 + 20 + a] = 1.0e-8;  // (structural_element_2d.cpp:1215-1221)
}

// The plate element's local frame (GeomElem::_init_local_elem /
// _init_local_elem_2d, src/mesh/geom_elem.cpp:330-380, 460-511): none when
// every node has node 0's z (the element is used in the global x-y plane);
// else x along node 1 - node 0, z along x cross (node 2 - node 0), y = z cross
// x, T = [x y z] (columns), local coordinates T^T (X - X0).
bool plate_local_frame(const double X[4][3], double T[9], double XL[4][3]) {
    bool use = false;
    for (int n = 1; n < 4; n++)
        if (std::fabs(X[0][2] - X[n][2]) != 0.) use = true;
    if (!use) return false;
    double v1[3], v2[3], v3[3];
    for (int i = 0; i < 3; i++) { v1[i] = X[1][i] - X[0][i]; v2[i] = X[2][i] - X[0][i]; }
    auto unit = [](double* v) {
        const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int i = 0; i < 3; i++) v[i] /= n;
    };
    auto cross = [](const double* a, const double* b, double* c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    unit(v1);
    unit(v2);
    cross(v1, v2, v3);
    unit(v3);
    cross(v3, v1, v2);
    unit(v2);
    for (int i = 0; i < 3; i++) { T[i * 3 + 0] = v1[i]; T[i * 3 + 1] = v2[i]; T[i * 3 + 2] = v3[i]; }
    for (int n = 0; n < 4; n++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += T[k * 3 + j] * (X[n][k] - X[0][k]);
            XL[n][j] = s;
        }
    return true;
}

// StructuralElementBase::transform_vector_to_local_system / _to_global_system
// and transform_matrix_to_global_system (src/elasticity/structural_element_base.
// cpp:1568-1660) for a 4-node element with 6 variables (variable-major,
// index v*4 + a): the same 3x3 T on (u, v, w) and on (tx, ty, tz) of every node
void plate_to_local(const double T[9], const double* g, double* l) {
    for (int a = 0; a < 4; a++)
        for (int blk = 0; blk < 2; blk++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += T[k * 3 + j] * g[(blk * 3 + k) * 4 + a];
                l[(blk * 3 + j) * 4 + a] = s;
            }
}
void plate_to_global(const double T[9], const double* l, double* g) {
    for (int a = 0; a < 4; a++)
        for (int blk = 0; blk < 2; blk++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += T[j * 3 + k] * l[(blk * 3 + k) * 4 + a];
                g[(blk * 3 + j) * 4 + a] = s;
            }
}
void plate_matrix_to_global(const double T[9], const double* KL, double* KG) {
    // KG = M KL M^T, M = blockdiag over (node, translation / rotation) of T
    static thread_local double W[24 * 24];
    for (int c = 0; c < 24; c++) {           // W = M KL, column by column
        double col[24], out[24];
        for (int r = 0; r < 24; r++) col[r] = KL[r * 24 + c];
        plate_to_global(T, col, out);
        for (int r = 0; r < 24; r++) W[r * 24 + c] = out[r];
    }
    for (int r = 0; r < 24; r++) {           // KG = W M^T = (M W^T)^T
        double row[24], out[24];
        for (int c = 0; c < 24; c++) row[c] = W[r * 24 + c];
        plate_to_global(T, row, out);
        for (int c = 0; c < 24; c++) KG[r * 24 + c] = out[c];
    }
}

void quad4_mindlin_sens(const double X[4][3], const PlateSection& ds, const double* u, double* df) {
    quad4_mindlin_impl(X, ds, u, df, nullptr);
}

void quad4_pressure(const double X[4][3], double p, double* f) {
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            double N[4], dN[4][2], g[4][3];
            quad4_shape(GP2[i], GP2[j], N, dN);
            const double JxW = quad4_map(X, dN, g) * 1.0;
            // force = p * normal, normal = (0,0,-1); f -= JxW N^T force
            const double fz = p * -1.0;
            for (int a = 0; a < 4; a++) f[2 * 4 + a] -= JxW * (N[a] * fz);
        }
}

// ------------------------------------------------------ hex8 conduction ----
void hex8_heat(const double X[8][3], double k0, double alpha, const double* T, double* f,
               double* K) {
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                double N[8], dN[8][3], g[8][3];
                hex8_shape(GP2[i], GP2[j], GP2[k], N, dN);
                const double JxW = hex8_map(X, dN, g) * 1.0;
                double Tq = 0, dT[3] = {0, 0, 0};
                for (int a = 0; a < 8; a++) {
                    Tq += N[a] * T[a];
                    for (int d = 0; d < 3; d++) dT[d] += g[a][d] * T[a];
                }
                const double kk = k0 * (1.0 + alpha * Tq);   // conductance k(T) I
                const double dk = k0 * alpha;                  // dk/dT
                double flux[3];
                for (int d = 0; d < 3; d++) flux[d] = dT[d] * kk;
                for (int d = 0; d < 3; d++)
                    for (int a = 0; a < 8; a++) f[a] += JxW * (g[a][d] * flux[d]);
                if (K) {
                    for (int d = 0; d < 3; d++)
                        for (int a = 0; a < 8; a++)
                            for (int b = 0; b < 8; b++) K[a * 8 + b] += JxW * kk * (g[a][d] * g[b][d]);
                    // dk/dT term (heat_conduction_elem_base.cpp:124-142)
                    for (int d = 0; d < 3; d++)
                        for (int a = 0; a < 8; a++)
                            for (int b = 0; b < 8; b++)
                                K[a * 8 + b] += JxW * dT[d] * dk * (g[a][d] * N[b]);
                }
            }
}

// side quadrature: the hex8 shape functions at 2x2 Gauss points of side s
static void hex8_side_point(int side, double s, double t, double xi[3]) {
    switch (side) {
        case 0: xi[0] = s; xi[1] = t; xi[2] = -1; break;
        case 1: xi[0] = s; xi[1] = -1; xi[2] = t; break;
        case 2: xi[0] = 1; xi[1] = s; xi[2] = t; break;
        case 3: xi[0] = s; xi[1] = 1; xi[2] = t; break;
        case 4: xi[0] = -1; xi[1] = s; xi[2] = t; break;
        default: xi[0] = s; xi[1] = t; xi[2] = 1; break;
    }
}

static double hex8_side_jxw(const double X[8][3], int side, const double dN[8][3]) {
    // surface Jacobian |dx/ds x dx/dt| with (s,t) the two in-face reference directions
    int r0, r1;
    switch (side) {
        case 0: case 5: r0 = 0; r1 = 1; break;
        case 1: case 3: r0 = 0; r1 = 2; break;
        default: r0 = 1; r1 = 2; break;
    }
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int n = 0; n < 8; n++)
        for (int i = 0; i < 3; i++) { a[i] += X[n][i] * dN[n][r0]; b[i] += X[n][i] * dN[n][r1]; }
    const double g11 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double g12 = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double g22 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    return std::sqrt(g11 * g22 - g12 * g12);
}

void hex8_convection(const double X[8][3], int side, double h, double Tinf, const double* T,
                     double* f, double* K) {
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            double xi[3], N[8], dN[8][3];
            hex8_side_point(side, GP2[i], GP2[j], xi);
            hex8_shape(xi[0], xi[1], xi[2], N, dN);
            const double JxW = hex8_side_jxw(X, side, dN) * 1.0;
            double temp = 0;
            for (int a = 0; a < 8; a++) temp += N[a] * T[a];
            const double th = 1.0;
            for (int a = 0; a < 8; a++) f[a] += JxW * N[a] * th * h * (temp - Tinf);
            if (K)
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++) K[a * 8 + b] += JxW * (N[a] * N[b]) * th * h;
        }
}

// surface_radiation_residual on side s (heat_conduction_elem_base.cpp:1402-1463):
// f += int N sigma eps th ((T + T0)^4 - (Ta + T0)^4),
// K += int N N^T sigma eps th 4 (T + T0)^3, th = 1 (a 3D element has no section,
// src/property_cards/isotropic_element_property_card_3D.h:164-167)
void hex8_radiation(const double X[8][3], int side, double sigma, double emiss, double Tamb,
                    double T0, const double* T, double* f, double* K) {
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            double xi[3], N[8], dN[8][3];
            hex8_side_point(side, GP2[i], GP2[j], xi);
            hex8_shape(xi[0], xi[1], xi[2], N, dN);
            const double JxW = hex8_side_jxw(X, side, dN) * 1.0;
            double temp = 0;
            for (int a = 0; a < 8; a++) temp += N[a] * T[a];
            const double th = 1.0;
            const double flux = sigma * emiss * th * (std::pow(temp + T0, 4.) - std::pow(Tamb + T0, 4.));
            for (int a = 0; a < 8; a++) f[a] += JxW * N[a] * flux;
            if (K) {
                const double dflux = sigma * emiss * th * 4. * std::pow(temp + T0, 3.);
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++) K[a * 8 + b] += JxW * (N[a] * N[b]) * dflux;
            }
        }
}

// StructuralElement3D::surface_pressure_residual on side s (solid_element_3d.
// cpp:815-876): f -= int N (p n) dA over the side (2x2 side Gauss), n the
// outward unit normal, f in the element's variable-major layout (u0..u7,
// v0..v7, w0..w7).  n dA = sign (dx/ds x dx/dt) with (s, t) the side's
// in-face reference directions of hex8_side_point; sign orients the cross
// product outward for libMesh's side numbering (side 0: zeta = -1, 1: eta =
// -1, 2: xi = +1, 3: eta = +1, 4: xi = -1, 5: zeta = +1).
void hex8_pressure(const double X[8][3], int side, double p, double* f) {
    static const double kSign[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};
    int r0, r1;
    switch (side) {
        case 0: case 5: r0 = 0; r1 = 1; break;
        case 1: case 3: r0 = 0; r1 = 2; break;
        default: r0 = 1; r1 = 2; break;
    }
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
            double xi[3], N[8], dN[8][3];
            hex8_side_point(side, GP2[i], GP2[j], xi);
            hex8_shape(xi[0], xi[1], xi[2], N, dN);
            double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            for (int n = 0; n < 8; n++)
                for (int d = 0; d < 3; d++) { a[d] += X[n][d] * dN[n][r0]; b[d] += X[n][d] * dN[n][r1]; }
            const double nda[3] = {kSign[side] * (a[1] * b[2] - a[2] * b[1]),
                                   kSign[side] * (a[2] * b[0] - a[0] * b[2]),
                                   kSign[side] * (a[0] * b[1] - a[1] * b[0])};
            for (int d = 0; d < 3; d++)
                for (int n = 0; n < 8; n++) f[d * 8 + n] -= N[n] * p * nda[d];
        }
}

void hex8_source(const double X[8][3], double q, double* f) {
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                double N[8], dN[8][3], g[8][3];
                hex8_shape(GP2[i], GP2[j], GP2[k], N, dN);
                const double JxW = hex8_map(X, dN, g) * 1.0;
                for (int a = 0; a < 8; a++) f[a] -= JxW * N[a] * q * 1.0;
            }
}

}  // namespace oracle
