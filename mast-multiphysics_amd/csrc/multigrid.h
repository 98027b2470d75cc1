// Geometric multigrid hierarchy on lattice meshes (host): hex8 elasticity (3D,
// 3 variables per node), the quad4 Mindlin plate (2D, 5 variables) and hex8
// heat conduction (3D, 1 variable).
//
// A mesh qualifies when its nodes form a complete lattice of axis-aligned
// boxes ((nx+1)(ny+1)(nz+1) points; nz = 0 for the plate) and every element is
// one lattice cell in libMesh node order.  Levels halve every dimension while
// all stay even; level l+1's nodes are level l's nodes at even lattice
// indices.  Prolongation is tri/bilinear interpolation (a fine node
// interpolates from the corners of the coarse cell containing it), coarse
// operators are Galerkin products P^T A P formed element by element (exact for
// nested cells), and a coarse DOF is constrained when the fine DOF at the same
// point is.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "structure.h"

namespace mastg {

struct MGLevelHost {
    int dim = 3, nav = 3, npe = 8;       // lattice dimension, variables and nodes per element
    int nx = 0, ny = 0, nz = 0;          // cells per dimension (nz = 0 in 2D)
    int64_t n_nodes = 0, n_elems = 0;
    // node rank of lattice point (i + (nx+1)(j + (ny+1) k)); identity on coarse levels
    std::vector<int32_t> rank_of_lat;
    // element of lattice cell (i + nx (j + ny k)); identity on coarse levels
    std::vector<int32_t> elem_of_cell;
    Structure S;                         // coarse levels only (level 0 is the context's)
};

struct MGHierarchy {
    bool ok = false;
    std::string why;                     // reason when !ok
    std::vector<MGLevelHost> levels;     // [0] = the problem mesh
    int64_t coarsest_dofs = 0;
};

// Largest coarsest-level system solved by the dense inverse.
constexpr int64_t kMGMaxCoarseDofs = 1536;
// Coarsening continues while a level has more DOFs than this (and its cell
// counts are even): the coarsest inverse then fits in LDS.
constexpr int64_t kMGCoarseTarget = 136;

// Detects the lattice and builds the coarse levels (hex8: nav 3, quad4: nav 5).
MGHierarchy build_mg_hierarchy(const Structure& S);

// One-class hex8 lattice: Gc[c] = Q_c^T Khat Q_c ([8][64 blocks][3x3], block
// (A, B) at (A 8 + B) 9) for the 8 children c of a coarse cell, made exactly
// symmetric.  Khat: [64][3x3] in the same layout.
void mg_class_child_products(const double* Khat, double* Gc);

// Coarse cells of level C (the first coarse level of the hex8 problem level F)
// with a constrained DOF among the 27 fine lattice points they cover:
// dirty[E] = 1 and E in list.  constrained: by node rank, 3 per node.
void mg_dirty_coarse(const MGLevelHost& F, const MGLevelHost& C, const std::vector<uint8_t>& constrained,
                     std::vector<uint8_t>& dirty, std::vector<int32_t>& list);

}  // namespace mastg
