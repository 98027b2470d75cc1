// Host-side structure of one mesh: active DOF numbering, node-block and
// scalar CSR patterns, element -> block positions and the contribution lists
// the segmented-reduction scatter walks.  Built once per mesh and uploaded;
// design changes never touch it.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mastg {

struct Structure {
    int npe = 8;                 // nodes per element
    int nav = 3;                 // active variables per node
    int n_vars = 6;              // caller variables per node
    int active_var[6] = {0, 1, 2, 0, 0, 0};
    int64_t n_nodes = 0, n_elems = 0, n_dofs = 0;
    int64_t n_active = 0, nnz = 0, n_blocks = 0;

    // node rank = position of the node in caller DOF order (active DOF of
    // (rank, v) = rank*nav + v); everything device-side is in rank space
    std::vector<int32_t> rank_of_node;     // [n_nodes]
    std::vector<int32_t> node_of_rank;     // [n_nodes]
    std::vector<int64_t> caller_dof_of_active;  // [n_active]
    std::vector<int64_t> node_base_r;      // [n_nodes] caller DOF of (rank, variable 0)
    std::vector<uint8_t> constrained_var;  // [n_nodes*n_vars] every caller variable, by rank

    std::vector<double> xyz_r;             // [n_nodes*3] coordinates by rank
    std::vector<int32_t> econn;            // [n_elems*npe] element node ranks

    // node-block CSR: row = node rank, columns = neighbour ranks (sorted)
    std::vector<int64_t> blk_ptr;          // [n_nodes+1]
    int64_t max_blocks_per_row = 0;        // largest node-row degree
    std::vector<int32_t> blk_col;          // [n_blocks]
    // scalar CSR over active DOFs (rows (rank,v), columns (nbr, w))
    std::vector<int64_t> row_ptr;          // [n_active+1]
    std::vector<int32_t> col;              // [nnz]
    std::vector<int64_t> diag_pos;         // [n_active]

    // segmented reduction: block k receives contributions
    // cidx[cptr[k] .. cptr[k+1]) = e*npe*npe + la*npe + lb
    std::vector<int32_t> cptr;             // [n_blocks+1]
    std::vector<int32_t> cidx;             // [n_elems*npe*npe]
    // for every block its transposed partner (for J^T)
    std::vector<int32_t> blk_transpose;    // [n_blocks]

    std::vector<uint8_t> constrained;      // [n_active]
    // node -> (element, local node) incidence, for residual / load assembly
    std::vector<int32_t> ninc_ptr;         // [n_nodes+1]
    std::vector<int32_t> ninc;             // [n_elems*npe] = e*npe + la (la ascending per node)

    // congruence classes: elements whose node coordinates relative to their
    // first node agree (to 1e-12 of the element size) share one reference
    // element matrix; n_classes = 0 when the mesh has more than kMaxClasses
    static constexpr int kMaxClasses = 8;
    int n_classes = 0;
    std::vector<uint8_t> elem_class;       // [n_elems]
    std::vector<int32_t> class_rep;        // [n_classes] representative element
};

// Builds the structure; returns an empty string on success, else a message.
std::string build_structure(int physics, int64_t n_nodes, int64_t n_elems,
                            const double* xyz, const int32_t* conn, int n_vars,
                            const int64_t* node_dof_base, int64_t n_constrained,
                            const int64_t* constrained_dofs, Structure& S);

}  // namespace mastg
