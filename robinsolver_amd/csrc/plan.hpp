// plan.hpp -- tree geometry, computed once on the host.
//
// Mirrors the reference's domain partition (bin_tree / node_divide,
// robin_tree_module.f90:288-337, 1369-1412), leaf grids (get_h,
// fd_module.f90:1060-1084), parent point boxes (robin_tree_module.f90:404-410)
// and face concatenation (robin_tree_module.f90:896-959, 1205-1230).
//
// Boundary vectors of every node are kept in "node order": the node's faces in
// face order (x1,x2,y1,y2,z1,z2) except its own interface face, which comes
// last.  A node's RtR map is then [T_NN T_NI; T_IN T_II] with contiguous blocks,
// and the merge and solve passes are plain strided matrix products.
#pragma once
#include <string>
#include <vector>

namespace rs {

struct PlanNode {
    double box[2][3];
    int ptbox[3];
    int iface;        // own interface face (1..6), 0 for the root
    int isbndry[7];
    int cut;          // axis (0..2) along which this node is divided; -1 for leaves
};

struct Level {
    int nodes = 0;
    int ptbox[3] = {0, 0, 0};
    int pf[7] = {0};      // points per face, 1..6
    int P = 0;            // boundary points of a node
    int cut = -1;         // axis along which this level's nodes are divided
    int own_iface[2] = {0, 0};
    int noff[2][7];       // start of face f in node order, per parity
    // merge of this level's nodes from the children at level+1
    int isz = 0;          // interface points
    int bc = 0;           // child boundary points
    int bn = 0;           // bc - isz
    std::vector<int> prow;  // [parent parity][child][bn] -> parent node-order row
    std::vector<int> imap;  // [2 isz] interface rows -> rows of the child pair
};

struct Plan {
    int d = 0, L = 0;
    double domain[2][3];
    double h_tgt = 0, kb = 0;
    double h[3] = {0, 0, 0};
    double bp[2][2][2];   // bndry_params (row, col, re/im)
    double nu[2], zeta[2];
    int n_leaves = 0, m = 0, n_int = 0, nb = 0;
    std::vector<std::vector<PlanNode>> nodes;  // per level, left to right
    std::vector<Level> lv;
    std::vector<int> root_leaf, root_loc;      // root boundary point -> (leaf, local boundary index)
    int leaf_foff[7];                          // face starts in leaf point order
    std::vector<int> leaf_nbr;                 // [leaf][7]: leaf across face f (1..6), -1 on the physical boundary
    // returns empty string on success, else the error message
    std::string build(int d, int L, const double* domain, double h_tgt, double kb, const double* bp);
    void leaf_points(int leaf, double* xyz, int* face) const;
    // the leaf grid index (g0, g1, g2) and face id of every point of a leaf, in
    // the order of leaf_points (the same for all leaves): [m][4]
    std::vector<int> point_pattern() const;
    // node-order index -> face-order index for a node of level lvl with given parity
    std::vector<int> node_to_face_order(int lvl, int parity) const;
};

}  // namespace rs
