// Host-side adaptive refinement behind the C ABI: the dependency closure and
// level buckets of AdaptiveRefineList (src/adap/RefineListAdaptive.h:8-126),
// the adjacency of ElementAssociation (src/adap/ElementAssociation.h:42-249)
// and the 1->4 / 1->8 split of MeshRefine2D (src/adap/MeshRefine2D.h:39-872).
//
// The decisions, orders and numbering are the reference's (element and node
// numbers are append order; a hanging local node keeps the parent's node and
// gets an Hnm row, as the reference keeps the pointer); the data structures
// are not: node->element lists and face/edge neighbour lists are built in
// O(n) by counting instead of the O(n_node * n_elem * nnod) scan, repeated
// node detection goes through a spatial hash instead of a scan over all
// nodes, and refinelist position lookups through a map.  Each of these
// returns what the reference's loop returns (first match by node number,
// neighbours in ascending element order with the same multiplicities).
//
// Where the reference prints and blocks on stdin, or recurses without end,
// the call returns CFDV_ERR_REFINE with the reason.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cfdv.h"

// bit-for-bit the reference's coordinate and distance arithmetic
#pragma clang fp contract(off)

namespace {

struct RefineError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// get_nodes_from_face_3d (FiniteElement/get_face.h:106-165), get_nodes_from_face (:80-89)
constexpr int FACE3[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {6, 7, 4, 5}};
constexpr int FACE2[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
// get_opposing_face_3d (get_face.h:201-226)
constexpr int OPP3[6] = {5, 3, 4, 1, 2, 0};
// edge_association_3d (ElementAssociation.h:140-156)
constexpr int EDGE3[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                              {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
// create_nodes_3d (MeshRefine2D.h:289-331): isoparametric position per candidate
constexpr double ZETA3[19][3] = {{0, 0, -1}, {0, -1, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, 0},
                                 {0, -1, -1}, {1, 0, -1}, {0, 1, -1}, {-1, 0, -1},
                                 {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0},
                                 {0, -1, 1}, {1, 0, 1}, {0, 1, 1}, {-1, 0, 1}};
// edge candidate and the two face candidates it needs (:317-330)
constexpr int EDGEFACES3[12][3] = {{7, 0, 1}, {8, 0, 2}, {9, 0, 3}, {10, 0, 4}, {11, 4, 1}, {12, 1, 2},
                                   {13, 2, 3}, {14, 3, 4}, {15, 5, 1}, {16, 5, 2}, {17, 5, 3}, {18, 5, 4}};
constexpr double ZETA2[5][2] = {{0, -1}, {1, 0}, {0, 1}, {-1, 0}, {0, 0}};
// corners interpolated by each candidate's Hnm row (createElements_3d
// :504-523: 0.25 on four corners for a face centre, 0.5 on two for an edge)
constexpr int HCOLS3[19][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {1, 2, 5, 6}, {2, 3, 6, 7}, {0, 3, 4, 7}, {4, 5, 6, 7},
                               {-1, -1, -1, -1},
                               {0, 1, -1, -1}, {1, 2, -1, -1}, {2, 3, -1, -1}, {0, 3, -1, -1},
                               {0, 4, -1, -1}, {1, 5, -1, -1}, {2, 6, -1, -1}, {3, 7, -1, -1},
                               {4, 5, -1, -1}, {5, 6, -1, -1}, {6, 7, -1, -1}, {4, 7, -1, -1}};
constexpr int HCOLS2[5][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}, {-1, -1}};
// set_Element_Node calls (local node, candidate) per new element, in call
// order; the last row is the refined element itself (:525-600; 2D :477-495)
constexpr int KIDS3[8][7][2] = {
    {{0, 7}, {2, 8}, {3, 0}, {4, 1}, {5, 12}, {6, 2}, {7, 6}},
    {{0, 0}, {1, 8}, {3, 9}, {5, 2}, {6, 13}, {7, 3}, {4, 6}},
    {{0, 10}, {1, 0}, {2, 9}, {4, 4}, {6, 3}, {7, 14}, {5, 6}},
    {{0, 11}, {1, 1}, {3, 4}, {5, 15}, {6, 5}, {7, 18}, {2, 6}},
    {{0, 1}, {1, 12}, {2, 2}, {4, 15}, {6, 16}, {7, 5}, {3, 6}},
    {{1, 2}, {2, 13}, {3, 3}, {4, 5}, {5, 16}, {7, 17}, {0, 6}},
    {{0, 4}, {2, 3}, {3, 14}, {4, 18}, {5, 5}, {6, 17}, {1, 6}},
    {{1, 7}, {2, 0}, {3, 10}, {4, 11}, {5, 1}, {7, 4}, {6, 6}}};
constexpr int KIDS2[4][3][2] = {{{0, 0}, {2, 1}, {3, 4}}, {{1, 1}, {3, 2}, {0, 4}},
                                {{0, 3}, {2, 2}, {1, 4}}, {{1, 0}, {3, 3}, {2, 4}}};
// addFaces_3d (:706-761): receivers of a face by index (-1 = the refined element)
constexpr int FACERECV3[6][4] = {{-1, 0, 1, 2}, {-1, 0, 3, 4}, {0, 1, 4, 5}, {1, 2, 5, 6}, {-1, 2, 3, 6}, {3, 4, 5, 6}};

// get_edge_from_nodes (get_face.h:232-292)
int edge_candidate(int a, int b) {
    if (a > b) std::swap(a, b);
    static const int tab[12][3] = {{0, 1, 7}, {1, 2, 8}, {2, 3, 9}, {0, 3, 10}, {0, 4, 11}, {1, 5, 12},
                                   {2, 6, 13}, {3, 7, 14}, {4, 5, 15}, {5, 6, 16}, {6, 7, 17}, {4, 7, 18}};
    for (auto& t : tab)
        if (t[0] == a && t[1] == b) return t[2];
    throw RefineError("get_edge_from_nodes: no edge through local nodes " + std::to_string(a) + " " +
                      std::to_string(b) + " (the reference prints and blocks on stdin, get_face.h:288-291)");
}
// get_face_3d (get_face.h:168-183) on a local node list
int face_of_nodes3(const int* n) {
    for (int f = 0; f < 6; f++)
        if (n[0] == FACE3[f][0] && n[1] == FACE3[f][1] && n[2] == FACE3[f][2] && n[3] == FACE3[f][3]) return f;
    if (n[0] == 4 && n[1] == 5 && n[2] == 6 && n[3] == 7) return 5;
    throw RefineError("get_face_3d: no face for the node list (get_face.h:184-189)");
}

struct FaceRec {
    int32_t face, bc;
};

// ElementAssociation's lists in CSR form
struct Association {
    std::vector<int64_t> f_ptr;      // [n_elem * nface + 1]
    std::vector<int32_t> f_list;
    std::vector<uint8_t> boundary;   // [n_elem * nface]
    std::vector<int64_t> e_ptr;      // [n_elem * 12 + 1]
    std::vector<int32_t> e_list;
    int nface = 0;
    size_t nf(long e, int j) const { return (size_t)(f_ptr[e * nface + j + 1] - f_ptr[e * nface + j]); }
    const int32_t* f(long e, int j) const { return f_list.data() + f_ptr[e * nface + j]; }
};

struct CellKey {
    int64_t a, b, c;
    bool operator==(const CellKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct CellHash {
    size_t operator()(const CellKey& k) const {
        uint64_t h = (uint64_t)k.a * 0x9E3779B97F4A7C15ULL;
        h ^= (uint64_t)k.b * 0xC2B2AE3D27D4EB4FULL + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.c * 0x165667B19E3779F9ULL + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

}  // namespace

struct cfdv_refine {
    int ndim = 0, nnod = 0, neqn = 0, nface = 0;
    std::vector<int32_t> conn;                 // [n_elem][nnod]
    std::vector<std::vector<FaceRec>> faces;   // Element::face order
    std::vector<double> hnm;                   // [n_elem][nnod][nnod], H(N, M) at N * nnod + M
    std::vector<uint8_t> adap;
    std::vector<int32_t> level;
    std::vector<double> x, U, U0;              // [n_node][ndim], [n_node][neqn]
    std::vector<uint8_t> dir;                  // [n_node][neqn]
    std::string err;

    // repeated-node lookup (createNodeRef): cells of 2 tol on the compared
    // coordinates; every node of the mesh is in its cell
    int cmp_dims = 2;
    double tol = 1e-3;
    std::unordered_map<CellKey, std::vector<int32_t>, CellHash> cells;

    long n_elem() const { return (long)level.size(); }
    long n_node() const { return (long)(x.size() / ndim); }
    double* H(long e) { return hnm.data() + (size_t)e * nnod * nnod; }

    // ---- ElementAssociation (ElementAssociation.h:55-249) ----------------
    Association associate() const {
        const long ne = n_elem(), nn = n_node();
        // node -> elements, element order, one entry per matching local node (:62-77)
        std::vector<int64_t> np(nn + 1, 0);
        for (long j = 0; j < ne; j++)
            for (int k = 0; k < nnod; k++) np[conn[j * nnod + k] + 1]++;
        for (long i = 0; i < nn; i++) np[i + 1] += np[i];
        std::vector<int32_t> nl(np[nn]);
        {
            std::vector<int64_t> fill(np.begin(), np.end() - 1);
            for (long j = 0; j < ne; j++)
                for (int k = 0; k < nnod; k++) nl[fill[conn[j * nnod + k]]++] = (int32_t)j;
        }
        auto list = [&](long node, const int32_t** p) {
            *p = nl.data() + np[node];
            return (size_t)(np[node + 1] - np[node]);
        };
        auto count = [](const int32_t* p, size_t n, int32_t v) {
            size_t c = 0;
            for (size_t i = 0; i < n; i++) c += (p[i] == v);
            return c;
        };
        Association a;
        a.nface = nface;
        a.f_ptr.assign((size_t)ne * nface + 1, 0);
        a.boundary.assign((size_t)ne * nface, 0);
        a.e_ptr.assign((size_t)ne * 12 + 1, 0);
        std::vector<int32_t> tmp;
        for (long i = 0; i < ne; i++) {
            const int32_t* cn = conn.data() + i * nnod;
            for (int j = 0; j < nface; j++) {
                tmp.clear();
                if (nnod == 4) {   // face_association_2d (:222-245)
                    const int32_t *L0, *L1;
                    const size_t n0 = list(cn[FACE2[j][0]], &L0), n1 = list(cn[FACE2[j][1]], &L1);
                    for (size_t k = 0; k < n0; k++)
                        for (size_t l = 0; l < n1; l++)
                            if (L0[k] == L1[l] && L1[l] != i) tmp.push_back(L0[k]);
                } else {           // face_association_3d (:100-136)
                    const int32_t* L[4];
                    size_t n[4];
                    for (int q = 0; q < 4; q++) n[q] = list(cn[FACE3[j][q]], &L[q]);
                    for (size_t k = 0; k < n[0]; k++) {
                        const int32_t a0 = L[0][k];
                        if (a0 == i) continue;
                        const size_t c = count(L[1], n[1], a0) * count(L[2], n[2], a0) * count(L[3], n[3], a0);
                        for (size_t r = 0; r < c; r++) tmp.push_back(a0);
                    }
                }
                a.f_list.insert(a.f_list.end(), tmp.begin(), tmp.end());
                a.f_ptr[i * nface + j + 1] = (int64_t)a.f_list.size();
            }
            // boundary flags from the element's faces (:87-98)
            for (const FaceRec& fr : faces[i]) a.boundary[i * nface + fr.face] = 1;
            // edge_association_3d (:139-213): common elements of an edge's
            // two nodes that are no face neighbour of i
            for (int j = 0; j < 12; j++) {
                if (nnod == 8) {
                    const int32_t *L0, *L1;
                    const size_t n0 = list(cn[EDGE3[j][0]], &L0), n1 = list(cn[EDGE3[j][1]], &L1);
                    for (size_t k = 0; k < n0; k++)
                        for (size_t l = 0; l < n1; l++) {
                            if (!(L0[k] == L1[l] && L1[l] != i)) continue;
                            bool match = false;
                            for (int q = 0; q < nface && !match; q++) {
                                const int32_t* fl = a.f_list.data() + a.f_ptr[i * nface + q];
                                const size_t fn = (size_t)(a.f_ptr[i * nface + q + 1] - a.f_ptr[i * nface + q]);
                                for (size_t r = 0; r < fn; r++)
                                    if (fl[r] == L0[k]) { match = true; break; }
                            }
                            if (!match) a.e_list.push_back(L0[k]);
                        }
                }
                a.e_ptr[i * 12 + j + 1] = (int64_t)a.e_list.size();
            }
        }
        return a;
    }

    // ---- AdaptiveRefineList (RefineListAdaptive.h:8-126) ------------------
    void refine_list(const uint8_t* flags, std::vector<int32_t>& elelist, std::vector<int32_t>& rl) const {
        const long ne = n_elem();
        int n = 0;
        for (long i = 0; i < ne; i++) n = std::max(n, (int)level[i]);
        rl.assign((size_t)n + 2, 0);
        const Association ea = associate();
        std::vector<int32_t> list;
        for (long i = 0; i < ne; i++)
            if (flags[i]) list.push_back((int32_t)i);
        // find_dependancies (:68-112); its recursion follows elements[k], k
        // the position within the face list, as written
        std::function<void(int, long, int)> deps = [&](int rlev, long elno, int depth) {
            if (depth > 256)
                throw RefineError("find_dependancies recurses without end: RefineListAdaptive.h:84-88 "
                                  "recurses on elements[k], k the position within a face list");
            for (int j = 0; j < nface; j++) {
                if (ea.boundary[elno * nface + j]) continue;
                const size_t nf = ea.nf(elno, j);
                for (size_t k = 0; k < nf; k++) {
                    const int32_t nb = ea.f(elno, j)[k];
                    if (level[nb] < rlev) {
                        list.push_back(nb);
                        if ((long)k < ne && level[k] > 0) deps(level[k], (long)k, depth + 1);
                    }
                }
            }
            for (int j = 0; j < 12; j++)
                for (int64_t p = ea.e_ptr[elno * 12 + j]; p < ea.e_ptr[elno * 12 + j + 1]; p++) {
                    const int32_t nb = ea.e_list[p];
                    if (level[nb] < rlev) list.push_back(nb);
                }
        };
        const size_t rsize = list.size();
        for (size_t i = 0; i < rsize; i++) {
            const int32_t el = list[i];
            if (level[el] != 0) deps(level[el], el, 0);
        }
        // remove_redundant_elements (:114-126): first occurrences kept
        std::vector<uint8_t> seen((size_t)ne, 0);
        std::vector<int32_t> dedup;
        for (int32_t v : list)
            if (!seen[v]) { seen[v] = 1; dedup.push_back(v); }
        elelist.clear();
        for (int i = 1; i <= n + 1; i++) {   // level buckets (:53-65)
            rl[i] = rl[i - 1];
            for (int32_t v : dedup)
                if (level[v] == i - 1) { elelist.push_back(v); rl[i]++; }
        }
    }

    // ---- MeshRefine2D (MeshRefine2D.h:39-872) -----------------------------
    CellKey cell_of(const double* p) const {
        const double w = 2.0 * tol;
        CellKey k{(int64_t)std::floor(p[0] / w), (int64_t)std::floor(p[1] / w), 0};
        if (cmp_dims == 3) k.c = (int64_t)std::floor(p[2] / w);
        return k;
    }
    void index_node(long i) { cells[cell_of(x.data() + i * ndim)].push_back((int32_t)i); }
    void build_index(int dims, double t) {
        cmp_dims = dims;
        tol = t;
        cells.clear();
        for (long i = 0; i < n_node(); i++) index_node(i);
    }
    // first node (lowest number) closer than tol, as the scan of createNodeRef (:337-361)
    long find_node(const double* p) const {
        const CellKey c = cell_of(p);
        long best = -1;
        const int zr = (cmp_dims == 3) ? 1 : 0;
        for (int64_t da = -1; da <= 1; da++)
            for (int64_t db = -1; db <= 1; db++)
                for (int64_t dc = -zr; dc <= zr; dc++) {
                    auto it = cells.find(CellKey{c.a + da, c.b + db, c.c + dc});
                    if (it == cells.end()) continue;
                    for (int32_t i : it->second) {
                        if (best >= 0 && i >= best) continue;
                        const double* q = x.data() + (size_t)i * ndim;
                        double s = (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]);
                        if (cmp_dims == 3) s = s + (q[2] - p[2]) * (q[2] - p[2]);
                        if (std::sqrt(s) < tol) best = i;
                    }
                }
        return best;
    }
    // createNode + addDirichlet + createNodeRef (:337-407, 627-657)
    int32_t create_node_ref(const double* zeta, long e) {
        double Phi[8];
        if (nnod == 8) {
            const double z0 = zeta[0], z1 = zeta[1], z2 = zeta[2];
            Phi[0] = 0.125 * (1 - z0) * (1 - z1) * (1 - z2);
            Phi[1] = 0.125 * (1 + z0) * (1 - z1) * (1 - z2);
            Phi[2] = 0.125 * (1 + z0) * (1 + z1) * (1 - z2);
            Phi[3] = 0.125 * (1 - z0) * (1 + z1) * (1 - z2);
            Phi[4] = 0.125 * (1 - z0) * (1 - z1) * (1 + z2);
            Phi[5] = 0.125 * (1 + z0) * (1 - z1) * (1 + z2);
            Phi[6] = 0.125 * (1 + z0) * (1 + z1) * (1 + z2);
            Phi[7] = 0.125 * (1 - z0) * (1 + z1) * (1 + z2);
        } else {
            const double z0 = zeta[0], z1 = zeta[1];
            Phi[0] = 0.25 * (1 - z0) * (1 - z1);
            Phi[1] = 0.25 * (1 + z0) * (1 - z1);
            Phi[2] = 0.25 * (1 + z0) * (1 + z1);
            Phi[3] = 0.25 * (1 - z0) * (1 + z1);
        }
        double nx[3] = {0, 0, 0}, nU[5] = {0, 0, 0, 0, 0}, nU0[5] = {0, 0, 0, 0, 0}, nd[5] = {0, 0, 0, 0, 0};
        const int32_t* cn = conn.data() + e * nnod;
        for (int N = 0; N < nnod; N++) {
            const long g = cn[N];
            for (int i = 0; i < neqn; i++) {
                nU[i] += Phi[N] * U[g * neqn + i];
                nU0[i] += Phi[N] * U0[g * neqn + i];
            }
            for (int i = 0; i < ndim; i++) nx[i] += Phi[N] * x[g * ndim + i];
        }
        for (int N = 0; N < nnod; N++) {
            const long g = cn[N];
            for (int i = 0; i < neqn; i++) nd[i] += Phi[N] * (double)dir[g * neqn + i];
        }
        const long hit = find_node(nx);
        if (hit >= 0) return (int32_t)hit;
        const long id = n_node();
        for (int i = 0; i < ndim; i++) x.push_back(nx[i]);
        for (int i = 0; i < neqn; i++) {
            U.push_back(nU[i]);
            U0.push_back(nU0[i]);
            dir.push_back(nd[i] < 1 ? 0 : 1);
        }
        index_node(id);
        return (int32_t)id;
    }

    void update_hanging(long el, int32_t node, int idx) {   // update_hanging_nodes (:258-275)
        double* h = H(el);
        for (int i = 0; i < nnod; i++) h[idx * nnod + i] = 0.0;
        h[idx * nnod + idx] = 1.0;
        bool ident = true;
        for (int a = 0; a < nnod && ident; a++)
            for (int b = 0; b < nnod; b++)
                if (h[a * nnod + b] != (a == b ? 1.0 : 0.0)) { ident = false; break; }
        if (ident) adap[el] = 0;
        conn[el * nnod + idx] = node;
    }

    struct Prior {
        int32_t elno, face;
        int local[4];
    };

    void refine(long e, const Association& ea, const std::vector<int32_t>& rlist,
                const std::unordered_map<int32_t, std::vector<int32_t>>& pos) {
        bool nodelist[19] = {};
        // findNeighbors (:827-862)
        int jlist = 0;
        {
            auto it = pos.find((int32_t)e);
            if (it != pos.end()) jlist = it->second.back();
        }
        for (int j = 0; j < nface; j++) {
            const size_t nf = ea.nf(e, j);
            for (size_t q = 0; q < nf; q++) {
                const int32_t k = ea.f(e, j)[q];
                auto it = pos.find(k);
                if (it == pos.end()) continue;
                for (int32_t l : it->second)
                    if (!(level[k] == level[e] && l < jlist)) nodelist[j] = true;
            }
        }
        // addPriorRefine (:786-821)
        std::vector<Prior> prior;
        const size_t bordering = (nnod == 4) ? 2 : 4;
        for (int j = 0; j < nface; j++) {
            const size_t nf = ea.nf(e, j);
            if (nf != bordering) continue;
            nodelist[j] = true;
            for (size_t q = 0; q < nf; q++) {
                Prior p{ea.f(e, j)[q], j, {0, 0, 0, 0}};
                if (nnod == 4) {
                    const int of = (j + 2) % 4;
                    p.local[0] = FACE2[of][0];
                    p.local[1] = FACE2[of][1];
                } else {
                    for (int t = 0; t < 4; t++) p.local[t] = FACE3[OPP3[j]][t];
                }
                prior.push_back(p);
            }
        }
        // addBoundaries (:766-781)
        for (const FaceRec& fr : faces[e]) nodelist[fr.face] = true;
        // new nodes (create_nodes_2d/3d, :277-331)
        int32_t new_node[19] = {};
        if (nnod == 4) {
            for (int c = 0; c < 4; c++)
                if (nodelist[c]) new_node[c] = create_node_ref(ZETA2[c], e);
            nodelist[4] = true;
            new_node[4] = create_node_ref(ZETA2[4], e);
        } else {
            for (int c = 0; c < 6; c++)
                if (nodelist[c]) new_node[c] = create_node_ref(ZETA3[c], e);
            nodelist[6] = true;
            new_node[6] = create_node_ref(ZETA3[6], e);
            for (auto& ef : EDGEFACES3)
                if (nodelist[ef[1]] && nodelist[ef[2]]) {
                    nodelist[ef[0]] = true;
                    new_node[ef[0]] = create_node_ref(ZETA3[ef[0]], e);
                }
        }
        // createElements (:411-462) + createElements_2d/3d + set_Element_Node
        const long ne0 = n_elem();
        const int nk = nnod - 1;
        for (int c = 0; c < nk; c++) {
            for (int j = 0; j < nnod; j++) conn.push_back(conn[e * nnod + j]);
            faces.emplace_back();
            for (int a = 0; a < nnod; a++)
                for (int b = 0; b < nnod; b++) hnm.push_back(a == b ? 1.0 : 0.0);
            adap.push_back(0);
            level.push_back(0);
        }
        {
            double* h = H(e);
            for (int a = 0; a < nnod; a++)
                for (int b = 0; b < nnod; b++) h[a * nnod + b] = (a == b) ? 1.0 : 0.0;
        }
        for (int ci = 0; ci < nnod; ci++) {
            const long el = (ci == nnod - 1) ? e : ne0 + ci;
            for (int t = 0; t < nk; t++) {
                const int elenode = (nnod == 8) ? KIDS3[ci][t][0] : KIDS2[ci][t][0];
                const int cand = (nnod == 8) ? KIDS3[ci][t][1] : KIDS2[ci][t][1];
                if (nodelist[cand]) {
                    conn[el * nnod + elenode] = new_node[cand];
                } else {
                    adap[el] = 1;
                    double* h = H(el) + elenode * nnod;
                    for (int i = 0; i < nnod; i++) h[i] = 0.0;
                    if (nnod == 8) {
                        const double w = (HCOLS3[cand][2] >= 0) ? 0.25 : 0.5;
                        for (int q = 0; q < 4; q++)
                            if (HCOLS3[cand][q] >= 0) h[HCOLS3[cand][q]] = w;
                    } else {
                        for (int q = 0; q < 2; q++)
                            if (HCOLS2[cand][q] >= 0) h[HCOLS2[cand][q]] = 0.5;
                    }
                }
            }
        }
        // faces (addFaces_3d :706-761, addFaces_2d :663-702)
        if (nnod == 8) {
            std::vector<FaceRec> old = faces[e];
            faces[e].clear();
            for (long i = (long)old.size() - 1; i >= 0; i--)
                for (int r : FACERECV3[old[i].face]) faces[r < 0 ? e : ne0 + r].push_back(old[i]);
        } else {
            std::vector<FaceRec>& fl = faces[e];
            for (long i = (long)fl.size() - 1; i >= 0; i--) {
                const FaceRec fr = fl[i];
                if (fr.face == 0) {
                    faces[ne0 + 0].push_back(fr);
                } else if (fr.face == 1) {
                    faces[ne0 + 0].push_back(fr);
                    faces[ne0 + 1].push_back(fr);
                    fl.erase(fl.begin() + i);
                } else if (fr.face == 2) {
                    faces[ne0 + 1].push_back(fr);
                    faces[ne0 + 2].push_back(fr);
                    fl.erase(fl.begin() + i);
                } else if (fr.face == 3) {
                    faces[ne0 + 2].push_back(fr);
                }
            }
        }
        level[e]++;
        for (int c = 0; c < nk; c++) level[ne0 + c] = level[e];
        // hanging nodes of elements refined before across a face
        if (nnod == 8) {   // do_hanging_nodes_3d (:129-221)
            for (const Prior& p : prior) {
                for (int N = 0; N < 8; N++) {
                    const double* h = H(p.elno) + N * 8;
                    if (h[p.local[0]] == 0.25 && h[p.local[1]] == 0.25 && h[p.local[2]] == 0.25 &&
                        h[p.local[3]] == 0.25 && new_node[p.face] != 0)
                        update_hanging(p.elno, new_node[p.face], N);
                }
                for (int N = 0; N < 8; N++)
                    for (int ni = 0; ni < 4; ni++)
                        for (int nj = ni + 1; nj < 4; nj++) {
                            const double* h = H(p.elno) + N * 8;
                            if (!(h[p.local[ni]] == 0.5 && h[p.local[nj]] == 0.5)) continue;
                            const int* n = FACE3[OPP3[face_of_nodes3(p.local)]];
                            const int ni0 = (ni == 0 || ni == 2) ? ni + 1 : ni - 1;
                            const int nj0 = (nj == 0 || nj == 2) ? nj + 1 : nj - 1;
                            const int edge = edge_candidate(n[ni0], n[nj0]);
                            if (new_node[edge] != 0) update_hanging(p.elno, new_node[edge], N);
                        }
            }
        } else {           // do_hanging_nodes_2d (:225-253)
            for (const Prior& p : prior) {
                const int32_t node_number = new_node[p.face];
                if (level[p.elno] != level[e]) continue;
                int indexi = -1;
                for (int j = 0; j < nnod; j++) {
                    const double* h = H(p.elno) + j * nnod;
                    if (h[p.local[0]] == 0.5 && h[p.local[1]] == 0.5) indexi = j;
                }
                if (indexi == -1)
                    throw RefineError("do_hanging_nodes_2d: no Hnm row interpolates the shared face of element " +
                                      std::to_string(p.elno) + " (the reference prints and blocks on stdin, "
                                      "MeshRefine2D.h:241-248)");
                update_hanging(p.elno, node_number, indexi);
            }
        }
    }

    void apply(const std::vector<int32_t>& elelist, const std::vector<int32_t>& rl) {
        std::unordered_map<int32_t, std::vector<int32_t>> pos;
        for (size_t l = 0; l < elelist.size(); l++) pos[elelist[l]].push_back((int32_t)l);
        const int n = (int)rl.size() - 2;
        Association ea = associate();
        for (int j = 0; j <= n; j++) {
            for (int i = rl[j]; i < rl[j + 1]; i++) refine(elelist[i], ea, elelist, pos);
            if (j != n) ea = associate();   // ea.reset between levels (:68-69)
        }
    }
};

namespace {
int fail(cfdv_refine* r, const std::string& msg, int code) {
    if (r) r->err = msg;
    return code;
}
}  // namespace

extern "C" {

int cfdv_refine_create(const cfdv_mesh* m, const double* U, const double* U0, const int32_t* refine_level,
                       cfdv_refine** out) {
    if (!out) return CFDV_ERR_ARG;
    *out = nullptr;
    if (!m || !U || !m->conn || !m->x || !m->bc || !m->dirichlet) return CFDV_ERR_ARG;
    if (!((m->ndim == 3 && m->nnod == 8 && m->neqn == 5) || (m->ndim == 2 && m->nnod == 4 && m->neqn == 4)))
        return CFDV_ERR_ARG;
    if (m->n_elem <= 0 || m->n_node <= 0 || m->n_elem > INT32_MAX / 8 || m->n_node > INT32_MAX / 2)
        return CFDV_ERR_ARG;
    if (m->adap && !m->hnm) return CFDV_ERR_ARG;
    cfdv_refine* r = new (std::nothrow) cfdv_refine();
    if (!r) return CFDV_ERR_ARG;
    try {
        r->ndim = m->ndim;
        r->nnod = m->nnod;
        r->neqn = m->neqn;
        r->nface = (m->ndim == 3) ? 6 : 4;
        const long ne = (long)m->n_elem, nn = (long)m->n_node;
        r->conn.assign(m->conn, m->conn + ne * r->nnod);
        for (int32_t v : r->conn)
            if (v < 0 || v >= nn) { delete r; return CFDV_ERR_ARG; }
        r->faces.resize(ne);
        for (long e = 0; e < ne; e++)   // loader order: ascending face index (LoadBinary.h:228-234)
            for (int f = 0; f < r->nface; f++) {
                const int32_t b = m->bc[e * r->nface + f];
                if (b != 0) r->faces[e].push_back(FaceRec{f, b});
            }
        const int nn2 = r->nnod * r->nnod;
        r->hnm.assign((size_t)ne * nn2, 0.0);
        r->adap.assign(ne, 0);
        for (long e = 0; e < ne; e++) {
            double* h = r->H(e);
            const bool a = m->adap && m->adap[e];
            r->adap[e] = a ? 1 : 0;
            for (int N = 0; N < r->nnod; N++)
                for (int M = 0; M < r->nnod; M++)
                    h[N * r->nnod + M] = (m->hnm && m->adap) ? m->hnm[(size_t)e * nn2 + N + r->nnod * M]
                                                             : (N == M ? 1.0 : 0.0);
        }
        r->level.assign(ne, 0);
        if (refine_level)
            for (long e = 0; e < ne; e++) {
                if (refine_level[e] < 0) { delete r; return CFDV_ERR_ARG; }
                r->level[e] = refine_level[e];
            }
        r->x.assign(m->x, m->x + nn * r->ndim);
        r->U.assign(U, U + nn * r->neqn);
        r->U0.assign(U0 ? U0 : U, (U0 ? U0 : U) + nn * r->neqn);
        r->dir.assign(m->dirichlet, m->dirichlet + nn * r->neqn);
    } catch (const std::exception&) {
        delete r;
        return CFDV_ERR_ARG;
    }
    *out = r;
    return CFDV_OK;
}

void cfdv_refine_destroy(cfdv_refine* r) { delete r; }

const char* cfdv_refine_last_error(const cfdv_refine* r) { return r ? r->err.c_str() : "null refine object"; }

int cfdv_refine_list(cfdv_refine* r, const uint8_t* flags, int32_t* elelist, int32_t* n_list, int32_t* level_ptr,
                     int32_t* n_level_ptr) {
    if (!r) return CFDV_ERR_ARG;
    if (!flags || !elelist || !n_list || !level_ptr || !n_level_ptr) return fail(r, "null argument", CFDV_ERR_ARG);
    try {
        std::vector<int32_t> el, rl;
        r->refine_list(flags, el, rl);
        std::memcpy(elelist, el.data(), sizeof(int32_t) * el.size());
        std::memcpy(level_ptr, rl.data(), sizeof(int32_t) * rl.size());
        *n_list = (int32_t)el.size();
        *n_level_ptr = (int32_t)rl.size();
    } catch (const RefineError& ex) {
        return fail(r, ex.what(), CFDV_ERR_REFINE);
    } catch (const std::exception& ex) {
        return fail(r, ex.what(), CFDV_ERR_ARG);
    }
    return CFDV_OK;
}

int cfdv_refine_apply(cfdv_refine* r, const int32_t* elelist, int32_t n_list, const int32_t* level_ptr,
                      int32_t n_level_ptr, int mode, double tol) {
    if (!r) return CFDV_ERR_ARG;
    if (n_list < 0 || n_level_ptr < 2 || !level_ptr || (n_list > 0 && !elelist))
        return fail(r, "bad list arguments", CFDV_ERR_ARG);
    if (mode != CFDV_REFINE_PARITY && mode != CFDV_REFINE_3D) return fail(r, "unknown mode", CFDV_ERR_ARG);
    std::vector<int32_t> el(elelist, elelist + n_list), rl(level_ptr, level_ptr + n_level_ptr);
    const long ne = r->n_elem();
    for (int32_t v : el)
        if (v < 0 || v >= ne) return fail(r, "element number out of range in elelist", CFDV_ERR_ARG);
    for (size_t i = 0; i + 1 < rl.size(); i++)
        if (rl[i] < 0 || rl[i] > rl[i + 1] || rl[i + 1] > n_list)
            return fail(r, "level_ptr must be non-decreasing within [0, n_list]", CFDV_ERR_ARG);
    try {
        r->build_index((mode == CFDV_REFINE_3D && r->ndim == 3) ? 3 : 2, tol > 0 ? tol : 10E-4);
        r->apply(el, rl);
    } catch (const RefineError& ex) {
        return fail(r, ex.what(), CFDV_ERR_REFINE);
    } catch (const std::exception& ex) {
        return fail(r, ex.what(), CFDV_ERR_ARG);
    }
    return CFDV_OK;
}

int cfdv_refine_size(const cfdv_refine* r, int64_t* n_elem, int64_t* n_node, int32_t* max_level) {
    if (!r) return CFDV_ERR_ARG;
    if (n_elem) *n_elem = r->n_elem();
    if (n_node) *n_node = r->n_node();
    if (max_level) {
        int32_t mx = 0;
        for (int32_t v : r->level) mx = std::max(mx, v);
        *max_level = mx;
    }
    return CFDV_OK;
}

int cfdv_refine_get(const cfdv_refine* r, int32_t* conn, int32_t* bc, double* x, double* U, double* U0,
                    uint8_t* dirichlet, uint8_t* adap, double* hnm, int32_t* refine_level) {
    if (!r) return CFDV_ERR_ARG;
    const long ne = r->n_elem();
    if (conn) std::memcpy(conn, r->conn.data(), sizeof(int32_t) * r->conn.size());
    if (bc) {   // SaveBinary.h:59-69: bc[face] from the element's face list
        std::memset(bc, 0, sizeof(int32_t) * ne * r->nface);
        for (long e = 0; e < ne; e++)
            for (const FaceRec& fr : r->faces[e]) bc[e * r->nface + fr.face] = fr.bc;
    }
    if (x) std::memcpy(x, r->x.data(), sizeof(double) * r->x.size());
    if (U) std::memcpy(U, r->U.data(), sizeof(double) * r->U.size());
    if (U0) std::memcpy(U0, r->U0.data(), sizeof(double) * r->U0.size());
    if (dirichlet) std::memcpy(dirichlet, r->dir.data(), r->dir.size());
    if (adap) std::memcpy(adap, r->adap.data(), r->adap.size());
    if (hnm) {
        const int nn = r->nnod;
        for (long e = 0; e < ne; e++)
            for (int N = 0; N < nn; N++)
                for (int M = 0; M < nn; M++)
                    hnm[(size_t)e * nn * nn + N + nn * M] = r->hnm[(size_t)e * nn * nn + N * nn + M];
    }
    if (refine_level) std::memcpy(refine_level, r->level.data(), sizeof(int32_t) * r->level.size());
    return CFDV_OK;
}

}  // extern "C"
