// C++ host adapter over the C ABI (include/cfdv.h) with the reference's
// functor/solver interface.
//
// Templated on the caller's element and node types; it reads only the
// members the reference's loader fills (LoadBinary.h:174-245, add_face at
// :255-318):
//   Element::node[j]            -> Node*            (FDV/Element.h:30)
//   Element::face[f]->face, bc  face index and code (FiniteElement/Face.h:6-22)
//   Node::number                index into the node vector (FDV/Node.h:43)
//   Node::x(i), U(r), U0(r)     Tensor<double,1> call syntax (FDV/Node.h:38-42)
//   Node::dirichlet(r)          Tensor<bool,1>          (FDV/Node.h:47)
//   Element::adap, Hnm(N, M), refine_level, number      (FDV/Element.h:24-41)
//
//   cfdv::Context      mesh + thermo on the device   (cfdv_create)
//   cfdv::FDVGalerkin  FDVGalerkin<Element*> over all elements + ApplyBC
//                      (FDV/FDVGalerkin.h:52-154, FDV/Function/ApplyBC.h)
//   cfdv::HipGMRES     GMRESbase::iterate / update (Utility/Solvers/
//                      GMRESbase.h:8-23, MPI_GMRES.h:125-301)
//   cfdv::refine_pass  the --adap branch (main.cpp:143-151): the flag pass of
//                      AdaptiveRefineList on the device, its closure and
//                      MeshRefine2D on the host (adap/RefineListAdaptive.h,
//                      adap/MeshRefine2D.h); cfdv::apply_refinement writes
//                      the result back into the element and node vectors
//
// Every negative return code of the C ABI is thrown as cfdv::Error.
#pragma once

#include <cstdint>
#include <algorithm>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfdv.h"

namespace cfdv {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc, const cfdv_ctx* h, const char* what) {
    if (rc < 0) throw Error(rc, std::string(what) + ": " + cfdv_last_error(h));
}

// Flat arrays of the element/node graph, in the caller's element and node
// order (node id = Node::number).
struct MeshArrays {
    int ndim = 0, nnod = 0, neqn = 0;
    int64_t n_elem = 0, n_node = 0;
    std::vector<int32_t> conn, bc;
    std::vector<double> x;
    std::vector<uint8_t> dirichlet;
    std::vector<uint8_t> adap;   // empty when no element is adapted
    std::vector<double> hnm;

    cfdv_mesh view() const {
        cfdv_mesh m;
        m.ndim = ndim;
        m.nnod = nnod;
        m.neqn = neqn;
        m.n_elem = n_elem;
        m.n_node = n_node;
        m.conn = conn.data();
        m.x = x.data();
        m.bc = bc.data();
        m.dirichlet = dirichlet.data();
        m.adap = adap.empty() ? nullptr : adap.data();
        m.hnm = adap.empty() ? nullptr : hnm.data();
        return m;
    }
};

template <class ElementT, class NodeT>
MeshArrays flatten(const std::vector<ElementT*>& elements, const std::vector<NodeT*>& nodes, int ndim,
                   int nnod, int neqn) {
    const int nface = ndim == 3 ? 6 : 4;
    MeshArrays a;
    a.ndim = ndim;
    a.nnod = nnod;
    a.neqn = neqn;
    a.n_elem = (int64_t)elements.size();
    a.n_node = (int64_t)nodes.size();
    a.conn.resize((size_t)a.n_elem * nnod);
    a.bc.assign((size_t)a.n_elem * nface, 0);
    a.x.resize((size_t)a.n_node * ndim);
    a.dirichlet.resize((size_t)a.n_node * neqn);
    bool any_adap = false;
    for (size_t e = 0; e < elements.size(); e++) {
        const ElementT* el = elements[e];
        for (int j = 0; j < nnod; j++) a.conn[e * nnod + j] = (int32_t)el->node[j]->number;
        for (const auto* f : el->face) {
            if (f->face < 0 || f->face >= nface) throw Error(CFDV_ERR_ARG, "face index out of range");
            a.bc[e * nface + f->face] = f->bc;
        }
        any_adap = any_adap || el->adap;
    }
    if (any_adap) {   // hanging-node elements: Element::adap and Element::Hnm
        a.adap.assign((size_t)a.n_elem, 0);
        a.hnm.assign((size_t)a.n_elem * nnod * nnod, 0.0);
        for (size_t e = 0; e < elements.size(); e++) {
            ElementT* el = elements[e];
            if (!el->adap) continue;
            a.adap[e] = 1;
            for (int N = 0; N < nnod; N++)
                for (int M = 0; M < nnod; M++) a.hnm[e * nnod * nnod + N + nnod * M] = el->Hnm(N, M);
        }
    }
    for (size_t n = 0; n < nodes.size(); n++) {
        const NodeT* nd = nodes[n];
        if ((size_t)nd->number != n) throw Error(CFDV_ERR_ARG, "Node::number is not the node's index");
        for (int i = 0; i < ndim; i++) a.x[n * ndim + i] = nd->x(i);
        for (int r = 0; r < neqn; r++) a.dirichlet[n * neqn + r] = nd->dirichlet(r) ? 1 : 0;
    }
    return a;
}

class Context {
public:
    template <class ElementT, class NodeT>
    Context(const std::vector<ElementT*>& elements, const std::vector<NodeT*>& nodes, int ndim, int nnod,
            int neqn, const cfdv_thermo& th, int blocks = CFDV_BLOCKS_FP64)
        : ndim_(ndim), neqn_(neqn), n_node_((int64_t)nodes.size()) {
        MeshArrays a = flatten(elements, nodes, ndim, nnod, neqn);
        cfdv_mesh m = a.view();
        int rc = cfdv_create(&m, &th, blocks, &h_);
        if (rc < 0) throw Error(rc, "cfdv_create failed (code " + std::to_string(rc) + ")");
        set_state(nodes);
    }
    Context(const cfdv_mesh& m, const cfdv_thermo& th, int blocks, const double* U)
        : ndim_(m.ndim), neqn_(m.neqn), n_node_(m.n_node) {
        int rc = cfdv_create(&m, &th, blocks, &h_);
        if (rc < 0) throw Error(rc, "cfdv_create failed (code " + std::to_string(rc) + ")");
        check(cfdv_set_state(h_, U), h_, "cfdv_set_state");
    }
    ~Context() { cfdv_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    cfdv_ctx* handle() const { return h_; }
    // flag pass of AdaptiveRefineList (RefineListAdaptive.h:28-38) at the
    // device state: flags[e] = s1 > 0.8
    std::vector<uint8_t> refine_flags(int64_t n_elem) const {
        std::vector<uint8_t> f((size_t)n_elem);
        int32_t nf = 0, nt = 0;
        check(cfdv_refine_flags(h_, f.data(), nullptr, &nf, &nt), h_, "cfdv_refine_flags");
        return f;
    }
    int neqn() const { return neqn_; }
    int64_t n_node() const { return n_node_; }

    // Node::U of every node -> device (what NodeUnpack reads each step)
    template <class NodeT>
    void set_state(const std::vector<NodeT*>& nodes) {
        std::vector<double> U((size_t)n_node_ * neqn_);
        for (size_t n = 0; n < nodes.size(); n++)
            for (int r = 0; r < neqn_; r++) U[n * neqn_ + r] = nodes[n]->U(r);
        check(cfdv_set_state(h_, U.data()), h_, "cfdv_set_state");
    }
    // device U, U0 -> Node::U, Node::U0 (what MPI_GMRES::update leaves)
    template <class NodeT>
    void get_state(std::vector<NodeT*>& nodes) const {
        std::vector<double> U((size_t)n_node_ * neqn_), U0(U.size());
        check(cfdv_get_state(h_, U.data(), U0.data()), h_, "cfdv_get_state");
        for (size_t n = 0; n < nodes.size(); n++)
            for (int r = 0; r < neqn_; r++) {
                nodes[n]->U(r) = U[n * neqn_ + r];
                nodes[n]->U0(r) = U0[n * neqn_ + r];
            }
    }
    void get_state(double* U, double* U0) const { check(cfdv_get_state(h_, U, U0), h_, "cfdv_get_state"); }

    double cfl(double cfl_number) const {
        double dt = 0;
        check(cfdv_cfl(h_, cfl_number, &dt), h_, "cfdv_cfl");
        return dt;
    }
    void build(double dt) const { check(cfdv_build(h_, dt), h_, "cfdv_build"); }
    cfdv_solve_stats solve(const cfdv_solve_opts& o) const {
        cfdv_solve_stats s;
        check(cfdv_solve(h_, &o, &s), h_, "cfdv_solve");
        return s;
    }
    void update() const { check(cfdv_update(h_), h_, "cfdv_update"); }
    std::vector<double> rms() const {
        double u[5] = {0, 0, 0, 0, 0};
        check(cfdv_rms(h_, u), h_, "cfdv_rms");
        return std::vector<double>(u, u + neqn_);
    }

private:
    cfdv_ctx* h_ = nullptr;
    int ndim_, neqn_;
    int64_t n_node_;
};

// FDVGalerkin<Element*> applied to every element, then ApplyBC: one call.
struct FDVGalerkin {
    FDVGalerkin(const Context& c, double dt) : ctx(c), dt(dt) {}
    void operator()() const { ctx.build(dt); }
    const Context& ctx;
    double dt;
};

// The GMRESbase interface.  Reference semantics by default; FP64 / MIXED
// modes are the restarted and mixed-precision solvers with a relative
// residual target.
template <class ElementT, class NodeT>
class HipGMRES {
public:
    HipGMRES(const Context& c, int restart, int m, int mode = CFDV_SOLVE_REFERENCE, double rtol = 1e-8)
        : ctx_(c) {
        opts_.mode = mode;
        opts_.restarts = restart;
        opts_.m = m;
        opts_.rtol = rtol;
        opts_.inner_m = 30;
        opts_.inner_rtol = 1e-3;
        opts_.inner_restarts = 4;
        opts_.transpose = 0;
        opts_.precond = CFDV_PRECOND_NONE;
        opts_.warm_start = 0;
    }
    virtual ~HipGMRES() = default;
    // blocks and rhs are already on the device (FDVGalerkin); the element
    // vector is accepted for interface compatibility only
    virtual void iterate(std::vector<ElementT*>&) { last_ = ctx_.solve(opts_); }
    virtual void update(std::vector<NodeT*>& nodes) {
        ctx_.update();
        ctx_.get_state(nodes);
    }
    const cfdv_solve_stats& stats() const { return last_; }
    cfdv_solve_opts& options() { return opts_; }

private:
    const Context& ctx_;
    cfdv_solve_opts opts_{};
    cfdv_solve_stats last_{};
};

// ---- the --adap branch (main.cpp:143-151) ------------------------------
struct RefineResult {
    MeshArrays mesh;                    // refined mesh; adap/hnm filled when any element hangs
    std::vector<double> U, U0;          // [n_node][neqn]
    std::vector<int32_t> level;         // Element::refine_level
    std::vector<int32_t> elelist, level_ptr;   // AdaptiveRefineList's outputs
};

// flags from the device (ctx must hold the state U), closure and split on
// the host; mode CFDV_REFINE_PARITY (the reference's x, y node match) or
// CFDV_REFINE_3D
inline RefineResult refine_pass(const Context& ctx, const MeshArrays& m, const std::vector<double>& U,
                                const std::vector<double>& U0, const std::vector<int32_t>& level,
                                int mode = CFDV_REFINE_PARITY) {
    const std::vector<uint8_t> flags = ctx.refine_flags(m.n_elem);
    cfdv_mesh view = m.view();
    cfdv_refine* r = nullptr;
    int rc = cfdv_refine_create(&view, U.data(), U0.empty() ? nullptr : U0.data(),
                                level.empty() ? nullptr : level.data(), &r);
    if (rc < 0) throw Error(rc, "cfdv_refine_create failed");
    std::unique_ptr<cfdv_refine, void (*)(cfdv_refine*)> guard(r, cfdv_refine_destroy);
    auto chk = [&](int code, const char* what) {
        if (code < 0) throw Error(code, std::string(what) + ": " + cfdv_refine_last_error(r));
    };
    RefineResult out;
    int32_t max_level = 0;
    chk(cfdv_refine_size(r, nullptr, nullptr, &max_level), "cfdv_refine_size");
    out.elelist.resize((size_t)m.n_elem);
    out.level_ptr.resize((size_t)max_level + 2);
    int32_t nl = 0, nr = 0;
    chk(cfdv_refine_list(r, flags.data(), out.elelist.data(), &nl, out.level_ptr.data(), &nr), "cfdv_refine_list");
    out.elelist.resize((size_t)nl);
    out.level_ptr.resize((size_t)nr);
    chk(cfdv_refine_apply(r, out.elelist.data(), nl, out.level_ptr.data(), nr, mode, 0.0), "cfdv_refine_apply");
    int64_t ne = 0, nn = 0;
    chk(cfdv_refine_size(r, &ne, &nn, nullptr), "cfdv_refine_size");
    MeshArrays& a = out.mesh;
    a.ndim = m.ndim;
    a.nnod = m.nnod;
    a.neqn = m.neqn;
    a.n_elem = ne;
    a.n_node = nn;
    const int nface = m.ndim == 3 ? 6 : 4;
    a.conn.resize((size_t)ne * m.nnod);
    a.bc.resize((size_t)ne * nface);
    a.x.resize((size_t)nn * m.ndim);
    a.dirichlet.resize((size_t)nn * m.neqn);
    a.adap.resize((size_t)ne);
    a.hnm.resize((size_t)ne * m.nnod * m.nnod);
    out.U.resize((size_t)nn * m.neqn);
    out.U0.resize(out.U.size());
    out.level.resize((size_t)ne);
    chk(cfdv_refine_get(r, a.conn.data(), a.bc.data(), a.x.data(), out.U.data(), out.U0.data(), a.dirichlet.data(),
                        a.adap.data(), a.hnm.data(), out.level.data()),
        "cfdv_refine_get");
    if (std::find(a.adap.begin(), a.adap.end(), (uint8_t)1) == a.adap.end()) {
        a.adap.clear();
        a.hnm.clear();
    }
    return out;
}

// Write a refined mesh back into the reference's vectors: existing elements
// and nodes keep their objects (numbers are stable: refinement appends),
// new ones are made with the reference constructors Element(nnod, neqn,
// ndim) and Node(ndim, neqn); faces come from make_face(face, bc).
template <class ElementT, class NodeT, class FaceT>
void apply_refinement(const RefineResult& res, std::vector<ElementT*>& elements, std::vector<NodeT*>& nodes,
                      const std::function<FaceT*(int face, int bc)>& make_face) {
    const MeshArrays& a = res.mesh;
    const int nnod = a.nnod, neqn = a.neqn, ndim = a.ndim, nface = ndim == 3 ? 6 : 4;
    for (int64_t n = (int64_t)nodes.size(); n < a.n_node; n++) {
        NodeT* nd = new NodeT(ndim, neqn);
        nd->number = (unsigned int)n;
        nodes.push_back(nd);
    }
    for (int64_t n = 0; n < a.n_node; n++) {
        NodeT* nd = nodes[(size_t)n];
        for (int i = 0; i < ndim; i++) nd->x(i) = a.x[(size_t)n * ndim + i];
        for (int q = 0; q < neqn; q++) {
            nd->U(q) = res.U[(size_t)n * neqn + q];
            nd->U0(q) = res.U0[(size_t)n * neqn + q];
            nd->dirichlet(q) = a.dirichlet[(size_t)n * neqn + q] != 0;
        }
    }
    for (int64_t e = (int64_t)elements.size(); e < a.n_elem; e++) {
        ElementT* el = new ElementT(nnod, neqn, ndim);
        el->number = (int)e;
        elements.push_back(el);
    }
    for (int64_t e = 0; e < a.n_elem; e++) {
        ElementT* el = elements[(size_t)e];
        for (int j = 0; j < nnod; j++) el->node[j] = nodes[(size_t)a.conn[(size_t)e * nnod + j]];
        el->face.clear();
        for (int f = 0; f < nface; f++) {
            const int32_t b = a.bc[(size_t)e * nface + f];
            if (b != 0) el->face.push_back(make_face(f, b));
        }
        el->adap = !a.adap.empty() && a.adap[(size_t)e];
        for (int N = 0; N < nnod; N++)
            for (int M = 0; M < nnod; M++)
                el->Hnm(N, M) = el->adap ? a.hnm[(size_t)e * nnod * nnod + N + nnod * M] : (N == M ? 1.0 : 0.0);
        el->refine_level = res.level[(size_t)e];
    }
}

}  // namespace cfdv
