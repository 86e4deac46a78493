// cfdv_run: the reference's time loop (main.cpp:162-206) on one MI355X,
// driven through the C++ adapter over the C ABI.
//
//   cfdv_run <deck_dir> [--iters N] [--mode reference|fp64|mixed]
//            [--rtol X] [--precond none|nodeblock] [--out DIR] [--dump FILE]
//            [--flatten FILE] [--adap parity|3d] [--tecplot N] [--tecplot-only]
//            [--storage rows|ebe] [--save N] [--warm-start 0|1|2]
//
// Reads a reference binary deck (control, aero, binarynodes, binaryelements;
// LoadBinary.h:156-245) into element/node objects shaped like the
// reference's (Element::node / face, Node::number / x / U / U0 / dirichlet),
// sets the Dirichlet flags as add_face does (LoadBinary.h:255-318), then per
// iteration: CFL, FDVGalerkin + ApplyBC, GMRES iterate + update, and appends
// the RMS line to <out>/RMSError.txt in MPI_RMSError's format
// (MPI_RMSError.h:55-70: "iter\t" then each Urms followed by a space).
// --dump writes the final node U as raw doubles [n_node][neqn].
// --flatten writes the arrays handed to cfdv_create (conn int32, bc int32,
// x double, dirichlet uint8, U double, in that order) and exits without
// touching the device.
// --adap runs the refinement branch instead of the time loop (main.cpp:
// 143-157): the s1 > 0.8 flag pass on the device, AdaptiveRefineList's
// closure and MeshRefine2D on the host ("parity": the reference's x, y node
// match, quirk Q14; "3d": every coordinate), NodeUnpack, then saves the
// refined deck into <out> (binarynodes, binaryelements, control with adap = 1,
// aero, and the adap file of SaveBinary::write_adap, SaveBinary.h:153-168).
// Records are written at the deck's own dimension: the reference's 3D
// SaveBinary writes 2D-sized records (quirk Q16).
// --tecplot N writes <out>/tecplot_<iter>.plt every N iterations (the
// reference's Output every 100, main.cpp:183) with the state the build of
// that iteration reads and the element s1..s4 of that build (ele_data, as
// main.cpp passes it); --tecplot-only writes <out>/tecplot_<iter>.plt of the
// loaded deck without s1..s4 and without touching the device.
// --save N writes a restartable deck <out>/<iter>/ (binarynodes,
// binaryelements, adap, control with iter, aero) every N iterations before
// the step and after the last one, as main.cpp:177 and :208 call
// config->Save every 1000 (the default; 0 disables); running cfdv_run on
// such a directory continues from its iteration.  <out>/time.txt gets
// IterationTimer's lines (Utility/Timer.h:7-35: header, then "iter
// run_time step_time t dt" per iteration), in wall-clock seconds.
// --storage: the FP64 and mixed solvers (--mode fp64|mixed) multiply over
// the element blocks summed into node rows (rows, the default) or element by
// element (ebe); the reference solve always runs element by element.
// --warm-start (fp64 and mixed modes; 0, the reference's zero start, by
// default): 1 starts each solve from the last increment scaled by dt / dt_prev,
// 2 from the last two increments' rate extrapolated (cfdv_solve_opts).
// A deck whose control has adap = 1 is read with its adap file
// (LoadBinary.h:247-261: "elno adap refine_level" then the serialised Hnm
// tensor, TensorStream.h:95-175, column-major).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "cfdv_reference_adapter.hpp"

namespace {

// Tensor<T,1>-like storage with the reference's call syntax
template <class T>
struct Vec {
    std::vector<T> d;
    explicit Vec(size_t n = 0) : d(n, T()) {}
    T& operator()(int i) { return d[(size_t)i]; }
    T operator()(int i) const { return d[(size_t)i]; }
};
struct BVec {   // Tensor<bool,1>: element access by reference
    std::deque<bool> d;
    explicit BVec(size_t n = 0) : d(n, false) {}
    void set(int i) { d[(size_t)i] = true; }
    bool& operator()(int i) { return d[(size_t)i]; }
    bool operator()(int i) const { return d[(size_t)i]; }
};

struct Node {
    Node(int ndim, int neqn) : x(ndim), U(neqn), U0(neqn), dirichlet(neqn) {}
    unsigned int number = 0;
    Vec<double> x, U, U0;
    BVec dirichlet;
};
struct Face {
    int face = 0;
    int bc = 0;
};
// Tensor<double,2>-like storage with the reference's call syntax,
// column-major as the reference's Tensor (Tensor.h:53-117)
struct Mat {
    int n = 0;
    std::vector<double> d;
    explicit Mat(int n_ = 0) : n(n_), d((size_t)n_ * n_, 0.0) {}
    double& operator()(int i, int j) { return d[(size_t)i + (size_t)n * j]; }
    double operator()(int i, int j) const { return d[(size_t)i + (size_t)n * j]; }
};
struct Element {
    Element() = default;
    Element(int nnod, int /*neqn*/, int /*ndim*/) : node((size_t)nnod), Hnm(nnod) {}
    int number = 0;
    std::vector<Node*> node;
    std::vector<Face*> face;
    Mat Hnm;
    bool adap = false;
    int refine_level = 0;
};

// face -> local node ids (get_face.h:80-165)
const int FACE3[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {6, 7, 4, 5}};
const int FACE2[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

std::map<std::string, double> read_keys(const std::string& path) {
    std::map<std::string, double> k;
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '!' || line[0] == '#') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = line.substr(0, eq), val = line.substr(eq + 1);
        key.erase(key.find_last_not_of(" \t") + 1);
        key.erase(0, key.find_first_not_of(" \t"));
        k[key] = std::atof(val.c_str());
    }
    return k;
}

// Thermo::set (Thermo.h:22-33), in its operation order
double thermo_cv(const cfdv_thermo& th) { return 1. / th.gamma / (th.gamma - 1.0) / std::pow(th.mach, 2); }

struct Deck {
    int ndim = 0, nnod = 0, neqn = 0;
    std::map<std::string, double> control, aero;
    cfdv_thermo th{};
    std::vector<std::unique_ptr<Node>> node_store;
    std::vector<std::unique_ptr<Element>> elem_store;
    std::vector<std::unique_ptr<Face>> face_store;
    std::vector<Node*> nodes;
    std::vector<Element*> elements;
};

Deck load(const std::string& dir) {
    Deck d;
    d.control = read_keys(dir + "/control");
    d.aero = read_keys(dir + "/aero");
    d.ndim = (int)d.control.at("ndim");
    d.nnod = (int)d.control.at("nnod");
    d.neqn = (int)d.control.at("neqn");
    d.th.gamma = d.aero.at("gamma");
    d.th.mach = d.aero.at("M");
    d.th.reynolds = d.aero.at("Re");
    d.th.prandtl = d.aero.at("Pr");
    d.th.tinf = d.aero.at("Tinf");
    d.th.twall = d.aero.at("Twall");
    d.th.adiabatic = d.aero.at("Adiabatic") != 0;
    const int ndim = d.ndim, nnod = d.nnod, neqn = d.neqn, nface = ndim == 3 ? 6 : 4;
    const double Cv = thermo_cv(d.th);

    std::ifstream nin(dir + "/binarynodes", std::ios::binary);
    std::ifstream ein(dir + "/binaryelements", std::ios::binary);
    if (!nin || !ein) throw std::runtime_error("missing binary deck files in " + dir);
    const long nn = (long)d.control.at("nodes"), ne = (long)d.control.at("elements");
    std::vector<double> rec(2 * ndim + 2);
    for (long i = 0; i < nn; i++) {
        nin.read((char*)rec.data(), (std::streamsize)(rec.size() * sizeof(double)));
        if (!nin) throw std::runtime_error("short binarynodes");
        auto n = std::make_unique<Node>(ndim, neqn);
        n->number = (unsigned)i;
        for (int j = 0; j < ndim; j++) n->x(j) = rec[j];
        // primitive -> conservative (LoadBinary.h:179-196)
        const double rho = rec[ndim];
        n->U(0) = rho;
        for (int j = 0; j < ndim; j++) n->U(1 + j) = rec[ndim + 1 + j];
        n->U(ndim + 1) = rec[2 * ndim + 1] * Cv;
        double vs = 0;
        for (int j = 0; j < ndim; j++) {
            vs += n->U(j + 1) * n->U(j + 1);
            n->U(j + 1) *= n->U(0);
        }
        n->U(ndim + 1) += 0.5 * vs;
        n->U(ndim + 1) *= n->U(0);
        d.nodes.push_back(n.get());
        d.node_store.push_back(std::move(n));
    }
    std::vector<int32_t> erec(nnod + nface);
    for (long i = 0; i < ne; i++) {
        ein.read((char*)erec.data(), (std::streamsize)(erec.size() * sizeof(int32_t)));
        if (!ein) throw std::runtime_error("short binaryelements");
        auto e = std::make_unique<Element>(nnod, neqn, ndim);
        e->number = (int)i;
        for (int j = 0; j < nnod; j++) {
            if (erec[j] < 0 || erec[j] >= nn) throw std::runtime_error("node id out of range");
            e->node[(size_t)j] = d.nodes[(size_t)erec[j]];
        }
        for (int f = 0; f < nface; f++) {
            const int bc = erec[nnod + f];
            if (bc == 0) continue;
            auto fc = std::make_unique<Face>();
            fc->face = f;
            fc->bc = bc;
            e->face.push_back(fc.get());
            d.face_store.push_back(std::move(fc));
            // Dirichlet flags of add_face (LoadBinary.h:276-316)
            const int nb = ndim == 3 ? 4 : 2;
            for (int i = 0; i < nb; i++) {
                Node* nd = e->node[(size_t)(ndim == 3 ? FACE3[f][i] : FACE2[f][i])];
                if (bc == -1 || bc == -12) {
                    for (int q = 0; q < neqn; q++) nd->dirichlet.set(q);
                } else if (bc == 2 || bc == 9) {
                    nd->dirichlet.set(ndim);
                } else if (bc == 22) {
                    nd->dirichlet.set(2);
                } else if (bc == 1 || bc == 11 || bc == 21 || bc == 31 || bc == 41 || bc == 4) {
                    for (int q = 0; q < ndim; q++) nd->dirichlet.set(1 + q);
                    if (!d.th.adiabatic) nd->dirichlet.set(1 + ndim);
                }
            }
        }
        d.elements.push_back(e.get());
        d.elem_store.push_back(std::move(e));
    }
    if (d.control.count("adap") && d.control.at("adap") != 0) {
        // adap file (LoadBinary.h:247-261)
        std::ifstream af(dir + "/adap");
        if (!af) throw std::runtime_error("control has adap = 1 but " + dir + "/adap is missing");
        long elno;
        while (af >> elno) {
            int adap = 0, level = 0, tdim = 0, imax = 0, jmax = 0;
            std::string type;
            if (!(af >> adap >> level >> type >> tdim >> imax >> jmax) || tdim != 2 || imax != nnod || jmax != nnod ||
                elno < 0 || elno >= ne)
                throw std::runtime_error("bad record in " + dir + "/adap");
            Element* e = d.elements[(size_t)elno];
            e->adap = adap != 0;
            e->refine_level = level;
            for (size_t q = 0; q < e->Hnm.d.size(); q++)
                if (!(af >> e->Hnm.d[q])) throw std::runtime_error("short Hnm in " + dir + "/adap");
        }
    }
    return d;
}

// the file-name number: to_string<double>(iter) (Utility/tostring.h:5-11,
// MPI_TecplotOut.h:189), i.e. the default stream format of a double
std::string iter_name(int iter) {
    std::ostringstream ss;
    ss << (double)iter;
    return ss.str();
}

// MPI_TecplotOut (Utility/IO/MPI_TecplotOut.h:43-372) on one rank: node
// primitives from NodeUnpack (FDV/Function/NodeUnpack.h:10-30), BLOCK data
// of 7-digit scientific values, a line break after value 100, 200, ...,
// p tab-separated, optional per-element s1..s4 (tab-separated, no line
// break between the four blocks), then the 1-based connectivity.
void write_tecplot(const Deck& d, const std::string& path, double time, const double* s /* [n_elem][4] or null */) {
    const int ndim = d.ndim, nnod = d.nnod;
    const double Cv = thermo_cv(d.th), gamma = d.th.gamma;
    const size_t nn = d.nodes.size(), ne = d.elements.size();
    std::vector<double> rho(nn), T(nn), p(nn), v(nn * ndim);
    for (size_t n = 0; n < nn; n++) {
        const Node* nd = d.nodes[n];
        rho[n] = nd->U(0);
        double vsqr = 0;
        for (int i = 0; i < ndim; i++) {
            v[n * ndim + i] = nd->U(1 + i) / nd->U(0);
            vsqr += (v[n * ndim + i] * v[n * ndim + i]);
        }
        const double E = nd->U(ndim + 1) / nd->U(0);
        const double e = E - 0.5 * vsqr;
        T[n] = e / Cv;
        p[n] = rho[n] * (gamma - 1.0) * e;
    }
    std::ofstream tout(path.c_str(), std::ios::out);
    const bool ele_data = s != nullptr;
    tout << "title = \"Output at time = " << time << " \"" << std::endl;
    tout << "variables = \"x\", \"y\",";
    if (ndim == 3) tout << " \"z\",";
    tout << "\"rho\", \"u\", \"v\",";
    if (ndim == 3) tout << " \"w\",";
    tout << "\"T\", \"p\",";
    if (ele_data) tout << "\"s1\", \"s2\", \"s3\", \"s4\", ";
    tout << std::endl;
    tout << "zone E=" << ne << ", N=" << nn;
    tout << ", DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL";
    if (ndim == 3) tout << ", DATAPACKING=BLOCK, ZONETYPE=FEBRICK";
    tout << std::endl;
    if (ele_data) {
        if (ndim == 2) tout << "VARLOCATION=([8,9,10,11]=CellCentered)" << std::endl;
        else tout << "VARLOCATION=([10,11,12,13]=CellCentered)" << std::endl;
    }
    tout << std::endl;
    tout << std::scientific << std::endl;
    tout << std::setprecision(7) << std::endl;
    auto block = [&](auto value, const char* sep, bool end_line) {
        for (size_t n = 0; n < nn; n++) {
            tout << value(n) << sep;
            if (n % 100 == 0 && n != 0) tout << std::endl;
        }
        if (end_line) tout << std::endl;
    };
    for (int i = 0; i < ndim; i++) block([&](size_t n) { return d.nodes[n]->x(i); }, " ", true);
    block([&](size_t n) { return rho[n]; }, " ", true);
    for (int i = 0; i < ndim; i++) block([&](size_t n) { return v[n * ndim + i]; }, " ", true);
    block([&](size_t n) { return T[n]; }, " ", true);
    block([&](size_t n) { return p[n]; }, "\t", true);
    if (ele_data)
        for (int q = 0; q < 4; q++)
            for (size_t f = 0; f < ne; f++) {
                tout << s[f * 4 + q] << "\t";
                if (f % 100 == 0 && f != 0) tout << std::endl;
            }
    for (size_t e = 0; e < ne; e++) {
        for (int n = 0; n < nnod; n++) tout << d.elements[e]->node[(size_t)n]->number + 1 << " ";
        tout << std::endl;
    }
    if (!tout) throw std::runtime_error("cannot write " + path);
}

// SaveBinaryData (SaveBinary.h:34-168) after NodeUnpack (NodeUnpack.h:
// 124-144): node x, rho, v, T; element node ids then bc codes; the adap file
// of the adapted elements.  The control keeps the source deck's keys with
// the new sizes, adap = 1 when any element hangs (always for a refined
// deck: force_adap) and, for a checkpoint, iter = the iteration it starts.
void save_deck(const Deck& d, const std::string& dir, const std::string& src, bool force_adap = true,
               int iter = -1) {
    const int ndim = d.ndim, nnod = d.nnod, nface = ndim == 3 ? 6 : 4;
    const double Cv = thermo_cv(d.th);
    std::ofstream nf(dir + "/binarynodes", std::ios::binary);
    for (const Node* n : d.nodes) {
        std::vector<double> rec(2 * ndim + 2);
        const double rho = n->U(0);
        double vs = 0;
        for (int j = 0; j < ndim; j++) {
            rec[j] = n->x(j);
            const double v = n->U(1 + j) / rho;
            rec[ndim + 1 + j] = v;
            vs += v * v;
        }
        rec[ndim] = rho;
        const double E = n->U(ndim + 1) / rho;
        rec[2 * ndim + 1] = (E - 0.5 * vs) / Cv;
        nf.write((const char*)rec.data(), (std::streamsize)(rec.size() * sizeof(double)));
    }
    std::ofstream ef(dir + "/binaryelements", std::ios::binary);
    for (const Element* e : d.elements) {
        std::vector<int32_t> rec(nnod + nface, 0);
        for (int j = 0; j < nnod; j++) rec[j] = (int32_t)e->node[(size_t)j]->number;
        for (const Face* f : e->face) rec[nnod + f->face] = f->bc;
        ef.write((const char*)rec.data(), (std::streamsize)(rec.size() * sizeof(int32_t)));
    }
    // control: the source deck's keys with the new sizes and adap = 1
    std::ifstream cin_(src + "/control");
    std::ofstream cf(dir + "/control");
    std::string line;
    bool any_adap = force_adap;
    for (const Element* e : d.elements) any_adap = any_adap || e->adap;
    bool have_adap = false, have_iter = false;
    while (std::getline(cin_, line)) {
        const size_t eq = line.find('=');
        std::string key = eq == std::string::npos ? "" : line.substr(0, eq);
        key.erase(key.find_last_not_of(" \t") + 1);
        key.erase(0, key.find_first_not_of(" \t"));
        if (key == "nodes") line = "nodes = " + std::to_string(d.nodes.size());
        else if (key == "elements") line = "elements = " + std::to_string(d.elements.size());
        else if (key == "adap") { line = any_adap ? "adap = 1" : "adap = 0"; have_adap = true; }
        else if (key == "iter" && iter >= 0) { line = "iter = " + std::to_string(iter); have_iter = true; }
        cf << line << "\n";
    }
    if (!have_adap && any_adap) cf << "adap = 1\n";
    if (!have_iter && iter >= 0) cf << "iter = " << iter << "\n";
    std::ifstream ain(src + "/aero");
    std::ofstream aout(dir + "/aero");
    aout << ain.rdbuf();
    // write_adap: adapted elements only, Hnm serialised column-major
    std::ofstream adf(dir + "/adap");
    adf.precision(17);
    for (const Element* e : d.elements) {
        if (!e->adap) continue;
        adf << e->number << " " << 1 << " " << e->refine_level << " double 2 " << nnod << " " << nnod;
        for (double v : e->Hnm.d) adf << " " << v;
        adf << "\n";
    }
    if (!nf || !ef || !cf || !adf) throw std::runtime_error("cannot write the refined deck into " + dir);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: cfdv_run <deck_dir> [--iters N] [--mode reference|fp64|mixed] "
                             "[--rtol X] [--precond none|nodeblock] [--out DIR] [--dump FILE] [--flatten FILE] "
                             "[--adap parity|3d] [--tecplot N] [--tecplot-only] [--storage rows|ebe] "
                             "[--save N] [--warm-start 0|1|2]\n");
        return 2;
    }
    std::string dir = argv[1], out = ".", dump, mode = "reference", flat, precond = "none", adap, storage = "rows";
    int tecplot = 0, save_every = 1000, warm = 0;
    bool tecplot_only = false;
    int iters = -1;
    double rtol = 1e-8;
    for (int i = 2; i < argc; i += 2) {
        std::string a = argv[i];
        if (a == "--tecplot-only") {
            tecplot_only = true;
            i--;
            continue;
        }
        if (i + 1 >= argc) {
            std::fprintf(stderr, "option %s needs a value\n", a.c_str());
            return 2;
        }
        std::string v = argv[i + 1];
        if (a == "--iters") iters = std::atoi(v.c_str());
        else if (a == "--mode") mode = v;
        else if (a == "--rtol") rtol = std::atof(v.c_str());
        else if (a == "--out") out = v;
        else if (a == "--dump") dump = v;
        else if (a == "--flatten") flat = v;
        else if (a == "--precond") precond = v;
        else if (a == "--adap") adap = v;
        else if (a == "--tecplot") tecplot = std::atoi(v.c_str());
        else if (a == "--storage") storage = v;
        else if (a == "--save") save_every = std::atoi(v.c_str());
        else if (a == "--warm-start") warm = std::atoi(v.c_str());
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    try {
        Deck d = load(dir);
        if (!flat.empty()) {
            cfdv::MeshArrays a = cfdv::flatten(d.elements, d.nodes, d.ndim, d.nnod, d.neqn);
            std::ofstream f(flat, std::ios::binary);
            f.write((const char*)a.conn.data(), (std::streamsize)(a.conn.size() * sizeof(int32_t)));
            f.write((const char*)a.bc.data(), (std::streamsize)(a.bc.size() * sizeof(int32_t)));
            f.write((const char*)a.x.data(), (std::streamsize)(a.x.size() * sizeof(double)));
            f.write((const char*)a.dirichlet.data(), (std::streamsize)a.dirichlet.size());
            for (const Node* n : d.nodes) f.write((const char*)n->U.d.data(), (std::streamsize)(d.neqn * sizeof(double)));
            return f ? 0 : 1;
        }
        if (tecplot_only) {
            const int it0 = d.control.count("iter") ? (int)d.control.at("iter") : 0;
            write_tecplot(d, out + "/tecplot_" + iter_name(it0) + ".plt", 0.0, nullptr);
            return 0;
        }
        const double cfl = d.control.at("cfl");
        const int restart = (int)d.control.at("gmresrestart");
        const int m = (int)d.control.at("gmresiter");
        int iter = (int)d.control.count("iter") ? (int)d.control.at("iter") : 0;
        const int itermax = iters >= 0 ? iter + iters : (int)d.control.at("itermax");
        const int md = mode == "fp64" ? CFDV_SOLVE_FP64 : mode == "mixed" ? CFDV_SOLVE_MIXED : CFDV_SOLVE_REFERENCE;
        if (storage != "rows" && storage != "ebe") throw std::runtime_error("--storage takes rows or ebe");
        if (warm < 0 || warm > 2) throw std::runtime_error("--warm-start takes 0, 1 or 2");
        // the FP64 and mixed solvers multiply over assembled node rows unless
        // --storage ebe; the reference solve always multiplies element by
        // element
        int blocks = CFDV_BLOCKS_FP64;
        if (md != CFDV_SOLVE_REFERENCE && storage == "rows") blocks |= CFDV_BLOCKS_ROWS;
        else if (md == CFDV_SOLVE_MIXED) blocks |= CFDV_BLOCKS_FP32;
        std::printf("nodes/eles: %zu %zu\n", d.nodes.size(), d.elements.size());
        cfdv::Context gpu(d.elements, d.nodes, d.ndim, d.nnod, d.neqn, d.th, blocks);
        if (!adap.empty()) {
            if (adap != "parity" && adap != "3d") throw std::runtime_error("--adap takes parity or 3d");
            cfdv::MeshArrays a = cfdv::flatten(d.elements, d.nodes, d.ndim, d.nnod, d.neqn);
            std::vector<double> U, U0;
            std::vector<int32_t> level;
            for (const Node* n : d.nodes)
                for (int q = 0; q < d.neqn; q++) {
                    U.push_back(n->U(q));
                    U0.push_back(n->U0(q));
                }
            for (const Element* e : d.elements) level.push_back(e->refine_level);
            cfdv::RefineResult res =
                cfdv::refine_pass(gpu, a, U, U0, level, adap == "3d" ? CFDV_REFINE_3D : CFDV_REFINE_PARITY);
            std::function<Face*(int, int)> make_face = [&d](int f, int bc) {
                auto fc = std::make_unique<Face>();
                fc->face = f;
                fc->bc = bc;
                d.face_store.push_back(std::move(fc));
                return d.face_store.back().get();
            };
            std::vector<Element*> els = d.elements;
            std::vector<Node*> nds = d.nodes;
            cfdv::apply_refinement<Element, Node, Face>(res, els, nds, make_face);
            for (size_t i = d.elements.size(); i < els.size(); i++) d.elem_store.emplace_back(els[i]);
            for (size_t i = d.nodes.size(); i < nds.size(); i++) d.node_store.emplace_back(nds[i]);
            d.elements = els;
            d.nodes = nds;
            std::printf("refined %zu elements in %zu levels -> %zu elements, %zu nodes\n", res.elelist.size(),
                        res.level_ptr.size() - 1, d.elements.size(), d.nodes.size());
            save_deck(d, out, dir);
            return 0;
        }
        cfdv::HipGMRES<Element, Node> gmres(gpu, restart, m, md, rtol);
        if (md != CFDV_SOLVE_REFERENCE) gmres.options().restarts = std::max(restart, 200);
        if (precond == "nodeblock") gmres.options().precond = CFDV_PRECOND_NODE_BLOCK;
        if (md != CFDV_SOLVE_REFERENCE) gmres.options().warm_start = warm;
        std::ofstream rms(out + "/RMSError.txt", std::ios::app);
        std::ofstream tfile(out + "/time.txt");
        tfile << "Iteration, Run_Time, Time_Elapsed, Sim_Time, Step_dt" << std::endl;
        auto checkpoint = [&](int it) {
            const std::string sd = out + "/" + std::to_string(it);
            ::mkdir(sd.c_str(), 0755);
            save_deck(d, sd, dir, false, it);
        };
        double t = 0, run_time = 0;
        bool first = true;
        for (; iter < itermax; iter++) {
            const auto t0 = std::chrono::steady_clock::now();
            if (save_every > 0 && iter % save_every == 0) checkpoint(iter);
            // after an update the device holds the nodes' U already; with a
            // warm start the upload is skipped, since cfdv_set_state forgets
            // the increments the next solve starts from
            if (first || warm == 0) gpu.set_state(d.nodes);
            first = false;
            const double dt = gpu.cfl(cfl);
            cfdv::FDVGalerkin fdv(gpu, dt);
            fdv();
            if (tecplot > 0 && iter % tecplot == 0) {   // Output (main.cpp:183)
                std::vector<double> sv(d.elements.size() * 4);
                cfdv::check(cfdv_get_blocks(gpu.handle(), nullptr, nullptr, sv.data()), gpu.handle(),
                            "cfdv_get_blocks");
                write_tecplot(d, out + "/tecplot_" + iter_name(iter) + ".plt", t, sv.data());
            }
            gmres.iterate(d.elements);
            gmres.update(d.nodes);
            // MPI_RMSError (MPI_RMSError.h:19-74), from the updated nodes
            std::vector<double> s(d.neqn, 0.0), s0(d.neqn, 0.0);
            for (const Node* n : d.nodes)
                for (int j = 0; j < d.neqn; j++) {
                    s[j] += n->U(j) * n->U(j);
                    s0[j] += n->U0(j) * n->U0(j);
                }
            rms << iter << "\t";
            for (int j = 0; j < d.neqn; j++) rms << (s0[j] == 0 ? 1.0 : std::sqrt(s[j] / s0[j]) - 1.0) << " ";
            rms << "\n";
            const cfdv_solve_stats& st = gmres.stats();
            std::printf(" t = %g, iteration %d dt: %g gmres iters %d rel resid %.3e\n", t, iter, dt, st.iters,
                        st.rel_resid);
            std::fflush(stdout);
            t += dt;
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            run_time += el;
            tfile << iter << " " << run_time << " " << el << " " << t << " " << dt << " " << std::endl;
        }
        if (save_every > 0) checkpoint(iter);
        if (!dump.empty()) {
            std::ofstream f(dump, std::ios::binary);
            for (const Node* n : d.nodes) f.write((const char*)n->U.d.data(), (std::streamsize)(d.neqn * sizeof(double)));
        }
    } catch (const cfdv::Error& e) {
        std::fprintf(stderr, "cfdv error %d: %s\n", e.code(), e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
