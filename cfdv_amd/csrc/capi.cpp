// Host runtime and C ABI of libcfdv_amd (see include/cfdv.h).
//
// Owns the device-resident mesh, state, element blocks and Krylov space, and
// drives the kernels of build.hip / solve.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/cfdv.h"
#include "cfdv_internal.h"

using namespace cfdv;

namespace {

constexpr int NBLK_RED = 1024;  // blocks of the reduction kernels (partials: NBLK_RED x 32 doubles)
constexpr int MAX_M1 = 501;     // Krylov dimension m <= 500: H holds (m+1) x m

// 2-point Gauss-Legendre by Newton iteration, the rule of
// GaussQuad::get_gauss (src/FiniteElement/GaussQuad.h:116-159).
void gauss2(double w[2], double z[2]) {
    const unsigned N = 2;
    const double eps = 3.E-14;
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    double zz = std::cos(3.141592654 * (1 - 0.25) / (N + 0.5));
    double z1 = zz + 10;
    while (std::fabs(zz - z1) > eps) {
        p1 = 1.0;
        p2 = 0.0;
        for (unsigned j = 1; j <= N; j++) {
            p3 = p2;
            p2 = p1;
            p1 = ((2.0 * j - 1.0) * zz * p2 - (j - 1.) * p3) / j;
        }
        p0 = N * (zz * p1 - p2) / (zz * zz - 1.0);
        z1 = zz;
        zz = z1 - p1 / p0;
    }
    z[0] = -zz;
    z[1] = zz;
    w[0] = 2.0 / ((1.0 - zz * zz) * p0 * p0);
    w[1] = w[0];
}

GaussTables make_tables() {
    GaussTables g;
    std::memset(&g, 0, sizeof(g));
    double w[2], z[2];
    gauss2(w, z);
    // 3D domain points, last index fastest (GaussQuad::three, :231-254)
    int gp = 0;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++, gp++) {
                const double zt[3] = {z[i], z[j], z[k]};
                g.w3[gp] = 1.0 * w[i] * w[j] * w[k];
                // testfunction3d / dtestfunction3d (TestFunctions3D.h:7-56)
                const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
                const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
                const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
                for (int n = 0; n < 8; n++) {
                    g.phi3[gp][n] = 0.125 * (1 + sx[n] * zt[0]) * (1 + sy[n] * zt[1]) * (1 + sz[n] * zt[2]);
                    g.dpxi3[gp][n][0] = 0.125 * sx[n] * (1 + sy[n] * zt[1]) * (1 + sz[n] * zt[2]);
                    g.dpxi3[gp][n][1] = 0.125 * sy[n] * (1 + sx[n] * zt[0]) * (1 + sz[n] * zt[2]);
                    g.dpxi3[gp][n][2] = 0.125 * sz[n] * (1 + sx[n] * zt[0]) * (1 + sy[n] * zt[1]);
                }
            }
    // face points (GaussQuad::two)
    gp = 0;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++, gp++) {
            g.wf[gp][0] = w[i];
            g.wf[gp][1] = w[j];
            g.zf[gp][0] = z[i];
            g.zf[gp][1] = z[j];
        }
    // 2D domain (testfunction / dtestfunction, TestFunctions2D.h:7-32)
    gp = 0;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++, gp++) {
            const double zt[2] = {z[i], z[j]};
            g.w2[gp] = 1.0 * w[i] * w[j];
            g.z2[gp][0] = zt[0];
            g.z2[gp][1] = zt[1];
            const double sx[4] = {-1, 1, 1, -1};
            const double sy[4] = {-1, -1, 1, 1};
            for (int n = 0; n < 4; n++) {
                g.phi2[gp][n] = 0.25 * (1 + sx[n] * zt[0]) * (1 + sy[n] * zt[1]);
                g.dpxi2[gp][n][0] = 0.25 * sx[n] * (1 + sy[n] * zt[1]);
                g.dpxi2[gp][n][1] = 0.25 * sy[n] * (1 + sx[n] * zt[0]);
            }
        }
    for (int i = 0; i < 2; i++) {
        g.we[i] = w[i];
        g.ze[i] = z[i];
    }
    return g;
}

const int FACE3[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {6, 7, 4, 5}};

template <typename T>
T* dmalloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    HIP_CHECK_THROW(hipMalloc(&p, n * sizeof(T)));
    return (T*)p;
}

}  // namespace

struct cfdv_ctx {
    std::string err;
    int ndim = 0, nnod = 0, neqn = 0, nface = 0, nbnod = 0, k = 0;
    long n_elem = 0, n_node = 0, n_pad = 0, ndof = 0;
    int ncolor = 0;
    std::vector<long> cstart;       // color ranges in slots, size ncolor+1
    std::vector<int> perm;          // slot -> original element (-1 padding)
    std::vector<int> perm_inv;      // original -> slot
    std::vector<int> color_orig;    // color per original element
    DevThermo th{};
    hipStream_t st = nullptr;
    int block_flags = 0;
    int r64_pairs = 0;   // FP64 element blocks in pair_offset order (contexts with rows)
    bool built = false, have_du = false;

    int *d_conn = nullptr, *d_bc = nullptr, *d_perm_inv = nullptr;
    unsigned long long* d_cmask = nullptr;
    int* d_bslots = nullptr;        // 3D slots with faces / constrained dofs
    long n_bslots = 0;
    int *d_bslots1 = nullptr, *d_bslotsm = nullptr;   // those with <= 1 face with a bc code, and more
    long n_bslots1 = 0, n_bslotsm = 0;
    unsigned char* d_adap = nullptr;  // hanging-node flags per slot (or none)
    double* d_hnm = nullptr;          // [n_pad][64] Hnm(N, M) at N + 8 M
    int* d_aslots = nullptr;
    long n_aslots = 0;
    double *d_x = nullptr, *d_U = nullptr, *d_U0 = nullptr, *d_area = nullptr, *d_cleng = nullptr;
    double *d_R64 = nullptr;
    float* d_R32 = nullptr;
    double *d_rhs = nullptr, *d_s = nullptr;
    int *d_node_ptr = nullptr, *d_node_ent = nullptr;
    double *d_b = nullptr, *d_du = nullptr, *d_r = nullptr, *d_w = nullptr, *d_y = nullptr;
    double* d_V = nullptr;
    int vcap = 0;
    float *d_V32 = nullptr, *d_w32 = nullptr;
    int vcap32 = 0;
    double *d_d64 = nullptr;
    // node-block Jacobi preconditioner (built lazily after each build)
    double *d_P64 = nullptr, *d_z64 = nullptr, *d_t64 = nullptr;
    float *d_P32 = nullptr, *d_z32 = nullptr;
    bool pinv_valid = false;
    // assembled node rows (CFDV_BLOCKS_ROWS), formed after each build
    RowsArgs ra{};
    long n_groups = 0, n_row_blocks = 0;
    long *d_goff = nullptr, *d_coff = nullptr, *d_diag = nullptr;
    int *d_cols = nullptr, *d_contrib = nullptr;
    double* d_B64 = nullptr;
    float* d_B32 = nullptr;
    bool rows_off = false;    // set while the reference solve runs (EBE only)
    bool rows_tr = false;     // the rows hold K^T (assembled for K^T x)
    // rows built in place (CFDV_BLOCKS_ROWS alone, hex8): the build adds
    // each element block into a staged copy of the FP64 rows, one color
    // after the other, and the rows (or those of K^T) are laid out from it
    bool rows_direct = false;
    long n_stage = 0;
    double* d_S64 = nullptr;                 // [n_stage][neqn^2] staged FP64 rows
    int* d_pslot = nullptr;                  // [n_pad][nnod^2] staged slot of each pair
    unsigned long long* d_pfirst = nullptr;  // [n_pad] pairs first into their slot
    int* d_mirror = nullptr;                 // [n_stage] slot of the mirrored pair
    double* d_Rb = nullptr;                  // [n_bslots][nnod^2][neqn^2] domain parts
    int* d_bidx = nullptr;                   // [n_pad] position in the face-kernel list
    std::vector<long> bstart, astart;        // per color: ranges of bslots, aslots
    // discrete adjoint: per-slot tangents [n_pad][41], host-given vectors
    double *d_gelem = nullptr, *d_lam = nullptr, *d_duv = nullptr, *d_gU = nullptr, *d_g = nullptr;
    // eight adjoint right-hand sides at once: FP32 basis of side-by-side
    // vectors [8 ndof] (vcap8 slots), FP64 correction and work, the
    // right-hand sides, solutions and residuals vector-major [8][ndof]
    float *d_V8 = nullptr, *d_z8 = nullptr;
    int vcap8 = 0;
    double *d_x8 = nullptr, *d_t8 = nullptr, *d_G8 = nullptr, *d_L8 = nullptr, *d_R8 = nullptr;
    double last_dt = 0.0;
    // warm start: d_du keeps the last update's increment, built with prev_dt;
    // with warm_start 2 or 3, d_old[k] the k+1-th before it (old_dt[k];
    // n_old of them valid), d_dut is scratch
    void* d_flag_scratch = nullptr;   // this device's flag scratch (cfdv_refine_flags_noalloc)
    std::string flag_scratch_err;
    bool have_prev = false;
    bool warm_consumed = false;   // a solve of the current build formed the warm start
    bool warm_had = false;        // ... and it is kept in d_x0
    double* d_x0 = nullptr;
    int n_old = 0, n_saved = 0;
    double prev_dt = 0.0, old_dt[2] = {0.0, 0.0};
    double *d_old[2] = {nullptr, nullptr}, *d_dut = nullptr;
    double* d_r0 = nullptr;
    double *d_part = nullptr, *d_small = nullptr, *d_hcol = nullptr, *d_cflpart = nullptr;
    // GMRES least squares on the device: Hessenberg H ((m+1) x m, m <= 500),
    // then g, cs, sn; a copy of H for the reference mode's rotations; the
    // gate of the Arnoldi step kernels ([0]) and the reference breakdown
    // flag ([1]); per-step records in host-visible pinned memory (h_steps,
    // device address d_steps) and two events that mark the steps
    double *d_hess = nullptr, *d_hessr = nullptr;
    int* d_gate = nullptr;
    GmresStep *h_steps = nullptr, *d_steps = nullptr;
    hipEvent_t ev_step[2] = {nullptr, nullptr};
    // the last reference-mode solve, per restart: H before the rotations
    // ((m+1) x m, column-major), the rotated g (m+1) and the Arnoldi steps
    int trace_m = 0;
    std::vector<double> trace_H, trace_g;
    std::vector<int> trace_mm;
    int *d_counters = nullptr;     // [0] ties, [1] status, [2] singular node block
    hipEvent_t ev[8];
    std::vector<void*> allocs;
    // matvec timing: event pairs recorded around every full matvec while a
    // step runs; summed once the step has synchronised
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, int>> mv_marks;   // (event index, precision)
    int ev_used = 0;

    ~cfdv_ctx() {
        for (auto& e : evpool) (void)hipEventDestroy(e);
        for (void* p : allocs) (void)hipFree(p);
        if (h_steps) (void)hipHostFree(h_steps);
        for (auto& e : ev_step)
            if (e) (void)hipEventDestroy(e);
        if (st) {
            for (auto& e : ev) (void)hipEventDestroy(e);
            if (ev_sync) (void)hipEventDestroy(ev_sync);
            (void)hipStreamDestroy(st);
        }
    }
    template <typename T>
    T* alloc(size_t n) {
        T* p = dmalloc<T>(n);
        allocs.push_back(p);
        return p;
    }
    // frees a buffer from alloc() before the context ends
    void release(void* p) {
        if (!p) return;
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree(p);
    }
    // wait for the stream by polling an event: the solvers wait once per
    // Arnoldi step for a few scalars, and a blocking wait adds its wake-up
    // latency to every step
    hipEvent_t ev_sync = nullptr;
    void sync() {
        HIP_CHECK_THROW(hipEventRecord(ev_sync, st));
        wait_event(ev_sync);
    }
    static void wait_event(hipEvent_t e) {
        hipError_t r;
        while ((r = hipEventQuery(e)) == hipErrorNotReady) {
        }
        HIP_CHECK_THROW(r);
    }
    void check_launch() { HIP_CHECK_THROW(hipGetLastError()); }

    // products K x and K^T x use the assembled rows when the context holds
    // them (not inside the reference solve)
    bool use_rows(int prec, bool tr) const {
        (void)tr;
        if (rows_off) return false;
        return prec == 64 ? d_B64 != nullptr : d_B32 != nullptr;
    }
    // the rows hold K (tr false) or K^T (tr true): re-sum them from the FP64
    // element blocks when the orientation changes
    void orient_rows(bool tr) {
        if (!d_B64 || rows_tr == tr) return;
        if (rows_direct) launch_rows_from_stage(ra, n_stage, d_S64, d_mirror, tr, d_B64, d_B32, st);
        else launch_rows_assemble(ra, d_R64, tr, d_B64, d_B32, st);
        check_launch();
        rows_tr = tr;
    }
    // blocks available for a product of this precision
    bool have_blocks(int prec, bool tr) const {
        if (use_rows(prec, tr)) return true;
        return prec == 64 ? d_R64 != nullptr : d_R32 != nullptr;
    }
    // ---- matvec y = K x (or K^T x): assembled rows, or EBE over all colors ----
    hipEvent_t next_event() {
        if (ev_used == (int)evpool.size()) {
            hipEvent_t e;
            HIP_CHECK_THROW(hipEventCreate(&e));
            evpool.push_back(e);
        }
        return evpool[ev_used++];
    }
    // gate: a device flag that turns the product into a no-op when set
    // (the Arnoldi step enqueued after the last one of a GMRES cycle)
    void matvec(const void* x, void* y, int prec, bool tr, const int* gate = nullptr) {
        if (use_rows(prec, tr)) orient_rows(tr);
        int mark = -1;
        if (timing) {
            mark = ev_used;
            HIP_CHECK_THROW(hipEventRecord(next_event(), st));
        }
        if (use_rows(prec, tr)) {
            launch_rows_matvec(ra, d_B64, d_B32, x, y, prec, st, gate);
        } else {
            if (prec == 32) vec_fill_f((float*)y, 0.0f, ndof, st, gate);
            else vec_fill((double*)y, 0.0, ndof, st, gate);
            for (int c = 0; c < ncolor; c++) {
                MatvecArgs a{ndim, nnod, neqn, cstart[c], cstart[c + 1], d_conn, r64_pairs};
                launch_matvec_color(a, d_R64, d_R32, x, y, prec, tr, st, gate);
            }
        }
        check_launch();
        if (timing) {
            HIP_CHECK_THROW(hipEventRecord(next_event(), st));
            mv_marks.push_back({mark, prec});
        }
    }
    // forget the last product's timing marks (a gated product that did nothing)
    void drop_last_matvec_mark() {
        if (timing && !mv_marks.empty()) mv_marks.pop_back();
    }
    // sum of the recorded matvec intervals (ms) per precision; call after a sync
    void collect_matvec_times(double* ms64, double* ms32, int* n64, int* n32) {
        *ms64 = *ms32 = 0;
        *n64 = *n32 = 0;
        for (auto& m : mv_marks) {
            float t = 0;
            HIP_CHECK_THROW(hipEventElapsedTime(&t, evpool[m.first], evpool[m.first + 1]));
            if (m.second == 32) { *ms32 += t; (*n32)++; }
            else { *ms64 += t; (*n64)++; }
        }
        mv_marks.clear();
        ev_used = 0;
    }
    // dot products: out[j] = <V_j, w> (device), optionally copied to host
    template <typename T>
    void multidot(const T* V, int nv, const T* w, double* out_dev) {
        vec_multidot<T>(V, ndof, nv, w, ndof, d_part, out_dev, NBLK_RED, st);
    }
    double norm(const double* v) {
        vec_multidot<double>(v, ndof, 1, v, ndof, d_part, d_small, NBLK_RED, st);
        double h;
        HIP_CHECK_THROW(hipMemcpyAsync(&h, d_small, sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
        return std::sqrt(h);
    }
    // Krylov basis stride: ndof rounded up to 64 elements, so that every
    // basis vector starts 256-byte aligned for 16-byte vector loads
    long ldv() const { return (ndof + 63) & ~63L; }
    void ensure_krylov(int m) {
        if (vcap >= m + 1) return;
        if (d_V) {
            auto it = std::find(allocs.begin(), allocs.end(), (void*)d_V);
            if (it != allocs.end()) allocs.erase(it);
            (void)hipFree(d_V);
        }
        d_V = alloc<double>((size_t)(m + 1) * ldv());
        vcap = m + 1;
    }
    void ensure_krylov32(int m) {
        if (vcap32 >= m + 1) return;
        if (d_V32) {
            auto it = std::find(allocs.begin(), allocs.end(), (void*)d_V32);
            if (it != allocs.end()) allocs.erase(it);
            (void)hipFree(d_V32);
        }
        d_V32 = alloc<float>((size_t)(m + 1) * ldv());
        vcap32 = m + 1;
    }
    // internal node order (contexts with node rows): internal node i is the
    // caller's node norder[i] (d_norder), caller node j is internal node
    // d_npos[j]; node arrays cross the C ABI in the caller's order
    bool renum = false;
    int *d_norder = nullptr, *d_npos = nullptr;
    double* d_ptmp = nullptr;    // [ndof] staging of a permuted node array
    float* d_ptmp32 = nullptr;   // [8 ndof] staging of a permuted 8-vector array
    // host node array with q values per node into device memory (internal order)
    void upload_nodes(double* dev, const double* host, int q) {
        const size_t n = (size_t)n_node * q;
        if (!renum) {
            HIP_CHECK_THROW(hipMemcpyAsync(dev, host, sizeof(double) * n, hipMemcpyHostToDevice, st));
            return;
        }
        if (!d_ptmp) d_ptmp = alloc<double>((size_t)ndof);
        if (n > (size_t)ndof) throw std::invalid_argument("node array wider than the staging buffer");
        HIP_CHECK_THROW(hipMemcpyAsync(d_ptmp, host, sizeof(double) * n, hipMemcpyHostToDevice, st));
        vec_gather_nodes<double>(dev, d_ptmp, d_norder, n_node, q, st);
        check_launch();
    }
    void download_nodes(double* host, const double* dev, int q) {
        const size_t n = (size_t)n_node * q;
        if (!renum) {
            HIP_CHECK_THROW(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, st));
            return;
        }
        if (!d_ptmp) d_ptmp = alloc<double>((size_t)ndof);
        if (n > (size_t)ndof) throw std::invalid_argument("node array wider than the staging buffer");
        vec_gather_nodes<double>(d_ptmp, dev, d_npos, n_node, q, st);
        check_launch();
        HIP_CHECK_THROW(hipMemcpyAsync(host, d_ptmp, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    }
    void upload_nodes32(float* dev, const float* host, int q) {
        const size_t n = (size_t)n_node * q;
        if (!renum) {
            HIP_CHECK_THROW(hipMemcpyAsync(dev, host, sizeof(float) * n, hipMemcpyHostToDevice, st));
            return;
        }
        if (!d_ptmp32) d_ptmp32 = alloc<float>((size_t)8 * ndof);
        if (n > (size_t)8 * ndof) throw std::invalid_argument("node array wider than the staging buffer");
        HIP_CHECK_THROW(hipMemcpyAsync(d_ptmp32, host, sizeof(float) * n, hipMemcpyHostToDevice, st));
        vec_gather_nodes<float>(dev, d_ptmp32, d_norder, n_node, q, st);
        check_launch();
    }
    void download_nodes32(float* host, const float* dev, int q) {
        const size_t n = (size_t)n_node * q;
        if (!renum) {
            HIP_CHECK_THROW(hipMemcpyAsync(host, dev, sizeof(float) * n, hipMemcpyDeviceToHost, st));
            return;
        }
        if (!d_ptmp32) d_ptmp32 = alloc<float>((size_t)8 * ndof);
        if (n > (size_t)8 * ndof) throw std::invalid_argument("node array wider than the staging buffer");
        vec_gather_nodes<float>(d_ptmp32, dev, d_npos, n_node, q, st);
        check_launch();
        HIP_CHECK_THROW(hipMemcpyAsync(host, d_ptmp32, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    }
    long ldv8() const { return (8 * ndof + 63) & ~63L; }
    void ensure_multi(int m) {
        if (!d_x8) {
            d_x8 = alloc<double>((size_t)8 * ndof);
            d_t8 = alloc<double>((size_t)8 * ndof);
            d_z8 = alloc<float>((size_t)8 * ndof);
            d_G8 = alloc<double>((size_t)8 * ndof);
            d_L8 = alloc<double>((size_t)8 * ndof);
            d_R8 = alloc<double>((size_t)8 * ndof);
        }
        if (vcap8 >= m + 1) return;
        if (d_V8) {
            auto it = std::find(allocs.begin(), allocs.end(), (void*)d_V8);
            if (it != allocs.end()) allocs.erase(it);
            (void)hipFree(d_V8);
        }
        d_V8 = alloc<float>((size_t)(m + 1) * ldv8());
        vcap8 = m + 1;
    }
    // node-block inverses from the current blocks (FP64 if held, else FP32)
    void ensure_precond() {
        if (!d_P64) {
            // chunks of 64 nodes (the last one padded)
            const size_t np = (size_t)((n_node + 63) / 64) * 64;
            d_P64 = alloc<double>(np * neqn * neqn);
            d_P32 = alloc<float>(np * neqn * neqn);
            d_z64 = alloc<double>((size_t)ndof);
            d_z32 = alloc<float>((size_t)ndof);
            d_t64 = alloc<double>((size_t)ndof);
        }
        if (pinv_valid) return;
        HIP_CHECK_THROW(hipMemsetAsync(d_counters + 2, 0, sizeof(int), st));
        if (d_B64) launch_nodeblock_rows(ra, d_diag, d_B64, rows_tr, d_P64, d_P32, d_counters + 2, st);
        else launch_nodeblock(d_node_ptr, d_node_ent, d_R64, d_R32, nnod, neqn, n_node, d_P64, d_P32, d_counters + 2, st);
        check_launch();
        int singular = 0;
        HIP_CHECK_THROW(hipMemcpyAsync(&singular, d_counters + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        sync();
        if (singular) throw std::runtime_error("singular node block in the node-block Jacobi preconditioner");
        pinv_valid = true;
    }
};

namespace {

int fail(cfdv_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

DevThermo make_thermo(const cfdv_thermo* t) {
    // Thermo::set (src/FDV/Thermo.h:12-33)
    DevThermo th{};
    th.gamma = t->gamma;
    th.gamm1 = (t->gamma - 1.0);
    th.gm1d2 = ((t->gamma - 1.0) / 2.0);
    th.gamm3 = (t->gamma - 3.0);
    th.gm3d2 = ((t->gamma - 3.0) / 2.0);
    th.Pr = t->prandtl;
    th.csuth = 110.0 / t->tinf;
    th.cmach = t->mach;
    th.Cv = 1. / t->gamma / (t->gamma - 1.0) / (t->mach * t->mach);
    th.cgas = 1.0 / t->gamma / (t->mach * t->mach);
    th.creyn = t->reynolds;
    th.Twall = t->twall;
    th.adiabatic = t->adiabatic;
    return th;
}

// Structure of the assembled node rows (see rows.hip): per node the
// neighbour nodes (nodes sharing an element, itself included) as slots, most
// contributing elements first, chunks of 64 rows with groups of 4 slots, and
// per slot the element blocks that sum into it, listed in element order.
// ptr/elem: node -> element CSR.
struct RowsHost {
    std::vector<long> goff, coff, diag;
    std::vector<int> cols, contrib;
    long n_blocks = 0;   // node pairs sharing an element (the real blocks)
    // rows built in place (contexts without element blocks): per element
    // slot and node pair the staged row slot it adds into (-1: none), per
    // element slot the pairs that come first into their row slot (element
    // slots ascend with the color, so the first is the lowest), and per
    // staged row slot the slot of the mirrored pair (-1: padding)
    std::vector<int> pslot, mirror;
    std::vector<unsigned long long> pfirst;
};

RowsHost rows_structure(int nn, long ne, long nnode, long n_pad, const int32_t* conn, const std::vector<int>& perm_inv,
                        const std::vector<int>& ptr, const std::vector<int>& elem, bool direct) {
    const int nn2 = nn * nn;
    if ((double)n_pad * nn2 >= 2147483647.0) throw std::invalid_argument("mesh too large for assembled rows");
    // neighbour lists
    std::vector<long> aptr(nnode + 1, 0);
    std::vector<int> adj;
    adj.reserve((size_t)nnode * (nn == 8 ? 27 : 9));
    std::vector<int> tmp;
    for (long n = 0; n < nnode; n++) {
        tmp.clear();
        for (int p = ptr[n]; p < ptr[n + 1]; p++) {
            const long e = elem[p];
            for (int l = 0; l < nn; l++) tmp.push_back(conn[e * nn + l]);
        }
        if (tmp.empty()) tmp.push_back((int)n);   // a node of no element: a zero row
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        adj.insert(adj.end(), tmp.begin(), tmp.end());
        aptr[n + 1] = (long)adj.size();
    }
    // contributions per adjacency entry (CSR over adj), element order
    std::vector<long> eptr(adj.size() + 1, 0);
    auto slot_of = [&](long row, int col) -> long {
        const int* b = adj.data() + aptr[row];
        const int* e = adj.data() + aptr[row + 1];
        const int* p = std::lower_bound(b, e, col);
        return aptr[row] + (p - b);
    };
    for (long e = 0; e < ne; e++)
        for (int N = 0; N < nn; N++)
            for (int M = 0; M < nn; M++)
                eptr[slot_of(conn[e * nn + N], conn[e * nn + M]) + 1]++;
    for (size_t i = 0; i < adj.size(); i++) eptr[i + 1] += eptr[i];
    std::vector<int> codes((size_t)eptr[adj.size()]);
    {
        std::vector<long> pos(eptr.begin(), eptr.end() - 1);
        for (long e = 0; e < ne; e++) {
            const long sl = perm_inv[e];
            for (int N = 0; N < nn; N++)
                for (int M = 0; M < nn; M++) {
                    const long a = slot_of(conn[e * nn + N], conn[e * nn + M]);
                    codes[pos[a]++] = (int)(sl * nn2 + N * nn + M);
                }
        }
    }
    // slot order within a row: most contributors first (ties by column), so
    // that slot k of the 64 rows of a chunk tends to have the same number of
    // contributors where refinement makes the rows irregular (the gather
    // then pads less); sperm[aptr[r] + k] is the adjacency entry of slot k
    std::vector<long> sperm(adj.size());
    for (long r = 0; r < nnode; r++) {
        const long b = aptr[r], e = aptr[r + 1];
        for (long a = b; a < e; a++) sperm[a] = a;
        std::stable_sort(sperm.begin() + b, sperm.begin() + e, [&](long x, long y) {
            return eptr[x + 1] - eptr[x] > eptr[y + 1] - eptr[y];
        });
    }
    // chunks and groups
    const long nch = (nnode + 63) / 64;
    RowsHost h;
    h.n_blocks = (long)adj.size();
    std::vector<long>& goff = h.goff;
    goff.assign(nch + 1, 0);
    for (long ch = 0; ch < nch; ch++) {
        long kmax = 0;
        for (long r = ch * 64; r < std::min(nnode, ch * 64 + 64); r++) kmax = std::max(kmax, aptr[r + 1] - aptr[r]);
        goff[ch + 1] = goff[ch] + (kmax + 3) / 4;
    }
    const long ng = goff[nch];
    std::vector<int>& cols = h.cols;
    std::vector<long>& coff = h.coff;
    cols.assign((size_t)ng * 4 * 64, 0);
    coff.assign((size_t)ng * 4 + 1, 0);
    for (long ch = 0; ch < nch; ch++)
        for (long k = goff[ch] * 4; k < goff[ch + 1] * 4; k++) {
            const long kl = k - goff[ch] * 4;
            long cmax = 0;
            for (int lane = 0; lane < 64; lane++) {
                const long r = ch * 64 + lane;
                if (r >= nnode) continue;
                const long deg = aptr[r + 1] - aptr[r];
                if (kl < deg) {
                    const long a = sperm[aptr[r] + kl];
                    cols[(size_t)k * 64 + lane] = adj[a];
                    cmax = std::max(cmax, eptr[a + 1] - eptr[a]);
                } else {
                    cols[(size_t)k * 64 + lane] = (int)r;
                }
            }
            coff[k + 1] = coff[k] + cmax;
        }
    // global slot of each row's own column (its diagonal block)
    h.diag.assign(nnode, 0);
    for (long r = 0; r < nnode; r++) {
        long kl = 0;
        while (adj[sperm[aptr[r] + kl]] != (int)r) kl++;
        h.diag[r] = goff[r / 64] * 4 + kl;
    }
    if (direct) {
        // staged row slot of row r's slot kl: (4 goff[r / 64] + kl) * 64 + r % 64
        const long nstage = ng * 4 * 64;
        std::vector<int> kl_of(adj.size());
        for (long r = 0; r < nnode; r++)
            for (long kl = 0; kl < aptr[r + 1] - aptr[r]; kl++) kl_of[sperm[aptr[r] + kl]] = (int)kl;
        h.pslot.assign((size_t)n_pad * nn2, -1);
        h.pfirst.assign((size_t)n_pad, 0ULL);
        h.mirror.assign((size_t)nstage, -1);
        for (long r = 0; r < nnode; r++)
            for (long kl = 0; kl < aptr[r + 1] - aptr[r]; kl++) {
                const long a = sperm[aptr[r] + kl];
                const long stg = (goff[r / 64] * 4 + kl) * 64 + r % 64;
                long best = -1;
                for (long i = eptr[a]; i < eptr[a + 1]; i++) {
                    h.pslot[(size_t)codes[i]] = (int)stg;
                    if (best < 0 || codes[i] / nn2 < codes[best] / nn2) best = i;
                }
                if (best >= 0) h.pfirst[(size_t)(codes[best] / nn2)] |= 1ULL << (codes[best] % nn2);
                const int col = adj[a];
                const long a2 = slot_of(col, (int)r);
                h.mirror[(size_t)stg] = (int)((goff[col / 64] * 4 + kl_of[a2]) * 64 + col % 64);
            }
        for (long e = 0; e < ne; e++)
            for (int q = 0; q < nn2; q++) {
                const int v = h.pslot[(size_t)perm_inv[e] * nn2 + q];
                if (v < 0 || v >= nstage) throw std::runtime_error("row structure: element pair without a row slot");
            }
        return h;   // the rows are not summed from element blocks: no lists
    }
    std::vector<int>& contrib = h.contrib;
    contrib.assign((size_t)coff[ng * 4] * 64, -1);
    for (long ch = 0; ch < nch; ch++)
        for (long k = goff[ch] * 4; k < goff[ch + 1] * 4; k++) {
            const long kl = k - goff[ch] * 4;
            for (int lane = 0; lane < 64; lane++) {
                const long r = ch * 64 + lane;
                if (r >= nnode || kl >= aptr[r + 1] - aptr[r]) continue;
                const long a = sperm[aptr[r] + kl];
                for (long i = eptr[a]; i < eptr[a + 1]; i++)
                    contrib[(size_t)(coff[k] + (i - eptr[a])) * 64 + lane] = codes[i];
            }
        }
    return h;
}

// The internal node order of contexts with node rows.  A candidate order:
// the nodes along a Morton curve of their coordinates (each axis quantized
// by rank to 2^21 levels, so that stretched directions keep their
// resolution), then within each window of 1024 nodes the ones with most
// neighbours first, so that the 64 rows of a chunk have about as many slots
// each.  It replaces the caller's order when it shrinks the vector footprint
// of the rows that run at the same time: for windows of 128 chunks (8192
// rows, what one XCD's 32 CUs hold at four waves each), the number of
// distinct 64-node blocks of x that the rows' columns touch, summed over
// every 8th window, by 5% or more.  A structured grid in its own x-fastest
// numbering keeps it (2.1M hex8 ramp: 408 blocks per window against 446;
// the FP32 product 0.72 of HBM in its order, 0.69 in the Morton one); the
// 2-level refined ramp, whose refined nodes are numbered in the order the
// refinement appends them, is renumbered (698 against 486 blocks; FP32
// product 0.65 -> 0.75 of HBM, assembly 25.5 -> 22.9 ms).  Returns
// order[i] = caller node of internal node i, or an empty vector to keep the
// caller's order.
std::vector<int> choose_node_order(int ndim, int nn, long nnode, long ne, const double* x, const int32_t* conn) {
    const int bits = 21;
    std::vector<uint64_t> key((size_t)nnode, 0);
    std::vector<int> idx((size_t)nnode);
    for (int a = 0; a < ndim; a++) {
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int p, int q) {
            return x[(size_t)p * ndim + a] < x[(size_t)q * ndim + a];
        });
        for (long r = 0; r < nnode; r++) {
            const uint64_t qv = (uint64_t)((double)r / (double)nnode * (double)(1u << bits));
            uint64_t k = 0;
            for (int b = 0; b < bits; b++) k |= ((qv >> b) & 1ULL) << (ndim * b + a);
            key[idx[r]] |= k;
        }
    }
    std::vector<int> order((size_t)nnode);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return key[p] < key[q]; });
    // node -> element lists and the neighbours of a node (sorted, unique)
    std::vector<int> ptr((size_t)nnode + 1, 0);
    for (long i = 0; i < ne * nn; i++) ptr[conn[i] + 1]++;
    for (long i = 0; i < nnode; i++) ptr[i + 1] += ptr[i];
    std::vector<int> elem((size_t)ne * nn);
    {
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (long e = 0; e < ne; e++)
            for (int l = 0; l < nn; l++) elem[pos[conn[e * nn + l]]++] = (int)e;
    }
    std::vector<int> tmp;
    auto neighbours = [&](long n) {
        tmp.clear();
        for (int p = ptr[n]; p < ptr[n + 1]; p++)
            for (int l = 0; l < nn; l++) tmp.push_back(conn[(long)elem[p] * nn + l]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    };
    std::vector<int> deg((size_t)nnode, 1);
    for (long n = 0; n < nnode; n++) {
        neighbours(n);
        deg[n] = std::max<int>(1, (int)tmp.size());
    }
    for (long w0 = 0; w0 < nnode; w0 += 1024) {
        const long w1 = std::min(nnode, w0 + 1024);
        std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int p, int q) { return deg[p] > deg[q]; });
    }
    std::vector<int> pos((size_t)nnode);
    for (long i = 0; i < nnode; i++) pos[order[i]] = (int)i;
    // footprint of both orders: distinct 64-node blocks per window of 8192 rows
    const long W = 8192;
    std::vector<int> blk;
    long fc = 0, fo = 0;
    for (long w0 = 0; w0 < nnode; w0 += 8 * W) {
        const long w1 = std::min(nnode, w0 + W);
        for (int which = 0; which < 2; which++) {
            blk.clear();
            for (long r = w0; r < w1; r++) {
                const long n = which ? order[r] : r;
                neighbours(n);
                for (int c : tmp) blk.push_back((which ? pos[c] : c) >> 6);
            }
            std::sort(blk.begin(), blk.end());
            const long u = std::unique(blk.begin(), blk.end()) - blk.begin();
            (which ? fo : fc) += u;
        }
    }
    if (!((double)fo < 0.95 * (double)fc)) return {};
    return order;
}

void setup_rows(cfdv_ctx* c, const cfdv_mesh* m, const std::vector<int>& ptr, const std::vector<int>& elem) {
    RowsHost h = rows_structure(c->nnod, c->n_elem, c->n_node, c->n_pad, m->conn, c->perm_inv, ptr, elem,
                                c->rows_direct);
    const long nch = (long)h.goff.size() - 1, ng = h.goff[nch], nnode = c->n_node;
    const auto &goff = h.goff, &coff = h.coff;
    const auto &cols = h.cols, &contrib = h.contrib;
    c->n_groups = ng;
    c->n_row_blocks = h.n_blocks;
    c->d_goff = c->alloc<long>(nch + 1);
    c->d_cols = c->alloc<int>(cols.size());
    c->d_diag = c->alloc<long>(h.diag.size());
    HIP_CHECK_THROW(hipMemcpy(c->d_diag, h.diag.data(), sizeof(long) * h.diag.size(), hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_goff, goff.data(), sizeof(long) * goff.size(), hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_cols, cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice));
    if (!c->rows_direct) {
        c->d_coff = c->alloc<long>(coff.size());
        c->d_contrib = c->alloc<int>(contrib.size());
        HIP_CHECK_THROW(hipMemcpy(c->d_coff, coff.data(), sizeof(long) * coff.size(), hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(c->d_contrib, contrib.data(), sizeof(int) * contrib.size(), hipMemcpyHostToDevice));
    }
    const size_t nval = (size_t)ng * 4 * c->neqn * c->neqn * 64;
    c->d_B64 = c->alloc<double>(nval);
    c->d_B32 = c->alloc<float>(nval);
    if (c->rows_direct) {
        const size_t nstage = (size_t)ng * 4 * 64;
        c->n_stage = (long)nstage;
        c->d_S64 = c->alloc<double>(nstage * c->neqn * c->neqn);
        // row slots no element adds into (padding, rows of no element) stay zero
        HIP_CHECK_THROW(hipMemset(c->d_S64, 0, sizeof(double) * nstage * c->neqn * c->neqn));
        c->d_pslot = c->alloc<int>(h.pslot.size());
        c->d_pfirst = c->alloc<unsigned long long>(h.pfirst.size());
        c->d_mirror = c->alloc<int>(h.mirror.size());
        HIP_CHECK_THROW(hipMemcpy(c->d_pslot, h.pslot.data(), sizeof(int) * h.pslot.size(), hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(c->d_pfirst, h.pfirst.data(), sizeof(unsigned long long) * h.pfirst.size(),
                                  hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(c->d_mirror, h.mirror.data(), sizeof(int) * h.mirror.size(), hipMemcpyHostToDevice));
    }
    c->ra = RowsArgs{nnode, nch, c->nnod, c->neqn, c->d_goff, c->d_cols, c->d_coff, c->d_contrib, c->r64_pairs};
}

void setup(cfdv_ctx* c, const cfdv_mesh* m) {
    const int nn = c->nnod, nf = c->nface, nq = c->neqn;
    const long ne = c->n_elem, nnode = c->n_node;
    for (long i = 0; i < ne * nn; i++)
        if (m->conn[i] < 0 || m->conn[i] >= nnode) throw std::invalid_argument("connectivity out of range");
    // node -> element CSR (element order), used by coloring and compose
    std::vector<int> ptr(nnode + 1, 0), ent((size_t)ne * nn), elem((size_t)ne * nn);
    for (long i = 0; i < ne * nn; i++) ptr[m->conn[i] + 1]++;
    for (long i = 0; i < nnode; i++) ptr[i + 1] += ptr[i];
    {
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (long e = 0; e < ne; e++)
            for (int l = 0; l < nn; l++) {
                const int g = m->conn[e * nn + l];
                elem[pos[g]] = (int)e;
                ent[pos[g]++] = (int)(e * nn + l);   // original ids, fixed below
            }
    }
    // ---- coloring on the device (Jones-Plassmann == greedy in priority order) ----
    std::vector<unsigned long long> prio(ne);
    // priority = element index: the coloring equals sequential first-fit
    // greedy in element order (8 colors on a structured hex mesh); rounds
    // grow with the mesh extent (~nx + 2ny + 4nz), a one-time setup cost
    for (long e = 0; e < ne; e++) prio[e] = (unsigned long long)e;
    {
        int* d_conn_o = dmalloc<int>((size_t)ne * nn);
        int* d_ptr = dmalloc<int>(nnode + 1);
        int* d_elem = dmalloc<int>((size_t)ne * nn);
        unsigned long long* d_prio = dmalloc<unsigned long long>(ne);
        int* d_c0 = dmalloc<int>(ne);
        int* d_c1 = dmalloc<int>(ne);
        int* d_chg = dmalloc<int>(1);
        HIP_CHECK_THROW(hipMemcpy(d_conn_o, m->conn, sizeof(int) * ne * nn, hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(d_ptr, ptr.data(), sizeof(int) * (nnode + 1), hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(d_elem, elem.data(), sizeof(int) * ne * nn, hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(d_prio, prio.data(), sizeof(unsigned long long) * ne, hipMemcpyHostToDevice));
        std::vector<int> init(ne, -1);
        HIP_CHECK_THROW(hipMemcpy(d_c0, init.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
        int rounds = 0;
        for (;;) {
            int zero = 0, chg = 0;
            HIP_CHECK_THROW(hipMemcpyAsync(d_chg, &zero, sizeof(int), hipMemcpyHostToDevice, c->st));
            launch_color_round(d_conn_o, nn, d_ptr, d_elem, ne, d_prio, d_c0, d_c1, d_chg, c->st);
            HIP_CHECK_THROW(hipGetLastError());
            HIP_CHECK_THROW(hipMemcpyAsync(&chg, d_chg, sizeof(int), hipMemcpyDeviceToHost, c->st));
            c->sync();
            std::swap(d_c0, d_c1);
            if (!chg) break;
            if (++rounds > 100000) throw std::runtime_error("coloring did not converge");
        }
        c->color_orig.resize(ne);
        HIP_CHECK_THROW(hipMemcpy(c->color_orig.data(), d_c0, sizeof(int) * ne, hipMemcpyDeviceToHost));
        for (void* p : {(void*)d_conn_o, (void*)d_ptr, (void*)d_elem, (void*)d_prio, (void*)d_c0, (void*)d_c1,
                        (void*)d_chg})
            (void)hipFree(p);
    }
    int ncol = 0;
    for (long e = 0; e < ne; e++) ncol = std::max(ncol, c->color_orig[e] + 1);
    c->ncolor = ncol;
    // ---- slots: elements sorted by (color, id), each color padded to 64 ----
    std::vector<long> cnt(ncol, 0);
    for (long e = 0; e < ne; e++) cnt[c->color_orig[e]]++;
    c->cstart.assign(ncol + 1, 0);
    for (int k = 0; k < ncol; k++) c->cstart[k + 1] = c->cstart[k] + (cnt[k] + 63) / 64 * 64;
    c->n_pad = c->cstart[ncol];
    c->perm.assign(c->n_pad, -1);
    c->perm_inv.assign(ne, -1);
    {
        // within a color, elements by their lowest node number (ties by id):
        // the elements around a run of consecutive nodes -- one chunk of node
        // rows -- then sit together in slot order, so that the assembly's
        // block reads share cache lines and neighbours build on one XCD.
        // Every sum over elements (rhs, rows, colored products) runs in
        // element or color order, so the slot order changes no result.
        std::vector<int> key(ne), ord(ne);
        for (long e = 0; e < ne; e++) {
            int k = m->conn[(long)e * nn];
            for (int l = 1; l < nn; l++) k = std::min(k, m->conn[(long)e * nn + l]);
            key[e] = k;
            ord[e] = (int)e;
        }
        std::sort(ord.begin(), ord.end(), [&](int a, int b) {
            const int ca = c->color_orig[a], cb = c->color_orig[b];
            if (ca != cb) return ca < cb;
            return key[a] != key[b] ? key[a] < key[b] : a < b;
        });
        std::vector<long> pos(c->cstart.begin(), c->cstart.end() - 1);
        for (const int e : ord) {
            const long s = pos[c->color_orig[e]]++;
            c->perm[s] = e;
            c->perm_inv[e] = (int)s;
        }
    }
    // ---- sorted element arrays ----
    const long np = c->n_pad;
    std::vector<int> conn_s((size_t)np * nn, -1), bc_s((size_t)np * nf, 0);
    std::vector<unsigned long long> cmask(np, 0);
    for (long s = 0; s < np; s++) {
        const int e = c->perm[s];
        if (e < 0) continue;
        for (int l = 0; l < nn; l++) conn_s[s * nn + l] = m->conn[(long)e * nn + l];
        for (int f = 0; f < nf; f++) bc_s[s * nf + f] = m->bc[(long)e * nf + f];
        // ApplyBC (ApplyBC.h:14-133): dof (ja, jeqn) of a face with bc != 0 is
        // constrained when the node is Dirichlet for jeqn, except energy on an
        // isothermal bc == 1 face (quirk Q5)
        unsigned long long mk = 0;
        for (int f = 0; f < nf; f++) {
            const int bc = m->bc[(long)e * nf + f];
            if (bc == 0) continue;
            for (int t = 0; t < c->nbnod; t++) {
                const int ja = (c->ndim == 3) ? FACE3[f][t] : ((t == 0) ? f : ((f + 1) & 3));
                const int g = m->conn[(long)e * nn + ja];
                for (int q = 0; q < nq; q++) {
                    if (!m->dirichlet[(long)g * nq + q]) continue;
                    if (bc == 1 && q == nq - 1 && !c->th.adiabatic) continue;
                    mk |= 1ULL << (nq * ja + q);
                }
            }
        }
        cmask[s] = mk;
    }
    // 3D slot lists: faces or constrained dofs (face kernel on top of the
    // domain part) and hanging-node elements (whole element per lane)
    std::vector<int> bslots, aslots;
    std::vector<unsigned char> adap_s;
    std::vector<double> hnm_s;
    const int nn2 = c->nnod * c->nnod;   // Hnm entries per element
    if (m->adap) {
        adap_s.assign((size_t)np, 0);
        hnm_s.assign((size_t)np * nn2, 0.0);
        for (long s = 0; s < np; s++) {
            const int e = c->perm[s];
            if (e < 0 || !m->adap[e]) continue;
            adap_s[s] = 1;
            std::memcpy(hnm_s.data() + (size_t)s * nn2, m->hnm + (size_t)e * nn2, sizeof(double) * nn2);
        }
    }
    if (c->ndim == 3)
        for (long s = 0; s < np; s++) {
            if (c->perm[s] < 0) continue;
            bool b = cmask[s] != 0ULL;
            for (int f = 0; f < nf; f++) b = b || bc_s[s * nf + f] != 0;
            if (!adap_s.empty() && adap_s[s]) {
                // hanging-node elements with faces: whole element per lane;
                // without, the interior kernel (Hnm-mapped tables)
                if (b) aslots.push_back((int)s);
                continue;
            }
            if (b) bslots.push_back((int)s);
        }
    c->n_bslots = (long)bslots.size();
    c->n_aslots = (long)aslots.size();
    // compose CSR entries point at slots; lists stay in original element order
    for (size_t i = 0; i < ent.size(); i++) {
        const int e = ent[i] / nn, l = ent[i] % nn;
        ent[i] = c->perm_inv[e] * nn + l;
    }
    c->d_conn = c->alloc<int>((size_t)np * nn);
    c->d_bc = c->alloc<int>((size_t)np * nf);
    c->d_cmask = c->alloc<unsigned long long>(np);
    c->d_x = c->alloc<double>((size_t)nnode * c->ndim);
    c->d_U = c->alloc<double>(c->ndof);
    c->d_U0 = c->alloc<double>(c->ndof);
    c->d_area = c->alloc<double>(np);
    c->d_cleng = c->alloc<double>(np);
    c->d_node_ptr = c->alloc<int>(nnode + 1);
    c->d_node_ent = c->alloc<int>((size_t)ne * nn);
    c->d_perm_inv = c->alloc<int>(ne);
    HIP_CHECK_THROW(hipMemcpy(c->d_conn, conn_s.data(), sizeof(int) * np * nn, hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_bc, bc_s.data(), sizeof(int) * np * nf, hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_cmask, cmask.data(), sizeof(unsigned long long) * np, hipMemcpyHostToDevice));
    if (c->n_bslots > 0) {
        c->d_bslots = c->alloc<int>((size_t)c->n_bslots);
        HIP_CHECK_THROW(hipMemcpy(c->d_bslots, bslots.data(), sizeof(int) * c->n_bslots, hipMemcpyHostToDevice));
        // split by the number of faces with a bc code 0 < bc < 100 (the face
        // kernel's points): one batch of points or more
        std::vector<int> b1, bm;
        for (const int sl : bslots) {
            int nfa = 0;
            for (int f = 0; f < nf; f++) nfa += (bc_s[(size_t)sl * nf + f] != 0 && bc_s[(size_t)sl * nf + f] < 100);
            (nfa <= 1 ? b1 : bm).push_back(sl);
        }
        c->n_bslots1 = (long)b1.size();
        c->n_bslotsm = (long)bm.size();
        if (!b1.empty()) {
            c->d_bslots1 = c->alloc<int>(b1.size());
            HIP_CHECK_THROW(hipMemcpy(c->d_bslots1, b1.data(), sizeof(int) * b1.size(), hipMemcpyHostToDevice));
        }
        if (!bm.empty()) {
            c->d_bslotsm = c->alloc<int>(bm.size());
            HIP_CHECK_THROW(hipMemcpy(c->d_bslotsm, bm.data(), sizeof(int) * bm.size(), hipMemcpyHostToDevice));
        }
    }
    if (!adap_s.empty()) {
        c->d_adap = c->alloc<unsigned char>((size_t)np);
        c->d_hnm = c->alloc<double>((size_t)np * nn2);
        HIP_CHECK_THROW(hipMemcpy(c->d_adap, adap_s.data(), (size_t)np, hipMemcpyHostToDevice));
        HIP_CHECK_THROW(hipMemcpy(c->d_hnm, hnm_s.data(), sizeof(double) * np * nn2, hipMemcpyHostToDevice));
    }
    if (c->n_aslots > 0) {
        c->d_aslots = c->alloc<int>((size_t)c->n_aslots);
        HIP_CHECK_THROW(hipMemcpy(c->d_aslots, aslots.data(), sizeof(int) * c->n_aslots, hipMemcpyHostToDevice));
    }
    HIP_CHECK_THROW(hipMemcpy(c->d_x, m->x, sizeof(double) * nnode * c->ndim, hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_node_ptr, ptr.data(), sizeof(int) * (nnode + 1), hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_node_ent, ent.data(), sizeof(int) * ne * nn, hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemcpy(c->d_perm_inv, c->perm_inv.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
    HIP_CHECK_THROW(hipMemsetAsync(c->d_U0, 0, sizeof(double) * c->ndof, c->st));
    // block storage
    const size_t kk = (size_t)c->k * c->k;
    if (c->block_flags & CFDV_BLOCKS_FP64) c->d_R64 = c->alloc<double>((size_t)np * kk);
    if (c->block_flags & CFDV_BLOCKS_FP32) c->d_R32 = c->alloc<float>((size_t)np * kk);
    if (c->block_flags & CFDV_BLOCKS_ROWS) setup_rows(c, m, ptr, elem);
    if (c->rows_direct) {
        // per color the sub-ranges of the face-kernel and per-lane lists
        // (both ascend by slot, the colors are slot ranges), and the domain
        // parts of the face kernel's elements between the two kernels
        c->bstart.assign(c->ncolor + 1, 0);
        c->astart.assign(c->ncolor + 1, 0);
        for (int k = 0; k <= c->ncolor; k++) {
            c->bstart[k] = std::lower_bound(bslots.begin(), bslots.end(), (int)c->cstart[k]) - bslots.begin();
            c->astart[k] = std::lower_bound(aslots.begin(), aslots.end(), (int)c->cstart[k]) - aslots.begin();
        }
        std::vector<int> bidx((size_t)np, -1);
        for (size_t i = 0; i < bslots.size(); i++) bidx[bslots[i]] = (int)i;
        c->d_bidx = c->alloc<int>((size_t)np);
        HIP_CHECK_THROW(hipMemcpy(c->d_bidx, bidx.data(), sizeof(int) * np, hipMemcpyHostToDevice));
        if (c->n_bslots > 0) c->d_Rb = c->alloc<double>((size_t)c->n_bslots * kk);
    }
    c->d_rhs = c->alloc<double>((size_t)np * c->k);
    c->d_s = c->alloc<double>((size_t)np * 4);
    c->d_b = c->alloc<double>(c->ndof);
    c->d_du = c->alloc<double>(c->ndof);
    c->d_r = c->alloc<double>(c->ndof);
    c->d_w = c->alloc<double>(c->ndof);
    c->d_y = c->alloc<double>(c->ndof);
    c->d_d64 = c->alloc<double>(c->ndof);
    c->d_w32 = c->alloc<float>(c->ndof);
    c->d_part = c->alloc<double>((size_t)NBLK_RED * 32 + 64);
    c->d_small = c->alloc<double>(1024);
    c->d_hcol = c->alloc<double>(2048);
    c->d_cflpart = c->alloc<double>((np + 255) / 256);
    c->d_counters = c->alloc<int>(4);
    c->d_hess = c->alloc<double>((size_t)MAX_M1 * (MAX_M1 - 1) + 3 * (size_t)MAX_M1);
    c->d_hessr = c->alloc<double>((size_t)MAX_M1 * (MAX_M1 - 1));
    c->d_gate = c->alloc<int>(4);
    HIP_CHECK_THROW(hipHostMalloc((void**)&c->h_steps, sizeof(GmresStep) * MAX_M1,
                                  hipHostMallocMapped | hipHostMallocCoherent));
    HIP_CHECK_THROW(hipHostGetDevicePointer((void**)&c->d_steps, c->h_steps, 0));
    for (auto& e : c->ev_step) HIP_CHECK_THROW(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // CalcLength once (main.cpp:120-121)
    launch_length(c->ndim, c->d_conn, c->d_x, np, c->d_area, c->d_cleng, c->st);
    c->check_launch();
    c->sync();
}

// ---------------------------------------------------------------------------
// GMRES variants
// ---------------------------------------------------------------------------
// MPI_GMRES::iterate at one rank (MPI_GMRES.h:125-220), quirks Q2-Q4.  The
// Hessenberg matrix, the early-break estimate, the rotations and the back
// substitution run on the device (launch_ref_arnoldi, launch_ref_finish);
// the host reads one GmresStep per Arnoldi step, as the reference's loop
// tests its estimate before the next product.
int solve_reference(cfdv_ctx* c, const cfdv_solve_opts* o, cfdv_solve_stats* s) {
    const int m_ = o->m, ldh = m_ + 1;
    const long n = c->ndof;
    c->ensure_krylov(m_);
    double* V = c->d_V;
    const long ldv = c->ldv();
    double* H = c->d_hess;
    double* g = H + (size_t)ldh * m_;
    double* y = c->d_small + 16;
    c->trace_m = m_;
    c->trace_H.assign((size_t)o->restarts * ldh * m_, 0.0);
    c->trace_g.assign((size_t)o->restarts * ldh, 0.0);
    c->trace_mm.assign(o->restarts, 0);
    vec_fill(c->d_du, 0.0, n, c->st);
    HIP_CHECK_THROW(hipMemsetAsync(c->d_gate + 1, 0, sizeof(int), c->st));
    launch_compose(c->d_node_ptr, c->d_node_ent, c->d_rhs, c->k, c->neqn, c->n_node, c->d_b, c->st);
    const double bnorm = c->norm(c->d_b);
    const double tol = bnorm * 1.0E-9;
    s->bnorm = bnorm;
    for (int gloop = 0; gloop < o->restarts; gloop++) {
        int mm = m_;
        // clear() (:821-831): h, g, y start from zero every restart
        vec_fill(H, 0.0, (long)ldh * m_, c->st);
        c->matvec(c->d_du, c->d_y, 64, false);
        s->matvecs64++;
        vec_sub(c->d_r, c->d_b, c->d_y, n, c->st);
        const double mag_r = c->norm(c->d_r);
        if (!(mag_r > 0)) s->breakdown = 1;
        launch_gmres_start(g, m_, mag_r, c->d_gate, c->st);
        // v_0 = r / mag_r
        vec_set(c->d_small + 8, mag_r, c->st);
        vec_scale_inv<double>(V, c->d_r, c->d_small + 8, n, c->st);
        for (int j = 0; j < mm; j++) {
            double* w = c->d_w;
            c->matvec(V + (size_t)j * ldv, w, 64, false);
            s->matvecs64++;
            s->iters++;
            // modified Gram-Schmidt, one inner product at a time (:166-174)
            for (int i = 0; i <= j; i++) {
                c->multidot<double>(V + (size_t)i * ldv, 1, w, c->d_hcol + i);
                vec_axpy(w, V + (size_t)i * ldv, c->d_hcol + i, -1.0, n, c->st);
            }
            c->multidot<double>(w, 1, w, c->d_hcol + j + 1);
            launch_ref_arnoldi(H, ldh, c->d_hcol, j, mag_r, tol, c->d_small + 9, c->d_gate + 1, c->d_steps, c->st);
            vec_scale_inv<double>(V + (size_t)(j + 1) * ldv, w, c->d_small + 9, n, c->st);
            c->check_launch();
            HIP_CHECK_THROW(hipEventRecord(c->ev_step[0], c->st));
            cfdv_ctx::wait_event(c->ev_step[0]);
            if (c->h_steps[j].stop) {
                mm = j + 1;
                break;
            }
        }
        launch_ref_finish(H, c->d_hessr, ldh, g, mm, y, c->st);
        vec_combine<double>(c->d_du, V, ldv, mm, y, n, c->st);
        c->check_launch();
        HIP_CHECK_THROW(hipMemcpyAsync(c->trace_H.data() + (size_t)gloop * ldh * m_, H, sizeof(double) * ldh * m_,
                                       hipMemcpyDeviceToHost, c->st));
        HIP_CHECK_THROW(hipMemcpyAsync(c->trace_g.data() + (size_t)gloop * ldh, g, sizeof(double) * ldh,
                                       hipMemcpyDeviceToHost, c->st));
        c->sync();
        c->trace_mm[gloop] = mm;
        s->est_resid = std::fabs(c->trace_g[(size_t)gloop * ldh + mm]);
        s->restarts++;   // Q4: no early exit from the restart loop
    }
    int brk = 0;
    HIP_CHECK_THROW(hipMemcpyAsync(&brk, c->d_gate + 1, sizeof(int), hipMemcpyDeviceToHost, c->st));
    // final true residual for reporting
    c->matvec(c->d_du, c->d_y, 64, false);
    vec_sub(c->d_r, c->d_b, c->d_y, n, c->st);
    s->resid = c->norm(c->d_r);
    s->rel_resid = bnorm > 0 ? s->resid / bnorm : 0.0;
    if (brk) s->breakdown = 1;
    c->have_du = true;
    if (s->breakdown) return CFDV_ERR_BREAKDOWN;
    if (!std::isfinite(s->resid)) return CFDV_ERR_NAN;
    return CFDV_OK;
}

// Restarted GMRES(m) with CGS2 orthogonalisation and Givens rotations.
// Solves K x = rhs_dev (or K^T x) into x_dev, starting from x_dev = 0.
// prec: 64 (FP64 blocks/vectors) or 32 (FP32 blocks and basis; dots and
// Hessenberg in FP64).  With Pinv (node-block inverses in T) the system is
// right-preconditioned: K Pinv u = rhs, x = Pinv u, so the Givens estimate
// is still the true residual; z is a work vector for Pinv V_j.  Returns
// Arnoldi iterations; *conv set when the Givens estimate met tol_abs.
//
// The Hessenberg column, its rotations, g and the back substitution stay on
// the device (launch_givens_step, launch_backsub).  The host enqueues step
// j + 1 before it reads step j's GmresStep record, so the stream never waits
// for the host; every kernel of a step is gated on the device flag that
// step j's rotation sets when the cycle ends there, so the step enqueued
// after the last one does no work.
template <typename T>
int gmres_core(cfdv_ctx* c, int prec, bool tr, const T* rhs, double* x, int m, int maxrestart, double tol_abs,
               double rnorm0, cfdv_solve_stats* s, bool* conv, T* V, T* w, const T* Pinv, T* z, bool cgs2) {
    const long n = c->ndof;
    const long ldv = c->ldv();
    const int ldh = m + 1;
    double* H = c->d_hess;
    double* g = H + (size_t)ldh * m;
    double* cs = g + ldh;
    double* sn = cs + m;
    double* y = c->d_small + 16;
    int* gate = c->d_gate;
    int iters = 0;
    *conv = false;
    double beta = rnorm0;
    // r0 = rhs (x = 0 on entry); later restarts recompute r = rhs - K x
    for (int rs = 0; rs < maxrestart; rs++) {
        if (rs > 0) {
            // r = rhs - K x, in the working precision of the basis
            if (prec == 64) {
                c->matvec(x, c->d_y, 64, tr);
                s->matvecs64++;
                vec_sub(c->d_r, (const double*)rhs, c->d_y, n, c->st);
                beta = c->norm(c->d_r);
                if (beta <= tol_abs) { *conv = true; break; }
                vec_set(c->d_small + 8, beta, c->st);
                vec_scale_inv<double>((double*)V, c->d_r, c->d_small + 8, n, c->st);
            } else {
                // FP32 inner: r = rhs - K32 x with x held in FP64 (d_d64)
                c->matvec(x, c->d_y, 3, tr);
                s->matvecs32++;
                // rhs is FP32 here; form r in FP64 then round into V_0
                vec_fill(c->d_r, 0.0, n, c->st);
                vec_add_f2d(c->d_r, (const float*)rhs, 1.0, n, c->st);
                vec_axpy(c->d_r, c->d_y, c->d_small + 10, -1.0, n, c->st);   // d_small[10] == 1
                beta = c->norm(c->d_r);
                if (beta <= tol_abs) { *conv = true; break; }
                vec_d2f_scaled((float*)V, c->d_r, 1.0 / beta, n, c->st);
            }
        } else {
            vec_set(c->d_small + 8, beta, c->st);
            vec_scale_inv<T>(V, rhs, c->d_small + 8, n, c->st);
        }
        launch_gmres_start(g, m, beta, gate, c->st);
        int jdone = 0;
        bool early = false;   // the end of the cycle was read one step late
        bool z_ready = false;   // z = P v_j formed with v_j by the previous step
        for (int j = 0; j < m; j++) {
            if (Pinv) {
                if (!z_ready)
                    vec_nodeblock_apply<T, T>(Pinv, V + (size_t)j * ldv, z, c->neqn, c->n_node, false, tr, c->st,
                                              gate);
                c->matvec(z, w, prec, tr, gate);
            } else {
                c->matvec(V + (size_t)j * ldv, w, prec, tr, gate);
            }
            // Classical Gram-Schmidt.  CGS2 (two passes: h1 = V^T w,
            // w -= V h1, h2 = V^T w, w -= V h2, |w|^2) keeps the basis
            // orthogonal to working precision; CGS1 (h1 = V^T w,
            // w -= V h1, |w|^2) saves a pass over the basis and serves the
            // FP32 inner solve, whose result the FP64 refinement checks and
            // corrects.  Up to 32 basis vectors each pass is one fused kernel.
            const int nv = j + 1;
            if (nv <= 32) {
                vec_cgs<T>(0, w, V, ldv, nv, nullptr, n, c->d_part, NBLK_RED, c->d_hcol, c->st, gate);
                if (cgs2)
                    vec_cgs<T>(1, w, V, ldv, nv, c->d_hcol, n, c->d_part, NBLK_RED, c->d_hcol + 1024, c->st, gate);
                vec_cgs<T>(2, w, V, ldv, nv, cgs2 ? c->d_hcol + 1024 : c->d_hcol, n, c->d_part, NBLK_RED,
                           c->d_hcol + 2040, c->st, gate);
            } else {
                vec_multidot<T>(V, ldv, nv, w, n, c->d_part, c->d_hcol, NBLK_RED, c->st, gate);
                vec_multi_axpy<T>(w, V, ldv, nv, c->d_hcol, 1.0, n, c->st, gate);
                if (cgs2) {
                    vec_multidot<T>(V, ldv, nv, w, n, c->d_part, c->d_hcol + 1024, NBLK_RED, c->st, gate);
                    vec_multi_axpy<T>(w, V, ldv, nv, c->d_hcol + 1024, 1.0, n, c->st, gate);
                }
                vec_multidot<T>(w, ldv, 1, w, n, c->d_part, c->d_hcol + 2040, NBLK_RED, c->st, gate);
            }
            launch_givens_step(H, ldh, g, cs, sn, c->d_hcol, cgs2 ? c->d_hcol + 1024 : nullptr, c->d_hcol + 2040, j,
                               tol_abs, c->d_small + 9, gate, c->d_steps, c->st);
            if (Pinv && j + 1 < m) {
                vec_nodeblock_scale_apply<T>(Pinv, w, c->d_small + 9, V + (size_t)(j + 1) * ldv, z, c->neqn,
                                             c->n_node, tr, c->st, gate);
                z_ready = true;
            } else {
                vec_scale_inv<T>(V + (size_t)(j + 1) * ldv, w, c->d_small + 9, n, c->st, gate);
            }
            c->check_launch();
            HIP_CHECK_THROW(hipEventRecord(c->ev_step[j & 1], c->st));
            if (prec == 64) s->matvecs64++; else s->matvecs32++;
            iters++;
            jdone = j + 1;
            if (j > 0) {
                cfdv_ctx::wait_event(c->ev_step[(j - 1) & 1]);
                if (c->h_steps[j - 1].stop) {
                    // step j was enqueued behind the last step and did nothing
                    if (prec == 64) s->matvecs64--; else s->matvecs32--;
                    iters--;
                    jdone = j;
                    early = true;
                    c->drop_last_matvec_mark();
                    break;
                }
            }
        }
        if (!early) cfdv_ctx::wait_event(c->ev_step[(jdone - 1) & 1]);
        const GmresStep last = c->h_steps[jdone - 1];
        s->est_resid = last.est;
        if (last.fail) s->breakdown = 1;
        launch_backsub(H, ldh, g, jdone, y, c->st);
        if (Pinv) {
            // x += Pinv (V y), the combination formed in FP64
            vec_fill(c->d_t64, 0.0, n, c->st);
            vec_combine<T>(c->d_t64, V, ldv, jdone, y, n, c->st);
            vec_nodeblock_apply<double, T>(Pinv, c->d_t64, x, c->neqn, c->n_node, true, tr, c->st);
        } else {
            vec_combine<T>(x, V, ldv, jdone, y, n, c->st);
        }
        c->check_launch();
        s->restarts++;
        if (last.est <= tol_abs) { *conv = true; break; }
        if (last.stop) break;   // invariant subspace or failure
    }
    return iters;
}

// K x = b (K^T x = b with o->transpose) for device b, x (both [ndof]).
// warm: x holds a starting guess x0 on entry; K d = b - K x0 is solved to
// the same absolute tolerance rtol ||b|| and x = x0 + d.  The Givens
// estimate ends a GMRES cycle; the true residual b - K x decides: while it is
// above rtol ||b|| the solve continues with K d = b - K x from the current x,
// within the same total of o->restarts restarts, and fails with NOCONV when
// they run out.
int solve_fp64_rhs(cfdv_ctx* c, const cfdv_solve_opts* o, cfdv_solve_stats* s, const double* b, double* x,
                   bool warm = false) {
    const long n = c->ndof;
    if (!c->have_blocks(64, o->transpose != 0)) return fail(c, CFDV_ERR_ARG, "FP64 solve needs CFDV_BLOCKS_FP64");
    c->ensure_krylov(o->m);
    if (!c->d_r0) c->d_r0 = c->alloc<double>((size_t)n);
    const double bnorm = c->norm(b);
    s->bnorm = bnorm;
    bool conv = false;
    const bool pc = o->precond == CFDV_PRECOND_NODE_BLOCK;
    if (pc) c->ensure_precond();
    const bool tr = o->transpose != 0;
    const double tol = o->rtol * bnorm;
    // r0 = b - K x and its norm (the product is counted by the caller when
    // a further cycle uses it)
    auto residual = [&]() {
        c->matvec(x, c->d_y, 64, tr);
        vec_sub(c->d_r0, b, c->d_y, n, c->st);
        return c->norm(c->d_r0);
    };
    double rnorm = bnorm;
    if (warm) {
        rnorm = residual();
        s->matvecs64++;
    } else {
        vec_fill(x, 0.0, n, c->st);
        if (bnorm > 0) {
            s->iters += gmres_core<double>(c, 64, tr, b, x, o->m, o->restarts, tol, bnorm, s, &conv, c->d_V, c->d_w,
                                           pc ? c->d_P64 : nullptr, pc ? c->d_z64 : nullptr, true);
            rnorm = residual();
        }
    }
    while (rnorm > tol && std::isfinite(rnorm) && s->restarts < o->restarts) {
        if (!warm) s->matvecs64++;   // this residual product starts a further cycle
        warm = false;
        vec_fill(c->d_d64, 0.0, n, c->st);
        s->iters += gmres_core<double>(c, 64, tr, c->d_r0, c->d_d64, o->m, o->restarts - s->restarts, tol, rnorm, s,
                                       &conv, c->d_V, c->d_w, pc ? c->d_P64 : nullptr, pc ? c->d_z64 : nullptr, true);
        vec_set(c->d_small + 10, 1.0, c->st);
        vec_axpy(x, c->d_d64, c->d_small + 10, 1.0, n, c->st);
        rnorm = residual();
    }
    s->resid = rnorm;
    s->rel_resid = bnorm > 0 ? s->resid / bnorm : 0.0;
    if (!std::isfinite(s->resid)) return fail(c, CFDV_ERR_NAN, "NaN in the FP64 solve");
    if (s->rel_resid > o->rtol && bnorm > 0) return fail(c, CFDV_ERR_NOCONV, "FP64 GMRES did not converge");
    return CFDV_OK;
}

// x0 in d_du from the increment rates r_k = du_k / dt_k of the last updates
// (du_1 the last): warm_start 1, x0 = dt r_1; 2, dt (2 r_1 - r_2); 3,
// dt (3 r_1 - 3 r_2 + r_3) -- the rate extrapolated with a polynomial of
// degree warm_start - 1, lowered while fewer updates are kept (after a
// reset).  Modes 2 and 3 keep du_1 (and du_2) for the next step.
// The first solve of a build forms the guess (and shifts the history); a
// repeated solve of the same build starts from the same guess, kept in
// d_x0, so it returns the same du and leaves the history as one solve does.
bool warm_guess(cfdv_ctx* c, const cfdv_solve_opts* o) {
    const long n = c->ndof;
    if (c->warm_consumed) {
        if (!c->warm_had) return false;
        vec_copy(c->d_du, c->d_x0, n, c->st);
        return true;
    }
    c->warm_consumed = true;
    c->warm_had = false;
    c->n_saved = 0;
    if (!o->warm_start || o->transpose || !c->have_prev || !(c->prev_dt > 0) || !(c->last_dt > 0)) return false;
    if (!c->d_x0) c->d_x0 = c->alloc<double>((size_t)n);
    const int ws = std::min(std::max(o->warm_start, 1), 3);
    if (ws == 1) {
        vec_set(c->d_small + 12, c->prev_dt / c->last_dt, c->st);
        vec_scale_to(c->d_du, c->d_du, c->d_small + 12, n, c->st);
        vec_copy(c->d_x0, c->d_du, n, c->st);
        c->warm_had = true;
        return true;
    }
    const int keep = ws - 1;   // older increments kept for the next step
    for (int k = 0; k < keep; k++)
        if (!c->d_old[k]) c->d_old[k] = c->alloc<double>((size_t)n);
    if (!c->d_dut) c->d_dut = c->alloc<double>((size_t)n);
    const int order = std::min(ws, 1 + c->n_old);
    static const double coef[3][3] = {{1, 0, 0}, {2, -1, 0}, {3, -3, 1}};
    const double* cf = coef[order - 1];
    vec_set(c->d_small + 12, c->prev_dt / (cf[0] * c->last_dt), c->st);
    vec_scale_to(c->d_dut, c->d_du, c->d_small + 12, n, c->st);
    for (int k = 1; k < order; k++) {
        vec_set(c->d_small + 12 + k, c->last_dt / c->old_dt[k - 1], c->st);
        vec_axpy(c->d_dut, c->d_old[k - 1], c->d_small + 12 + k, cf[k], n, c->st);
    }
    // shift the history: du_1 becomes the first older increment
    double* spare = c->d_old[keep - 1];
    for (int k = keep - 1; k > 0; k--) {
        c->d_old[k] = c->d_old[k - 1];
        c->old_dt[k] = c->old_dt[k - 1];
    }
    c->d_old[0] = c->d_du;
    c->old_dt[0] = c->prev_dt;
    c->d_du = c->d_dut;
    c->d_dut = spare;
    vec_copy(c->d_x0, c->d_du, n, c->st);
    c->check_launch();
    c->n_saved = std::min(c->n_old + 1, keep);
    c->warm_had = true;
    return true;
}

int solve_fp64(cfdv_ctx* c, const cfdv_solve_opts* o, cfdv_solve_stats* s) {
    if (!c->have_blocks(64, o->transpose != 0)) return fail(c, CFDV_ERR_ARG, "FP64 solve needs CFDV_BLOCKS_FP64");
    launch_compose(c->d_node_ptr, c->d_node_ent, c->d_rhs, c->k, c->neqn, c->n_node, c->d_b, c->st);
    const bool warm = warm_guess(c, o);
    const int rc = solve_fp64_rhs(c, o, s, c->d_b, c->d_du, warm);
    c->have_du = rc == CFDV_OK || rc == CFDV_ERR_NOCONV;
    return rc;
}

// FP64 iterative refinement around an FP32 GMRES: r = b - K64 du (FP64),
// solve K32 d = r/||r|| in FP32, du += ||r|| d (FP64).  warm: du starts
// from the guess held in x.
int solve_mixed_rhs(cfdv_ctx* c, const cfdv_solve_opts* o, cfdv_solve_stats* s, const double* b, double* x,
                    bool warm = false) {
    const long n = c->ndof;
    const bool tr = o->transpose != 0;
    if (!c->have_blocks(32, tr))
        return fail(c, CFDV_ERR_ARG, "mixed solve needs CFDV_BLOCKS_FP32 or CFDV_BLOCKS_ROWS");
    const int outer_prec = c->have_blocks(64, tr) ? 64 : 3;   // FP64 residual with FP64 (or widened FP32) blocks
    c->ensure_krylov32(o->inner_m);
    const bool pc = o->precond == CFDV_PRECOND_NODE_BLOCK;
    if (pc) c->ensure_precond();
    const double one = 1.0;
    vec_set(c->d_small + 10, one, c->st);
    const double bnorm = c->norm(b);
    s->bnorm = bnorm;
    double rnorm = bnorm;
    if (!warm) {
        vec_fill(x, 0.0, n, c->st);
        vec_copy(c->d_r, b, n, c->st);
    } else {
        c->matvec(x, c->d_y, outer_prec, tr);
        if (outer_prec == 64) s->matvecs64++; else s->matvecs32++;
        vec_sub(c->d_r, b, c->d_y, n, c->st);
        rnorm = c->norm(c->d_r);
    }
    int sweeps = 0;
    while (rnorm > o->rtol * bnorm && sweeps < o->restarts) {
        // inner right-hand side: FP32 copy of r / ||r||
        vec_d2f_scaled(c->d_w32, c->d_r, 1.0 / rnorm, n, c->st);
        vec_fill(c->d_d64, 0.0, n, c->st);
        bool conv = false;
        cfdv_solve_stats si{};
        // the FP32 right-hand side stays in d_w32; the last basis slot is the
        // Arnoldi work vector
        float* wbuf = c->d_V32 + (size_t)o->inner_m * c->ldv();   // last slot as the Arnoldi work vector
        // the inner solve reduces r by inner_rtol, or only by what is left to
        // reach the target (half of rtol ||b|| / ||r||) when that is less
        const double need = 0.5 * o->rtol * bnorm / rnorm;
        const double itol = std::min(0.5, std::max(o->inner_rtol, need));
        s->iters += gmres_core<float>(c, 32, tr, c->d_w32, c->d_d64, o->inner_m - 1, o->inner_restarts, itol, 1.0,
                                      &si, &conv, c->d_V32, wbuf, pc ? c->d_P32 : nullptr, pc ? c->d_z32 : nullptr,
                                      false);
        s->matvecs32 += si.matvecs32;
        s->matvecs64 += si.matvecs64;
        // du += ||r|| d  (FP64)
        vec_set(c->d_small + 11, rnorm, c->st);
        vec_axpy(x, c->d_d64, c->d_small + 11, 1.0, n, c->st);
        // r = b - K x in FP64
        c->matvec(x, c->d_y, outer_prec, tr);
        if (outer_prec == 64) s->matvecs64++; else s->matvecs32++;
        vec_sub(c->d_r, b, c->d_y, n, c->st);
        rnorm = c->norm(c->d_r);
        sweeps++;
        if (!std::isfinite(rnorm)) break;
    }
    s->restarts = sweeps;
    s->resid = rnorm;
    s->rel_resid = bnorm > 0 ? rnorm / bnorm : 0.0;
    if (!std::isfinite(rnorm)) return fail(c, CFDV_ERR_NAN, "NaN in the mixed solve");
    if (s->rel_resid > o->rtol && bnorm > 0) return fail(c, CFDV_ERR_NOCONV, "mixed refinement did not converge");
    return CFDV_OK;
}

int solve_mixed(cfdv_ctx* c, const cfdv_solve_opts* o, cfdv_solve_stats* s) {
    if (!c->have_blocks(32, o->transpose != 0))
        return fail(c, CFDV_ERR_ARG, "mixed solve needs CFDV_BLOCKS_FP32 or CFDV_BLOCKS_ROWS");
    launch_compose(c->d_node_ptr, c->d_node_ent, c->d_rhs, c->k, c->neqn, c->n_node, c->d_b, c->st);
    const bool warm = warm_guess(c, o);
    const int rc = solve_mixed_rhs(c, o, s, c->d_b, c->d_du, warm);
    c->have_du = rc == CFDV_OK || rc == CFDV_ERR_NOCONV;
    return rc;
}

// The 8-vector product K X / K^T X on side-by-side FP32 vectors
void matvec8(cfdv_ctx* c, const float* X, float* Y, bool tr) {
    if (c->d_B32) {
        c->orient_rows(tr);
        launch_rows_matvec8(c->ra, c->d_B32, X, Y, c->st);
    } else {
        vec_fill_f(Y, 0.0f, c->ndof * 8, c->st);
        for (int k = 0; k < c->ncolor; k++) {
            MatvecArgs a{c->ndim, c->nnod, c->neqn, c->cstart[k], c->cstart[k + 1], c->d_conn, 0};
            launch_matvec_multi_color(a, c->d_R32, X, Y, tr, c->st);
        }
    }
    c->check_launch();
}

// One cycle of global GMRES (Jbilou, Messaoudi and Sadok, 1999) for the 8
// systems K^T u_s = v_s at once: the Krylov vectors are the 8 systems'
// columns side by side, inner products and norms are Frobenius ones over
// all of them, so every Arnoldi step is one 8-vector product and one
// Hessenberg column (one polynomial in K^T for all 8).  rhs8 in V_0 on
// entry ([8 ndof] FP32, Frobenius norm beta); x8 += Pinv^T V y in FP64.
// Right-preconditioned by node-block Jacobi (Pinv) like the single solve;
// stops when the Givens estimate is at or below tol_abs.  The Hessenberg
// column, its rotations and the back substitution run on the device; the
// host reads each step's record before the next.
int gmres_global8(cfdv_ctx* c, bool tr, double* x8, int m, double tol_abs, double beta, cfdv_solve_stats* s,
                  bool* conv, const float* Pinv) {
    const long n = 8 * c->ndof;
    const long ldv = c->ldv8();
    float* V = c->d_V8;
    float* w = V + (size_t)m * ldv;   // the last slot is the Arnoldi work vector
    const int ldh = m + 1;
    double* H = c->d_hess;
    double* g = H + (size_t)ldh * m;
    double* cs = g + ldh;
    double* sn = cs + m;
    double* y = c->d_small + 16;
    *conv = false;
    launch_gmres_start(g, m, beta, c->d_gate, c->st);
    int jdone = 0;
    for (int j = 0; j < m - 1; j++) {
        float* vj = V + (size_t)j * ldv;
        if (Pinv) {
            vec_nodeblock_apply8<float, float>(Pinv, vj, c->d_z8, c->neqn, c->n_node, false, tr, c->st);
            matvec8(c, c->d_z8, w, tr);
        } else {
            matvec8(c, vj, w, tr);
        }
        s->matvecs32++;
        const int nv = j + 1;
        // classical Gram-Schmidt, one pass (h = V^T w; w -= V h, |w|^2):
        // the FP64 refinement around it checks and corrects
        if (nv <= 32) {
            vec_cgs<float>(0, w, V, ldv, nv, nullptr, n, c->d_part, NBLK_RED, c->d_hcol, c->st);
            vec_cgs<float>(2, w, V, ldv, nv, c->d_hcol, n, c->d_part, NBLK_RED, c->d_hcol + 2040, c->st);
        } else {
            vec_multidot<float>(V, ldv, nv, w, n, c->d_part, c->d_hcol, NBLK_RED, c->st);
            vec_multi_axpy<float>(w, V, ldv, nv, c->d_hcol, 1.0, n, c->st);
            vec_multidot<float>(w, ldv, 1, w, n, c->d_part, c->d_hcol + 2040, NBLK_RED, c->st);
        }
        launch_givens_step(H, ldh, g, cs, sn, c->d_hcol, nullptr, c->d_hcol + 2040, j, tol_abs, c->d_small + 9,
                           c->d_gate, c->d_steps, c->st);
        vec_scale_inv<float>(V + (size_t)(j + 1) * ldv, w, c->d_small + 9, n, c->st, c->d_gate);
        c->check_launch();
        HIP_CHECK_THROW(hipEventRecord(c->ev_step[0], c->st));
        cfdv_ctx::wait_event(c->ev_step[0]);
        s->iters++;
        jdone = j + 1;
        const GmresStep st = c->h_steps[j];
        s->est_resid = st.est;
        if (st.fail) s->breakdown = 1;
        if (st.stop) break;
    }
    launch_backsub(H, ldh, g, jdone, y, c->st);
    if (Pinv) {
        vec_fill(c->d_t8, 0.0, n, c->st);
        vec_combine<float>(c->d_t8, V, ldv, jdone, y, n, c->st);
        vec_nodeblock_apply8<double, float>(Pinv, c->d_t8, x8, c->neqn, c->n_node, true, tr, c->st);
    } else {
        vec_combine<float>(x8, V, ldv, jdone, y, n, c->st);
    }
    c->check_launch();
    *conv = s->est_resid <= tol_abs;
    return jdone;
}

// K^T lambda_s = g_s (or K lambda_s = g_s) for 8 right-hand sides at once:
// per right-hand side FP64 iterative refinement (r_s = g_s - K64^T lambda_s,
// lambda_s += ||r_s|| d_s) around one FP32 global GMRES cycle per sweep on
// the normalised residuals r_s / ||r_s|| of the systems not yet converged
// (converged ones enter as zero columns).  Ends when every relative residual
// is at or below o->rtol, or after o->restarts sweeps.  G, L: device [8][ndof].
int solve_multi8(cfdv_ctx* c, const cfdv_solve_opts* o, bool tr, const double* G, double* L, double rel[8],
                 cfdv_solve_stats* s) {
    const long n = c->ndof;
    if (!c->have_blocks(32, tr))
        return fail(c, CFDV_ERR_ARG, "the 8-vector solve needs CFDV_BLOCKS_FP32 or CFDV_BLOCKS_ROWS");
    const int outer_prec = c->have_blocks(64, tr) ? 64 : 3;
    const int m = std::min(o->inner_m, MAX_M1 - 1);
    c->ensure_multi(m);
    const bool pc = o->precond == CFDV_PRECOND_NODE_BLOCK;
    if (pc) c->ensure_precond();
    double gn[8], rn[8];
    for (int k = 0; k < 8; k++) {
        gn[k] = c->norm(G + (size_t)k * n);
        rn[k] = gn[k];
    }
    s->bnorm = 0.0;
    for (int k = 0; k < 8; k++) s->bnorm = std::max(s->bnorm, gn[k]);
    vec_fill(L, 0.0, 8 * n, c->st);
    vec_copy(c->d_R8, G, 8 * n, c->st);
    int sweeps = 0;
    for (;;) {
        Scale8 inv{}, amp{};
        int active = 0;
        bool finite = true;
        for (int k = 0; k < 8; k++) {
            finite = finite && std::isfinite(rn[k]);
            if (gn[k] > 0 && rn[k] > o->rtol * gn[k]) {
                inv.v[k] = 1.0 / rn[k];
                amp.v[k] = rn[k];
                active++;
            }
        }
        if (!finite || active == 0 || sweeps >= o->restarts) break;
        // the inner right-hand side: the active residuals, each of unit norm,
        // side by side (Frobenius norm beta); V_0 is it over beta
        const double beta = std::sqrt((double)active);
        for (int k = 0; k < 8; k++) inv.v[k] /= beta;
        vec_multivec_pack(c->d_V8, c->d_R8, inv, n, c->st);
        vec_fill(c->d_x8, 0.0, 8 * n, c->st);
        // reduce the Frobenius residual by inner_rtol (each column then by at
        // most sqrt(active) inner_rtol)
        bool conv = false;
        cfdv_solve_stats si{};
        gmres_global8(c, tr, c->d_x8, m, o->inner_rtol * beta, beta, &si, &conv, pc ? c->d_P32 : nullptr);
        s->iters += si.iters;
        s->matvecs32 += si.matvecs32;
        if (si.breakdown) s->breakdown = 1;
        // lambda_s += ||r_s|| d_s, r_s = g_s - K lambda_s in FP64
        vec_multivec_unpack_axpy(L, c->d_x8, amp, n, c->st);
        for (int k = 0; k < 8; k++) {
            if (amp.v[k] == 0.0) continue;
            c->matvec(L + (size_t)k * n, c->d_y, outer_prec, tr);
            if (outer_prec == 64) s->matvecs64++; else s->matvecs32++;
            vec_sub(c->d_R8 + (size_t)k * n, G + (size_t)k * n, c->d_y, n, c->st);
            rn[k] = c->norm(c->d_R8 + (size_t)k * n);
        }
        sweeps++;
    }
    s->restarts = sweeps;
    double worst = 0.0;
    for (int k = 0; k < 8; k++) {
        rel[k] = gn[k] > 0 ? rn[k] / gn[k] : 0.0;
        if (!std::isfinite(rel[k])) return fail(c, CFDV_ERR_NAN, "NaN in the 8-vector solve");
        worst = std::max(worst, rel[k]);
    }
    s->rel_resid = worst;
    s->resid = 0.0;
    for (int k = 0; k < 8; k++) s->resid = std::max(s->resid, rn[k]);
    if (worst > o->rtol) return fail(c, CFDV_ERR_NOCONV, "8-vector refinement did not converge");
    return CFDV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define GUARD_BEGIN try {
#define GUARD_END                                                    \
    }                                                                \
    catch (const std::exception& ex) {                               \
        return fail(ctx, CFDV_ERR_HIP, ex.what());                   \
    }

extern "C" {

int cfdv_create(const cfdv_mesh* mesh, const cfdv_thermo* thermo, int block_flags, cfdv_ctx** out) {
    if (!out || !mesh || !thermo) return CFDV_ERR_ARG;
    *out = nullptr;
    if (!((mesh->ndim == 3 && mesh->nnod == 8 && mesh->neqn == 5) ||
          (mesh->ndim == 2 && mesh->nnod == 4 && mesh->neqn == 4)))
        return CFDV_ERR_ARG;
    if (mesh->n_elem <= 0 || mesh->n_node <= 0 || !mesh->conn || !mesh->x || !mesh->bc || !mesh->dirichlet)
        return CFDV_ERR_ARG;
    if (mesh->adap && !mesh->hnm) return CFDV_ERR_ARG;
    if (!(block_flags & (CFDV_BLOCKS_FP64 | CFDV_BLOCKS_FP32 | CFDV_BLOCKS_ROWS))) return CFDV_ERR_ARG;
    if (block_flags & ~(CFDV_BLOCKS_FP64 | CFDV_BLOCKS_FP32 | CFDV_BLOCKS_ROWS)) return CFDV_ERR_ARG;
    // rows with FP32 element blocks need the FP64 ones; rows alone (built in
    // place, no element blocks) are implemented for hex8
    if ((block_flags & CFDV_BLOCKS_ROWS) && (block_flags & CFDV_BLOCKS_FP32) && !(block_flags & CFDV_BLOCKS_FP64))
        return CFDV_ERR_ARG;
    if (block_flags == CFDV_BLOCKS_ROWS && mesh->ndim != 3) return CFDV_ERR_ARG;
    cfdv_ctx* ctx = new (std::nothrow) cfdv_ctx();
    if (!ctx) return CFDV_ERR_HIP;
    try {
        // an error an earlier, unrelated HIP call left on this thread is not
        // this context's (its own launches are checked as they are made)
        (void)hipGetLastError();
        ctx->ndim = mesh->ndim;
        ctx->nnod = mesh->nnod;
        ctx->neqn = mesh->neqn;
        ctx->nface = (mesh->ndim == 3) ? 6 : 4;
        ctx->nbnod = (mesh->ndim == 3) ? 4 : 2;
        ctx->k = ctx->nnod * ctx->neqn;
        ctx->n_elem = mesh->n_elem;
        ctx->n_node = mesh->n_node;
        ctx->ndof = mesh->n_node * mesh->neqn;
        ctx->block_flags = block_flags;
        ctx->r64_pairs = (block_flags & CFDV_BLOCKS_ROWS) ? 1 : 0;
        ctx->rows_direct = block_flags == CFDV_BLOCKS_ROWS;
        ctx->th = make_thermo(thermo);
        HIP_CHECK_THROW(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
        for (auto& e : ctx->ev) HIP_CHECK_THROW(hipEventCreate(&e));
        HIP_CHECK_THROW(hipEventCreateWithFlags(&ctx->ev_sync, hipEventDisableTiming));
        upload_gauss_tables(make_tables());
        try {
            ctx->d_flag_scratch = flags_scratch();
        } catch (const std::exception& ex) {
            ctx->flag_scratch_err = ex.what();   // reported by cfdv_refine_flags_noalloc
        }
        if (block_flags & CFDV_BLOCKS_ROWS) {
            // the context works in its own node order (locality_order); the
            // mesh's node arrays and connectivity are renumbered here and node
            // vectors are permuted at every call that passes one
            for (long i = 0; i < mesh->n_elem * mesh->nnod; i++)
                if (mesh->conn[i] < 0 || mesh->conn[i] >= mesh->n_node)
                    throw std::invalid_argument("connectivity out of range");
            const long nnode = mesh->n_node, nd = mesh->ndim, nq = mesh->neqn;
            std::vector<int> order = choose_node_order(mesh->ndim, mesh->nnod, nnode, mesh->n_elem, mesh->x, mesh->conn);
            if (order.empty()) {
                order.resize((size_t)nnode);
                std::iota(order.begin(), order.end(), 0);
            }
            std::vector<int> pos((size_t)nnode);
            for (long i = 0; i < nnode; i++) pos[order[i]] = (int)i;
            std::vector<int32_t> conn((size_t)mesh->n_elem * mesh->nnod);
            for (size_t i = 0; i < conn.size(); i++) conn[i] = pos[mesh->conn[i]];
            std::vector<double> xs((size_t)nnode * nd);
            std::vector<uint8_t> ds((size_t)nnode * nq);
            for (long i = 0; i < nnode; i++) {
                for (int a = 0; a < nd; a++) xs[(size_t)i * nd + a] = mesh->x[(size_t)order[i] * nd + a];
                for (int q = 0; q < nq; q++) ds[(size_t)i * nq + q] = mesh->dirichlet[(size_t)order[i] * nq + q];
            }
            cfdv_mesh mi = *mesh;
            mi.conn = conn.data();
            mi.x = xs.data();
            mi.dirichlet = ds.data();
            setup(ctx, &mi);
            ctx->d_norder = ctx->alloc<int>((size_t)nnode);
            ctx->d_npos = ctx->alloc<int>((size_t)nnode);
            HIP_CHECK_THROW(hipMemcpy(ctx->d_norder, order.data(), sizeof(int) * nnode, hipMemcpyHostToDevice));
            HIP_CHECK_THROW(hipMemcpy(ctx->d_npos, pos.data(), sizeof(int) * nnode, hipMemcpyHostToDevice));
            bool identity = true;
            for (long i = 0; i < nnode && identity; i++) identity = order[i] == (int)i;
            ctx->renum = !identity;
        } else {
            setup(ctx, mesh);
        }
    } catch (const std::invalid_argument& ex) {
        delete ctx;
        return CFDV_ERR_ARG;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "cfdv_create: %s\n", ex.what());
        delete ctx;
        return CFDV_ERR_HIP;
    }
    *out = ctx;
    return CFDV_OK;
}

void cfdv_destroy(cfdv_ctx* ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->st);
    delete ctx;
}

const char* cfdv_last_error(const cfdv_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cfdv_mesh_info(const cfdv_ctx* ctx, int64_t* n_elem, int64_t* n_node, int* ncolors, int64_t* n_slots) {
    if (!ctx) return CFDV_ERR_ARG;
    if (n_elem) *n_elem = ctx->n_elem;
    if (n_node) *n_node = ctx->n_node;
    if (ncolors) *ncolors = ctx->ncolor;
    if (n_slots) *n_slots = ctx->n_pad;
    return CFDV_OK;
}

int cfdv_rows_info(const cfdv_ctx* ctx, int64_t* n_blocks, int64_t* n_slots) {
    if (!ctx) return CFDV_ERR_ARG;
    if (n_blocks) *n_blocks = ctx->d_B64 ? ctx->n_row_blocks : 0;
    if (n_slots) *n_slots = ctx->d_B64 ? ctx->n_groups * 4 * 64 : 0;
    return CFDV_OK;
}

int cfdv_set_state(cfdv_ctx* ctx, const double* U) {
    if (!ctx || !U) return CFDV_ERR_ARG;
    GUARD_BEGIN
    ctx->upload_nodes(ctx->d_U, U, ctx->neqn);
    ctx->sync();
    ctx->built = false;
    ctx->pinv_valid = false;
    ctx->have_prev = false;
    ctx->n_old = 0;
    return CFDV_OK;
    GUARD_END
}

int cfdv_get_state(cfdv_ctx* ctx, double* U, double* U0) {
    if (!ctx) return CFDV_ERR_ARG;
    GUARD_BEGIN
    if (U) {
        ctx->download_nodes(U, ctx->d_U, ctx->neqn);
        ctx->sync();
    }
    if (U0) {
        ctx->download_nodes(U0, ctx->d_U0, ctx->neqn);
        ctx->sync();
    }
    return CFDV_OK;
    GUARD_END
}

int cfdv_cfl(cfdv_ctx* ctx, double cfl, double* dt) {
    if (!ctx || !dt) return CFDV_ERR_ARG;
    GUARD_BEGIN
    const int nb = (int)((ctx->n_pad + 255) / 256);
    launch_cfl(ctx->ndim, ctx->th, ctx->d_conn, ctx->d_U, ctx->d_cleng, ctx->n_pad, ctx->d_cflpart, nb, ctx->st);
    min_reduce(ctx->d_cflpart, nb, ctx->d_small, ctx->st);
    ctx->check_launch();
    double mn;
    HIP_CHECK_THROW(hipMemcpyAsync(&mn, ctx->d_small, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    if (!(mn < 1000)) return fail(ctx, CFDV_ERR_CFL, "no element set dt (CFL)");
    *dt = cfl * mn;
    return CFDV_OK;
    GUARD_END
}

int cfdv_build(cfdv_ctx* ctx, double dt) {
    if (!ctx) return CFDV_ERR_ARG;
    GUARD_BEGIN
    HIP_CHECK_THROW(hipMemsetAsync(ctx->d_counters, 0, sizeof(int) * 4, ctx->st));
    BuildArgs a{};
    a.th = ctx->th;
    a.ndim = ctx->ndim;
    a.n_elem_pad = ctx->n_pad;
    a.conn = ctx->d_conn;
    a.bc = ctx->d_bc;
    a.cmask = ctx->d_cmask;
    a.x = ctx->d_x;
    a.U = ctx->d_U;
    a.dt = dt;
    ctx->last_dt = dt;
    a.R64 = ctx->d_R64;
    a.R32 = ctx->d_R32;
    a.rhs = ctx->d_rhs;
    a.svals = ctx->d_s;
    a.ties = ctx->d_counters;
    a.status = ctx->d_counters + 1;
    a.bslots = ctx->d_bslots;
    a.n_bslots = ctx->n_bslots;
    a.bslots1 = ctx->d_bslots1;
    a.n_bslots1 = ctx->n_bslots1;
    a.bslotsm = ctx->d_bslotsm;
    a.n_bslotsm = ctx->n_bslotsm;
    a.adap = ctx->d_adap;
    a.hnm = ctx->d_hnm;
    a.aslots = ctx->d_aslots;
    a.n_aslots = ctx->n_aslots;
    a.r64_pairs = ctx->r64_pairs;
    a.e_lo = 0;
    a.e_hi = ctx->n_pad;
    if (ctx->rows_direct) {
        a.S64 = ctx->d_S64;
        a.pslot = ctx->d_pslot;
        a.pfirst = ctx->d_pfirst;
        a.Rb = ctx->d_Rb;
        a.bidx = ctx->d_bidx;
        launch_build_direct(a, ctx->ncolor, ctx->cstart.data(), ctx->bstart.data(), ctx->astart.data(), ctx->st);
    } else {
        launch_build(a, ctx->d_R64 != nullptr, ctx->d_R32 != nullptr, ctx->st);
    }
    if (ctx->timing) HIP_CHECK_THROW(hipEventRecord(ctx->ev[5], ctx->st));
    if (ctx->rows_direct)
        launch_rows_from_stage(ctx->ra, ctx->n_stage, ctx->d_S64, ctx->d_mirror, false, ctx->d_B64, ctx->d_B32, ctx->st);
    else if (ctx->d_B64)
        launch_rows_assemble(ctx->ra, ctx->d_R64, false, ctx->d_B64, ctx->d_B32, ctx->st);
    ctx->rows_tr = false;
    ctx->pinv_valid = false;   // the node-block inverses belong to the previous blocks
    ctx->warm_consumed = false;
    if (ctx->timing) HIP_CHECK_THROW(hipEventRecord(ctx->ev[6], ctx->st));
    ctx->check_launch();
    int cnt[2];
    HIP_CHECK_THROW(hipMemcpyAsync(cnt, ctx->d_counters, sizeof(int) * 2, hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    ctx->built = true;
    if (cnt[1] & 1) return fail(ctx, CFDV_ERR_BC, "unknown boundary-condition code");
    if (cnt[1] & 2) return fail(ctx, CFDV_ERR_GEOM, "face surface Jacobian below 1e-12");
    if (cnt[1] & 4) return fail(ctx, CFDV_ERR_HIP, "internal: an element of the one-face list has more faces");
    return CFDV_OK;
    GUARD_END
}

int cfdv_get_blocks(cfdv_ctx* ctx, double* R, double* rhs, double* s) {
    if (!ctx || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_get_blocks before cfdv_build");
    GUARD_BEGIN
    const int k = ctx->k;
    const size_t kk = (size_t)k * k;
    if (R) {
        if (!ctx->d_R64) return fail(ctx, CFDV_ERR_ARG, "no FP64 blocks");
        std::vector<double> h((size_t)ctx->n_pad * kk);
        HIP_CHECK_THROW(hipMemcpyAsync(h.data(), ctx->d_R64, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->st));
        ctx->sync();
        for (long e = 0; e < ctx->n_elem; e++) {
            const long sl = ctx->perm_inv[e];
            for (int i = 0; i < k; i++)
                for (int j = 0; j < k; j++)
                    R[(size_t)e * kk + i + (size_t)k * j] =
                        h[!ctx->r64_pairs ? blk_offset<double>(sl, i * k + j, (int)kk)
                          : ctx->ndim == 3 ? pair_offset<8, 5>(sl, i * k + j) : pair_offset<4, 4>(sl, i * k + j)];
        }
    }
    if (rhs || s) {
        std::vector<double> hr((size_t)ctx->n_pad * k), hs((size_t)ctx->n_pad * 4);
        HIP_CHECK_THROW(hipMemcpyAsync(hr.data(), ctx->d_rhs, sizeof(double) * hr.size(), hipMemcpyDeviceToHost, ctx->st));
        HIP_CHECK_THROW(hipMemcpyAsync(hs.data(), ctx->d_s, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, ctx->st));
        ctx->sync();
        for (long e = 0; e < ctx->n_elem; e++) {
            const long sl = ctx->perm_inv[e];
            if (rhs) std::memcpy(rhs + (size_t)e * k, hr.data() + (size_t)sl * k, sizeof(double) * k);
            if (s) std::memcpy(s + (size_t)e * 4, hs.data() + (size_t)sl * 4, sizeof(double) * 4);
        }
    }
    return CFDV_OK;
    GUARD_END
}

int cfdv_get_blocks32(cfdv_ctx* ctx, float* R) {
    if (!ctx || !ctx->built || !R || !ctx->d_R32) return fail(ctx, CFDV_ERR_ARG, "no FP32 blocks built");
    GUARD_BEGIN
    const int k = ctx->k;
    const size_t kk = (size_t)k * k;
    std::vector<float> h((size_t)ctx->n_pad * kk);
    HIP_CHECK_THROW(hipMemcpyAsync(h.data(), ctx->d_R32, sizeof(float) * h.size(), hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    for (long e = 0; e < ctx->n_elem; e++) {
        const long sl = ctx->perm_inv[e];
        for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++)
                R[(size_t)e * kk + i + (size_t)k * j] = h[blk_offset<float>(sl, i * k + j, (int)kk)];
    }
    return CFDV_OK;
    GUARD_END
}

int cfdv_get_rhs_vector(cfdv_ctx* ctx, double* b) {
    if (!ctx || !b || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_get_rhs_vector before cfdv_build");
    GUARD_BEGIN
    launch_compose(ctx->d_node_ptr, ctx->d_node_ent, ctx->d_rhs, ctx->k, ctx->neqn, ctx->n_node, ctx->d_b, ctx->st);
    ctx->check_launch();
    ctx->download_nodes(b, ctx->d_b, ctx->neqn);
    ctx->sync();
    return CFDV_OK;
    GUARD_END
}

int cfdv_solve(cfdv_ctx* ctx, const cfdv_solve_opts* o, cfdv_solve_stats* s) {
    if (!ctx || !o || !s) return CFDV_ERR_ARG;
    if (!ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_solve before cfdv_build");
    if (o->m < 1 || o->m > 500 || o->restarts < 1) return fail(ctx, CFDV_ERR_ARG, "bad Krylov dimension/restarts");
    if (o->precond != CFDV_PRECOND_NONE && o->precond != CFDV_PRECOND_NODE_BLOCK)
        return fail(ctx, CFDV_ERR_ARG, "unknown preconditioner");
    if (o->precond != CFDV_PRECOND_NONE && o->mode == CFDV_SOLVE_REFERENCE)
        return fail(ctx, CFDV_ERR_ARG, "the reference solve is unpreconditioned");
    if (o->warm_start < 0 || o->warm_start > 3) return fail(ctx, CFDV_ERR_ARG, "warm_start takes 0 to 3");
    std::memset(s, 0, sizeof(*s));
    GUARD_BEGIN
    switch (o->mode) {
    case CFDV_SOLVE_REFERENCE:
        if (!ctx->d_R64) return fail(ctx, CFDV_ERR_ARG, "reference solve needs FP64 blocks");
        if (o->transpose) return fail(ctx, CFDV_ERR_ARG, "reference solve has no transpose");
        {
            // the reference multiplies element by element
            struct RowsOff {
                cfdv_ctx* c;
                ~RowsOff() { c->rows_off = false; }
            } guard{ctx};
            ctx->rows_off = true;
            return solve_reference(ctx, o, s);
        }
    case CFDV_SOLVE_FP64:
        return solve_fp64(ctx, o, s);
    case CFDV_SOLVE_MIXED:
        if (o->inner_m < 2 || o->inner_m > 500 || o->inner_restarts < 1)
            return fail(ctx, CFDV_ERR_ARG, "bad inner Krylov parameters");
        return solve_mixed(ctx, o, s);
    default:
        return fail(ctx, CFDV_ERR_ARG, "unknown solve mode");
    }
    GUARD_END
}

int cfdv_reference_trace(cfdv_ctx* ctx, int32_t* m, int32_t* n_restarts, double* H, double* g, int32_t* mm) {
    if (!ctx) return CFDV_ERR_ARG;
    if (ctx->trace_mm.empty()) return fail(ctx, CFDV_ERR_ARG, "no reference-mode solve to trace");
    const size_t nr = ctx->trace_mm.size();
    if (m) *m = ctx->trace_m;
    if (n_restarts) *n_restarts = (int32_t)nr;
    if (H) std::memcpy(H, ctx->trace_H.data(), sizeof(double) * ctx->trace_H.size());
    if (g) std::memcpy(g, ctx->trace_g.data(), sizeof(double) * ctx->trace_g.size());
    if (mm) std::memcpy(mm, ctx->trace_mm.data(), sizeof(int32_t) * nr);
    return CFDV_OK;
}

int cfdv_get_du(cfdv_ctx* ctx, double* du) {
    if (!ctx || !du || !ctx->have_du) return fail(ctx, CFDV_ERR_ARG, "no solution");
    GUARD_BEGIN
    ctx->download_nodes(du, ctx->d_du, ctx->neqn);
    ctx->sync();
    return CFDV_OK;
    GUARD_END
}

int cfdv_update(cfdv_ctx* ctx) {
    if (!ctx || !ctx->have_du) return fail(ctx, CFDV_ERR_ARG, "cfdv_update before cfdv_solve");
    GUARD_BEGIN
    vec_update_state(ctx->d_U, ctx->d_U0, ctx->d_du, ctx->ndof, ctx->st);
    ctx->check_launch();
    ctx->sync();
    ctx->built = false;
    ctx->pinv_valid = false;
    ctx->have_du = false;
    ctx->n_old = ctx->n_saved;   // d_old[] hold the increments before d_du's
    ctx->n_saved = 0;
    ctx->have_prev = true;   // d_du keeps the increment for a warm start
    ctx->prev_dt = ctx->last_dt;
    return CFDV_OK;
    GUARD_END
}

int cfdv_rms(cfdv_ctx* ctx, double* urms) {
    if (!ctx || !urms) return CFDV_ERR_ARG;
    GUARD_BEGIN
    double su[5], su0[5];
    vec_sumsq_eqn(ctx->d_U, ctx->neqn, ctx->n_node, ctx->d_part, ctx->d_small, NBLK_RED, ctx->st);
    vec_sumsq_eqn(ctx->d_U0, ctx->neqn, ctx->n_node, ctx->d_part + NBLK_RED * 5, ctx->d_small + 5, NBLK_RED, ctx->st);
    ctx->check_launch();
    HIP_CHECK_THROW(hipMemcpyAsync(su, ctx->d_small, sizeof(double) * ctx->neqn, hipMemcpyDeviceToHost, ctx->st));
    HIP_CHECK_THROW(hipMemcpyAsync(su0, ctx->d_small + 5, sizeof(double) * ctx->neqn, hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    for (int i = 0; i < ctx->neqn; i++) urms[i] = (su0[i] == 0) ? 1 : std::sqrt(su[i] / su0[i]) - 1.0;
    return CFDV_OK;
    GUARD_END
}

int cfdv_refine_flags(cfdv_ctx* ctx, uint8_t* flags, double* s1, int32_t* nflagged, int32_t* nties) {
    if (!ctx || !flags) return CFDV_ERR_ARG;
    GUARD_BEGIN
    uint8_t* d_f = ctx->alloc<uint8_t>(ctx->n_elem);
    double* d_s1 = s1 ? ctx->alloc<double>(ctx->n_elem) : nullptr;
    HIP_CHECK_THROW(hipMemsetAsync(ctx->d_counters, 0, sizeof(int) * 4, ctx->st));
    launch_flags(ctx->ndim, ctx->th, ctx->d_conn, ctx->d_U, ctx->d_perm_inv, ctx->n_elem, d_f, d_s1, ctx->d_counters,
                 ctx->st);
    ctx->check_launch();
    int ties;
    HIP_CHECK_THROW(hipMemcpyAsync(flags, d_f, ctx->n_elem, hipMemcpyDeviceToHost, ctx->st));
    if (s1) HIP_CHECK_THROW(hipMemcpyAsync(s1, d_s1, sizeof(double) * ctx->n_elem, hipMemcpyDeviceToHost, ctx->st));
    HIP_CHECK_THROW(hipMemcpyAsync(&ties, ctx->d_counters, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    ctx->release(d_s1);
    ctx->release(d_f);
    if (nflagged) {
        long cnt = 0;
        for (long e = 0; e < ctx->n_elem; e++) cnt += flags[e];
        *nflagged = (int32_t)cnt;
    }
    if (nties) *nties = ties;
    return CFDV_OK;
    GUARD_END
}

int cfdv_refine_flags_noalloc(cfdv_ctx* ctx, uint8_t* flags, double* s1, int32_t* nflagged, int32_t* nties) {
    if (!ctx || !flags) return CFDV_ERR_ARG;
    if (ctx->n_elem > CFDV_FLAGS_NOALLOC_MAX_ELEM)
        return fail(ctx, CFDV_ERR_ARG,
                    "cfdv_refine_flags_noalloc: " + std::to_string(ctx->n_elem) +
                        " elements exceed CFDV_FLAGS_NOALLOC_MAX_ELEM (" +
                        std::to_string((long)CFDV_FLAGS_NOALLOC_MAX_ELEM) + "); use cfdv_refine_flags");
    if (!ctx->d_flag_scratch)
        return fail(ctx, CFDV_ERR_HIP, "cfdv_refine_flags_noalloc: no flag scratch on this device: " +
                                           ctx->flag_scratch_err);
    GUARD_BEGIN
    uint8_t* d_f = nullptr;
    double* d_s1 = nullptr;
    int* d_ties = nullptr;
    launch_flags_scratch(ctx->ndim, ctx->th, ctx->d_conn, ctx->d_U, ctx->d_perm_inv, ctx->n_elem,
                         ctx->d_flag_scratch, s1 != nullptr, ctx->st, &d_f, &d_s1, &d_ties);
    ctx->check_launch();
    int ties = 0;
    HIP_CHECK_THROW(hipMemcpyAsync(flags, d_f, ctx->n_elem, hipMemcpyDeviceToHost, ctx->st));
    if (s1) HIP_CHECK_THROW(hipMemcpyAsync(s1, d_s1, sizeof(double) * ctx->n_elem, hipMemcpyDeviceToHost, ctx->st));
    HIP_CHECK_THROW(hipMemcpyAsync(&ties, d_ties, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    if (nflagged) {
        long cnt = 0;
        for (long e = 0; e < ctx->n_elem; e++) cnt += flags[e];
        *nflagged = (int32_t)cnt;
    }
    if (nties) *nties = ties;
    return CFDV_OK;
    GUARD_END
}

int cfdv_color(cfdv_ctx* ctx, int32_t* color, int32_t* ncolors) {
    if (!ctx) return CFDV_ERR_ARG;
    if (color) std::memcpy(color, ctx->color_orig.data(), sizeof(int32_t) * ctx->n_elem);
    if (ncolors) *ncolors = ctx->ncolor;
    return CFDV_OK;
}

int cfdv_matvec(cfdv_ctx* ctx, const double* x, double* y, int precision, int transpose) {
    if (!ctx || !x || !y || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_matvec before cfdv_build");
    if (precision != 64 && precision != 32 && precision != 3) return fail(ctx, CFDV_ERR_ARG, "bad precision");
    if (!ctx->have_blocks(precision, transpose != 0))
        return fail(ctx, CFDV_ERR_ARG, precision == 64 ? "no FP64 blocks" : "no FP32 blocks");
    GUARD_BEGIN
    const long n = ctx->ndof;
    ctx->upload_nodes(ctx->d_r, x, ctx->neqn);
    if (precision == 32) {
        vec_d2f_scaled(ctx->d_w32, ctx->d_r, 1.0, n, ctx->st);
        float* yf = ctx->alloc<float>(n);
        ctx->matvec(ctx->d_w32, yf, 32, transpose != 0);
        vec_fill(ctx->d_y, 0.0, n, ctx->st);
        vec_add_f2d(ctx->d_y, yf, 1.0, n, ctx->st);
        ctx->sync();
        ctx->release(yf);
    } else {
        ctx->matvec(ctx->d_r, ctx->d_y, precision, transpose != 0);
    }
    ctx->download_nodes(y, ctx->d_y, ctx->neqn);
    ctx->sync();
    return CFDV_OK;
    GUARD_END
}

int cfdv_profile_matvec(cfdv_ctx* ctx, int precision, int reps, double* ms) {
    if (!ctx || !ms || reps < 1 || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "profile before build");
    if (precision != 64 && precision != 32) return fail(ctx, CFDV_ERR_ARG, "bad precision");
    if (!ctx->have_blocks(precision, false))
        return fail(ctx, CFDV_ERR_ARG, precision == 64 ? "no FP64 blocks" : "no FP32 blocks");
    GUARD_BEGIN
    void* x = precision == 32 ? (void*)ctx->d_w32 : (void*)ctx->d_r;
    void* y = precision == 32 ? (void*)(ctx->d_V32 ? ctx->d_V32 : nullptr) : (void*)ctx->d_y;
    float* ytmp = nullptr;
    if (precision == 32) {
        vec_fill_f(ctx->d_w32, 1.0f, ctx->ndof, ctx->st);
        if (!y) { ytmp = ctx->alloc<float>(ctx->ndof); y = ytmp; }
    } else {
        vec_fill(ctx->d_r, 1.0, ctx->ndof, ctx->st);
    }
    ctx->matvec(x, y, precision, false);   // warm
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[0], ctx->st));
    for (int i = 0; i < reps; i++) ctx->matvec(x, y, precision, false);
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[1], ctx->st));
    HIP_CHECK_THROW(hipEventSynchronize(ctx->ev[1]));
    float t = 0;
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    *ms = t / reps;
    ctx->release(ytmp);
    return CFDV_OK;
    GUARD_END
}

int cfdv_set_thermo(cfdv_ctx* ctx, const cfdv_thermo* thermo) {
    if (!ctx || !thermo) return CFDV_ERR_ARG;
    if ((thermo->adiabatic != 0) != (ctx->th.adiabatic != 0))
        return fail(ctx, CFDV_ERR_ARG, "the adiabatic flag is fixed at create (it shapes ApplyBC)");
    if (!(thermo->mach > 0) || !(thermo->gamma > 1) || !(thermo->reynolds > 0) || !(thermo->prandtl > 0) ||
        !(thermo->tinf > 0))
        return fail(ctx, CFDV_ERR_ARG, "bad thermo constants");
    ctx->th = make_thermo(thermo);
    ctx->built = false;
    ctx->pinv_valid = false;
    ctx->have_du = false;
    ctx->have_prev = false;
    ctx->n_old = 0;
    return CFDV_OK;
}

int cfdv_adjoint_solve(cfdv_ctx* ctx, const cfdv_solve_opts* o, const double* g, double* lambda,
                       cfdv_solve_stats* s) {
    if (!ctx || !o || !g || !lambda || !s) return CFDV_ERR_ARG;
    if (!ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_adjoint_solve before cfdv_build");
    if (o->mode != CFDV_SOLVE_FP64 && o->mode != CFDV_SOLVE_MIXED)
        return fail(ctx, CFDV_ERR_ARG, "the adjoint needs the FP64 or mixed solver");
    if (o->m < 1 || o->m > 500 || o->restarts < 1) return fail(ctx, CFDV_ERR_ARG, "bad Krylov dimension/restarts");
    if (o->precond != CFDV_PRECOND_NONE && o->precond != CFDV_PRECOND_NODE_BLOCK)
        return fail(ctx, CFDV_ERR_ARG, "unknown preconditioner");
    if (o->mode == CFDV_SOLVE_MIXED && (o->inner_m < 2 || o->inner_m > 500 || o->inner_restarts < 1))
        return fail(ctx, CFDV_ERR_ARG, "bad inner Krylov parameters");
    std::memset(s, 0, sizeof(*s));
    GUARD_BEGIN
    if (!ctx->d_g) {
        ctx->d_g = ctx->alloc<double>(ctx->ndof);
        ctx->d_lam = ctx->alloc<double>(ctx->ndof);
    }
    ctx->upload_nodes(ctx->d_g, g, ctx->neqn);
    cfdv_solve_opts ot = *o;
    ot.transpose = 1;
    const int rc = (o->mode == CFDV_SOLVE_FP64) ? solve_fp64_rhs(ctx, &ot, s, ctx->d_g, ctx->d_lam)
                                                : solve_mixed_rhs(ctx, &ot, s, ctx->d_g, ctx->d_lam);
    ctx->download_nodes(lambda, ctx->d_lam, ctx->neqn);
    ctx->sync();
    return rc;
    GUARD_END
}

int cfdv_adjoint_solve_multi(cfdv_ctx* ctx, const cfdv_solve_opts* o, const double* G, double* Lambda, int nvec,
                             double* rel_resid, cfdv_solve_stats* s) {
    if (!ctx || !o || !G || !Lambda || !rel_resid || !s) return CFDV_ERR_ARG;
    if (nvec != 8) return fail(ctx, CFDV_ERR_ARG, "cfdv_adjoint_solve_multi takes 8 right-hand sides");
    if (!ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_adjoint_solve_multi before cfdv_build");
    if (ctx->ndim != 3) return fail(ctx, CFDV_ERR_ARG, "the 8-vector product is implemented for hex8 meshes");
    if (o->mode != CFDV_SOLVE_MIXED) return fail(ctx, CFDV_ERR_ARG, "the 8-vector solve is the mixed solver");
    if (o->restarts < 1 || o->inner_m < 2 || o->inner_m > MAX_M1 - 1 || !(o->rtol > 0) || !(o->inner_rtol > 0))
        return fail(ctx, CFDV_ERR_ARG, "bad sweeps, inner Krylov dimension or tolerances");
    if (o->precond != CFDV_PRECOND_NONE && o->precond != CFDV_PRECOND_NODE_BLOCK)
        return fail(ctx, CFDV_ERR_ARG, "unknown preconditioner");
    std::memset(s, 0, sizeof(*s));
    GUARD_BEGIN
    ctx->ensure_multi(std::min(o->inner_m, MAX_M1 - 1));
    const size_t n = (size_t)ctx->ndof;
    for (int k = 0; k < 8; k++) ctx->upload_nodes(ctx->d_G8 + k * n, G + k * n, ctx->neqn);
    const int rc = solve_multi8(ctx, o, true, ctx->d_G8, ctx->d_L8, rel_resid, s);
    for (int k = 0; k < 8; k++) {
        ctx->download_nodes(Lambda + k * n, ctx->d_L8 + k * n, ctx->neqn);
        ctx->sync();
    }
    return rc;
    GUARD_END
}

int cfdv_build_vjp(cfdv_ctx* ctx, const double* lambda, const double* du, double* gU, double* gmach) {
    if (!ctx || !lambda || !gU || !gmach) return CFDV_ERR_ARG;
    if (ctx->ndim != 3) return fail(ctx, CFDV_ERR_ARG, "build tangents are implemented for hex8 meshes");
    if (!ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_build_vjp before cfdv_build");
    if (!du && !ctx->have_du) return fail(ctx, CFDV_ERR_ARG, "no du given and none solved");
    GUARD_BEGIN
    if (!ctx->d_gelem) {
        ctx->d_gelem = ctx->alloc<double>((size_t)ctx->n_pad * 41);
        ctx->d_duv = ctx->alloc<double>(ctx->ndof);
        ctx->d_gU = ctx->alloc<double>(ctx->ndof);
        if (!ctx->d_lam) {
            ctx->d_g = ctx->alloc<double>(ctx->ndof);
            ctx->d_lam = ctx->alloc<double>(ctx->ndof);
        }
    }
    ctx->upload_nodes(ctx->d_lam, lambda, ctx->neqn);
    if (du)
        ctx->upload_nodes(ctx->d_duv, du, ctx->neqn);
    else
        vec_copy(ctx->d_duv, ctx->d_du, ctx->ndof, ctx->st);
    BuildArgs a{};
    a.th = ctx->th;
    a.ndim = ctx->ndim;
    a.n_elem_pad = ctx->n_pad;
    a.conn = ctx->d_conn;
    a.bc = ctx->d_bc;
    a.cmask = ctx->d_cmask;
    a.x = ctx->d_x;
    a.U = ctx->d_U;
    a.dt = ctx->last_dt;
    a.adap = ctx->d_adap;
    a.hnm = ctx->d_hnm;
    launch_build_vjp(a, make_tables(), ctx->d_lam, ctx->d_duv, ctx->d_gelem, ctx->st);
    launch_vjp_gather(ctx->d_node_ptr, ctx->d_node_ent, ctx->d_gelem, ctx->n_node, ctx->d_gU, ctx->st);
    launch_column_sum(ctx->d_gelem, ctx->n_pad, 41, 40, ctx->d_part, NBLK_RED, ctx->d_small + 1000, ctx->st);
    ctx->check_launch();
    ctx->download_nodes(gU, ctx->d_gU, ctx->neqn);
    HIP_CHECK_THROW(hipMemcpyAsync(gmach, ctx->d_small + 1000, sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    ctx->sync();
    return CFDV_OK;
    GUARD_END
}

// Y = K X (or K^T X) for 8 vectors at once with the FP32 blocks
static void matvec_multi_dev(cfdv_ctx* ctx, const float* X, float* Y, bool tr) {
    if (ctx->d_B32) {   // over the assembled rows
        ctx->orient_rows(tr);
        launch_rows_matvec8(ctx->ra, ctx->d_B32, X, Y, ctx->st);
        ctx->check_launch();
        return;
    }
    vec_fill_f(Y, 0.0f, ctx->ndof * 8, ctx->st);
    for (int c = 0; c < ctx->ncolor; c++) {
        MatvecArgs a{ctx->ndim, ctx->nnod, ctx->neqn, ctx->cstart[c], ctx->cstart[c + 1], ctx->d_conn, 0};
        launch_matvec_multi_color(a, ctx->d_R32, X, Y, tr, ctx->st);
    }
    ctx->check_launch();
}

int cfdv_matvec_multi(cfdv_ctx* ctx, const float* X, float* Y, int nvec, int transpose) {
    if (!ctx || !X || !Y || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "cfdv_matvec_multi before cfdv_build");
    if (nvec != 8) return fail(ctx, CFDV_ERR_ARG, "cfdv_matvec_multi takes 8 vectors");
    if (ctx->ndim != 3 || !(ctx->d_R32 || ctx->d_B32))
        return fail(ctx, CFDV_ERR_ARG, "needs hex8 FP32 blocks or node rows");
    GUARD_BEGIN
    const size_t n = (size_t)ctx->ndof * 8;
    float* dX = ctx->alloc<float>(n);
    float* dY = ctx->alloc<float>(n);
    ctx->upload_nodes32(dX, X, 8 * ctx->neqn);
    matvec_multi_dev(ctx, dX, dY, transpose != 0);
    ctx->download_nodes32(Y, dY, 8 * ctx->neqn);
    ctx->sync();
    ctx->release(dX);
    ctx->release(dY);
    return CFDV_OK;
    GUARD_END
}

int cfdv_profile_matvec_multi(cfdv_ctx* ctx, int reps, double* ms) {
    if (!ctx || !ms || reps < 1 || !ctx->built) return fail(ctx, CFDV_ERR_ARG, "profile before build");
    if (ctx->ndim != 3 || !(ctx->d_R32 || ctx->d_B32))
        return fail(ctx, CFDV_ERR_ARG, "needs hex8 FP32 blocks or node rows");
    GUARD_BEGIN
    const size_t n = (size_t)ctx->ndof * 8;
    float* dX = ctx->alloc<float>(n);
    float* dY = ctx->alloc<float>(n);
    vec_fill_f(dX, 1.0f, (long)n, ctx->st);
    matvec_multi_dev(ctx, dX, dY, false);   // warm
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[0], ctx->st));
    for (int i = 0; i < reps; i++) matvec_multi_dev(ctx, dX, dY, false);
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[1], ctx->st));
    HIP_CHECK_THROW(hipEventSynchronize(ctx->ev[1]));
    float t = 0;
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    *ms = t / reps;
    ctx->release(dX);
    ctx->release(dY);
    return CFDV_OK;
    GUARD_END
}

int cfdv_step(cfdv_ctx* ctx, double cfl, const cfdv_solve_opts* o, cfdv_step_stats* st) {
    if (!ctx || !o || !st) return CFDV_ERR_ARG;
    std::memset(st, 0, sizeof(*st));
    GUARD_BEGIN
    // product timing is on for this step only, whichever way it returns
    struct TimingOff {
        cfdv_ctx* c;
        ~TimingOff() {
            c->timing = false;
            c->mv_marks.clear();
            c->ev_used = 0;
        }
    } timing_guard{ctx};
    ctx->timing = true;
    ctx->mv_marks.clear();
    ctx->ev_used = 0;
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[0], ctx->st));
    int rc = cfdv_cfl(ctx, cfl, &st->dt);
    if (rc) return rc;
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[1], ctx->st));
    rc = cfdv_build(ctx, st->dt);
    if (rc) return rc;
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[2], ctx->st));
    rc = cfdv_solve(ctx, o, &st->solve);
    ctx->timing = false;   // update and RMS are not products
    if (rc) return rc;
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[3], ctx->st));
    rc = cfdv_update(ctx);
    if (rc) return rc;
    rc = cfdv_rms(ctx, st->urms);
    if (rc) return rc;
    HIP_CHECK_THROW(hipEventRecord(ctx->ev[4], ctx->st));
    HIP_CHECK_THROW(hipEventSynchronize(ctx->ev[4]));
    float t;
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    st->ms_cfl = t;
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[1], ctx->ev[2]));
    st->ms_build = t;
    if (ctx->d_B64) {
        HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[5], ctx->ev[6]));
        st->ms_assemble = t;
        st->ms_build -= t;
    }
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
    st->ms_solve = t;
    HIP_CHECK_THROW(hipEventElapsedTime(&t, ctx->ev[3], ctx->ev[4]));
    st->ms_update = t;
    ctx->collect_matvec_times(&st->ms_matvec64, &st->ms_matvec32, &st->n_matvec64, &st->n_matvec32);
    return CFDV_OK;
    GUARD_END
}

}  // extern "C"
