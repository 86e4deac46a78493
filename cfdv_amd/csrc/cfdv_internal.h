// Internal declarations shared by the HIP translation units of libcfdv_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "fdv_device.h"

#define HIP_CHECK_THROW(x)                                                               \
    do {                                                                                 \
        hipError_t err_ = (x);                                                           \
        if (err_ != hipSuccess)                                                          \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(err_)); \
    } while (0)

namespace cfdv {

struct GaussTables;

struct BuildArgs {
    DevThermo th;
    int ndim;
    long n_elem_pad;          // color-sorted, padded element count
    const int* conn;          // [n_elem_pad][nnod], -1 in padding slots
    const int* bc;            // [n_elem_pad][nface]
    const unsigned long long* cmask;  // [n_elem_pad] constrained (Dirichlet) dofs
    const double* x;          // [n_node][ndim]
    const double* U;          // [n_node][neqn]
    double dt;
    double* R64;              // chunked SoA blocks (pair_offset layout with
                              // r64_pairs) or nullptr
    float* R32;
    double* rhs;              // [n_elem_pad][k]
    double* svals;            // [n_elem_pad][4] or nullptr
    int* ties;                // count of FDVParam r == 0.01 ties
    int* status;              // bit 0: unknown bc code, bit 1: bad face detjac
    const int* bslots;        // 3D: slots with faces or constrained dofs and no
    long n_bslots;            //     Hnm (face kernel), ascending
    const int* bslots1;       // the same with at most one face with a bc code
    long n_bslots1;
    const int* bslotsm;       // and with more
    long n_bslotsm;
    const unsigned char* adap;  // [n_elem_pad] hanging-node element flags or nullptr
    const double* hnm;        // [n_elem_pad][64] Hnm(N, M) at N + 8 M (adapted slots)
    const int* aslots;        // 3D adapted slots (full per-lane kernel), ascending
    long n_aslots;
    int r64_pairs;            // R64 in pair_offset order (contexts with rows)
    long e_lo, e_hi;          // interior kernel: the slot range it builds
    // rows built in place (no element blocks; R64 and R32 null): each
    // block is added into the staged FP64 row slot S64 + pslot[e * 64 + pair]
    // * 25 (stored when bit pair of pfirst[e] is set), and the face kernel's
    // elements keep their domain part in Rb at bidx[e]
    double* S64;
    const int* pslot;
    const unsigned long long* pfirst;
    double* Rb;
    const int* bidx;
};

void upload_gauss_tables(const GaussTables& g);
// discrete adjoint of the hex8 build: gelem [n_elem_pad][41] = d phi_e /
// d (U of the element's nodes, Mach), phi_e = lam_e^T (rhs_e - R_e du_e)
void launch_build_vjp(const BuildArgs& a, const GaussTables& G, const double* lam, const double* du, double* gelem,
                      hipStream_t st);
void launch_vjp_gather(const int* node_ptr, const int* node_ent, const double* gelem, long n_node, double* gU,
                       hipStream_t st);
// out = sum_i a[i * stride + offset], deterministic (part: nblk doubles)
void launch_column_sum(const double* a, long n, int stride, int offset, double* part, int nblk, double* out,
                       hipStream_t st);
void launch_build(const BuildArgs& a, bool w64, bool w32, hipStream_t st);
// hex8 build into the staged rows, color after color (cstart: ncolor + 1
// slot bounds; bstart, astart: the colors' ranges of a.bslots, a.aslots)
void launch_build_direct(const BuildArgs& a, int ncolor, const long* cstart, const long* bstart, const long* astart,
                         hipStream_t st);
void launch_flags(int ndim, const DevThermo& th, const int* conn, const double* U, const int* perm_inv,
                  long n_elem, uint8_t* flags, double* s1, int* ties, hipStream_t st);
// The allocation-free flag pass works in a fixed device array (build.hip):
// FLAGS_SCRATCH_BYTES per GPU, held for the life of the process; layout per
// call: [0, 256) counters, n_elem flag bytes at 256, n_elem doubles of s1
// after the flags rounded up to 256 bytes.
constexpr size_t FLAGS_SCRATCH_BYTES = (size_t)32 << 20;
constexpr long FLAGS_NOALLOC_MAX_ELEM = 3700000;
static_assert(256 + ((FLAGS_NOALLOC_MAX_ELEM + 255) / 256) * 256 + 8 * FLAGS_NOALLOC_MAX_ELEM <=
                  (long)FLAGS_SCRATCH_BYTES,
              "flag scratch too small for FLAGS_NOALLOC_MAX_ELEM");
// device address of the scratch array on the current device (resolved once
// per device with hipGetSymbolAddress); throws if it cannot be resolved
void* flags_scratch();
// the flag pass into the scratch at scratch_dev (n_elem <=
// FLAGS_NOALLOC_MAX_ELEM): counters cleared by a kernel, flags and (want_s1)
// s1 written there; returns their device addresses and the ties counter's
void launch_flags_scratch(int ndim, const DevThermo& th, const int* conn, const double* U, const int* perm_inv,
                          long n_elem, void* scratch_dev, bool want_s1, hipStream_t st, uint8_t** flags_dev,
                          double** s1_dev, int** ties_dev);
void launch_cfl(int ndim, const DevThermo& th, const int* conn, const double* U, const double* cleng,
                long n_elem_pad, double* part, int nblocks, hipStream_t st);
void launch_length(int ndim, const int* conn, const double* x, long n_elem_pad, double* area, double* cleng,
                   hipStream_t st);

// ---- solve kernels (solve.hip) ----
struct MatvecArgs {
    int ndim, nnod, neqn;
    long e_begin, e_end;      // color range (multiple of 64)
    const int* conn;          // [n_elem_pad][nnod]
    int r64_pairs;            // FP64 blocks in pair_offset order
};
// y += sum over elements of one color of R_e x_e (or R_e^T x_e)
// Kernels that take a gate (a device int) return at once when *gate != 0:
// the Arnoldi step kernels of the restarted GMRES, launched one step ahead of
// the host's convergence test, do no work after the step that converged.
void launch_matvec_color(const MatvecArgs& a, const double* R64, const float* R32, const void* x, void* y,
                         int prec, bool transpose, hipStream_t st, const int* gate = nullptr);
// hex8 EBE product with 8 right-hand sides, FP32 blocks/vectors, vectors
// [n_node*neqn][8]; one color range per launch, Y accumulated
void launch_matvec_multi_color(const MatvecArgs& a, const float* R32, const float* X, float* Y, bool tr,
                               hipStream_t st);
void launch_compose(const int* node_ptr, const int* node_ent, const double* rhs, int k, int neqn, long n_node,
                    double* b, hipStream_t st);

// vector kernels; all reductions write per-block partials in a fixed order
// and finish with a single-block reduction, so results are deterministic.
void vec_fill(double* y, double v, long n, hipStream_t st, const int* gate = nullptr);
void vec_set(double* p, double v, hipStream_t st);   // *p = v, stream-ordered
void vec_fill_f(float* y, float v, long n, hipStream_t st, const int* gate = nullptr);
void vec_copy(double* y, const double* x, long n, hipStream_t st);
void vec_axpy(double* y, const double* x, const double* alpha_dev, double scale, long n, hipStream_t st);
void vec_sub(double* r, const double* b, const double* y, long n, hipStream_t st);  // r = b - y
void vec_scale_to(double* y, const double* x, const double* denom_dev, long n, hipStream_t st);  // y = x / *d
void vec_d2f_scaled(float* y, const double* x, double s, long n, hipStream_t st);
void vec_add_f2d(double* y, const float* x, double s, long n, hipStream_t st);
void vec_update_state(double* U, double* U0, const double* du, long n, hipStream_t st);
// dots: out[j] = <V_j, w> for j < nv (V rows of length n, stride ldv)
template <typename T>
void vec_multidot(const T* V, long ldv, int nv, const T* w, long n, double* part, double* out, int nblk,
                  hipStream_t st, const int* gate = nullptr);
// Classical Gram-Schmidt passes on 16-byte loads (V_j = V + j ldv, ldv a
// multiple of 64, nv <= 32): mode 0 out[j] = <V_j, w>; mode 1 w -= sum h_j
// V_j then out[j] = <V_j, w>; mode 2 w -= sum h_j V_j then out[0] = <w, w>.
// part holds nblk * nv doubles.
template <typename T>
void vec_cgs(int mode, T* w, const T* V, long ldv, int nv, const double* h, long n, double* part, int nblk,
             double* out, hipStream_t st, const int* gate = nullptr);
template <typename T>
void vec_multi_axpy(T* w, const T* V, long ldv, int nv, const double* h, double sign, long n, hipStream_t st,
                    const int* gate = nullptr);
template <typename T>
void vec_scale_inv(T* y, const T* x, const double* denom_dev, long n, hipStream_t st, const int* gate = nullptr);
template <typename T>
void vec_combine(double* x, const T* V, long ldv, int nv, const double* y, long n, hipStream_t st);
// node-block Jacobi: Pinv per node (entry r neqn + c), FP64 and FP32 copies,
// in chunks of 64 nodes: node n entry q at (n / 64) 64 neqn^2 + q 64 + n % 64
// (buffers hold round_up(n_node, 64) nodes); *status set to 1 when a node
// block is singular
void launch_nodeblock(const int* node_ptr, const int* node_ent, const double* R64, const float* R32, int nnod,
                      int neqn, long n_node, double* P64, float* P32, int* status, hipStream_t st);
// z = Pinv v (accumulate: z += Pinv v; trans: Pinv^T)
template <typename TV, typename TP>
void vec_nodeblock_apply(const TP* P, const TV* v, TV* z, int neqn, long n_node, bool accumulate, bool trans,
                         hipStream_t st, const int* gate = nullptr);
// v = w / *d and z = P v (node blocks) in one pass
template <typename T>
void vec_nodeblock_scale_apply(const T* P, const T* w, const double* d, T* v, T* z, int neqn, long n_node,
                               bool trans, hipStream_t st, const int* gate = nullptr);
void vec_sumsq_eqn(const double* U, int neqn, long n_node, double* part, double* out, int nblk, hipStream_t st);

// ---- GMRES least squares on the device (solve.hip) ----
// H: (m+1) x m Hessenberg, column-major, leading dimension ldh = m+1; g: the
// rotated right-hand side [m+1]; cs, sn: Givens rotations [m]; y [m].
// Per Arnoldi step j the device writes a GmresStep to host-visible memory.
struct GmresStep {
    double est;   // |g_{j+1}|: the residual estimate after step j
    int stop;     // the cycle ends after step j (converged, invariant space or failure)
    int fail;     // |w| not finite
};
// g = (beta, 0, ..., 0), *gate = 0 (start of a cycle)
void launch_gmres_start(double* g, int m, double beta, int* gate, hipStream_t st);
// step j: column j of H = h1 (+ h2) with H(j+1, j) = sqrt(*nrm2), rotated by the
// previous rotations; rotation j; g rotated; *hn_out = sqrt(*nrm2); *gate and
// out[j].stop = est <= tol, or |w| == 0, or |w| not finite.  Gated itself.
void launch_givens_step(double* H, int ldh, double* g, double* cs, double* sn, const double* h1, const double* h2,
                        const double* nrm2, int j, double tol, double* hn_out, int* gate, GmresStep* out,
                        hipStream_t st);
// y = R^-1 g for the leading n x n upper triangle of the rotated H
void launch_backsub(const double* H, int ldh, const double* g, int n, double* y, hipStream_t st);
// The reference GMRES (MPI_GMRES.h at one rank, quirks Q2-Q3) with H on the
// device: column j of H from the modified Gram-Schmidt dots hc[0..j] and
// |w|^2 = hc[j+1] (:154-181; *breakdown = 1 when |w| is not positive), then
// rotate_and_check's estimate (:696-730, angles from the unrotated H):
// out[j].est = |g0 prod_i (-s_i)|, out[j].stop = est < tol; *hn_out = |w|.
void launch_ref_arnoldi(double* H, int ldh, const double* hc, int j, double g0, double tol, double* hn_out,
                        int* breakdown, GmresStep* out, hipStream_t st);
// rotate (:733-752) on a copy Hr of H's leading (mm+1) x mm block and g in
// place, then the back substitution reading h transposed (:756-780, Q2)
// into y [mm]
void launch_ref_finish(const double* H, double* Hr, int ldh, double* g, int mm, double* y, hipStream_t st);
void min_reduce(const double* part, int n, double* out, hipStream_t st);

// ---- assembled node rows (rows.hip) ----
// Sliced block rows: chunks of 64 nodes, groups of 4 slots (neighbour
// nodes) per chunk; block values 16 bytes per lane at a time.
struct RowsArgs {
    long n_node, n_chunks;
    int nnod, neqn;
    const long* goff;      // [n_chunks + 1] first group of each chunk
    const int* cols;       // [n_groups * 4][64] column node of each slot
    const long* coff;      // [n_groups * 4 + 1] contributor rows per slot
    const int* contrib;    // [coff[end]][64] element block codes, -1 = none
    int r64_pairs;         // the FP64 element blocks are in pair_offset order
};
// y = K x over the rows; prec 64 (FP64 rows and vectors), 32 (FP32), 3 (FP32
// rows, FP64 vectors).  y is written, not accumulated.
void launch_rows_matvec(const RowsArgs& a, const double* B64, const float* B32, const void* x, void* y, int prec,
                        hipStream_t st, const int* gate = nullptr);
// Y = K X for 8 vectors over the FP32 rows, X and Y [n_node * neqn][8]
void launch_rows_matvec8(const RowsArgs& a, const float* B32, const float* X, float* Y, hipStream_t st);
// rows (FP64 and/or FP32, either may be null) from the FP64 element blocks:
// those of K, or with tr those of K^T (the same slots and sums, each block
// transposed), over which launch_rows_matvec then forms K^T x
void launch_rows_assemble(const RowsArgs& a, const double* R64, bool tr, double* B64, float* B32, hipStream_t st);
// the rows (tr: those of K^T, through the mirrored slots) laid out from the
// staged FP64 rows [n_stage][neqn^2]
void launch_rows_from_stage(const RowsArgs& a, long n_stage, const double* S64, const int* mirror, bool tr,
                            double* B64, float* B32, hipStream_t st);
// node-block inverses (as launch_nodeblock) from the diagonal slots of the
// FP64 rows (diag[n]: global slot index of node n's own column; tr: the
// rows hold K^T); the slot sums are the element blocks' in element order,
// so P equals launch_nodeblock's bit for bit
void launch_nodeblock_rows(const RowsArgs& a, const long* diag, const double* B64, bool tr, double* P64, float* P32,
                           int* status, hipStream_t st);

// ---- eight vectors side by side ([n][8]) (solve.hip) ----
struct Scale8 { double v[8]; };
// V[i][s] = (float)(R[s][i] scale[s]), R vector-major [8][n]
void vec_multivec_pack(float* V, const double* R, const Scale8& scale, long n, hipStream_t st);
// L[s][i] += a[s] X[i][s]
void vec_multivec_unpack_axpy(double* L, const double* X, const Scale8& a, long n, hipStream_t st);
// node-block Jacobi on each of the 8 vectors of v [n_node * neqn][8]
template <typename TV, typename TP>
void vec_nodeblock_apply8(const TP* P, const TV* v, TV* z, int neqn, long n_node, bool accumulate, bool trans,
                          hipStream_t st);

// out[i q + c] = in[map[i] q + c] (node permutation, q values per node)
template <typename T>
void vec_gather_nodes(T* out, const T* in, const int* map, long n_node, int q, hipStream_t st);

// ---- coloring (solve.hip) ----
void launch_color_round(const int* conn, int nnod, const int* node_ptr, const int* node_elem, long n_elem,
                        const unsigned long long* prio, const int* color_in, int* color_out, int* changed,
                        hipStream_t st);

// node-block Jacobi storage: node n entry q in chunks of 64 nodes
template <int NEQ>
__host__ __device__ __forceinline__ size_t pblk_offset(long n, int q) {
    return (size_t)(n >> 6) * 64 * NEQ * NEQ + (size_t)q * 64 + (size_t)(n & 63);
}

// Gauss-Jordan inverse with partial pivoting of one node block A (FP64),
// stored in FP64 and FP32 at node n; a singular block sets *status = 1 and
// stores the identity
template <int NEQ>
__device__ __forceinline__ void nodeblock_invert_store(double (&A)[NEQ][NEQ], long n, double* __restrict__ P64,
                                                       float* __restrict__ P32, int* __restrict__ status) {
    double B[NEQ][NEQ];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int c = 0; c < NEQ; c++) B[r][c] = (r == c) ? 1.0 : 0.0;
    bool singular = false;
#pragma unroll
    for (int c = 0; c < NEQ; c++) {
        // pivot: largest |A[r][c]| for r >= c (rows swapped by value)
        int piv = c;
        double best = fabs(A[c][c]);
#pragma unroll
        for (int r = c + 1; r < NEQ; r++)
            if (fabs(A[r][c]) > best) { best = fabs(A[r][c]); piv = r; }
        if (!(best > 0.0)) singular = true;
#pragma unroll
        for (int r = c + 1; r < NEQ; r++)
            if (r == piv) {
#pragma unroll
                for (int t = 0; t < NEQ; t++) {
                    double a = A[c][t]; A[c][t] = A[r][t]; A[r][t] = a;
                    double b = B[c][t]; B[c][t] = B[r][t]; B[r][t] = b;
                }
            }
        const double inv = 1.0 / A[c][c];
#pragma unroll
        for (int t = 0; t < NEQ; t++) { A[c][t] *= inv; B[c][t] *= inv; }
#pragma unroll
        for (int r = 0; r < NEQ; r++)
            if (r != c) {
                const double f = A[r][c];
#pragma unroll
                for (int t = 0; t < NEQ; t++) { A[r][t] -= f * A[c][t]; B[r][t] -= f * B[c][t]; }
            }
    }
    if (singular) *status = 1;
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int c = 0; c < NEQ; c++) {
            const double v = singular ? ((r == c) ? 1.0 : 0.0) : B[r][c];
            P64[pblk_offset<NEQ>(n, r * NEQ + c)] = v;
            P32[pblk_offset<NEQ>(n, r * NEQ + c)] = (float)v;
        }
}

}  // namespace cfdv
