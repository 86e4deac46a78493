// EBE matvec and Krylov vector kernels (gfx950).
//
// The EBE matvec restates MPI_GMRES::sum_Av / sum_Ax at one rank
// (src/Utility/Solvers/MPI_GMRES.h:442-693): y[gx*neqn+xx] +=
// R(x*neqn+xx, y*neqn+yy) v[gy*neqn+yy].  Elements are sorted by color; a
// launch covers one color, in which no two elements share a node, so the
// scatter into y needs no atomics and the result is deterministic.
#include "cfdv_internal.h"

namespace cfdv {

// ---------------------------------------------------------------------------
// EBE matvec, one thread per element.  Each lane gathers its element's k
// inputs, streams the k*k block as 16-byte vectors (blk_offset layout: one
// 1 KB contiguous read per wavefront per vector group), accumulates all k
// outputs in registers and scatters them once.  The loop is fully unrolled
// so every register index is static and loads can be issued far ahead of
// their use.
// ---------------------------------------------------------------------------
typedef double v2d_t __attribute__((ext_vector_type(2)));
typedef float v4f_t __attribute__((ext_vector_type(4)));
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef v2d_t type; };
template <> struct Vec16<float> { typedef v4f_t type; };

template <typename TB, typename TV, typename TA, int NNOD, int NEQ, bool TR, bool PAIRS = false>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 2))) void matvec_kernel(long e0, long e1, const int* __restrict__ conn,
                                                     const TB* __restrict__ R, const TV* __restrict__ x,
                                                     TV* __restrict__ y, const int* __restrict__ gate) {
    constexpr int K = NNOD * NEQ, KK = K * K, V = 16 / (int)sizeof(TB), GPR = K / V;   // vector groups per row
    typedef typename Vec16<TB>::type vec_t;
    if (gate && *gate) return;
    const long e = e0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e1) return;
    const int* ce = conn + e * NNOD;
    if (ce[0] < 0) return;
    if constexpr (PAIRS) {
        // blocks in node-pair order (pair_base): scalar loads, the same
        // products summed in the same order as the chunked path
        // pair (n, m) of this slot at pair_base(e, n NNOD + m)
        if (!TR) {
            TA xv[K];
#pragma unroll
            for (int m = 0; m < NNOD; m++) {
                const int gm = ce[m];
#pragma unroll
                for (int q = 0; q < NEQ; q++) xv[m * NEQ + q] = (TA)x[(size_t)gm * NEQ + q];
            }
            for (int n = 0; n < NNOD; n++) {
                const int gn = ce[n];
#pragma unroll
                for (int r = 0; r < NEQ; r++) {
                    TA acc = 0;
#pragma unroll
                    for (int m = 0; m < NNOD; m++)
#pragma unroll
                        for (int s = 0; s < NEQ; s++)
                            acc += (TA)R[pair_base<NNOD, NEQ>(e, n * NNOD + m) + r * NEQ + s] * xv[m * NEQ + s];
                    TV* yp = y + (size_t)gn * NEQ + r;
                    *yp = (TV)((TA)*yp + acc);
                }
            }
        } else {
            TA yv[K];
#pragma unroll
            for (int j = 0; j < K; j++) yv[j] = 0;
            for (int n = 0; n < NNOD; n++) {
                const int gn = ce[n];
                TA xi[NEQ];
#pragma unroll
                for (int r = 0; r < NEQ; r++) xi[r] = (TA)x[(size_t)gn * NEQ + r];
#pragma unroll
                for (int r = 0; r < NEQ; r++)
#pragma unroll
                    for (int m = 0; m < NNOD; m++)
#pragma unroll
                        for (int s = 0; s < NEQ; s++)
                            yv[m * NEQ + s] += (TA)R[pair_base<NNOD, NEQ>(e, n * NNOD + m) + r * NEQ + s] * xi[r];
            }
#pragma unroll
            for (int m = 0; m < NNOD; m++) {
                const int gm = ce[m];
#pragma unroll
                for (int q = 0; q < NEQ; q++) {
                    TV* yp = y + (size_t)gm * NEQ + q;
                    *yp = (TV)((TA)*yp + yv[m * NEQ + q]);
                }
            }
        }
        return;
    }
    const vec_t* Rb = reinterpret_cast<const vec_t*>(R + (size_t)(e >> 6) * 64 * KK) + (e & 63);
    if (!TR) {
        // y_block(n) = sum_j R(5n+r, j) x_j : all k inputs in registers, one
        // node's NEQ rows (NEQ*GPR independent 16-byte loads) per iteration
        TA xv[K];
#pragma unroll
        for (int m = 0; m < NNOD; m++) {
            const int gm = ce[m];
#pragma unroll
            for (int q = 0; q < NEQ; q++) xv[m * NEQ + q] = (TA)x[(size_t)gm * NEQ + q];
        }
        for (int n = 0; n < NNOD; n++) {
            vec_t rv[NEQ][GPR];
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int c = 0; c < GPR; c++)
                    rv[r][c] = __builtin_nontemporal_load(Rb + (size_t)((n * NEQ + r) * GPR + c) * 64);
            const int gn = ce[n];
#pragma unroll
            for (int r = 0; r < NEQ; r++) {
                TA acc = 0;
#pragma unroll
                for (int c = 0; c < GPR; c++)
#pragma unroll
                    for (int v = 0; v < V; v++) acc += (TA)rv[r][c][v] * xv[c * V + v];
                TV* yp = y + (size_t)gn * NEQ + r;
                *yp = (TV)((TA)*yp + acc);
            }
        }
    } else {
        // y_j += sum_i R(i, j) x_i : all k outputs in registers
        TA yv[K];
#pragma unroll
        for (int j = 0; j < K; j++) yv[j] = 0;
        for (int n = 0; n < NNOD; n++) {
            vec_t rv[NEQ][GPR];
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int c = 0; c < GPR; c++)
                    rv[r][c] = __builtin_nontemporal_load(Rb + (size_t)((n * NEQ + r) * GPR + c) * 64);
            const int gn = ce[n];
            TA xi[NEQ];
#pragma unroll
            for (int r = 0; r < NEQ; r++) xi[r] = (TA)x[(size_t)gn * NEQ + r];
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int c = 0; c < GPR; c++)
#pragma unroll
                    for (int v = 0; v < V; v++) yv[c * V + v] += (TA)rv[r][c][v] * xi[r];
        }
#pragma unroll
        for (int m = 0; m < NNOD; m++) {
            const int gm = ce[m];
#pragma unroll
            for (int q = 0; q < NEQ; q++) {
                TV* yp = y + (size_t)gm * NEQ + q;
                *yp = (TV)((TA)*yp + yv[m * NEQ + q]);
            }
        }
    }
}

void launch_matvec_color(const MatvecArgs& a, const double* R64, const float* R32, const void* x, void* y,
                         int prec, bool tr, hipStream_t st, const int* gate) {
    const long n = a.e_end - a.e_begin;
    if (n <= 0) return;
    const int bs = 128;
    const int nb = (int)((n + bs - 1) / bs);
#define L(TB, TV, TA, NN, NQ, T)                                                                      \
    hipLaunchKernelGGL((matvec_kernel<TB, TV, TA, NN, NQ, T>), dim3(nb), dim3(bs), 0, st, a.e_begin, \
                       a.e_end, a.conn, (const TB*)(prec == 64 ? (const void*)R64 : (const void*)R32),  \
                       (const TV*)x, (TV*)y, gate)
#define LP(NN, NQ, T)                                                                                      \
    hipLaunchKernelGGL((matvec_kernel<double, double, double, NN, NQ, T, true>), dim3(nb), dim3(bs), 0, st, \
                       a.e_begin, a.e_end, a.conn, R64, (const double*)x, (double*)y, gate)
    if (prec == 64 && a.r64_pairs) {
        if (a.ndim == 3) { if (tr) LP(8, 5, true); else LP(8, 5, false); }
        else { if (tr) LP(4, 4, true); else LP(4, 4, false); }
        return;
    }
#undef LP
    if (a.ndim == 3) {
        if (prec == 64) { if (tr) L(double, double, double, 8, 5, true); else L(double, double, double, 8, 5, false); }
        else if (prec == 32) { if (tr) L(float, float, float, 8, 5, true); else L(float, float, float, 8, 5, false); }
        else { if (tr) L(float, double, double, 8, 5, true); else L(float, double, double, 8, 5, false); }
    } else {
        if (prec == 64) { if (tr) L(double, double, double, 4, 4, true); else L(double, double, double, 4, 4, false); }
        else if (prec == 32) { if (tr) L(float, float, float, 4, 4, true); else L(float, float, float, 4, 4, false); }
        else { if (tr) L(float, double, double, 4, 4, true); else L(float, double, double, 4, 4, false); }
    }
#undef L
}

// ---------------------------------------------------------------------------
// Y = K X (or K^T X) for 8 vectors over the FP32 element blocks of one color
// (the EBE form, for contexts without node rows), on the FP32 matrix cores.
// Per element it is the dense product of the 40 x 40 block with the 40 x 8
// gathered vectors, computed transposed, Y_e^T = X_e^T R_e^T (or X_e^T R_e),
// as rank-1 updates of 4 x 4 tiles: v_mfma_f32_4x4x1f32 runs 16 independent
// 4 x 4 x 1 products, one per element, so a wavefront holds 16 elements of
// the color, 4 lanes each (block b = lane / 4; lane 4 b + q supplies A row q
// = vector 4 vb + q of X_e^T, B column q = block row 4 cb + q, and receives
// D column q = the 4 vectors 4 vb .. 4 vb + 3 of block row 4 cb + q).  Every
// 4 x 4 tile is full (40 and 8 are multiples of 4), 20 tiles per element,
// accumulated over the 40 inner indices.  K x reads each lane's block row as
// ten 16-byte pieces of the chunked layout (four inner indices per load); K^T
// x reads block column 4 cb + q, one entry per load.  Same-color elements
// share no node, so the Y rows are read, added to and written without
// atomics.  The summation order differs from the single product's (tiles of
// four vectors accumulate over the inner index in order).
// ---------------------------------------------------------------------------
template <bool TR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void matvec_multi_kernel(long e0, long e1, const int* __restrict__ conn,
                                                           const float* __restrict__ R,
                                                           const float* __restrict__ X, float* __restrict__ Y) {
    constexpr int NNOD = 8, NEQ = 5, K = NNOD * NEQ, KK = K * K, NV = 8;
    const int lane = threadIdx.x & 63;
    const int b = lane >> 2, q = lane & 3;
    const long e = e0 + ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16 + b;   // 16-aligned within a chunk
    if (e0 + ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16 >= e1) return;   // whole wavefront
    const bool live = e < e1 && conn[e * NNOD] >= 0;   // padding slots hold zero blocks
    int cn[NNOD];
#pragma unroll
    for (int m = 0; m < NNOD; m++) cn[m] = live ? conn[e * NNOD + m] : 0;
    const float* Re = R + (size_t)(e >> 6) * 64 * KK + (size_t)(e & 63) * 4;   // piece p at Re + 256 p
    v4f_t acc[10][2];
#pragma unroll
    for (int cb = 0; cb < 10; cb++)
#pragma unroll
        for (int vb = 0; vb < 2; vb++) acc[cb][vb] = v4f_t{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int kp = 0; kp < 10; kp++) {
        // B: entries (4 cb + q, 4 kp + kk) of R_e (K x) or (4 kp + kk, 4 cb + q) (K^T x)
        v4f_t bv[10];
        if (!TR) {
#pragma unroll
            for (int cb = 0; cb < 10; cb++)
                bv[cb] = __builtin_nontemporal_load(reinterpret_cast<const v4f_t*>(Re + (size_t)256 * ((4 * cb + q) * 10 + kp)));
        } else {
#pragma unroll
            for (int cb = 0; cb < 10; cb++)
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
                    bv[cb][kk] = __builtin_nontemporal_load(Re + (size_t)256 * ((4 * kp + kk) * 10 + cb) + q);
        }
        // A: vectors 4 vb + q of the gathered X_e at inner index 4 kp + kk
        float av[4][2];
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int k = 4 * kp + kk;
            const int m = k / NEQ, d = k - NEQ * (k / NEQ);
            const float* xp = X + ((size_t)cn[m] * NEQ + d) * NV + q;
#pragma unroll
            for (int vb = 0; vb < 2; vb++) av[kk][vb] = live ? xp[4 * vb] : 0.0f;
        }
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int cb = 0; cb < 10; cb++)
#pragma unroll
                for (int vb = 0; vb < 2; vb++)
                    acc[cb][vb] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[kk][vb], bv[cb][kk], acc[cb][vb], 0, 0, 0);
    }
    if (!live) return;
    // lane q holds row 4 cb + q of Y_e, vectors 4 vb .. 4 vb + 3
#pragma unroll
    for (int cb = 0; cb < 10; cb++) {
        const int r = 4 * cb + q;
        const int m = r / NEQ, d = r - NEQ * (r / NEQ);
        v4f_t* yp = reinterpret_cast<v4f_t*>(Y + ((size_t)cn[m] * NEQ + d) * NV);
#pragma unroll
        for (int vb = 0; vb < 2; vb++) yp[vb] += acc[cb][vb];
    }
}

void launch_matvec_multi_color(const MatvecArgs& a, const float* R32, const float* X, float* Y, bool tr,
                               hipStream_t st) {
    // 64 elements (four wavefronts of 16) per workgroup; color ranges start
    // on a multiple of 64
    const long n = a.e_end - a.e_begin;
    if (n <= 0) return;
    const int nb = (int)((n + 63) / 64);
    if (tr)
        hipLaunchKernelGGL((matvec_multi_kernel<true>), dim3(nb), dim3(256), 0, st, a.e_begin, a.e_end, a.conn, R32, X, Y);
    else
        hipLaunchKernelGGL((matvec_multi_kernel<false>), dim3(nb), dim3(256), 0, st, a.e_begin, a.e_end, a.conn, R32, X, Y);
}

// ---------------------------------------------------------------------------
// compose_rhs (MPI_GMRES.h:369-389) as a node-centric gather: for each node
// the contributing (element, local node) pairs are listed in element order,
// so b is summed in the same order as the reference's element loop.
// ---------------------------------------------------------------------------
__global__ void compose_kernel(const int* __restrict__ node_ptr, const int* __restrict__ node_ent,
                               const double* __restrict__ rhs, int k, int neqn, long n_node,
                               double* __restrict__ b) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_node) return;
    const int nnod = k / neqn;
    double acc[5] = {0, 0, 0, 0, 0};
    const int p0 = node_ptr[n], p1 = node_ptr[n + 1];
    // the first 8 entries (every node of a hex8 mesh without hanging
    // nodes) are fetched before any is summed, so their loads overlap
    constexpr int PF = 8;
    long off[PF];
#pragma unroll
    for (int i = 0; i < PF; i++) {
        const long ent = (p0 + i < p1) ? node_ent[p0 + i] : 0;   // e * nnod + local
        off[i] = (ent / nnod) * k + (ent % nnod) * neqn;
    }
    double v[PF][5];
#pragma unroll
    for (int i = 0; i < PF; i++)
#pragma unroll
        for (int q = 0; q < 5; q++) v[i][q] = (p0 + i < p1 && q < neqn) ? rhs[off[i] + q] : 0.0;
#pragma unroll
    for (int i = 0; i < PF; i++)
        if (p0 + i < p1) {
#pragma unroll
            for (int q = 0; q < 5; q++) acc[q] += v[i][q];
        }
    for (int p = p0 + PF; p < p1; p++) {
        const long ent = node_ent[p];
        const long e = ent / nnod;
        const int l = (int)(ent % nnod);
        for (int q = 0; q < neqn; q++) acc[q] += rhs[e * k + l * neqn + q];
    }
    for (int q = 0; q < neqn; q++) b[n * neqn + q] = acc[q];
}

void launch_compose(const int* node_ptr, const int* node_ent, const double* rhs, int k, int neqn, long n_node,
                    double* b, hipStream_t st) {
    const int nb = (int)((n_node + 255) / 256);
    hipLaunchKernelGGL(compose_kernel, dim3(nb), dim3(256), 0, st, node_ptr, node_ent, rhs, k, neqn, n_node, b);
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
static inline int grid_for(long n, int bs = 256, int cap = 8192) {
    long nb = (n + bs - 1) / bs;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return (int)nb;
}

__global__ void fill_kernel(double* y, double v, long n, const int* gate) {
    if (gate && *gate) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}
__global__ void fill_f_kernel(float* y, float v, long n, const int* gate) {
    if (gate && *gate) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}
__global__ void copy_kernel(double* y, const double* x, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = x[i];
}
__global__ void axpy_kernel(double* y, const double* x, const double* alpha, double scale, long n) {
    const double a = *alpha * scale;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}
__global__ void sub_kernel(double* r, const double* b, const double* y, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r[i] = b[i] - y[i];
}
__global__ void scale_to_kernel(double* y, const double* x, const double* d, long n) {
    const double dd = *d;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = x[i] / dd;
}
__global__ void d2f_kernel(float* y, const double* x, double s, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (float)(x[i] * s);
}
__global__ void add_f2d_kernel(double* y, const float* x, double s, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] += s * (double)x[i];
}
__global__ void update_kernel(double* U, double* U0, const double* du, long n) {
    // MPI_GMRES::update at one rank (MPI_GMRES.h:246-253)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double u = U[i];
        U0[i] = u;
        U[i] = u + du[i];
    }
}

__global__ void set_kernel(double* p, double v) { *p = v; }
void vec_set(double* p, double v, hipStream_t st) { hipLaunchKernelGGL(set_kernel, dim3(1), dim3(1), 0, st, p, v); }
void vec_fill(double* y, double v, long n, hipStream_t st, const int* gate) {
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, v, n, gate);
}
void vec_fill_f(float* y, float v, long n, hipStream_t st, const int* gate) {
    hipLaunchKernelGGL(fill_f_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, v, n, gate);
}
void vec_copy(double* y, const double* x, long n, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, x, n);
}
void vec_axpy(double* y, const double* x, const double* alpha, double scale, long n, hipStream_t st) {
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, x, alpha, scale, n);
}
void vec_sub(double* r, const double* b, const double* y, long n, hipStream_t st) {
    hipLaunchKernelGGL(sub_kernel, dim3(grid_for(n)), dim3(256), 0, st, r, b, y, n);
}
void vec_scale_to(double* y, const double* x, const double* d, long n, hipStream_t st) {
    hipLaunchKernelGGL(scale_to_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, x, d, n);
}
void vec_d2f_scaled(float* y, const double* x, double s, long n, hipStream_t st) {
    hipLaunchKernelGGL(d2f_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, x, s, n);
}
void vec_add_f2d(double* y, const float* x, double s, long n, hipStream_t st) {
    hipLaunchKernelGGL(add_f2d_kernel, dim3(grid_for(n)), dim3(256), 0, st, y, x, s, n);
}
void vec_update_state(double* U, double* U0, const double* du, long n, hipStream_t st) {
    hipLaunchKernelGGL(update_kernel, dim3(grid_for(n)), dim3(256), 0, st, U, U0, du, n);
}

// multi-dot: part[blk*nv + j] = sum over this block's slice of V_j . w
template <typename T, int NV>
__global__ __launch_bounds__(256) void multidot_kernel(const T* __restrict__ V, long ldv, int nv,
                                                       const T* __restrict__ w, long n, double* __restrict__ part,
                                                       const int* __restrict__ gate) {
    if (gate && *gate) return;
    __shared__ double sm[NV][256 / 64];
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double wi = (double)w[i];
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) acc[j] += (double)V[(size_t)j * ldv + i] * wi;
    }
    // wavefront reduction (64 lanes) then across the 4 waves of the block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double v = acc[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sm[j][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV && (int)threadIdx.x < nv) {
        const int j = threadIdx.x;
        part[(size_t)blockIdx.x * nv + j] = ((sm[j][0] + sm[j][1]) + sm[j][2]) + sm[j][3];
    }
}

// out[j] = sum over blocks of part[blk*nv + j], fixed order
// out[j] = sum_b part[b nv + j]: one block per j, each thread a strided
// partial sum, then a fixed tree
__global__ void finish_dot_kernel(const double* __restrict__ part, int nblk, int nv, double* __restrict__ out,
                                  const int* __restrict__ gate) {
    if (gate && *gate) return;
    __shared__ double sm[256];
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += part[(size_t)b * nv + j];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sm[0];
}

template <typename T>
void vec_multidot(const T* V, long ldv, int nv, const T* w, long n, double* part, double* out, int nblk,
                  hipStream_t st, const int* gate) {
    for (int j0 = 0; j0 < nv; j0 += 16) {
        const int c = (nv - j0 < 16) ? nv - j0 : 16;
        if (c <= 1)
            hipLaunchKernelGGL((multidot_kernel<T, 1>), dim3(nblk), dim3(256), 0, st, V + (size_t)j0 * ldv, ldv, c, w, n, part, gate);
        else if (c <= 4)
            hipLaunchKernelGGL((multidot_kernel<T, 4>), dim3(nblk), dim3(256), 0, st, V + (size_t)j0 * ldv, ldv, c, w, n, part, gate);
        else
            hipLaunchKernelGGL((multidot_kernel<T, 16>), dim3(nblk), dim3(256), 0, st, V + (size_t)j0 * ldv, ldv, c, w, n, part, gate);
        hipLaunchKernelGGL(finish_dot_kernel, dim3(c), dim3(256), 0, st, part, nblk, c, out + j0, gate);
    }
}
template void vec_multidot<double>(const double*, long, int, const double*, long, double*, double*, int, hipStream_t,
                                   const int*);
template void vec_multidot<float>(const float*, long, int, const float*, long, double*, double*, int, hipStream_t,
                                  const int*);

// ---------------------------------------------------------------------------
// Classical Gram-Schmidt kernels on 16-byte vector loads (V_j = V + j ldv,
// ldv a multiple of 64 elements, w 256-byte aligned; the n % VL tail is done
// by scalar loads).  Each block writes its partial sums (fixed grid, fixed
// order) for finish_dot_kernel, so results do not depend on timing.
//   mode 0: part[b][j] = sum_i V_j[i] w[i]
//   mode 1: w -= sum_j h[j] V_j,  part[b][j] = sum_i V_j[i] w_new[i]
//   mode 2: w -= sum_j h[j] V_j,  part[b][0] = sum_i w_new[i]^2
// ---------------------------------------------------------------------------
// entries per load: 16 bytes, except in the FP32 fused projection (mode 1)
// over 5..16 vectors, which holds every basis entry of the thread in
// registers and loads 8 bytes to stay at 3 waves per SIMD (145 VGPRs; 16-byte
// loads there need 247)
template <typename T, int L> struct VecT { typedef T type __attribute__((ext_vector_type(L))); };
template <typename T, int NV, int MODE> constexpr int cgs_lanes() {
    return sizeof(T) == 8 ? 2 : ((MODE == 1 && NV > 4 && NV <= 16) ? 2 : 4);
}

// part[b stride + j] = block b's sum of acc[j], j < nv
template <int NV>
__device__ __forceinline__ void block_partials(double (&acc)[NV], int nv, int stride, double* __restrict__ part) {
    __shared__ double sm[NV][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        double v = acc[j];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sm[j][wv] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv && threadIdx.x < NV) {
        const int j = threadIdx.x;
        part[(size_t)blockIdx.x * stride + j] = ((sm[j][0] + sm[j][1]) + sm[j][2]) + sm[j][3];
    }
}

template <typename T, int NV, int MODE>   // MODE 0: dot, 1: axpy + dot, 2: axpy + norm
__global__ __launch_bounds__(256) void cgs_kernel(T* __restrict__ w, const T* __restrict__ V, long ldv, int nv,
                                                  const double* __restrict__ h, long n, double* __restrict__ part,
                                                  const int* __restrict__ gate) {
    if (gate && *gate) return;
    constexpr int L = cgs_lanes<T, NV, MODE>();
    typedef typename VecT<T, L>::type vt;
    double hh[NV];
    if (MODE > 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) hh[j] = (j < nv) ? h[j] : 0.0;
    }
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.0;
    const long nvec = n / L;
    const long stride = (long)gridDim.x * blockDim.x;
#pragma unroll 1
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
        vt wv = reinterpret_cast<const vt*>(w)[i];
        if constexpr (MODE == 1) {
            // the basis entries are loaded once and held in registers for
            // both the projection and the dot products
            vt vv[NV];
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) vv[j] = reinterpret_cast<const vt*>(V + (size_t)j * ldv)[i];
            double a[L];
#pragma unroll
            for (int l = 0; l < L; l++) a[l] = 0.0;
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) {
#pragma unroll
                    for (int l = 0; l < L; l++) a[l] += hh[j] * (double)vv[j][l];
                }
#pragma unroll
            for (int l = 0; l < L; l++) wv[l] = (T)((double)wv[l] - a[l]);
            reinterpret_cast<vt*>(w)[i] = wv;
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) {
#pragma unroll
                    for (int l = 0; l < L; l++) acc[j] += (double)vv[j][l] * (double)wv[l];
                }
        } else if constexpr (MODE == 2) {
            double a[L];
#pragma unroll
            for (int l = 0; l < L; l++) a[l] = 0.0;
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) {
                    const vt vj = reinterpret_cast<const vt*>(V + (size_t)j * ldv)[i];
#pragma unroll
                    for (int l = 0; l < L; l++) a[l] += hh[j] * (double)vj[l];
                }
#pragma unroll
            for (int l = 0; l < L; l++) wv[l] = (T)((double)wv[l] - a[l]);
            reinterpret_cast<vt*>(w)[i] = wv;
#pragma unroll
            for (int l = 0; l < L; l++) acc[0] += (double)wv[l] * (double)wv[l];
        } else {
#pragma unroll
            for (int j = 0; j < NV; j++)
                if (j < nv) {
                    const vt vj = reinterpret_cast<const vt*>(V + (size_t)j * ldv)[i];
#pragma unroll
                    for (int l = 0; l < L; l++) acc[j] += (double)vj[l] * (double)wv[l];
                }
        }
    }
    // tail elements (n % L) by the first threads of block 0
    if (blockIdx.x == 0) {
        const long i = nvec * L + threadIdx.x;
        if (i < n) {
            double wi = (double)w[i];
            if (MODE > 0) {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < NV; j++)
                    if (j < nv) a += hh[j] * (double)V[(size_t)j * ldv + i];
                const T wn = (T)(wi - a);
                w[i] = wn;
                wi = (double)wn;
            }
            if (MODE == 2) acc[0] += wi * wi;
            else {
#pragma unroll
                for (int j = 0; j < NV; j++)
                    if (j < nv) acc[j] += (double)V[(size_t)j * ldv + i] * wi;
            }
        }
    }
    block_partials<NV>(acc, MODE == 2 ? 1 : nv, MODE == 2 ? 1 : nv, part);
}

template <typename T>
static void launch_cgs(int mode, T* w, const T* V, long ldv, int nv, const double* h, long n, double* part,
                       int nblk, double* out, hipStream_t st, const int* gate) {
#define CGS(NVV, MODE) \
    hipLaunchKernelGGL((cgs_kernel<T, NVV, MODE>), dim3(nblk), dim3(256), 0, st, w, V, ldv, nv, h, n, part, gate)
#define CGS_NV(MODE)                                   \
    if (nv <= 1) CGS(1, MODE);                         \
    else if (nv <= 4) CGS(4, MODE);                    \
    else if (nv <= 8) CGS(8, MODE);                    \
    else if (nv <= 16) CGS(16, MODE);                  \
    else CGS(32, MODE);
    if (mode == 0) { CGS_NV(0) }
    else if (mode == 1) { CGS_NV(1) }
    else { CGS_NV(2) }
#undef CGS_NV
#undef CGS
    const int cnt = mode == 2 ? 1 : nv;
    hipLaunchKernelGGL(finish_dot_kernel, dim3(cnt), dim3(256), 0, st, part, nblk, cnt, out, gate);
}

template <typename T>
void vec_cgs(int mode, T* w, const T* V, long ldv, int nv, const double* h, long n, double* part, int nblk,
             double* out, hipStream_t st, const int* gate) {
    if (nv > 32) throw std::runtime_error("vec_cgs: at most 32 basis vectors per call");
    launch_cgs<T>(mode, w, V, ldv, nv, h, n, part, nblk, out, st, gate);
}
template void vec_cgs<double>(int, double*, const double*, long, int, const double*, long, double*, int, double*,
                              hipStream_t, const int*);
template void vec_cgs<float>(int, float*, const float*, long, int, const double*, long, double*, int, double*,
                             hipStream_t, const int*);

// w -= sign * sum_j h[j] V_j  (h on device)
template <typename T>
__global__ __launch_bounds__(256) void multi_axpy_kernel(T* __restrict__ w, const T* __restrict__ V, long ldv, int nv,
                                                         const double* __restrict__ h, double sign, long n,
                                                         const int* __restrict__ gate) {
    if (gate && *gate) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < nv; j++) acc += h[j] * (double)V[(size_t)j * ldv + i];
        w[i] = (T)((double)w[i] - sign * acc);
    }
}
template <typename T>
void vec_multi_axpy(T* w, const T* V, long ldv, int nv, const double* h, double sign, long n, hipStream_t st,
                    const int* gate) {
    hipLaunchKernelGGL((multi_axpy_kernel<T>), dim3(grid_for(n)), dim3(256), 0, st, w, V, ldv, nv, h, sign, n, gate);
}
template void vec_multi_axpy<double>(double*, const double*, long, int, const double*, double, long, hipStream_t,
                                     const int*);
template void vec_multi_axpy<float>(float*, const float*, long, int, const double*, double, long, hipStream_t,
                                    const int*);

template <typename T>
__global__ void scale_inv_kernel(T* y, const T* x, const double* d, long n, const int* gate) {
    if (gate && *gate) return;
    const double dd = *d;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (T)((double)x[i] / dd);
}
template <typename T>
void vec_scale_inv(T* y, const T* x, const double* d, long n, hipStream_t st, const int* gate) {
    hipLaunchKernelGGL((scale_inv_kernel<T>), dim3(grid_for(n)), dim3(256), 0, st, y, x, d, n, gate);
}
template void vec_scale_inv<double>(double*, const double*, const double*, long, hipStream_t, const int*);
template void vec_scale_inv<float>(float*, const float*, const double*, long, hipStream_t, const int*);

// x += sum_j y[j] V_j
template <typename T>
__global__ void combine_kernel(double* __restrict__ x, const T* __restrict__ V, long ldv, int nv,
                               const double* __restrict__ y, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < nv; j++) acc += y[j] * (double)V[(size_t)j * ldv + i];
        x[i] += acc;
    }
}
template <typename T>
void vec_combine(double* x, const T* V, long ldv, int nv, const double* y, long n, hipStream_t st) {
    hipLaunchKernelGGL((combine_kernel<T>), dim3(grid_for(n)), dim3(256), 0, st, x, V, ldv, nv, y, n);
}
template void vec_combine<double>(double*, const double*, long, int, const double*, long, hipStream_t);
template void vec_combine<float>(double*, const float*, long, int, const double*, long, hipStream_t);

// ---------------------------------------------------------------------------
// GMRES least squares on the device, in FP64 whatever the basis precision.
// One workgroup; the serial parts run on lane 0 (at most a few hundred
// flops per Arnoldi step).  The host reads only the per-step GmresStep
// records, from host-visible memory, one step behind the stream.
// ---------------------------------------------------------------------------
__global__ void gmres_start_kernel(double* __restrict__ g, int m, double beta, int* __restrict__ gate) {
    for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = (i == 0) ? beta : 0.0;
    if (threadIdx.x == 0) *gate = 0;
}
void launch_gmres_start(double* g, int m, double beta, int* gate, hipStream_t st) {
    hipLaunchKernelGGL(gmres_start_kernel, dim3(1), dim3(64), 0, st, g, m, beta, gate);
}

__global__ void givens_step_kernel(double* __restrict__ H, int ldh, double* __restrict__ g, double* __restrict__ cs,
                                   double* __restrict__ sn, const double* __restrict__ h1,
                                   const double* __restrict__ h2, const double* __restrict__ nrm2, int j,
                                   double tol, double* __restrict__ hn_out, int* __restrict__ gate,
                                   GmresStep* __restrict__ out) {
    if (*gate) return;
    double* Hj = H + (size_t)ldh * j;
    for (int i = threadIdx.x; i <= j; i += blockDim.x) Hj[i] = h2 ? h1[i] + h2[i] : h1[i];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double hn = sqrt(*nrm2);
    const int fail = !isfinite(hn);
    Hj[j + 1] = hn;
    for (int i = 0; i < j; i++) {
        const double a = Hj[i], b = Hj[i + 1];
        Hj[i] = cs[i] * a + sn[i] * b;
        Hj[i + 1] = -sn[i] * a + cs[i] * b;
    }
    const double mag = sqrt(Hj[j] * Hj[j] + Hj[j + 1] * Hj[j + 1]);
    cs[j] = Hj[j] / mag;
    sn[j] = Hj[j + 1] / mag;
    Hj[j] = mag;
    Hj[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    const double est = fabs(g[j + 1]);
    const int stop = (est <= tol) || fail || !(hn > 0);
    *hn_out = hn;
    *gate = stop;
    out[j].est = est;
    out[j].stop = stop;
    out[j].fail = fail;
    __threadfence_system();
}
void launch_givens_step(double* H, int ldh, double* g, double* cs, double* sn, const double* h1, const double* h2,
                        const double* nrm2, int j, double tol, double* hn_out, int* gate, GmresStep* out,
                        hipStream_t st) {
    hipLaunchKernelGGL(givens_step_kernel, dim3(1), dim3(64), 0, st, H, ldh, g, cs, sn, h1, h2, nrm2, j, tol, hn_out,
                       gate, out);
}

__global__ void backsub_kernel(const double* __restrict__ H, int ldh, const double* __restrict__ g, int n,
                               double* __restrict__ y) {
    if (threadIdx.x != 0) return;
    for (int j = n - 1; j >= 0; j--) {
        double v = g[j];
        for (int i = j + 1; i < n; i++) v -= H[j + (size_t)ldh * i] * y[i];
        y[j] = v / H[j + (size_t)ldh * j];
    }
}
void launch_backsub(const double* H, int ldh, const double* g, int n, double* y, hipStream_t st) {
    hipLaunchKernelGGL(backsub_kernel, dim3(1), dim3(64), 0, st, H, ldh, g, n, y);
}

// The reference's arithmetic, unfused as the host compiler leaves it.
__global__ void ref_arnoldi_kernel(double* __restrict__ H, int ldh, const double* __restrict__ hc, int j, double g0,
                                   double tol, double* __restrict__ hn_out, int* __restrict__ breakdown,
                                   GmresStep* __restrict__ out) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    double* Hj = H + (size_t)ldh * j;
    for (int i = 0; i <= j; i++) Hj[i] = hc[i];
    const double hn = sqrt(hc[j + 1]);
    Hj[j + 1] = hn;
    if (!(hn > 0)) *breakdown = 1;
    *hn_out = hn;
    // gg[i+1] = -s_i gg[i] with c_i, s_i from the unrotated H (Q3)
    double gg = g0;
    for (int i = 0; i <= j; i++) {
        const double hii = H[i + (size_t)ldh * i], hi1 = H[i + 1 + (size_t)ldh * i];
        const double mag = sqrt(hii * hii + hi1 * hi1);
        const double ss = hi1 / mag;
        gg = -ss * gg;
    }
    const double est = fabs(gg);
    out[j].est = est;
    out[j].stop = est < tol;
    out[j].fail = 0;
    __threadfence_system();
}
void launch_ref_arnoldi(double* H, int ldh, const double* hc, int j, double g0, double tol, double* hn_out,
                        int* breakdown, GmresStep* out, hipStream_t st) {
    hipLaunchKernelGGL(ref_arnoldi_kernel, dim3(1), dim3(64), 0, st, H, ldh, hc, j, g0, tol, hn_out, breakdown, out);
}

__global__ void ref_finish_kernel(const double* __restrict__ H, double* __restrict__ Hr, int ldh,
                                  double* __restrict__ g, int mm, double* __restrict__ y) {
#pragma clang fp contract(off)
    __shared__ double cssn[2];
    for (int jj = 0; jj < mm; jj++)
        for (int i = threadIdx.x; i <= mm; i += blockDim.x) Hr[i + (size_t)ldh * jj] = H[i + (size_t)ldh * jj];
    __syncthreads();
    for (int i = 0; i < mm; i++) {
        if (threadIdx.x == 0) {
            const double hii = Hr[i + (size_t)ldh * i], hi1 = Hr[i + 1 + (size_t)ldh * i];
            const double mag = sqrt(hii * hii + hi1 * hi1);
            cssn[0] = hii / mag;
            cssn[1] = hi1 / mag;
        }
        __syncthreads();
        const double cc = cssn[0], ss = cssn[1];
        for (int jj = threadIdx.x; jj < mm; jj += blockDim.x) {
            const double h0 = Hr[i + (size_t)ldh * jj], h1 = Hr[i + 1 + (size_t)ldh * jj];
            Hr[i + (size_t)ldh * jj] = cc * h0 + ss * h1;
            Hr[i + 1 + (size_t)ldh * jj] = -ss * h0 + cc * h1;
        }
        if (threadIdx.x == 0) {
            g[i + 1] = -ss * g[i];
            g[i] = cc * g[i];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int j = 0; j < mm; j++) y[j] = 0.0;
    for (int j = mm - 1; j >= 0; j--) {
        double v = g[j];
        for (int i = mm - 1; i >= 0; i--)
            if (i != j) v -= Hr[i + (size_t)ldh * j] * y[i];
        y[j] = v / Hr[j + (size_t)ldh * j];
    }
}
void launch_ref_finish(const double* H, double* Hr, int ldh, double* g, int mm, double* y, hipStream_t st) {
    hipLaunchKernelGGL(ref_finish_kernel, dim3(1), dim3(64), 0, st, H, Hr, ldh, g, mm, y);
}

// ---------------------------------------------------------------------------
// Node-block Jacobi preconditioner.  P_n = sum over the (element, local node)
// pairs at node n of the element's diagonal block R_e(l, l) (neqn x neqn),
// summed in element order; Pinv_n by Gauss-Jordan with partial pivoting in
// FP64.  One thread per node.  The same node CSR as compose_kernel.
// ---------------------------------------------------------------------------
// Inverse blocks are stored in chunks of 64 nodes, entry-major inside a
// chunk: entry q = r NEQ + c of node n at (n / 64) 64 NEQ^2 + q 64 + n % 64,
// so that a wavefront's load of one entry for 64 consecutive nodes is one
// contiguous 512-byte (FP64) or 256-byte (FP32) segment.

template <typename TB, int NEQ>
__global__ __launch_bounds__(256) void nodeblock_kernel(const int* __restrict__ node_ptr,
                                                        const int* __restrict__ node_ent,
                                                        const TB* __restrict__ R, int nnod, long n_node,
                                                        double* __restrict__ P64, float* __restrict__ P32,
                                                        int* __restrict__ status) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_node) return;
    const int k = nnod * NEQ, KK = k * k;
    double A[NEQ][NEQ];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int c = 0; c < NEQ; c++) A[r][c] = 0.0;
    for (int p = node_ptr[n]; p < node_ptr[n + 1]; p++) {
        const long ent = node_ent[p];   // slot * nnod + local node
        const long e = ent / nnod;
        const int l = (int)(ent % nnod);
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int c = 0; c < NEQ; c++) {
                const int q = (l * NEQ + r) * k + (l * NEQ + c);   // R(i, j) at i k + j
                A[r][c] += (double)R[blk_offset<TB>(e, q, KK)];
            }
    }
    nodeblock_invert_store<NEQ>(A, n, P64, P32, status);
}

void launch_nodeblock(const int* node_ptr, const int* node_ent, const double* R64, const float* R32, int nnod,
                      int neqn, long n_node, double* P64, float* P32, int* status, hipStream_t st) {
    const int nb = (int)((n_node + 255) / 256);
    if (neqn == 5) {
        if (R64) hipLaunchKernelGGL((nodeblock_kernel<double, 5>), dim3(nb), dim3(256), 0, st, node_ptr, node_ent, R64, nnod, n_node, P64, P32, status);
        else hipLaunchKernelGGL((nodeblock_kernel<float, 5>), dim3(nb), dim3(256), 0, st, node_ptr, node_ent, R32, nnod, n_node, P64, P32, status);
    } else {
        if (R64) hipLaunchKernelGGL((nodeblock_kernel<double, 4>), dim3(nb), dim3(256), 0, st, node_ptr, node_ent, R64, nnod, n_node, P64, P32, status);
        else hipLaunchKernelGGL((nodeblock_kernel<float, 4>), dim3(nb), dim3(256), 0, st, node_ptr, node_ent, R32, nnod, n_node, P64, P32, status);
    }
}

// z[n] = Pinv_n v[n] (or z[n] += Pinv_n v[n]; Pinv_n^T with trans);
// arithmetic in the type of z
template <typename TV, typename TP, int NEQ>
__global__ __launch_bounds__(256) void nodeblock_apply_kernel(const TP* __restrict__ P, const TV* __restrict__ v,
                                                              TV* __restrict__ z, long n_node, int accumulate,
                                                              int trans, const int* __restrict__ gate) {
    if (gate && *gate) return;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < n_node; n += (long)gridDim.x * blockDim.x) {
        TV x[NEQ];
#pragma unroll
        for (int c = 0; c < NEQ; c++) x[c] = v[n * NEQ + c];
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            TV acc = accumulate ? z[n * NEQ + r] : (TV)0;
#pragma unroll
            for (int c = 0; c < NEQ; c++)
                acc += (TV)P[pblk_offset<NEQ>(n, trans ? c * NEQ + r : r * NEQ + c)] * x[c];
            z[n * NEQ + r] = acc;
        }
    }
}

template <typename TV, typename TP>
void vec_nodeblock_apply(const TP* P, const TV* v, TV* z, int neqn, long n_node, bool accumulate, bool trans,
                         hipStream_t st, const int* gate) {
    const int nb = grid_for(n_node);
    if (neqn == 5)
        hipLaunchKernelGGL((nodeblock_apply_kernel<TV, TP, 5>), dim3(nb), dim3(256), 0, st, P, v, z, n_node,
                           (int)accumulate, (int)trans, gate);
    else
        hipLaunchKernelGGL((nodeblock_apply_kernel<TV, TP, 4>), dim3(nb), dim3(256), 0, st, P, v, z, n_node,
                           (int)accumulate, (int)trans, gate);
}
template void vec_nodeblock_apply<double, double>(const double*, const double*, double*, int, long, bool, bool,
                                                  hipStream_t, const int*);
template void vec_nodeblock_apply<float, float>(const float*, const float*, float*, int, long, bool, bool,
                                                hipStream_t, const int*);
template void vec_nodeblock_apply<double, float>(const float*, const double*, double*, int, long, bool, bool,
                                                 hipStream_t, const int*);

// The next Arnoldi vector and its preconditioned image in one pass:
// v = w / d (as scale_inv_kernel) and z = P v (as nodeblock_apply_kernel,
// same order), so the right-preconditioned step reads w once instead of w
// and then v.
template <typename T, int NEQ>
__global__ __launch_bounds__(256) void nodeblock_scale_apply_kernel(const T* __restrict__ P, const T* __restrict__ w,
                                                                    const double* __restrict__ d, T* __restrict__ v,
                                                                    T* __restrict__ z, long n_node, int trans,
                                                                    const int* __restrict__ gate) {
    if (gate && *gate) return;
    const double dd = *d;
    for (long n = (long)blockIdx.x * blockDim.x + threadIdx.x; n < n_node; n += (long)gridDim.x * blockDim.x) {
        T x[NEQ];
#pragma unroll
        for (int c = 0; c < NEQ; c++) {
            x[c] = (T)((double)w[n * NEQ + c] / dd);
            v[n * NEQ + c] = x[c];
        }
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            T acc = (T)0;
#pragma unroll
            for (int c = 0; c < NEQ; c++) acc += P[pblk_offset<NEQ>(n, trans ? c * NEQ + r : r * NEQ + c)] * x[c];
            z[n * NEQ + r] = acc;
        }
    }
}

template <typename T>
void vec_nodeblock_scale_apply(const T* P, const T* w, const double* d, T* v, T* z, int neqn, long n_node,
                               bool trans, hipStream_t st, const int* gate) {
    const int nb = grid_for(n_node);
    if (neqn == 5)
        hipLaunchKernelGGL((nodeblock_scale_apply_kernel<T, 5>), dim3(nb), dim3(256), 0, st, P, w, d, v, z, n_node,
                           (int)trans, gate);
    else
        hipLaunchKernelGGL((nodeblock_scale_apply_kernel<T, 4>), dim3(nb), dim3(256), 0, st, P, w, d, v, z, n_node,
                           (int)trans, gate);
}
template void vec_nodeblock_scale_apply<double>(const double*, const double*, const double*, double*, double*, int,
                                                long, bool, hipStream_t, const int*);
template void vec_nodeblock_scale_apply<float>(const float*, const float*, const double*, float*, float*, int, long,
                                               bool, hipStream_t, const int*);

// ---------------------------------------------------------------------------
// Eight vectors side by side ([n][8], the 8 values of entry i contiguous), as
// the 8-vector product reads and writes them, and the node-block
// preconditioner applied to each of them.
// ---------------------------------------------------------------------------
// V[i][s] = (float)(R[s][i] * scale[s]) for the vector-major FP64 R [8][n]
__global__ void multivec_pack_kernel(float* __restrict__ V, const double* __restrict__ R, Scale8 scale, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float v[8];
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = (float)(R[(size_t)s * n + i] * scale.v[s]);
        v4f_t* o = reinterpret_cast<v4f_t*>(V + (size_t)i * 8);
        o[0] = v4f_t{v[0], v[1], v[2], v[3]};
        o[1] = v4f_t{v[4], v[5], v[6], v[7]};
    }
}
void vec_multivec_pack(float* V, const double* R, const Scale8& scale, long n, hipStream_t st) {
    hipLaunchKernelGGL(multivec_pack_kernel, dim3(grid_for(n)), dim3(256), 0, st, V, R, scale, n);
}

// L[s][i] += a[s] X[i][s] (X FP64, side by side; L vector-major)
__global__ void multivec_unpack_axpy_kernel(double* __restrict__ L, const double* __restrict__ X, Scale8 a, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int s = 0; s < 8; s++) L[(size_t)s * n + i] += a.v[s] * X[(size_t)i * 8 + s];
    }
}
void vec_multivec_unpack_axpy(double* L, const double* X, const Scale8& a, long n, hipStream_t st) {
    hipLaunchKernelGGL(multivec_unpack_axpy_kernel, dim3(grid_for(n)), dim3(256), 0, st, L, X, a, n);
}

// z[n][r][s] = sum_c Pinv_n(r, c) v[n][c][s] (Pinv_n^T with trans; += with
// accumulate), one thread per (node, vector); arithmetic in the type of z
template <typename TV, typename TP, int NEQ>
__global__ __launch_bounds__(256) void nodeblock_apply8_kernel(const TP* __restrict__ P, const TV* __restrict__ v,
                                                               TV* __restrict__ z, long n_node, int accumulate,
                                                               int trans) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n_node * 8; t += (long)gridDim.x * blockDim.x) {
        const long n = t >> 3;
        const int s = (int)(t & 7);
        TV x[NEQ];
#pragma unroll
        for (int c = 0; c < NEQ; c++) x[c] = v[((size_t)n * NEQ + c) * 8 + s];
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            TV acc = accumulate ? z[((size_t)n * NEQ + r) * 8 + s] : (TV)0;
#pragma unroll
            for (int c = 0; c < NEQ; c++)
                acc += (TV)P[pblk_offset<NEQ>(n, trans ? c * NEQ + r : r * NEQ + c)] * x[c];
            z[((size_t)n * NEQ + r) * 8 + s] = acc;
        }
    }
}
template <typename TV, typename TP>
void vec_nodeblock_apply8(const TP* P, const TV* v, TV* z, int neqn, long n_node, bool accumulate, bool trans,
                          hipStream_t st) {
    const int nb = grid_for(n_node * 8);
    if (neqn == 5)
        hipLaunchKernelGGL((nodeblock_apply8_kernel<TV, TP, 5>), dim3(nb), dim3(256), 0, st, P, v, z, n_node,
                           (int)accumulate, (int)trans);
    else
        hipLaunchKernelGGL((nodeblock_apply8_kernel<TV, TP, 4>), dim3(nb), dim3(256), 0, st, P, v, z, n_node,
                           (int)accumulate, (int)trans);
}
template void vec_nodeblock_apply8<float, float>(const float*, const float*, float*, int, long, bool, bool,
                                                 hipStream_t);
template void vec_nodeblock_apply8<double, float>(const float*, const double*, double*, int, long, bool, bool,
                                                  hipStream_t);
template void vec_nodeblock_apply8<double, double>(const double*, const double*, double*, int, long, bool, bool,
                                                   hipStream_t);

// out[i q + c] = in[map[i] q + c] for node i < n_node, c < q (a node
// permutation of an array with q values per node)
template <typename T>
__global__ void gather_nodes_kernel(T* __restrict__ out, const T* __restrict__ in, const int* __restrict__ map,
                                    long n_node, int q) {
    const long n = n_node * q;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = t / q;
        const int c = (int)(t - i * q);
        out[t] = in[(size_t)map[i] * q + c];
    }
}
template <typename T>
void vec_gather_nodes(T* out, const T* in, const int* map, long n_node, int q, hipStream_t st) {
    hipLaunchKernelGGL((gather_nodes_kernel<T>), dim3(grid_for(n_node * q)), dim3(256), 0, st, out, in, map, n_node, q);
}
template void vec_gather_nodes<double>(double*, const double*, const int*, long, int, hipStream_t);
template void vec_gather_nodes<float>(float*, const float*, const int*, long, int, hipStream_t);

// per-equation sums of squares of a node array (MPI_RMSError.h:19-26)
__global__ void sumsq_eqn_kernel(const double* __restrict__ U, int neqn, long n_node, double* __restrict__ part) {
    __shared__ double sm[5][256];
    double acc[5] = {0, 0, 0, 0, 0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_node; i += (long)gridDim.x * blockDim.x)
        for (int q = 0; q < neqn; q++) {
            const double u = U[i * neqn + q];
            acc[q] += u * u;
        }
    for (int q = 0; q < 5; q++) sm[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 5; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < (unsigned)neqn) part[(size_t)blockIdx.x * neqn + threadIdx.x] = sm[threadIdx.x][0];
}
void vec_sumsq_eqn(const double* U, int neqn, long n_node, double* part, double* out, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(sumsq_eqn_kernel, dim3(nblk), dim3(256), 0, st, U, neqn, n_node, part);
    hipLaunchKernelGGL(finish_dot_kernel, dim3(neqn), dim3(256), 0, st, part, nblk, neqn, out, nullptr);
}

__global__ void min_reduce_kernel(const double* part, int n, double* out) {
    __shared__ double sm[256];
    double m = 1000.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmin(m, part[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}
void min_reduce(const double* part, int n, double* out, hipStream_t st) {
    hipLaunchKernelGGL(min_reduce_kernel, dim3(1), dim3(256), 0, st, part, n, out);
}

// ---------------------------------------------------------------------------
// Greedy coloring, Jones-Plassmann form: an element is colored once every
// neighbour (element sharing a node) of higher priority is colored, with
// the smallest color none of those neighbours uses.  This is exactly the
// sequential first-fit greedy coloring visiting elements in priority order.
// ---------------------------------------------------------------------------
__global__ void color_round_kernel(const int* __restrict__ conn, int nnod, const int* __restrict__ node_ptr,
                                   const int* __restrict__ node_elem, long n_elem,
                                   const unsigned long long* __restrict__ prio, const int* __restrict__ cin,
                                   int* __restrict__ cout, int* changed) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem) return;
    if (cin[e] >= 0) { cout[e] = cin[e]; return; }
    const unsigned long long pe = prio[e];
    unsigned long long used0 = 0, used1 = 0;
    bool ready = true;
    for (int n = 0; n < nnod && ready; n++) {
        const int g = conn[e * nnod + n];
        for (int p = node_ptr[g]; p < node_ptr[g + 1]; p++) {
            const int f = node_elem[p];
            if (f == e) continue;
            if (prio[f] < pe) {
                const int c = cin[f];
                if (c < 0) { ready = false; break; }
                if (c < 64) used0 |= 1ULL << c;
                else used1 |= 1ULL << (c - 64);
            }
        }
    }
    if (!ready) { cout[e] = -1; *changed = 1; return; }   // still waiting
    int c = 0;
    while (c < 64 && ((used0 >> c) & 1ULL)) c++;
    if (c == 64)
        while (c < 128 && ((used1 >> (c - 64)) & 1ULL)) c++;
    cout[e] = c;
}

void launch_color_round(const int* conn, int nnod, const int* node_ptr, const int* node_elem, long n_elem,
                        const unsigned long long* prio, const int* cin, int* cout, int* changed, hipStream_t st) {
    const int nb = (int)((n_elem + 255) / 256);
    hipLaunchKernelGGL(color_round_kernel, dim3(nb), dim3(256), 0, st, conn, nnod, node_ptr, node_elem, n_elem, prio,
                       cin, cout, changed);
}

}  // namespace cfdv
