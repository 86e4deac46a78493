// Discrete adjoint of the FDV element build (hex8): forward-mode tangents
// of the element blocks through the per-lane element code of element3d.h,
// contracted with the adjoint and the update.
#include "cfdv_internal.h"
#include "dual.h"
#include "element3d.h"
#include "fdv_device.h"

namespace cfdv {

// ---------------------------------------------------------------------------
// Build tangents for the discrete adjoint (hex8).  For one element slot and
// one batch of ND of its 41 inputs (U of its 8 nodes x 5 equations, then
// the Mach number of the Thermo constants), the derivatives of
//   phi_e = sum_N lam_N^T (rhs_N - sum_M R_NM du_M)
// with R, rhs the element's blocks after faces and ApplyBC at the fixed dt
// (FDVGalerkin.h:121-154 + ApplyBC.h:14-133, differentiated in forward
// mode through the same per-lane code as build3d_kernel).  gelem[slot][k]
// receives d phi_e / d input k.
// ---------------------------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(64) void build3d_vjp_kernel(BuildArgs A, GaussTables G, const double* __restrict__ lam,
                                                         const double* __restrict__ du, int batch,
                                                         double* __restrict__ gelem) {
    typedef Dual<ND> D;
    constexpr int NNOD = 8, NEQ = 5, NIN = 41;
    // node rows as raw LDS (Dual has constructors, which __shared__ forbids)
    __shared__ double s_raw[NNOD * ND_STRIDE * (ND + 1)];
    __shared__ double s_red[ND][64];
    D (*s_nd)[ND_STRIDE] = reinterpret_cast<D (*)[ND_STRIDE]>(s_raw);
    const int lane = threadIdx.x;
    const long e = blockIdx.x;
    if (e >= A.n_elem_pad) return;
    const int* cn = A.conn + e * NNOD;
    if (cn[0] < 0) {   // padding slot
        for (int j = lane; j < ND; j += 64) {
            const int k = batch * ND + j;
            if (k < NIN) gelem[e * NIN + k] = 0.0;
        }
        return;
    }
    // Thermo with d/dMach seeded where the batch holds input 40:
    // Cv = 1/(gamma (gamma-1) M^2), cgas = 1/(gamma M^2), cmach = M
    ThermoT<D> th;
    th.gamma = A.th.gamma; th.gamm1 = A.th.gamm1; th.gm1d2 = A.th.gm1d2;
    th.gamm3 = A.th.gamm3; th.gm3d2 = A.th.gm3d2; th.Pr = A.th.Pr; th.csuth = A.th.csuth;
    th.creyn = A.th.creyn; th.Twall = A.th.Twall; th.adiabatic = A.th.adiabatic;
    th.Cv = D(A.th.Cv);
    th.cgas = D(A.th.cgas);
    th.cmach = D(A.th.cmach);
    {
        const int jm = NIN - 1 - batch * ND;
        if (jm >= 0 && jm < ND) {
            th.Cv.d[jm] = -2.0 * A.th.Cv / A.th.cmach;
            th.cgas.d[jm] = -2.0 * A.th.cgas / A.th.cmach;
            th.cmach.d[jm] = 1.0;
        }
    }
    if (lane < NNOD) {
        const int g = cn[lane];
        D Ul[NEQ];
#pragma unroll
        for (int q = 0; q < NEQ; q++) {
            Ul[q] = D(A.U[(size_t)g * NEQ + q]);
            const int j = lane * NEQ + q - batch * ND;
            if (j >= 0 && j < ND) Ul[q].d[j] = 1.0;
        }
        load_node3d_t(th, A.x + (size_t)g * 3, Ul, s_nd[lane]);
    }
    __syncthreads();
    const int N = lane >> 3, M = lane & 7;
    D R[NEQ][NEQ], rhs[NEQ], sp[4];
    int tie = 0, status = 0;
    build3d_elem(A, th, G, s_nd, e, lane, R, rhs, sp, &tie, &status);
    // this lane's share of phi_e
    const int gN = cn[N], gM = cn[M];
    D c(0.0);
#pragma unroll
    for (int r = 0; r < NEQ; r++) {
        D t = (M == 0) ? rhs[r] : D(0.0);
#pragma unroll
        for (int s = 0; s < NEQ; s++) t -= R[r][s] * du[(size_t)gM * NEQ + s];
        c += t * lam[(size_t)gN * NEQ + r];
    }
    // fixed-order reduction over the 64 lanes
#pragma unroll
    for (int j = 0; j < ND; j++) s_red[j][lane] = c.d[j];
    __syncthreads();
    if (lane < ND) {
        const int k = batch * ND + lane;
        double acc = 0.0;
        for (int t = 0; t < 64; t++) acc += s_red[lane][t];
        if (k < NIN) gelem[e * NIN + k] = acc;
    }
}

void launch_build_vjp(const BuildArgs& a, const GaussTables& G, const double* lam, const double* du, double* gelem,
                      hipStream_t st) {
    constexpr int ND = 8;
    for (int b = 0; b * ND < 41; b++)
        hipLaunchKernelGGL((build3d_vjp_kernel<ND>), dim3((int)a.n_elem_pad), dim3(64), 0, st, a, G, lam, du, b,
                           gelem);
}

// gU[n][q] = sum over the (slot, local node l) pairs of node n of
// gelem[slot][5 l + q], in element order (the compose_kernel CSR)
__global__ void vjp_gather_kernel(const int* __restrict__ node_ptr, const int* __restrict__ node_ent,
                                  const double* __restrict__ gelem, long n_node, double* __restrict__ gU) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_node) return;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int p = node_ptr[n]; p < node_ptr[n + 1]; p++) {
        const long ent = node_ent[p];
        const long e = ent / 8;
        const int l = (int)(ent % 8);
        for (int q = 0; q < 5; q++) acc[q] += gelem[e * 41 + l * 5 + q];
    }
    for (int q = 0; q < 5; q++) gU[n * 5 + q] = acc[q];
}

void launch_vjp_gather(const int* node_ptr, const int* node_ent, const double* gelem, long n_node, double* gU,
                       hipStream_t st) {
    hipLaunchKernelGGL(vjp_gather_kernel, dim3((int)((n_node + 255) / 256)), dim3(256), 0, st, node_ptr, node_ent,
                       gelem, n_node, gU);
}

// out = sum_i a[i * stride + offset], i < n: fixed grid, fixed-order
// two-stage sum (deterministic)
__global__ __launch_bounds__(256) void column_partial_kernel(const double* __restrict__ a, long n, int stride,
                                                             int offset, double* __restrict__ part) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc += a[i * stride + offset];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}
__global__ __launch_bounds__(256) void column_finish_kernel(const double* __restrict__ part, int nblk,
                                                            double* __restrict__ out) {
    __shared__ double sm[256];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) acc += part[b];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}
void launch_column_sum(const double* a, long n, int stride, int offset, double* part, int nblk, double* out,
                       hipStream_t st) {
    hipLaunchKernelGGL(column_partial_kernel, dim3(nblk), dim3(256), 0, st, a, n, stride, offset, part);
    hipLaunchKernelGGL(column_finish_kernel, dim3(1), dim3(256), 0, st, part, nblk, out);
}

}  // namespace cfdv
