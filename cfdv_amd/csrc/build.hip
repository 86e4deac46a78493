// FDV element build on gfx950: FDVParam + domain + boundary + ApplyBC in
// one pass, blocks written structure-of-arrays to HBM.
//
// Restates FDVGalerkin::operator() (src/FDV/FDVGalerkin.h:121-154) applied
// by std::for_each over all elements (src/main.cpp:185-186) followed by
// ApplyBC (src/FDV/Function/ApplyBC.h:14-133, src/main.cpp:194).
//
// Mapping: one 64-lane wavefront per hex8 element; lane (N, M) = (lane/8,
// lane%8) owns the 5x5 sub-block R(5N+r, 5M+s).  Every lane evaluates the
// Gauss-point state; only the node-pair contraction differs per lane.  The
// quad4 kernel packs four elements per wavefront (16 lanes each).
//
// Block layout (both precisions): blk_offset() in fdv_device.h -- chunks of
// 64 elements, 16-byte entry groups interleaved across the chunk, so that a
// thread-per-element matvec reads one 16-byte vector per lane from 1 KB of
// contiguous memory.
#include "cfdv_internal.h"
#include "element3d.h"
#include "fdv_device.h"

#include <mutex>
#include <vector>

namespace cfdv {

__constant__ GaussTables c_gauss;

void upload_gauss_tables(const GaussTables& g) {
    HIP_CHECK_THROW(hipMemcpyToSymbol(HIP_SYMBOL(c_gauss), &g, sizeof(g)));
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that
// consecutive logical blocks (neighbouring elements, i.e. the same 64-element
// output chunk) execute on one XCD and share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb / 8, r = nb % 8;
    const int x = b % 8, i = b / 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

__device__ __forceinline__ double readlane_f64(double x, int l) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// FDVParam (src/FDV/Function/FDVParam.h:15-111) of one hex8 element over
// the 64 lanes of its wavefront, every lane active: lane l evaluates node
// l % 8's Mach and Reynolds numbers, each octet reduces its extremes (lanes
// 0-31 the Mach numbers, 32-63 the Reynolds numbers), and both halves run the
// r / s tail on the same instructions (one pow instead of two).  Min and max
// are exact, so for finite states s equals fdvparam<3, 8>'s bit for bit
// (element3d.h), which evaluates the eight nodes one after the other on
// every lane.
__device__ void fdvparam_wave(const DevThermo& th, const double (*nd)[ND_STRIDE], int lane, double s[4],
                              int* tie) {
#pragma clang fp contract(off)
    const int beta = lane & 7;
    double dd = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) dd += nd[beta][ND_V + i] * nd[beta][ND_V + i];
    const double V = sqrt(dd);
    double T = nd[beta][ND_T];
    if (T < 1.0E-12) T = 0;
    const double a = sqrt(th.gamma * th.cgas * T);
    double M = 0;
    if (!(a < 1.0E-12)) M = V / a;
    const double mu = calc_mu(th, T);
    const double Re = (nd[beta][ND_RHO] * V) / mu;
    const bool re_half = lane >= 32;
    double hi = re_half ? Re : M, lo = hi;
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
        const double oh = __shfl_xor(hi, d, 64), ol = __shfl_xor(lo, d, 64);
        hi = (hi < oh) ? oh : hi;
        lo = (ol < lo) ? ol : lo;
    }
    const double eps = 0.0000001;
    const double r = sqrt(hi * hi - lo * lo) / (lo + eps);
    double sv = 0.0;
    int t = 0;
    if (r > 0.01) sv = (1.0 < r) ? 1.0 : r;
    else if (r < 0.01 && lo >= 10E-10) sv = 0.0;
    else if (re_half ? (lo <= 10E-10) : (lo < 10E-10)) sv = 1.0;
    else t = 1;
    // pow(0, 0.2) = 0 and pow(1, 0.2) = 1 exactly: the library pow runs only
    // on lanes that need it (none, for most elements)
    double pw = 0.5 * (1 + sv);
    if (sv != 0.0 && sv != 1.0) pw = 0.5 * (1 + pow(sv, 0.2));
    s[0] = readlane_f64(sv, 0);
    s[1] = readlane_f64(pw, 0);
    s[2] = readlane_f64(sv, 32);
    s[3] = readlane_f64(pw, 32);
    *tie = __builtin_amdgcn_readlane(t, 0) | __builtin_amdgcn_readlane(t, 32);
}

// eval1dboundary (src/FDV/Function/eval1dboundary.h:9-251)
CFDV_HD bool eval1dboundary(int bc, GPState<2>& f, const double x0[2], const double x1[2],
                               const double en[2]) {
    const double len = sqrt(en[0] * en[0] + en[1] * en[1]);
    const double n[2] = {en[0] / len, en[1] / len};
    switch (bc) {
    case 2: case 12: {
        double d = f.v[0] * n[0] + f.v[1] * n[1];
        f.v[0] -= d * n[0];
        d = f.v[0] * n[0] + f.v[1] * n[1];
        f.v[1] -= d * n[1];
        d = f.dT[0] * n[0] + f.dT[1] * n[1];
        f.dT[0] -= d * n[0];
        d = f.dT[0] * n[0] + f.dT[1] * n[1];
        f.dT[1] -= d * n[1];
        break;
    }
    case 1: case 51: {
        double d = f.dT[0] * n[0] + f.dT[1] * n[1];
        f.dT[0] -= d * n[0];
        d = f.dT[0] * n[0] + f.dT[1] * n[1];
        f.dT[1] -= d * n[1];
        break;
    }
    case -1:
        f.dV[0][0] = 0.0; f.dV[1][0] = 0.0; f.dV[0][1] = 0.0; f.dV[1][1] = 0.0;
        f.dT[0] = 0.0;
        break;
    case 6:
        f.dV[0][0] = 0.0; f.dV[1][0] = 0.0; f.dV[0][1] = 0.0; f.dV[1][1] = 0.0;
        f.dT[0] = 0.0;
        for (int i = 0; i < 4; i++) f.dU[i][0] = 0.0;
        break;
    case -2: case 7:
        f.dV[0][1] = 0.0; f.dV[1][1] = 0.0;
        f.dT[1] = 0.0;
        for (int i = 0; i < 4; i++) f.dU[i][1] = 0.0;
        break;
    case 4: case 8: {
        double nn[2];
        if (bc == 4) {
            f.v[0] = f.v[1] = 0;
            nn[0] = en[0]; nn[1] = en[1];
        } else {
            const double alpha = 10 * (3.14159 / 180.0);
            const double dl = fabs(x1[1] - x0[1]);
            nn[0] = cos(alpha) * dl;
            nn[1] = sin(alpha) * dl;
        }
        const double dudot = f.dV[0][0] * nn[0] + f.dV[0][1] * nn[1];
        const double dvdot = f.dV[1][0] * nn[0] + f.dV[1][1] * nn[1];
        const double dtdot = f.dT[0] * nn[0] + f.dT[1] * nn[1];
        for (int j = 0; j < 2; j++) {
            f.dV[0][j] -= dudot * nn[j];
            f.dV[1][j] -= dvdot * nn[j];
            f.dT[j] -= dtdot * nn[j];
        }
        double dUdot[4];
        for (int i = 0; i < 4; i++) dUdot[i] = f.dU[i][0] * nn[0] + f.dU[i][1] * nn[1];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 2; j++) f.dU[i][j] -= dUdot[i] * nn[j];
        break;
    }
    case -12:
        f.dV[0][1] = 0.0; f.dV[1][1] = 0.0;
        f.dT[1] = 0.0;
        f.v[1] = 0.0;
        for (int i = 0; i < 4; i++) f.dU[i][1] = 0.0;
        break;
    case 11: case 21:
        f.v[0] = f.v[1] = 0;
        f.dV[0][0] = 0.0; f.dV[1][0] = 0.0;
        f.dT[1] = 0.0;
        for (int i = 0; i < 4; i++) f.dU[i][1] = 0.0;
        break;
    case 0:
        break;
    default:
        return false;
    }
    calc_tau<2>(f);
    if (bc == 2 || bc == 12 || bc == 7) { f.tau[0][1] = 0.0; f.tau[1][0] = 0.0; }
    if (bc == 6) { f.tau[0][0] = 0.; f.tau[1][0] = 0.; }
    return true;
}

// ---------------------------------------------------------------------------
// Node-pair contraction shared by domain and boundary terms (see header of
// fdv_device.h for the factorised form).
// ---------------------------------------------------------------------------
template <int NDIM, int NEQ>
CFDV_HD void contract(double R[NEQ][NEQ], double cI, const double wg[NDIM],
                                         const double wa[NDIM], const double ws[NDIM], double cA,
                                         const double a[NDIM][NEQ][NEQ], const double b[NDIM][NEQ][NEQ],
                                         double s1, double s2, double s3, double s4) {
    double Ah[NEQ][NEQ], Sh[NEQ][NEQ];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) {
            double acc = (r == s) ? cI : 0.0;
            double ah = 0.0, sh = 0.0;
#pragma unroll
            for (int i = 0; i < NDIM; i++) {
                acc += wg[i] * (s1 * a[i][r][s] + s3 * b[i][r][s]);
                ah += wa[i] * (a[i][r][s] + b[i][r][s]);
                sh += ws[i] * (s2 * a[i][r][s] + s4 * b[i][r][s]);
            }
            R[r][s] += acc;
            Ah[r][s] = ah;
            Sh[r][s] = sh;
        }
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) {
            double p = 0.0;
#pragma unroll
            for (int t = 0; t < NEQ; t++) p += Ah[r][t] * Sh[t][s];
            R[r][s] += cA * p;
        }
}

template <typename T>
CFDV_HD size_t blk_index(long e, int q, int KK) {
    return blk_offset<T>(e, q, KK);
}
// entry q = i k + j of element e's FP64 block in the context's layout
template <int NNOD, int NEQ>
__device__ __forceinline__ size_t r64_at(const BuildArgs& A, long e, int q) {
    constexpr int K = NNOD * NEQ;
    return A.r64_pairs ? pair_offset<NNOD, NEQ>(e, q) : blk_offset<double>(e, q, K * K);
}

// Lane (N, M)'s NEQ x NEQ block R of element e to / from the FP64 blocks:
// with the node-pair layout one contiguous run (no per-entry index
// arithmetic), otherwise entry by entry in the chunked layout.
template <int NNOD, int NEQ>
__device__ __forceinline__ void r64_store(const BuildArgs& A, long e, int N, int M, const double R[NEQ][NEQ]) {
    constexpr int K = NNOD * NEQ;
    if (A.r64_pairs) {
        double* dst = A.R64 + pair_base<NNOD, NEQ>(e, N * NNOD + M);
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) dst[r * NEQ + s] = R[r][s];
    } else {
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++)
                A.R64[blk_offset<double>(e, (NEQ * N + r) * K + NEQ * M + s, K * K)] = R[r][s];
    }
}
template <int NNOD, int NEQ>
__device__ __forceinline__ void r64_load(const BuildArgs& A, long e, int N, int M, double R[NEQ][NEQ]) {
    constexpr int K = NNOD * NEQ;
    if (A.r64_pairs) {
        const double* src = A.R64 + pair_base<NNOD, NEQ>(e, N * NNOD + M);
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) R[r][s] = src[r * NEQ + s];
    } else {
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++)
                R[r][s] = A.R64[blk_offset<double>(e, (NEQ * N + r) * K + NEQ * M + s, K * K)];
    }
}

// Rows built in place: lane (N, M)'s block of element e added into its
// staged FP64 row slot, or stored there when the element is the slot's first
// contributor (the lowest element slot; colors run in slot order, so every
// slot sums its blocks in one fixed order).  No two elements of one color
// share a node, so no two lanes of a color's launches meet in one slot.
__device__ __forceinline__ void rows_put(const BuildArgs& A, long e, int N, int M, const double R[5][5]) {
    const int pr = N * 8 + M;
    double* dst = A.S64 + (size_t)A.pslot[e * 64 + pr] * 25;
    if ((A.pfirst[e] >> pr) & 1ULL) {
#pragma unroll
        for (int r = 0; r < 5; r++)
#pragma unroll
            for (int s = 0; s < 5; s++) dst[r * 5 + s] = R[r][s];
    } else {
        double old[25];
#pragma unroll
        for (int q = 0; q < 25; q++) old[q] = dst[q];
#pragma unroll
        for (int r = 0; r < 5; r++)
#pragma unroll
            for (int s = 0; s < 5; s++) dst[r * 5 + s] = old[r * 5 + s] + R[r][s];
    }
}
// the domain part of a face-kernel element between the two kernels
__device__ __forceinline__ double* rb_block(const BuildArgs& A, long e, int N, int M) {
    return A.Rb + ((size_t)A.bidx[e] * 64 + N * 8 + M) * 25;
}

// ---------------------------------------------------------------------------
// Interior hex8 elements (no face with a bc code, no constrained dof): the
// Gauss-point state is the same for all 64 node pairs, so it is evaluated
// once per point (phase 1: lanes 0..7, one point each) into LDS, and the 64
// lanes then contract it for their own node pair (phase 2).  Same arithmetic
// as build3d_elem's domain loop, in the same order.
// ---------------------------------------------------------------------------
struct GPShared {
    double wJ, SPhi;     // w * det J, sum_n Phi_n
    double D[5];         // sum_jb dPhi_bj G_b(j, s)
    double P[3][5];      // P(i, r) for r = 1..4 (r = 0 unused)
    GPState<3> f;
};

// The hex8 volume tables the interior kernel reads per point and node pair,
// staged in LDS once per element (lane-dependent indices would otherwise be
// vector loads from the constant buffer in every loop iteration)
// Per point: Phi, dPhi/dxi and (from phase 1 on) w det J, 33 doubles, so
// that the 8 lanes of phase 1 (one point each) read 8 different LDS banks.
struct GaussHex {
    struct Point {
        double phi[8];
        double dpxi[8][3];
        double wJ;
    } p[8];
};

// Phase 1 of the interior kernel is spread over the 64 lanes in three
// steps, each quantity formed once per point:
//   gp_geom  (lane = 8 gp + n): Jacobian of point gp (its 8 lanes alike),
//            dPhi_n/dx of node n and w det J;
//   gp_sums  (lane = 8 gp + k): for node column src(k) (U, v, E, T, or the
//            constant 1), u(gp) = sum_n Phi_n u_n and du/dx_j(gp) =
//            sum_n dPhi_n/dx_j u_n -- every sum over n = 0..7 in order, one
//            fused multiply-add per node, as a serial loop forms it;
//   gp_state (lane = 8 part + gp): the state and D, P of the point from
//            those sums.
// GP_NSRC node columns feed sums: ND_U .. ND_T (U, v, E, T) and the
// constant 1, whose sums are SPhi and sum_n dPhi_n/dx_j.
constexpr int GP_NSRC = 11;
CFDV_HD constexpr int gp_src_col(int k) { return k < 10 ? ND_U + k : -1; }

CFDV_HD void gp_geom(const GaussHex::Point& P, double w3, const double (*nd)[ND_STRIDE], int n, double dphi_n[3],
                     double* wJ) {
    double jac[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < 8; m++) acc += P.dpxi[m][j] * nd[m][ND_X + i];
            jac[i][j] = acc;
        }
    const double J = jac[0][0] * jac[1][1] * jac[2][2] + jac[0][1] * jac[1][2] * jac[2][0] +
                     jac[0][2] * jac[1][0] * jac[2][1] - jac[0][2] * jac[1][1] * jac[2][0] -
                     jac[0][1] * jac[1][0] * jac[2][2] - jac[0][0] * jac[1][2] * jac[2][1];
    double ji[3][3];
    ji[0][0] = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
    ji[0][1] = -(jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]);
    ji[0][2] = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
    ji[1][0] = -(jac[0][1] * jac[2][2] - jac[0][2] * jac[2][1]);
    ji[1][1] = jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0];
    ji[1][2] = -(jac[0][0] * jac[2][1] - jac[0][1] * jac[2][0]);
    ji[2][0] = jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1];
    ji[2][1] = -(jac[0][0] * jac[1][2] - jac[0][2] * jac[1][0]);
    ji[2][2] = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
    const double Jinv = 1.0 / J;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) acc += Jinv * (ji[i][j] * P.dpxi[n][j]);
        dphi_n[i] = acc;
    }
    *wJ = w3 * J;
}

// The four sums of node column col (-1: the constant 1) at point P, whose
// dPhi_n/dx are dphi[n]: out = {sum Phi_n u_n, sum dPhi_n/dx_j u_n (j = 0..2)}
CFDV_HD void gp_sums(const GaussHex::Point& P, const double (*dphi)[3], const double (*nd)[ND_STRIDE], int col,
                     double out[4]) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const double u = col >= 0 ? nd[n][col] : 1.0;
        a0 += P.phi[n] * u;
        a1 += dphi[n][0] * u;
        a2 += dphi[n][1] * u;
        a3 += dphi[n][2] * u;
    }
    out[0] = a0;
    out[1] = a1;
    out[2] = a2;
    out[3] = a3;
}

// The point's state from the sums S[4 k + c] of source k (gp_sums), its
// viscosity, conductivity and tau, and the viscous vectors D and P
CFDV_HD void gp_state(const DevThermo& th, const double* S, double dt, GPShared& o) {
    constexpr int NEQ = 5;
    const double reinv = 1.0 / th.creyn;
    GPState<3> f;
#pragma unroll
    for (int q = 0; q < NEQ; q++) {
        f.U[q] = S[4 * q];
#pragma unroll
        for (int j = 0; j < 3; j++) f.dU[q][j] = S[4 * q + 1 + j];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f.v[i] = S[4 * (ND_V - ND_U + i)];
#pragma unroll
        for (int j = 0; j < 3; j++) f.dV[i][j] = S[4 * (ND_V - ND_U + i) + 1 + j];
    }
    f.E = S[4 * (ND_E - ND_U)];
    f.T = S[4 * (ND_T - ND_U)];
#pragma unroll
    for (int j = 0; j < 3; j++) f.dT[j] = S[4 * (ND_T - ND_U) + 1 + j];
    const double SPhi = S[4 * 10];
    double SD[3];
#pragma unroll
    for (int j = 0; j < 3; j++) SD[j] = S[4 * 10 + 1 + j];
    f.mu = calc_mu(th, f.T);
    f.k = calc_k(th, f.mu);
    calc_tau<3>(f);
    o.SPhi = SPhi;
    o.D[0] = 0.0;
#pragma unroll
    for (int s = 1; s < 4; s++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) acc += (-f.tau[j][s - 1] * reinv) * SD[j];
        o.D[s] = acc;
    }
    {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double td = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) td += f.tau[j][k] * f.dV[k][j];
            acc += (-td - f.k * f.dT[j] * SD[j]) * reinv;
        }
        o.D[4] = acc;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        o.P[i][0] = 0.0;
#pragma unroll
        for (int r = 1; r < 5; r++) {
            double P;
            if (r < 4) P = dt * (-f.tau[i][r - 1] * reinv) * SPhi;
            else {
                double tv = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++) tv += f.tau[i][k] * f.v[k];
                P = dt * ((-tv - f.k * f.dT[i] * SPhi) * reinv);
            }
            o.P[i][r] = P;
        }
    }
    o.f = f;
}

// Structural nonzeros of the direction-I Jacobians a_I and b_I
// (NavierStokes3D.h:74-343): row 0 {1+I}, row 1+I all, rows 1+j (j != I)
// {0, 1+j, 1+I}, row 4 all -- 17 entries, row-major.
constexpr int JPAT_R[3][17] = {{0, 1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 4, 4},
                               {0, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 4, 4},
                               {0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4}};
constexpr int JPAT_S[3][17] = {{1, 0, 1, 2, 3, 4, 0, 1, 2, 0, 1, 3, 0, 1, 2, 3, 4},
                               {2, 0, 1, 2, 0, 1, 2, 3, 4, 0, 2, 3, 0, 1, 2, 3, 4},
                               {3, 0, 1, 3, 0, 2, 3, 0, 1, 2, 3, 4, 0, 1, 2, 3, 4}};

// Per Gauss point, what the node-pair contraction reads: the direction
// combinations M1_I = s1 a_I + s3 b_I on their 17 nonzeros (less the first,
// a_I(0, 1 + I) = 1 and b_I(0, .) = 0 make it s1), A_I = a_I + b_I and
// S_I = s2 a_I + s4 b_I as full row-major 5x5 (operands of the matrix-core
// product), the rhs vectors PQ_I = P_I - dt^2/2 A_I D and the distinct
// entries of the c_ij / Re blocks (CT_* below).
enum {
    CT_R0 = 0,    // [k] c_ii(1+k, 0), k == i:  mu_R Uc_k / rho^2
    CT_M0 = 3,    // [k] c_ii(1+k, 0), k != i, and c_ij(1+j, 0) (k = i):  mu Uc_k / rho^2
    CT_RR = 6,    // c_ii(1+k, 1+k), k == i:  -mu_R / rho
    CT_MR = 7,    // c_ii(1+k, 1+k), k != i, and c_ij(1+j, 1+i):  -mu / rho
    CT_40 = 8,    // [i] c_ii(4, 0)
    CT_4R = 11,   // [k] c_ii(4, 1+k), k == i
    CT_4M = 14,   // [k] c_ii(4, 1+k), k != i
    CT_44 = 17,   // c_ii(4, 4)
    CT_L0 = 18,   // [j] c_ij(1+i, 0):  lambda Uc_j / rho^2
    CT_LR = 21,   // c_ij(1+i, 1+j):  -lambda / rho
    CT_O40 = 22,  // [ct_pair(i, j)] c_ij(4, 0), i != j
    CT_O4I = 28,  // [j] c_ij(4, 1+i)
    CT_O4J = 31,  // [i] c_ij(4, 1+j)
    CT_N = 34
};
// (i, j), i != j, numbered 0..5 in row-major order
CFDV_HD constexpr int ct_pair(int i, int j) { return 2 * i + (j > i ? j - 1 : j); }
struct GPTab {
    double dphi[8][3];
    double M1[3][16];   // pattern slots 1..16
    double Af[3][25], Sf[3][25];
    double PQ[3][5];
    double C[CT_N];
    // 271 doubles (odd): the tables of the 8 points, which 8 lanes write at
    // once in phase 1, start in 8 different LDS banks
};

// Equation index x of direction 0 seen from direction I: momentum rows 1
// and 1 + I swap (I = 0: identity); pi_ax does the same for the axes 0 and I.
CFDV_HD constexpr int sig_eq(int I, int x) { return x == 1 ? 1 + I : (x == 1 + I ? 1 : x); }
CFDV_HD constexpr int pi_ax(int I, int k) { return k == 0 ? I : (k == I ? 0 : k); }
template <typename T>
CFDV_HD T pick3(int I, T x0, T x1, T x2) { return I == 1 ? x1 : (I == 2 ? x2 : x0); }

// Phase 1b, direction I (0..2, at run time) of one Gauss point (one lane),
// so that the lanes of the three directions run the same instructions.
// a_I and b_I are direction 0's formulas (calc_a3_dir<0>, calc_b3_dir<0>)
// on the state with axes 0 and I swapped -- velocities, momenta, derivative
// directions and tau -- whose entry (r, c) is entry (sig_eq(I, r),
// sig_eq(I, c)) of a_I, b_I (the Jacobians are covariant under that swap,
// NavierStokes3D.h:74-343).  Stored rotated, read with the swap folded into
// the (compile-time) indices: M1[I][p] is entry (sig_eq(I, JPAT_R[0][p]),
// sig_eq(I, JPAT_S[0][p])) of M1_I (slot p - 1), PQ[I][r] is row sig_eq(I, r)
// of PQ_I.
// Af and Sf, which the matrix-core product reads at lane-dependent indices,
// are stored in their own orientation.
// The lane also forms direction I's share of the point's c_ij / Re table
// (gp_tab_c_dir), so that no lane group runs a branch of its own.
CFDV_HD void gp_tab_c_dir(const DevThermo& th, const GPState<3>& f0, const GPState<3>& f, int I, GPTab& t);
CFDV_HD void gp_tab_dir_rot(const DevThermo& th, const GPShared& g, const double sp[4], double dt, int I,
                            GPTab& t) {
    const GPState<3>& f0 = g.f;
    GPState<3> f;   // the members calc_a3_dir<0> and calc_b3_dir<0> read
    f.E = f0.E;
    f.mu = f0.mu;
    f.k = f0.k;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        f.v[k] = pick3(I, f0.v[pi_ax(0, k)], f0.v[pi_ax(1, k)], f0.v[pi_ax(2, k)]);
        f.tau[0][k] = pick3(I, f0.tau[0][pi_ax(0, k)], f0.tau[1][pi_ax(1, k)], f0.tau[2][pi_ax(2, k)]);
    }
#pragma unroll
    for (int q = 0; q < 5; q++) {
        f.U[q] = pick3(I, f0.U[sig_eq(0, q)], f0.U[sig_eq(1, q)], f0.U[sig_eq(2, q)]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (q == 4 && k > 0) continue;   // calc_b3_dir<0> reads dU[4][0] only
            f.dU[q][k] = pick3(I, f0.dU[sig_eq(0, q)][pi_ax(0, k)], f0.dU[sig_eq(1, q)][pi_ax(1, k)],
                               f0.dU[sig_eq(2, q)][pi_ax(2, k)]);
        }
    }
    gp_tab_c_dir(th, f0, f, I, t);
    double a[5][5], b[5][5];
    calc_a3_dir<0>(th, f, a);
    calc_b3_dir<0>(th, f, b);
    const double s1 = sp[0], s2 = sp[1], s3 = sp[2], s4 = sp[3];
#pragma unroll
    for (int p = 1; p < 17; p++) {
        const int r = JPAT_R[0][p], c = JPAT_S[0][p];
        t.M1[I][p - 1] = s1 * a[r][c] + s3 * b[r][c];
    }
    const int sg[5] = {0, 1 + I, I == 1 ? 1 : 2, I == 2 ? 1 : 3, 4};   // sig_eq(I, .)
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 5; c++) {
            t.Af[I][sg[r] * 5 + sg[c]] = a[r][c] + b[r][c];
            t.Sf[I][sg[r] * 5 + sg[c]] = s2 * a[r][c] + s4 * b[r][c];
        }
    double D[5];
#pragma unroll
    for (int c = 0; c < 5; c++) D[c] = pick3(I, g.D[sig_eq(0, c)], g.D[sig_eq(1, c)], g.D[sig_eq(2, c)]);
    const double dtsq_2 = (dt * dt) / 2.0;
#pragma unroll
    for (int r = 0; r < 5; r++) {
        double q = 0.0;
#pragma unroll
        for (int c = 0; c < 5; c++) q += (a[r][c] + b[r][c]) * D[c];
        const double P = r == 0 ? 0.0 : pick3(I, g.P[0][sig_eq(0, r)], g.P[1][sig_eq(1, r)], g.P[2][sig_eq(2, r)]);
        t.PQ[I][r] = P - dtsq_2 * q;
    }
}

// Phase 1b, the c_ij / Re table of one Gauss point: accum_c3's coefficients
// (NavierStokes3D.h:347-477), each distinct value once, formed by the
// point's three direction lanes.  Lane I forms the entries indexed by I --
// velocity component I's (CT_R0, CT_M0, CT_L0), the diagonal block c_II's
// row 4 (CT_40, CT_4R, CT_4M), the pairs (I, j) (CT_O40, CT_O4I, CT_O4J) --
// and each lane the four scalars (CT_RR, CT_MR, CT_LR, CT_44: the same value
// to the same address).  Every entry is the same expression of the same
// values whichever lane forms it.  f0 is the point's state, f the state
// rotated for direction I (gp_tab_dir_rot), whose momentum components 1, 2,
// 3 are U_{1+I} and the other two, U_{1+j} for j = (I == 1 ? 0 : 1) and
// j = (I == 2 ? 0 : 2).
CFDV_HD void gp_tab_c_dir(const DevThermo& th, const GPState<3>& f0, const GPState<3>& f, int I, GPTab& t) {
    const double reinv = 1.0 / th.creyn;
    const double mu = f0.mu;
    const double lambda = -2. * mu / 3.;
    const double mu_R = 2. * mu + lambda;
    const double rho = f0.U[0];
    const double rinv = 1.0 / rho;
    const double denom = 1. / (rho * rho);
    const double Uc[3] = {f0.U[1], f0.U[2], f0.U[3]};
    const double re = f0.U[4];
    const double hkocv = f0.k / th.Cv;
    const double TERM1 = f0.k / th.Cv * denom;
    const double en = TERM1 * (-re + (Uc[0] * Uc[0] + Uc[1] * Uc[1] + Uc[2] * Uc[2]) * rinv);
    double* C = t.C;
    const double UcI = f.U[1], UrI = UcI * rinv;
    const double cR0I = mu_R * UcI * denom, cM0I = mu * UcI * denom;
    double cM0[3];
#pragma unroll
    for (int k = 0; k < 3; k++) cM0[k] = mu * Uc[k] * denom;
    C[CT_R0 + I] = cR0I * reinv;
    C[CT_M0 + I] = cM0I * reinv;
    C[CT_L0 + I] = lambda * UcI * denom * reinv;
    const double cRR = -mu_R * rinv, cMR = -mu * rinv, cLR = -lambda * rinv;
    C[CT_RR] = cRR * reinv;
    C[CT_MR] = cMR * reinv;
    C[CT_LR] = cLR * reinv;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) acc += Uc[k] * (k == I ? cR0I : cM0[k]);
    C[CT_40 + I] = (acc * rinv - en) * reinv;
    C[CT_4R + I] = (UrI * cRR + TERM1 * UcI) * reinv;
    C[CT_4M + I] = (UrI * cMR + TERM1 * UcI) * reinv;
    C[CT_O4I + I] = UrI * cMR * reinv;
    C[CT_O4J + I] = UrI * cLR * reinv;
#pragma unroll
    for (int o = 0; o < 2; o++) {
        const int j = o == 0 ? (I == 1 ? 0 : 1) : (I == 2 ? 0 : 2);
        const double Ucj = f.U[2 + o];
        C[CT_O40 + ct_pair(I, j)] = (UrI * (lambda * Ucj * denom) + (Ucj * rinv) * cM0I) * reinv;
    }
    C[CT_44] = -hkocv * rinv * reinv;
}

// accum_c3 (NavierStokes3D.h:347-477) contracted with the weights W from
// the tabulated coefficients:  out[r][s] += sum_ij W[i][j] c_ij[r][s] / Re,
// the diagonal blocks first, then the off-diagonal pairs, as accum_c3.
CFDV_HD void accum_c3_ctab(const double* C, const double W[3][3], double out[5][5]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double w = W[i][i];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            out[1 + k][0] += w * C[(k == i ? CT_R0 : CT_M0) + k];
            out[1 + k][1 + k] += w * C[k == i ? CT_RR : CT_MR];
        }
        out[4][0] += w * C[CT_40 + i];
#pragma unroll
        for (int k = 0; k < 3; k++) out[4][1 + k] += w * C[(k == i ? CT_4R : CT_4M) + k];
        out[4][4] += w * C[CT_44];
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (i == j) continue;
            const double w = W[i][j];
            out[1 + i][0] += w * C[CT_L0 + j];
            out[1 + i][1 + j] += w * C[CT_LR];
            out[1 + j][0] += w * C[CT_M0 + i];
            out[1 + j][1 + i] += w * C[CT_MR];
            out[4][0] += w * C[CT_O40 + ct_pair(i, j)];
            out[4][1 + i] += w * C[CT_O4I + j];
            out[4][1 + j] += w * C[CT_O4J + i];
        }
}

// Phase 2 for lane (N, M): the domain contraction of every Gauss point
// without the c_A A S product (mfma_as_kstep below) and the rhs
// (interior_rhs): diagonal, direction terms, c_ij, in contract3's order.
// before: called at the start of point gp's iteration (the kernel issues
// the matrix-core steps of the A S product there, so that they execute while
// the lanes contract the point on the VALU).
template <typename Before>
CFDV_HD void interior_phase2(const DevThermo& th, const GaussHex& G, const GPTab* tabs, int lane,
                             double dt, const double sp[4], double R[5][5], Before&& before) {
    constexpr int NEQ = 5;
    const int N = lane >> 3, M = lane & 7;
    const double s3 = sp[2];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) R[r][s] = 0.0;
#pragma unroll 1
    for (int gp = 0; gp < 8; gp++) {
        before(gp);
        const GPTab& g = tabs[gp];
        const double PN = G.p[gp].phi[N], PM = G.p[gp].phi[M];
        const double wJ = G.p[gp].wJ;
        double DN[3], DM[3];
#pragma unroll
        for (int i = 0; i < 3; i++) { DN[i] = g.dphi[N][i]; DM[i] = g.dphi[M][i]; }
        double wg[3], W[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            wg[i] = -dt * wJ * PM * DN[i];
#pragma unroll
            for (int j = 0; j < 3; j++) W[i][j] = -dt * s3 * wJ * DN[i] * DM[j];
        }
        const double cI = wJ * PN * PM;
#pragma unroll
        for (int r = 0; r < NEQ; r++) R[r][r] += cI;
#pragma unroll
        for (int I = 0; I < 3; I++)
#pragma unroll
            for (int p = 0; p < 17; p++)
                R[sig_eq(I, JPAT_R[0][p])][sig_eq(I, JPAT_S[0][p])] += wg[I] * (p == 0 ? sp[0] : g.M1[I][p - 1]);
        accum_c3_ctab(g.C, W, R);
    }
}

// The rhs entry (N, r) = (lane / 5, lane % 5) of lane < 40: P and the A D
// product (contract3's AQ) of every point, as PQ = P - dt^2/2 A D, summed
// point by point in the contraction's order
CFDV_HD double interior_rhs(const GaussHex& G, const GPTab* tabs, int lane) {
    const int N = lane / 5, r = lane - 5 * (lane / 5);
    double rhs = 0.0;
#pragma unroll 1
    for (int gp = 0; gp < 8; gp++) {
        const GPTab& g = tabs[gp];
        const double wJ = G.p[gp].wJ;
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) acc += (wJ * g.dphi[N][i]) * g.PQ[i][sig_eq(i, r)];
        rhs += acc;
    }
    return rhs;
}

// The c_A (sum_i dPhi_N,i A_i)(sum_j dPhi_M,j S_j) term of every node pair
// and Gauss point at once, on the FP64 matrix cores.  Over the element it is
// one 40 x 40 x 40 product: rows (N, r), columns (M, s), inner index
// (point g, u):
//   R_AS[5N + r][5M + s] = sum_g sum_u AH_g[5N + r][u] SH_g[u][5M + s],
//   AH_g[5N + r][u] = c_A,g sum_i dPhi_N,i(g) A_i(g)[r][u],
//   SH_g[u][5M + s] = sum_j dPhi_M,j(g) S_j(g)[u][s],
// c_A,g = dt^2 / 2 w_g det J_g.  Tiles of v_mfma_f64_16x16x4f64: 3 x 3 output
// tiles (rows and columns padded to 48), 10 steps of 4 along the inner
// index.  Operand layout of that instruction on gfx950: A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in register i / 4 of lane
// j + 16 (i % 4).  Each operand entry is formed once, by the lane that feeds
// it to the matrix core.  The 40 x 40 result goes to out[i * 40 + j].
typedef double dvec4_t __attribute__((ext_vector_type(4)));
// Inner step ks (inner indices 4 ks .. 4 ks + 3) of the product into acc
__device__ __forceinline__ void mfma_as_kstep(const GaussHex& G, const GPTab* tabs, double dtsq_2, int ks, int lane,
                                              dvec4_t acc[3][3]) {
    const int li = lane & 15, kk = lane >> 4;
    const int k = 4 * ks + kk, g = k / 5, u = k - 5 * g;
    const GPTab& t = tabs[g];
    const double cA = dtsq_2 * G.p[g].wJ;
    double av[3], bv[3];
#pragma unroll
    for (int mt = 0; mt < 3; mt++) {
        const int i = 16 * mt + li;       // row (N, r) of the element block
        double v = 0.0;
        if (i < 40) {
            const int N = i / 5, r = i - 5 * (i / 5);
            v = cA * (t.dphi[N][0] * t.Af[0][r * 5 + u] + t.dphi[N][1] * t.Af[1][r * 5 + u] +
                      t.dphi[N][2] * t.Af[2][r * 5 + u]);
        }
        av[mt] = v;
    }
#pragma unroll
    for (int nt = 0; nt < 3; nt++) {
        const int j = 16 * nt + li;       // column (M, s)
        double v = 0.0;
        if (j < 40) {
            const int M = j / 5, sc = j - 5 * (j / 5);
            v = t.dphi[M][0] * t.Sf[0][u * 5 + sc] + t.dphi[M][1] * t.Sf[1][u * 5 + sc] +
                t.dphi[M][2] * t.Sf[2][u * 5 + sc];
        }
        bv[nt] = v;
    }
#pragma unroll
    for (int mt = 0; mt < 3; mt++)
#pragma unroll
        for (int nt = 0; nt < 3; nt++)
            acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt], bv[nt], acc[mt][nt], 0, 0, 0);
}
// The 40 x 40 result to out[i * 40 + j] (out may alias the point tables:
// every lane is done reading them first)
__device__ __forceinline__ void mfma_as_store(const dvec4_t acc[3][3], int lane, double* out) {
    const int li = lane & 15, kk = lane >> 4;
    __syncthreads();
#pragma unroll
    for (int mt = 0; mt < 3; mt++)
#pragma unroll
        for (int nt = 0; nt < 3; nt++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = 16 * mt + 4 * q + kk, j = 16 * nt + li;
                if (i < 40 && j < 40) out[i * 40 + j] = acc[mt][nt][q];
            }
    __syncthreads();
}

template <bool W64, bool W32>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void build3d_interior_kernel(BuildArgs A) {
    // one wavefront (workgroup) per element slot
    constexpr int NNOD = 8, NEQ = 5, K = 40, KK = 1600;
    __shared__ double s_nd[NNOD][ND_STRIDE];
    __shared__ GPTab s_tab[8];
    __shared__ GaussHex s_g;
    const int lane = threadIdx.x;
    const long e = A.e_lo + (long)xcd_remap(blockIdx.x, gridDim.x);
    const DevThermo th = A.th;
    if (e >= A.e_hi) return;   // whole workgroup leaves together
    const int* cn = A.conn + e * NNOD;
    const bool real = cn[0] >= 0;
    bool boundary = false, hang = false;
    if (real) {
        const int* bcs = A.bc + e * 6;
        boundary = A.cmask[e] != 0ULL;
#pragma unroll
        for (int fc = 0; fc < 6; fc++) boundary |= (bcs[fc] != 0);
        // with FP64 blocks, a boundary element's domain part is written here
        // (FP64 R and rhs) and build3d_face_kernel adds faces and ApplyBC;
        // without, the per-lane kernel builds it whole.  Hanging-node
        // elements without faces are built here on Hnm-mapped tables;
        // those with faces are built whole by the per-lane kernel.
        hang = A.adap && A.adap[e];
        if (hang && boundary) return;
        if (boundary && !W64) return;
    } else {   // padding slot of a color range: zero block
        if (W64 && A.R64)
            for (int q = lane; q < KK; q += 64) A.R64[r64_at<8, 5>(A, e, q)] = 0.0;
        if (W32)
            for (int q = lane; q < KK; q += 64) A.R32[blk_index<float>(e, q, KK)] = 0.0f;
        if (lane < K) A.rhs[e * K + lane] = 0.0;
        return;
    }
    if (lane < NNOD) load_node3d(A, cn[lane], s_nd[lane]);
    if (!hang) {
        s_g.p[lane >> 3].phi[lane & 7] = c_gauss.phi3[lane >> 3][lane & 7];
#pragma unroll
        for (int j = 0; j < 3; j++) s_g.p[lane >> 3].dpxi[lane & 7][j] = c_gauss.dpxi3[lane >> 3][lane & 7][j];
    } else {
        // adap() (MeshRefine2D.h:877-900): Phi <- Phi Hnm, dPxi <- dPxi Hnm,
        // lane (gp, M) one column, summed over N in order
        const int gp = lane >> 3, m = lane & 7;
        const double* Hn = A.hnm + e * 64;
        double ph = 0.0, dp[3] = {0.0, 0.0, 0.0};
        for (int n = 0; n < 8; n++) {
            const double h = Hn[n + 8 * m];
            ph += c_gauss.phi3[gp][n] * h;
#pragma unroll
            for (int j = 0; j < 3; j++) dp[j] += c_gauss.dpxi3[gp][n][j] * h;
        }
        s_g.p[gp].phi[m] = ph;
#pragma unroll
        for (int j = 0; j < 3; j++) s_g.p[gp].dpxi[m][j] = dp[j];
    }
    __syncthreads();
    double sp[4];
    int tie = 0;
    fdvparam_wave(th, s_nd, lane, sp, &tie);
    // phase 1: (a) lane (gp, n) = (lane / 8, lane % 8): geometry of point
    // gp and dPhi_n/dx; (b) lane (gp, k): the interpolation sums of sources k
    // and k + 8, into the point's A/S table area (written only after every
    // lane has read them back); (c) lanes (part, gp) = (lane / 8, lane % 8),
    // part < 3: the point's state from the sums (the three lanes of a point
    // alike), then direction part's Jacobian combinations and share of the
    // c_ij table
    {
        const int gp = lane >> 3, n = lane & 7;
        double wJ;
        gp_geom(s_g.p[gp], c_gauss.w3[gp], s_nd, n, s_tab[gp].dphi[n], &wJ);
        if (n == 0) s_g.p[gp].wJ = wJ;
    }
    __syncthreads();
    static_assert(sizeof(GPTab::Af) >= 4 * GP_NSRC * sizeof(double), "sums fit the A table");
    {
        const int gp = lane >> 3, k = lane & 7;
        double* S = &s_tab[gp].Af[0][0];
        double o4[4];
        gp_sums(s_g.p[gp], s_tab[gp].dphi, s_nd, gp_src_col(k), o4);
#pragma unroll
        for (int c = 0; c < 4; c++) S[4 * k + c] = o4[c];
        if (k + 8 < GP_NSRC) {
            gp_sums(s_g.p[gp], s_tab[gp].dphi, s_nd, gp_src_col(k + 8), o4);
#pragma unroll
            for (int c = 0; c < 4; c++) S[4 * (k + 8) + c] = o4[c];
        }
    }
    __syncthreads();
    GPShared g;
    const int gp1 = lane & 7, part = lane >> 3;
    if (lane < 24) gp_state(th, &s_tab[gp1].Af[0][0], A.dt, g);
    __syncthreads();
    if (lane < 24) gp_tab_dir_rot(th, g, sp, A.dt, part, s_tab[gp1]);
    __syncthreads();
    // phase 2: lane (N, M) contracts every point for its node pair; the
    // c_A A S term of all pairs runs on the matrix cores, its 40 x 40 result
    // lands in LDS over the (then consumed) point tables
    const int N = lane >> 3, M = lane & 7;
    double R[NEQ][NEQ];
    // the A S product's 10 inner steps are spread over the 8 points of the
    // VALU contraction (steps 8 gp / 10 .. 8 (gp + 1) / 10 at point gp)
    const double dtsq_2 = (A.dt * A.dt) / 2.0;
    dvec4_t acc[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) acc[a][b] = dvec4_t{0.0, 0.0, 0.0, 0.0};
    interior_phase2(th, s_g, s_tab, lane, A.dt, sp, R, [&](int gp) {
        for (int ks = (gp * 10) >> 3; ks < ((gp + 1) * 10) >> 3; ks++) mfma_as_kstep(s_g, s_tab, dtsq_2, ks, lane, acc);
    });
    // the rhs, entry (N, r) per lane < 40, before the product overwrites
    // the point tables
    const double rhs_l = lane < K ? interior_rhs(s_g, s_tab, lane) : 0.0;
    double* s_as = reinterpret_cast<double*>(s_tab);
    static_assert(sizeof(s_tab) >= 1600 * sizeof(double), "point tables hold the 40 x 40 product");
    mfma_as_store(acc, lane, s_as);
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) R[r][s] += s_as[(NEQ * N + r) * K + NEQ * M + s];
    if (A.S64) {
        if (boundary) {
            double* dst = rb_block(A, e, N, M);
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int s = 0; s < NEQ; s++) dst[r * NEQ + s] = R[r][s];
        } else {
            rows_put(A, e, N, M, R);
        }
    } else if (W64) {
        r64_store<8, 5>(A, e, N, M, R);
    }
    if (W32 && !boundary) {
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++)
                A.R32[blk_index<float>(e, (NEQ * N + r) * K + NEQ * M + s, KK)] = (float)R[r][s];
    }
    if (lane < K) A.rhs[e * K + lane] = rhs_l;
    if (lane == 0) {
        if (A.svals) {
#pragma unroll
            for (int i = 0; i < 4; i++) A.svals[e * 4 + i] = sp[i];
        }
        if (tie) atomicAdd(A.ties, 1);
    }
}

// ---------------------------------------------------------------------------
// Boundary faces of non-hanging hex8 elements, in two phases like the
// interior kernel: the state and Jacobian combinations of each face Gauss
// point are formed once, by a lane group of their own, into LDS; then lane
// (N, M) contracts them for its node pair.  Same terms and point order as
// build3d_faces (element3d.h), which the per-lane kernel still uses for
// hanging-node elements.
// ---------------------------------------------------------------------------
struct FaceTab {
    GPTab t;             // M1, Af, Sf, PQ and the c_ij table (dphi, wJ unused)
    double en[3];        // w0 w1 det J times the unit face normal (TFNorm3D)
    double phib[4];      // face shape functions at the point
    double dpx[4][3];    // their x derivatives
    int fl[4];           // element-local nodes of the face
    int ok;              // 0: unknown bc code, the point adds nothing
};

// What lane (N, M) adds for one face point, with t = the face position of
// node M: R_NM += dt phi_N phi_M G + c_A,N PT[t] + dt s3 phi_N CT[t] and
// rhs_N += phi_N Q (c_A,N = -dt^2/2 phi_N), where
//   G     = sum_i en_i M1_i,
//   PT[t] = (sum_i en_i A_i)(sum_j dphi_t,j S_j),
//   CT[t] = sum_ij en_i dphi_t,j c_ij / Re,
//   Q     = -sum_i en_i P_i + dt^2/2 sum_i en_i (A_i D);
// none of them depends on N, so each is formed once per point.
struct FaceMat {
    double G[25];
    double PT[4][25];
    double CT[4][25];
    double Q[5];
};

// State of face point fgp on face fc (TFNorm3D, calc_TF_norm3D.h:9-167, the
// face interpolation and eval2dboundary's edits, FDVGalerkin.h:455-533),
// as build3d_faces computes it, into g (f, D, P, SPhi) and the geometry.
// Returns false (and sets *status bit 0) for an unknown bc code.
__device__ bool face_point(const DevThermo& th, const GaussTables& G, const double (*nd)[ND_STRIDE], int fc, int fgp,
                           int bc, double dt, GPShared& g, double en[3], double phib[4], double dpx[4][3], int fl[4],
                           int* status) {
    constexpr int NEQ = 5;
    const double reinv = 1.0 / th.creyn;
    const int FN[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {6, 7, 4, 5}};
    const double w0 = G.wf[fgp][0], w1 = G.wf[fgp][1];
    const double z0 = G.zf[fgp][0], z1 = G.zf[fgp][1];
    double nl[4][3];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        fl[t] = FN[fc][t];
#pragma unroll
        for (int i = 0; i < 3; i++) nl[t][i] = nd[fl[t]][ND_X + i];
    }
    phib[0] = 0.25 * (1 - z0) * (1 - z1);
    phib[1] = 0.25 * (1 + z0) * (1 - z1);
    phib[2] = 0.25 * (1 + z0) * (1 + z1);
    phib[3] = 0.25 * (1 - z0) * (1 + z1);
    // the Lambda/Mu magnitude is recomputed after each component
    // (calc_TF_norm3D.h:74-78)
    double Lm[3], Mu[3];
#pragma unroll
    for (int i = 0; i < 3; i++) { Lm[i] = nl[1][i] - nl[0][i]; Mu[i] = nl[3][i] - nl[0][i]; }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Lm[j] /= sqrt(Lm[0] * Lm[0] + Lm[1] * Lm[1] + Lm[2] * Lm[2]);
        Mu[j] /= sqrt(Mu[0] * Mu[0] + Mu[1] * Mu[1] + Mu[2] * Mu[2]);
    }
    double en_[3], ga[3];
    en_[0] = Lm[1] * Mu[2] - Lm[2] * Mu[1];
    en_[1] = Lm[2] * Mu[0] - Lm[0] * Mu[2];
    en_[2] = Lm[0] * Mu[1] - Lm[1] * Mu[0];
    ga[0] = en_[1] * Lm[2] - en_[2] * Lm[1];
    ga[1] = en_[2] * Lm[0] - en_[0] * Lm[2];
    ga[2] = en_[0] * Lm[1] - en_[1] * Lm[0];
    double dp[4][2];
    dp[0][0] = -0.25 * (1 - z1); dp[1][0] = 0.25 * (1 - z1);
    dp[2][0] = 0.25 * (1 + z1);  dp[3][0] = -0.25 * (1 + z1);
    dp[0][1] = -0.25 * (1 - z0); dp[1][1] = -0.25 * (1 + z0);
    dp[2][1] = 0.25 * (1 + z0);  dp[3][1] = 0.25 * (1 - z0);
    double dxd[3][2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < 4; t++) acc += dp[t][j] * nl[t][i];
            dxd[i][j] = acc;
        }
    double dtr[2][2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        dtr[0][j] = Lm[0] * dxd[0][j] + Lm[1] * dxd[1][j] + Lm[2] * dxd[2][j];
        dtr[1][j] = ga[0] * dxd[0][j] + ga[1] * dxd[1][j] + ga[2] * dxd[2][j];
    }
    const double detjac = dtr[0][0] * dtr[1][1] - dtr[1][0] * dtr[0][1];
    const double djinv = 1. / detjac;
    if (detjac < 1.0E-12) *status |= 2;
#pragma unroll
    for (int i = 0; i < 3; i++) en[i] = en_[i] * w0 * w1 * detjac;
    GPState<3>& f = g.f;
    f.E = 0; f.T = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f.v[i] = 0; f.dT[i] = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) { f.dV[i][j] = 0; f.tau[i][j] = 0; }
    }
#pragma unroll
    for (int q = 0; q < NEQ; q++) {
        f.U[q] = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) f.dU[q][j] = 0;
    }
    double Sph = 0.0, Sdp[3] = {0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = fl[t];
        const double d0 = (dtr[1][1] * dp[t][0] - dtr[1][0] * dp[t][1]) * djinv;
        const double d1 = (-dtr[0][1] * dp[t][0] + dtr[0][0] * dp[t][1]) * djinv;
#pragma unroll
        for (int i = 0; i < 3; i++) dpx[t][i] = Lm[i] * d0 + ga[i] * d1;
        const double ph = phib[t];
        Sph += ph;
        f.E += ph * nd[n][ND_E];
        const double Tt = nd[n][ND_T];
        f.T += ph * Tt;
#pragma unroll
        for (int q = 0; q < NEQ; q++) {
            const double u_ = nd[n][ND_U + q];
            f.U[q] += ph * u_;
#pragma unroll
            for (int j = 0; j < 3; j++) f.dU[q][j] += dpx[t][j] * u_;
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double v_ = nd[n][ND_V + i];
            Sdp[i] += dpx[t][i];
            f.v[i] += ph * v_;
            f.dT[i] += dpx[t][i] * Tt;
#pragma unroll
            for (int j = 0; j < 3; j++) f.dV[i][j] += dpx[t][j] * v_;
        }
    }
    // the unedited sums feed D and P
    double vph[3], dVp[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        vph[i] = f.v[i];
#pragma unroll
        for (int j = 0; j < 3; j++) dVp[i][j] = f.dV[i][j];
    }
    f.mu = calc_mu(th, f.T);
    f.k = calc_k(th, f.mu);
    if (!eval2dboundary(th, bc, f, en)) {
        *status |= 1;
        return false;
    }
    g.wJ = 0.0;
    g.SPhi = Sph;
    g.D[0] = 0.0;
#pragma unroll
    for (int s = 1; s < 4; s++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) acc += (-f.tau[j][s - 1] * reinv) * Sdp[j];
        g.D[s] = acc;
    }
    {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double td = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) td += f.tau[j][k] * dVp[k][j];
            acc += (-td - f.k * f.dT[j] * Sdp[j]) * reinv;
        }
        g.D[4] = acc;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        g.P[i][0] = 0.0;
#pragma unroll
        for (int r = 1; r < NEQ; r++) {
            double P;
            if (r < 4) P = dt * (-f.tau[i][r - 1] * reinv) * Sph;
            else {
                double tv = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++) tv += f.tau[i][k] * vph[k];
                P = dt * ((-tv - f.k * f.dT[i] * Sph) * reinv);
            }
            g.P[i][r] = P;
        }
    }
    return true;
}

// Part `part` of face point F's node-pair matrices: part t < 4 forms PT[t]
// and CT[t], part 4 forms G and Q; s1 is the first M1 slot of every direction.
__device__ void face_mats(const FaceTab& F, int part, double s1, FaceMat& out) {
    constexpr int NEQ = 5;
    const GPTab& t = F.t;
    if (part < 4) {
        double Ah[5][5];
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int u = 0; u < NEQ; u++)
                Ah[r][u] = F.en[0] * t.Af[0][r * 5 + u] + F.en[1] * t.Af[1][r * 5 + u] + F.en[2] * t.Af[2][r * 5 + u];
        const double d0 = F.dpx[part][0], d1 = F.dpx[part][1], d2 = F.dpx[part][2];
#pragma unroll
        for (int s = 0; s < NEQ; s++) {
            double Sh[5];
#pragma unroll
            for (int u = 0; u < NEQ; u++) Sh[u] = d0 * t.Sf[0][u * 5 + s] + d1 * t.Sf[1][u * 5 + s] + d2 * t.Sf[2][u * 5 + s];
#pragma unroll
            for (int r = 0; r < NEQ; r++) {
                double pr = 0.0;
#pragma unroll
                for (int u = 0; u < NEQ; u++) pr += Ah[r][u] * Sh[u];
                out.PT[part][r * 5 + s] = pr;
            }
        }
        double W[3][3], C[5][5];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) W[i][j] = F.en[i] * F.dpx[part][j];
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) C[r][s] = 0.0;
        accum_c3_ctab(t.C, W, C);
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) out.CT[part][r * 5 + s] = C[r][s];
    } else {
        double G[5][5];
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) G[r][s] = 0.0;
#pragma unroll
        for (int I = 0; I < 3; I++)
#pragma unroll
            for (int p = 0; p < 17; p++)
                G[sig_eq(I, JPAT_R[0][p])][sig_eq(I, JPAT_S[0][p])] += F.en[I] * (p == 0 ? s1 : t.M1[I][p - 1]);
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++) out.G[r * 5 + s] = G[r][s];
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) acc += F.en[i] * t.PQ[i][sig_eq(i, r)];
            out.Q[r] = -acc;
        }
    }
}

// Face point p's contribution for lane (N, M) from its matrices.
__device__ __forceinline__ void face_contract(const FaceTab& F, const FaceMat& Mx, int N, int M, double dt,
                                              double s3, double R[5][5], double rhs[5]) {
    constexpr int NEQ = 5;
    int fN = -1, fM = -1;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (F.fl[t] == N) fN = t;
        if (F.fl[t] == M) fM = t;
    }
    if (fN < 0) return;   // this lane's row is not on the face
    const double phN = F.phib[fN];
    if (M == 0) {
#pragma unroll
        for (int r = 0; r < NEQ; r++) rhs[r] += phN * Mx.Q[r];
    }
    if (fM < 0) return;   // column node off the face: no block terms
    const double cg = dt * phN * F.phib[fM], cA = -(dt * dt) / 2.0 * phN, cc = dt * s3 * phN;
    const double* PT = Mx.PT[fM];
    const double* CT = Mx.CT[fM];
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) R[r][s] += cg * Mx.G[r * 5 + s] + cA * PT[r * 5 + s] + cc * CT[r * 5 + s];
}

// Boundary elements whose domain part is already in R64/rhs: face integrals
// and ApplyBC on top.  One wavefront per listed slot (no hanging-node
// elements: those are built whole by build3d_kernel).  Face points come in
// build3d_faces' order (point-major over the faces with 0 < bc < 100), four
// at a time: lanes (point, part), part < 4, form the point tables (part
// 0..2 one direction's Jacobian combinations and share of the c_ij table
// each, part 3 the geometry), then lanes (point, part), part < 5, the node-pair
// matrices, then lane (N, M) adds them to its block.
// ONE: every listed element has at most one face with a bc code (one batch
// of points): the domain block is read after the point tables are formed,
// so that it holds no registers while they are (the loop over batches keeps
// it live across them, which costs the general kernel half its occupancy).
template <bool W32, bool ONE>
__global__ __launch_bounds__(64) void build3d_face_kernel(BuildArgs A) {
    constexpr int NNOD = 8, NEQ = 5, K = 40, KK = 1600, PB = 4;
    __shared__ double s_nd[NNOD][ND_STRIDE];
    __shared__ FaceTab s_ft[PB];
    __shared__ FaceMat s_fm[PB];
    __shared__ int s_status;
    const int lane = threadIdx.x;
    if ((long)blockIdx.x >= A.n_bslots) return;
    const long e = A.bslots[blockIdx.x];
    const int* cn = A.conn + e * NNOD;
    if (lane < NNOD) load_node3d(A, cn[lane], s_nd[lane]);
    if (lane == 0) s_status = 0;
    __syncthreads();
    const DevThermo th = A.th;
    double sp[4];
    int tie = 0;
    fdvparam_wave(th, s_nd, lane, sp, &tie);
    const int N = lane >> 3, M = lane & 7;
    const int* bcs = A.bc + e * 6;
    int nfa = 0;
#pragma unroll
    for (int fc = 0; fc < 6; fc++) nfa += (bcs[fc] != 0 && bcs[fc] < 100);
    const int npts = 4 * nfa;
    if (ONE && npts > PB) {   // not an element of this list (the host sorts them)
        if (lane == 0) atomicOr(A.status, 4);
        return;
    }
    // the domain part of R and rhs (written by the interior kernel), read
    // before the point tables are formed (so that its HBM latency overlaps
    // them) or, ONE, after them; written once after the last batch of points
    double R[NEQ][NEQ], rhs[NEQ];
    auto load_domain = [&]() {
#pragma unroll
        for (int r = 0; r < NEQ; r++) rhs[r] = (M == 0) ? A.rhs[e * K + NEQ * N + r] : 0.0;
        if (A.S64) {
            const double* src = rb_block(A, e, N, M);
#pragma unroll
            for (int r = 0; r < NEQ; r++)
#pragma unroll
                for (int s = 0; s < NEQ; s++) R[r][s] = src[r * NEQ + s];
        } else {
            r64_load<8, 5>(A, e, N, M, R);
        }
    };
    if (!ONE) load_domain();
    for (int b0 = 0; b0 < npts || b0 == 0; b0 += PB) {
        const int nb = (npts - b0 < PB) ? npts - b0 : PB;
        const int p = lane & 7, part = lane >> 3;
        if (p < nb && part < 4) {
            const int i = b0 + p, fgp = i / nfa, k = i - fgp * nfa;
            int fc = 0, seen = 0;
            for (int c = 0; c < 6; c++)
                if (bcs[c] != 0 && bcs[c] < 100) {
                    if (seen == k) fc = c;
                    seen++;
                }
            GPShared g;
            double en[3], phib[4], dpx[4][3];
            int fl[4], st = 0;
            const bool ok = face_point(th, c_gauss, s_nd, fc, fgp, bcs[fc], A.dt, g, en, phib, dpx, fl, &st);
            FaceTab& F = s_ft[p];
            if (part == 3) {
                F.ok = ok ? 1 : 0;
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    F.fl[t] = fl[t];
                    F.phib[t] = phib[t];
#pragma unroll
                    for (int j = 0; j < 3; j++) F.dpx[t][j] = dpx[t][j];
                }
#pragma unroll
                for (int j = 0; j < 3; j++) F.en[j] = en[j];
                if (st) atomicOr(&s_status, st);
            }
            if (ok && part < 3) gp_tab_dir_rot(th, g, sp, A.dt, part, F.t);
        }
        __syncthreads();
        if (p < nb && part < 5 && s_ft[p].ok) face_mats(s_ft[p], part, sp[0], s_fm[p]);
        if (ONE) load_domain();
        __syncthreads();
#pragma unroll 1
        for (int q = 0; q < nb; q++)
            if (s_ft[q].ok) face_contract(s_ft[q], s_fm[q], N, M, A.dt, sp[2], R, rhs);
        if (ONE || b0 + PB >= npts) break;
        __syncthreads();   // the tables are rewritten by the next batch
    }
    apply_bc3(A, e, lane, R, rhs);
    if (A.S64) rows_put(A, e, N, M, R);
    else r64_store<8, 5>(A, e, N, M, R);
    if (W32 && A.R32) {
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++)
                A.R32[blk_index<float>(e, (NEQ * N + r) * K + NEQ * M + s, KK)] = (float)R[r][s];
    }
    if (M == 0) {
#pragma unroll
        for (int r = 0; r < NEQ; r++) A.rhs[e * K + NEQ * N + r] = rhs[r];
    }
    if (lane == 0 && s_status) atomicOr(A.status, s_status);
}

template <bool W64, bool W32>
__global__ __launch_bounds__(256) void build3d_kernel(BuildArgs A, const int* list, long n_list) {
    constexpr int NNOD = 8, NEQ = 5, K = 40, KK = 1600;
    __shared__ double s_nd[4][NNOD][ND_STRIDE];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // with a slot list, entry i of the list; otherwise every slot
    const long i_ = (long)blockIdx.x * 4 + wv;
    const long e = list ? (i_ < n_list ? (long)list[i_] : A.n_elem_pad)
                        : (long)xcd_remap(blockIdx.x, gridDim.x) * 4 + wv;
    const DevThermo th = A.th;
    const bool in_range = e < A.n_elem_pad;
    const int* cn = A.conn + (in_range ? e : 0) * NNOD;
    const bool real = in_range && cn[0] >= 0;
    double (*nd)[ND_STRIDE] = s_nd[wv];
    if (real && lane < NNOD) load_node3d(A, cn[lane], nd[lane]);
    __syncthreads();
    if (!in_range) return;
    if (!real) {   // padding slot of a color range: zero block
        if (W64 && A.R64)
            for (int q = lane; q < KK; q += 64) A.R64[r64_at<8, 5>(A, e, q)] = 0.0;
        if (W32 && A.R32)
            for (int q = lane; q < KK; q += 64) A.R32[blk_index<float>(e, q, KK)] = 0.0f;
        if (lane < K) A.rhs[e * K + lane] = 0.0;
        return;
    }
    const int N = lane >> 3, M = lane & 7;
    double R[NEQ][NEQ], rhs[NEQ], sp[4];
    int tie = 0, status = 0;
    build3d_elem(A, A.th, c_gauss, nd, e, lane, R, rhs, sp, &tie, &status);
    if (status && lane == 0) atomicOr(A.status, status);
    // ---------------- write ----------------
    if (A.S64) rows_put(A, e, N, M, R);
    else if (W64) r64_store<8, 5>(A, e, N, M, R);
    if (W32 && A.R32) {
#pragma unroll
        for (int r = 0; r < NEQ; r++)
#pragma unroll
            for (int s = 0; s < NEQ; s++)
                A.R32[blk_index<float>(e, (NEQ * N + r) * K + NEQ * M + s, KK)] = (float)R[r][s];
    }
    if (M == 0) {
#pragma unroll
        for (int r = 0; r < NEQ; r++) A.rhs[e * K + NEQ * N + r] = rhs[r];
    }
    if (lane == 0) {
        if (A.svals) {
#pragma unroll
            for (int i = 0; i < 4; i++) A.svals[e * 4 + i] = sp[i];
        }
        if (tie) atomicAdd(A.ties, 1);
    }
}

// ---------------------------------------------------------------------------
// quad4 build kernel: four elements per wavefront, 16 lanes each
// ---------------------------------------------------------------------------
// One quad4 element seen from lane (N, M) = (l16/4, l16%4).
CFDV_HD void build2d_elem(const BuildArgs& A, const GaussTables& G, long e, int l16, double R[4][4],
                          double rhs[4], double sp[4], int* tie, int* status) {
    // No FMA contraction here: the discontinuity-capturing branch
    // (FDVGalerkin.h:349-379) tests gamma = |v||A| == 0, and on elements with
    // a uniform velocity along a side A is an exact zero only when the
    // Jacobian, dPhi/dx and dV sums round as in the reference's non-FMA
    // arithmetic; a contracted residue flips tau_b from 0 to tau.
#pragma clang fp contract(off)
    constexpr int NEQ = 4;
    const int N = l16 >> 2, M = l16 & 3;
    const DevThermo th = A.th;
    const int* cn = A.conn + e * 4;
    double xn[4][2], Un[4][4], vn[4][2], En[4], Tn[4], rhon[4], pn[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int g = cn[n];
        xn[n][0] = A.x[(size_t)g * 2 + 0];
        xn[n][1] = A.x[(size_t)g * 2 + 1];
#pragma unroll
        for (int q = 0; q < 4; q++) Un[n][q] = A.U[(size_t)g * 4 + q];
        unpack<2>(th, Un[n], rhon[n], vn[n], En[n], Tn[n]);
        const double vs = vn[n][0] * vn[n][0] + vn[n][1] * vn[n][1];
        pn[n] = rhon[n] * (th.gamma - 1.0) * (En[n] - 0.5 * vs);
    }
    fdvparam<2, 4>(th, vn, Tn, rhon, sp, tie);
    const double s1 = sp[0], s2 = sp[1], s3 = sp[2], s4 = sp[3];
    const double dt = A.dt, dtsq_2 = (dt * dt) / 2.0;
    const double reinv = 1.0 / th.creyn;

    // node fluxes F (NavierStokes2D::calc_F, :11-29) are state-only
    double Fn[4][2][4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const double rho = Un[n][0], l = Un[n][1], m = Un[n][2], en = Un[n][3], p = pn[n];
        Fn[n][0][0] = l; Fn[n][0][1] = p + (l * l) / rho; Fn[n][0][2] = (l * m) / rho; Fn[n][0][3] = (p + en) * (l / rho);
        Fn[n][1][0] = m; Fn[n][1][1] = (l * m) / rho; Fn[n][1][2] = p + (m * m) / rho; Fn[n][1][3] = (p + en) * (m / rho);
    }
    // element sizes for the numerical diffusion (FDVGalerkin.h:284-297, Q9),
    // summed node by node as the reference does: alpha = coth(R/2) - 2/R
    // cancels for small R, so the last bit of h matters
    double h_xi = 0.0, h_eta = 0.0;
    {
        constexpr double XR[4] = {0.0, 0.5, 0.5, 0.0}, XL[4] = {0.5, 0.0, 0.0, 0.5};
        constexpr double XT[4] = {0.5, 0.5, 0.0, 0.0}, XB[4] = {0.0, 0.0, 0.5, 0.5};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            h_xi += XR[n] * xn[n][0] - XL[n] * xn[n][0];
            h_eta += XT[n] * xn[n][1] - XB[n] * xn[n][1];
        }
    }

#pragma unroll
    for (int r = 0; r < NEQ; r++) {
        rhs[r] = 0.0;
#pragma unroll
        for (int s = 0; s < NEQ; s++) R[r][s] = 0.0;
    }
    // hanging-node element: Phi <- Phi Hnm, dPhi/dxi <- dPhi/dxi Hnm
    // (adap(), MeshRefine2D.h:877-900; FDVGalerkin.h:183-186)
    const double* H = (A.adap && A.adap[e]) ? A.hnm + e * 16 : nullptr;
    for (int gp = 0; gp < 4; gp++) {
        double phl[4], dpl[4][2];
#pragma unroll
        for (int n = 0; n < 4; n++) {
            phl[n] = G.phi2[gp][n];
            dpl[n][0] = G.dpxi2[gp][n][0];
            dpl[n][1] = G.dpxi2[gp][n][1];
        }
        // the numerical diffusion's Jacobian is always the plain one
        // (a fresh dtestfunction, FDVGalerkin.h:237-241)
        double jn[2][2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < 4; n++) acc += G.dpxi2[gp][n][j] * xn[n][i];
                jn[i][j] = acc;
            }
        double jac[2][2] = {{jn[0][0], jn[0][1]}, {jn[1][0], jn[1][1]}};
        if (H) {
            double pa[4] = {0, 0, 0, 0}, da[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
#pragma unroll
            for (int Nn = 0; Nn < 4; Nn++)
#pragma unroll
                for (int Mm = 0; Mm < 4; Mm++) {
                    const double h = H[Nn + 4 * Mm];
                    pa[Mm] += phl[Nn] * h;
                    da[Mm][0] += dpl[Nn][0] * h;
                    da[Mm][1] += dpl[Nn][1] * h;
                }
#pragma unroll
            for (int n = 0; n < 4; n++) {
                phl[n] = pa[n];
                dpl[n][0] = da[n][0];
                dpl[n][1] = da[n][1];
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    double acc = 0.0;
#pragma unroll
                    for (int n = 0; n < 4; n++) acc += dpl[n][j] * xn[n][i];
                    jac[i][j] = acc;
                }
        }
        const double J = jac[0][0] * jac[1][1] - jac[1][0] * jac[0][1];
        const double Jinv = 1.0 / J;
        double dphi[4][2];
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const double d0 = dpl[n][0], d1 = dpl[n][1];
            dphi[n][0] = Jinv * (d0 * jac[1][1] - jac[1][0] * d1);
            dphi[n][1] = Jinv * (jac[0][0] * d1 - d0 * jac[0][1]);
        }
        GPState<2> f;
        f.E = 0; f.T = 0;
#pragma unroll
        for (int i = 0; i < 2; i++) { f.v[i] = 0; f.dT[i] = 0; f.dV[i][0] = f.dV[i][1] = 0; }
#pragma unroll
        for (int q = 0; q < 4; q++) { f.U[q] = 0; f.dU[q][0] = f.dU[q][1] = 0; }
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const double ph = phl[n];
            f.E += ph * En[n];
            f.T += ph * Tn[n];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                f.U[q] += ph * Un[n][q];
#pragma unroll
                for (int j = 0; j < 2; j++) f.dU[q][j] += dphi[n][j] * Un[n][q];
            }
#pragma unroll
            for (int i = 0; i < 2; i++) {
                f.v[i] += ph * vn[n][i];
                f.dT[i] += dphi[n][i] * Tn[n];
#pragma unroll
                for (int j = 0; j < 2; j++) f.dV[i][j] += dphi[n][j] * vn[n][i];
            }
        }
        f.mu = calc_mu(th, f.T);
        f.k = calc_k(th, f.mu);
        calc_tau<2>(f);
        // numerical diffusion and discontinuity capturing (:225-387)
        double Psi[4];
        {
            const double J_xi = jn[0][0] * jn[0][0] + jn[1][0] * jn[1][0];
            const double J_eta = jn[0][1] * jn[0][1] + jn[1][1] * jn[1][1];
            const double ixi = 1. / pow(J_xi, 0.5), ieta = 1. / pow(J_eta, 0.5);
            const double v_xi = f.v[0] * (ixi * jn[0][0]) + f.v[1] * (ixi * jn[1][0]);
            const double v_eta = f.v[0] * (ieta * jn[0][1]) + f.v[1] * (ieta * jn[1][1]);
            const double S = f.v[0] * f.v[0] + f.v[1] * f.v[1];
            const double R_xi = f.U[0] * v_xi * h_xi / f.mu;
            const double R_eta = f.U[0] * v_eta * h_eta / f.mu;
            // coth by the host libm's method (libm_cosh / libm_sinh), bit for bit
            const double al_xi = (R_xi == 0) ? 0.0 : libm_cosh(R_xi / 2.) / libm_sinh(R_xi / 2.) - 2. / R_xi;
            const double al_eta = (R_eta == 0) ? 0.0 : libm_cosh(R_eta / 2.) / libm_sinh(R_eta / 2.) - 2. / R_eta;
            const double taun = (S == 0) ? 0.0 : (1. / 4.) * (al_xi * h_xi * v_xi + al_eta * h_eta * v_eta) / S;
            double Av[2] = {0, 0};
#pragma unroll
            for (int k = 0; k < 2; k++)
#pragma unroll
                for (int j = 0; j < 2; j++) Av[j] += f.dV[j][k] * f.v[k];
            const double gam = sqrt(f.v[0] * f.v[0] + f.v[1] * f.v[1]) * sqrt(Av[0] * Av[0] + Av[1] * Av[1]);
            double tau_b = 0;
#pragma unroll
            for (int j = 0; j < 2; j++) tau_b += taun * f.v[j] * Av[j] / gam;
            if ((tau_b - taun) < 0) tau_b = taun;
            if (gam == 0) tau_b = 0;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                double ps = 0.0;
#pragma unroll
                for (int i = 0; i < 2; i++) ps += taun * f.v[i] * dphi[n][i];
#pragma unroll
                for (int i = 0; i < 2; i++) ps += tau_b * f.v[i] * dphi[n][i];
                Psi[n] = ps;
            }
        }
        double a[2][4][4], b[2][4][4];
        calc_a2(th, f, a);
        calc_b2(th, f, b);
        double PN = 0, PM = 0, DN[2] = {0, 0}, DM[2] = {0, 0}, PsiN = 0;
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const double ph = phl[n];
            if (n == N) { PN = ph; DN[0] = dphi[n][0]; DN[1] = dphi[n][1]; PsiN = Psi[n]; }
            if (n == M) { PM = ph; DM[0] = dphi[n][0]; DM[1] = dphi[n][1]; }
        }
        const double wJ = G.w2[gp] * J;
        // rhs with F + G at nodes
        {
            double P[2][4], D[4];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int r = 0; r < 4; r++) P[i][r] = 0.0;
#pragma unroll
            for (int s = 0; s < 4; s++) D[s] = 0.0;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                double Gv[2][4];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    Gv[i][0] = 0.0;
                    Gv[i][1] = -f.tau[i][0] * reinv;
                    Gv[i][2] = -f.tau[i][1] * reinv;
                    Gv[i][3] = (-f.tau[i][0] * vn[n][0] - f.tau[i][1] * vn[n][1] - f.k * f.dT[i]) * reinv;
                }
                const double ph = phl[n];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const double fg = Fn[n][i][r] + Gv[i][r];
                        P[i][r] += dt * fg * ph;
                        D[r] += dphi[n][i] * fg;
                    }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double acc = 0.0;
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    double Q = 0.0;
#pragma unroll
                    for (int s = 0; s < 4; s++) Q += (a[i][r][s] + b[i][r][s]) * D[s];
                    acc += DN[i] * (P[i][r] - dtsq_2 * Q);
                }
                rhs[r] += wJ * acc;
            }
        }
        {
            double wg[2], W[2][2];
            const double sdm = DM[0] + DM[1];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                wg[i] = -dt * wJ * PM * DN[i] + dt * wJ * PsiN * sdm;   // g and the nd term (Q10)
#pragma unroll
                for (int j = 0; j < 2; j++) W[i][j] = -dt * s3 * wJ * DN[i] * DM[j];
            }
            contract<2, 4>(R, wJ * PN * PM, wg, DN, DM, dtsq_2 * wJ, a, b, s1, s2, s3, s4);
            accum_c2(th, f, W, R);
        }
    }
    // boundary edges
    const int* bcs = A.bc + e * 4;
    for (int egp = 0; egp < 2; egp++) {
        const double w0 = G.we[egp], z0 = G.ze[egp];
        for (int fc = 0; fc < 4; fc++) {
            const int bc = bcs[fc];
            if (!(bc != 0 && bc < 100)) continue;
            const int n0 = fc, n1 = (fc + 1) & 3;
            double x0[2], x1[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                double a0 = 0, a1 = 0;
#pragma unroll
                for (int n = 0; n < 4; n++) { a0 = (n == n0) ? xn[n][i] : a0; a1 = (n == n1) ? xn[n][i] : a1; }
                x0[i] = a0; x1[i] = a1;
            }
            const double phib[2] = {0.5 * (1 - z0), 0.5 * (1 + z0)};
            const double dxds = -0.5 * x0[0] + 0.5 * x1[0];
            const double dyds = -0.5 * x0[1] + 0.5 * x1[1];
            const double en[2] = {dyds * w0, -dxds * w0};
            const double delx = x1[0] - x0[0], dely = x1[1] - x0[1];
            double dpidx[2][2] = {{0, 0}, {0, 0}};
            if (!(fabs(delx) < 1.0E-12)) { dpidx[0][0] = -1.0 / delx; dpidx[1][0] = 1.0 / delx; }
            if (!(fabs(dely) < 1.0E-12)) { dpidx[0][1] = -1.0 / dely; dpidx[1][1] = 1.0 / dely; }
            const int fln[2] = {n0, n1};
            GPState<2> f;
            f.E = 0; f.T = 0;
#pragma unroll
            for (int i = 0; i < 2; i++) { f.v[i] = 0; f.dT[i] = 0; f.dV[i][0] = f.dV[i][1] = 0; f.tau[i][0] = f.tau[i][1] = 0; }
#pragma unroll
            for (int q = 0; q < 4; q++) { f.U[q] = 0; f.dU[q][0] = f.dU[q][1] = 0; }
            double vb[2][2], Ub[2][4], Fb[2][2][4];
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int r = 0; r < 4; r++) Fb[t][i][r] = 0.0;
#pragma unroll
            for (int t = 0; t < 2; t++) {
                double Et = 0, Tt = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) Ub[t][q] = 0;
                vb[t][0] = vb[t][1] = 0;
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    const bool sel = (n == fln[t]);
                    Et = sel ? En[n] : Et;
                    Tt = sel ? Tn[n] : Tt;
#pragma unroll
                    for (int q = 0; q < 4; q++) Ub[t][q] = sel ? Un[n][q] : Ub[t][q];
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        vb[t][i] = sel ? vn[n][i] : vb[t][i];
#pragma unroll
                        for (int r = 0; r < 4; r++) Fb[t][i][r] = sel ? Fn[n][i][r] : Fb[t][i][r];
                    }
                }
                const double ph = phib[t];
                f.E += ph * Et;
                f.T += ph * Tt;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    f.U[q] += ph * Ub[t][q];
#pragma unroll
                    for (int j = 0; j < 2; j++) f.dU[q][j] += dpidx[t][j] * Ub[t][q];
                }
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    f.v[i] += ph * vb[t][i];
                    f.dT[i] += dpidx[t][i] * Tt;
#pragma unroll
                    for (int j = 0; j < 2; j++) f.dV[i][j] += dpidx[t][j] * vb[t][i];
                }
            }
            f.mu = calc_mu(th, f.T);
            f.k = calc_k(th, f.mu);
            if (!eval1dboundary(bc, f, x0, x1, en)) {
                *status |= 1;
                continue;
            }
            double a[2][4][4], b[2][4][4];
            calc_a2(th, f, a);
            calc_b2(th, f, b);
            int fN = (N == n0) ? 0 : ((N == n1) ? 1 : -1);
            int fM = (M == n0) ? 0 : ((M == n1) ? 1 : -1);
            const double phN = (fN == 0) ? phib[0] : ((fN == 1) ? phib[1] : 0.0);
            const double phM = (fM == 0) ? phib[0] : ((fM == 1) ? phib[1] : 0.0);
            double dpM[2] = {0, 0};
            if (fM >= 0) { dpM[0] = dpidx[fM][0]; dpM[1] = dpidx[fM][1]; }
            if (fN >= 0 && M == 0) {
                double P[2][4], D[4];
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int r = 0; r < 4; r++) P[i][r] = 0.0;
#pragma unroll
                for (int s = 0; s < 4; s++) D[s] = 0.0;
#pragma unroll
                for (int t = 0; t < 2; t++) {
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        double Gv[4];
                        Gv[0] = 0.0;
                        Gv[1] = -f.tau[i][0] * reinv;
                        Gv[2] = -f.tau[i][1] * reinv;
                        Gv[3] = (-f.tau[i][0] * vb[t][0] - f.tau[i][1] * vb[t][1] - f.k * f.dT[i]) * reinv;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const double fg = Fb[t][i][r] + Gv[r];
                            P[i][r] += dt * fg * phib[t];
                            D[r] += dpidx[t][i] * fg;
                        }
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    double acc = 0.0;
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        double Q = 0.0;
#pragma unroll
                        for (int s = 0; s < 4; s++) Q += (a[i][r][s] + b[i][r][s]) * D[s];
                        acc += en[i] * (-P[i][r] + dtsq_2 * Q);
                    }
                    rhs[r] += phN * acc;
                }
            }
            if (fN >= 0 && fM >= 0) {
                double wg[2], W[2][2];
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    wg[i] = dt * phN * phM * en[i];
#pragma unroll
                    for (int j = 0; j < 2; j++) W[i][j] = dt * s3 * phN * en[i] * dpM[j];
                }
                contract<2, 4>(R, 0.0, wg, en, dpM, -dtsq_2 * phN, a, b, s1, s2, s3, s4);
                accum_c2(th, f, W, R);
            }
        }
    }
    const unsigned long long cm = A.cmask[e];
#pragma unroll
    for (int r = 0; r < NEQ; r++) {
        if ((cm >> (NEQ * N + r)) & 1ULL) {
            rhs[r] = 0.0;
#pragma unroll
            for (int s = 0; s < NEQ; s++) R[r][s] = (N == M && r == s) ? 0.5 : 0.0;
        }
    }
}

template <bool W64, bool W32>
__global__ __launch_bounds__(256) void build2d_kernel(BuildArgs A) {
    constexpr int NNOD = 4, NEQ = 4, K = 16, KK = 256;
    const int lane = threadIdx.x & 63;
    const long e = ((long)blockIdx.x * 256 + threadIdx.x) >> 4;
    if (e >= A.n_elem_pad) return;
    const int l16 = lane & 15;
    const int N = l16 >> 2, M = l16 & 3;
    const DevThermo th = A.th;
    const int* cn = A.conn + e * NNOD;
    if (cn[0] < 0) {
        for (int q = l16; q < KK; q += 16) {
            if (W64) A.R64[r64_at<4, 4>(A, e, q)] = 0.0;
            if (W32) A.R32[blk_index<float>(e, q, KK)] = 0.0f;
        }
        A.rhs[e * K + l16] = 0.0;
        return;
    }
    double R[NEQ][NEQ], rhs[NEQ], sp[4];
    int tie = 0, status = 0;
    build2d_elem(A, c_gauss, e, l16, R, rhs, sp, &tie, &status);
    if (status) atomicOr(A.status, status);
#pragma unroll
    for (int r = 0; r < NEQ; r++)
#pragma unroll
        for (int s = 0; s < NEQ; s++) {
            const int q = (NEQ * N + r) * K + NEQ * M + s;
            if (W64) A.R64[r64_at<4, 4>(A, e, q)] = R[r][s];
            if (W32) A.R32[blk_index<float>(e, q, KK)] = (float)R[r][s];
        }
    if (M == 0) {
#pragma unroll
        for (int r = 0; r < NEQ; r++) A.rhs[e * K + NEQ * N + r] = rhs[r];
    }
    if (l16 == 0) {
        if (A.svals)
            for (int i = 0; i < 4; i++) A.svals[e * 4 + i] = sp[i];
        if (tie) atomicAdd(A.ties, 1);
    }
}

void launch_build(const BuildArgs& a, bool w64, bool w32, hipStream_t st) {
    if (a.ndim == 3) {
        // interior elements (and padding) with the shared Gauss-point table,
        // one 64-lane workgroup per slot
        const int nb = (int)a.n_elem_pad;
        if (w64 && w32) hipLaunchKernelGGL((build3d_interior_kernel<true, true>), dim3(nb), dim3(64), 0, st, a);
        else if (w64) hipLaunchKernelGGL((build3d_interior_kernel<true, false>), dim3(nb), dim3(64), 0, st, a);
        else hipLaunchKernelGGL((build3d_interior_kernel<false, true>), dim3(nb), dim3(64), 0, st, a);
        // elements with faces or constrained dofs: faces + ApplyBC on top of
        // the FP64 domain part, or (FP32 blocks only) the full per-lane kernel
        // (with FP64 blocks one instantiation whatever w32, the FP32 copy
        // gated on R32 at run time, so that contexts with and without FP32
        // blocks run the same element arithmetic, bit for bit)
        BuildArgs b = a;
        if (!w32) b.R32 = nullptr;
        if (a.n_bslots > 0) {
            if (w64) {
                // elements with one face (most) on the two-waves-per-SIMD
                // instance, the others on the general one
                if (a.n_bslots1 > 0) {
                    BuildArgs b1 = b;
                    b1.bslots = a.bslots1;
                    b1.n_bslots = a.n_bslots1;
                    hipLaunchKernelGGL((build3d_face_kernel<true, true>), dim3((int)a.n_bslots1), dim3(64), 0, st, b1);
                }
                if (a.n_bslotsm > 0) {
                    BuildArgs bm = b;
                    bm.bslots = a.bslotsm;
                    bm.n_bslots = a.n_bslotsm;
                    hipLaunchKernelGGL((build3d_face_kernel<true, false>), dim3((int)a.n_bslotsm), dim3(64), 0, st, bm);
                }
            } else {
                hipLaunchKernelGGL((build3d_kernel<false, true>), dim3((int)((a.n_bslots + 3) / 4)), dim3(256), 0, st,
                                   a, a.bslots, a.n_bslots);
            }
        }
        // hanging-node elements (Hnm-transformed test functions): whole
        // element on the per-lane kernel
        if (a.n_aslots > 0) {
            const int nba = (int)((a.n_aslots + 3) / 4);
            if (w64) hipLaunchKernelGGL((build3d_kernel<true, true>), dim3(nba), dim3(256), 0, st, b, a.aslots, a.n_aslots);
            else hipLaunchKernelGGL((build3d_kernel<false, true>), dim3(nba), dim3(256), 0, st, a, a.aslots, a.n_aslots);
        }
    } else {
        const int nb = (int)((a.n_elem_pad * 16 + 255) / 256);
        if (w64 && w32) hipLaunchKernelGGL((build2d_kernel<true, true>), dim3(nb), dim3(256), 0, st, a);
        else if (w64) hipLaunchKernelGGL((build2d_kernel<true, false>), dim3(nb), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((build2d_kernel<false, true>), dim3(nb), dim3(256), 0, st, a);
    }
}

void launch_build_direct(const BuildArgs& a, int ncolor, const long* cstart, const long* bstart, const long* astart,
                         hipStream_t st) {
    // per color: the interior kernel over its slots (the face kernel's
    // elements leave their domain part in Rb), the face kernel over its
    // boundary elements, the per-lane kernel over its hanging-node elements
    // with faces; each adds into the staged rows
    for (int c = 0; c < ncolor; c++) {
        BuildArgs b = a;
        b.e_lo = cstart[c];
        b.e_hi = cstart[c + 1];
        const long nb = b.e_hi - b.e_lo;
        if (nb > 0) hipLaunchKernelGGL((build3d_interior_kernel<true, false>), dim3((unsigned)nb), dim3(64), 0, st, b);
        const long nbs = bstart[c + 1] - bstart[c];
        if (nbs > 0) {
            BuildArgs f = a;
            f.bslots = a.bslots + bstart[c];
            f.n_bslots = nbs;
            hipLaunchKernelGGL((build3d_face_kernel<true, false>), dim3((unsigned)nbs), dim3(64), 0, st, f);
        }
        const long nas = astart[c + 1] - astart[c];
        if (nas > 0)
            hipLaunchKernelGGL((build3d_kernel<true, true>), dim3((unsigned)((nas + 3) / 4)), dim3(256), 0, st, a,
                               a.aslots + astart[c], nas);
    }
}

// ---------------------------------------------------------------------------
// Refinement flags: AdaptiveRefineList's s1 > 0.8 test
// (src/adap/RefineListAdaptive.h:28-38), thread per element, no FMA
// contraction so that the flags are bit-identical to the CPU restatement.
// ---------------------------------------------------------------------------
template <int NDIM, int NNOD>
__global__ void flags_kernel(const DevThermo th, const int* __restrict__ conn, const double* __restrict__ U,
                             const int* __restrict__ perm_inv, long n_elem, uint8_t* __restrict__ flags,
                             double* __restrict__ s1_out, int* ties) {
#pragma clang fp contract(off)
    const long eo = (long)blockIdx.x * blockDim.x + threadIdx.x;   // original element id
    if (eo >= n_elem) return;
    const long e = perm_inv[eo];
    const int* cn = conn + e * NNOD;
    double vn[NNOD][NDIM], Tn[NNOD], rhon[NNOD];
    for (int n = 0; n < NNOD; n++) {
        double Ul[NDIM + 2], E;
        const int g = cn[n];
        for (int q = 0; q < NDIM + 2; q++) Ul[q] = U[(size_t)g * (NDIM + 2) + q];
        rhon[n] = Ul[0];
        double vsqr = 0;
        for (int i = 0; i < NDIM; i++) {
            vn[n][i] = Ul[1 + i] / Ul[0];
            vsqr += (vn[n][i] * vn[n][i]);
        }
        E = Ul[NDIM + 1] / Ul[0];
        const double ee = E - 0.5 * vsqr;
        Tn[n] = ee / th.Cv;
    }
    double s[4];
    int tie;
    fdvparam<NDIM, NNOD>(th, vn, Tn, rhon, s, &tie);
    flags[eo] = (s[0] > 0.8) ? 1 : 0;
    if (s1_out) s1_out[eo] = s[0];
    if (tie) atomicAdd(ties, 1);
}

void launch_flags(int ndim, const DevThermo& th, const int* conn, const double* U, const int* perm_inv,
                  long n_elem, uint8_t* flags, double* s1, int* ties, hipStream_t st) {
    const int nb = (int)((n_elem + 255) / 256);
    if (ndim == 3)
        hipLaunchKernelGGL((flags_kernel<3, 8>), dim3(nb), dim3(256), 0, st, th, conn, U, perm_inv, n_elem, flags, s1, ties);
    else
        hipLaunchKernelGGL((flags_kernel<2, 4>), dim3(nb), dim3(256), 0, st, th, conn, U, perm_inv, n_elem, flags, s1, ties);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free flag pass (cfdv_refine_flags_noalloc): a
// file-scope device array that every GPU the library runs on holds until the
// process ends.  In host code `scratch` only names the symbol; kernels get
// the device address that flags_scratch() resolves with hipGetSymbolAddress
// once per device, with that device current.  Calls on one GPU share the
// array: overlapping calls race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FLAGS_SCRATCH_BYTES];

void* flags_scratch() {
    static std::mutex mu;
    static std::vector<void*> addr;
    int dev = 0;
    HIP_CHECK_THROW(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
    if (!addr[dev]) {
        void* p = nullptr;
        HIP_CHECK_THROW(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
        addr[dev] = p;
    }
    return addr[dev];
}

__global__ void zero_ints_kernel(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

void launch_flags_scratch(int ndim, const DevThermo& th, const int* conn, const double* U, const int* perm_inv,
                          long n_elem, void* scratch_dev, bool want_s1, hipStream_t st, uint8_t** flags_dev,
                          double** s1_dev, int** ties_dev) {
    if (n_elem < 0 || n_elem > FLAGS_NOALLOC_MAX_ELEM) throw std::invalid_argument("flag scratch too small");
    unsigned char* base = static_cast<unsigned char*>(scratch_dev);
    int* counters = reinterpret_cast<int*>(base);
    uint8_t* f = base + 256;
    double* s1 = want_s1 ? reinterpret_cast<double*>(base + 256 + ((n_elem + 255) / 256) * 256) : nullptr;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, st, counters, 4);
    launch_flags(ndim, th, conn, U, perm_inv, n_elem, f, s1, counters, st);
    *flags_dev = f;
    *s1_dev = s1;
    *ties_dev = counters;
}

// ---------------------------------------------------------------------------
// CFL (src/FDV/Function/MPI_CFL.h:9-56): per-element dxele, then a
// block min; min is order-independent, so the result is exact.
// ---------------------------------------------------------------------------
template <int NDIM, int NNOD>
__global__ void cfl_kernel(const DevThermo th, const int* __restrict__ conn, const double* __restrict__ U,
                           const double* __restrict__ cleng, long n_elem_pad, double* __restrict__ part) {
    __shared__ double sm[256];
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dxele = 1000;
    if (e < n_elem_pad && conn[e * NNOD] >= 0) {
        for (int n = 0; n < NNOD; n++) {
            const int g = conn[e * NNOD + n];
            double v[NDIM], rho, E, T;
            double Ul[NDIM + 2];
            for (int q = 0; q < NDIM + 2; q++) Ul[q] = U[(size_t)g * (NDIM + 2) + q];
            unpack<NDIM>(th, Ul, rho, v, E, T);
            const double Temp = (T > 10E-10) ? T : 0.0;
            double vv = 0;
            for (int i = 0; i < NDIM; i++) vv += v[i] * v[i];
            const double vm = sqrt(vv);
            const double a = sqrt(1.4 * th.cgas * Temp);
            if (vm + a > 10E-10) {
                const double dxnode = cleng[e] / (vm + a);
                if (dxnode < dxele) dxele = dxnode;
            }
        }
    }
    sm[threadIdx.x] = dxele;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

void launch_cfl(int ndim, const DevThermo& th, const int* conn, const double* U, const double* cleng,
                long n_elem_pad, double* part, int nblocks, hipStream_t st) {
    if (ndim == 3)
        hipLaunchKernelGGL((cfl_kernel<3, 8>), dim3(nblocks), dim3(256), 0, st, th, conn, U, cleng, n_elem_pad, part);
    else
        hipLaunchKernelGGL((cfl_kernel<2, 4>), dim3(nblocks), dim3(256), 0, st, th, conn, U, cleng, n_elem_pad, part);
}

// CalcLength (src/FDV/Function/CalcLength.h:34-110): area by 2-point Gauss,
// cleng = area^(1/ndim)
template <int NDIM, int NNOD>
__global__ void length_kernel(const int* __restrict__ conn, const double* __restrict__ x, long n_elem_pad,
                              double* __restrict__ area, double* __restrict__ cleng) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_elem_pad) return;
    if (conn[e * NNOD] < 0) { area[e] = 0; cleng[e] = 0; return; }
    double xn[NNOD][NDIM];
    for (int n = 0; n < NNOD; n++)
        for (int i = 0; i < NDIM; i++) xn[n][i] = x[(size_t)conn[e * NNOD + n] * NDIM + i];
    double ar = 0;
    if (NDIM == 3) {
        for (int gp = 0; gp < 8; gp++) {
            double j[3][3];
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) {
                    double acc = 0;
                    for (int n = 0; n < 8; n++) acc += c_gauss.dpxi3[gp][n][k] * xn[n][i];
                    j[i][k] = acc;
                }
            const double J = j[0][0] * j[1][1] * j[2][2] + j[0][1] * j[1][2] * j[2][0] + j[0][2] * j[1][0] * j[2][1] -
                             j[0][2] * j[1][1] * j[2][0] - j[0][1] * j[1][0] * j[2][2] - j[0][0] * j[1][2] * j[2][1];
            ar += c_gauss.w3[gp] * J;
        }
    } else {
        for (int gp = 0; gp < 4; gp++) {
            double j[2][2];
            for (int i = 0; i < 2; i++)
                for (int k = 0; k < 2; k++) {
                    double acc = 0;
                    for (int n = 0; n < 4; n++) acc += c_gauss.dpxi2[gp][n][k] * xn[n][i];
                    j[i][k] = acc;
                }
            ar += c_gauss.w2[gp] * (j[0][0] * j[1][1] - j[1][0] * j[0][1]);
        }
    }
    area[e] = ar;
    cleng[e] = pow(ar, 1.0 / (double)NDIM);
}

void launch_length(int ndim, const int* conn, const double* x, long n_elem_pad, double* area, double* cleng,
                   hipStream_t st) {
    const int nb = (int)((n_elem_pad + 255) / 256);
    if (ndim == 3) hipLaunchKernelGGL((length_kernel<3, 8>), dim3(nb), dim3(256), 0, st, conn, x, n_elem_pad, area, cleng);
    else hipLaunchKernelGGL((length_kernel<2, 4>), dim3(nb), dim3(256), 0, st, conn, x, n_elem_pad, area, cleng);
}

}  // namespace cfdv
