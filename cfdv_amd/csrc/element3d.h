// Per-lane hex8 element physics of the FDV build, templated on the scalar
// type (double for the build kernels, Dual<N> for the adjoint tangents):
// FDVParam, NodeUnpack, the face-state edits of eval2dboundary, the
// factorised Gauss-point contraction, the face integrals, ApplyBC and the
// whole-element build seen from one lane of a 64-lane wavefront.
#pragma once
#include "cfdv_internal.h"
#include "fdv_device.h"

namespace cfdv {

// ---------------------------------------------------------------------------
// FDVParam (src/FDV/Function/FDVParam.h:15-111).  The r == 0.01 tie
// (quirk Q11) leaves s1/s3 as the previous element's value in the
// reference; here the tie is reported through *tie and s keeps 0.
// ---------------------------------------------------------------------------
template <int NDIM, int NNOD, typename TH, typename S>
CFDV_HD void fdvparam(const TH& th, const S vn[][NDIM], const S* Tn, const S* rhon, S s[4], int* tie) {
#pragma clang fp contract(off)
    S M_min = 0, M_max = 0, Re_min = 0, Re_max = 0;
#pragma unroll
    for (int beta = 0; beta < NNOD; beta++) {
        S dd = 0;
#pragma unroll
        for (int i = 0; i < NDIM; i++) dd += vn[beta][i] * vn[beta][i];
        const S V = sqrt(dd);
        S T = Tn[beta];
        if (T < 1.0E-12) T = 0;
        const S a = sqrt(th.gamma * th.cgas * T);
        S M = 0;
        if (!(a < 1.0E-12)) M = V / a;
        const S mu = calc_mu(th, T);
        const S Re = (rhon[beta] * V) / mu;
        if (beta == 0) { M_max = M; M_min = M; Re_max = Re; Re_min = Re; }
        M_max = (M_max < M) ? M : M_max;
        M_min = (M < M_min) ? M : M_min;
        Re_max = (Re_max < Re) ? Re : Re_max;
        Re_min = (Re < Re_min) ? Re : Re_min;
    }
    const double eps = 0.0000001;
    S r = sqrt(M_max * M_max - M_min * M_min) / (M_min + eps);
    *tie = 0;
    s[0] = 0.0;
    if (r > 0.01) s[0] = (1.0 < r) ? S(1.0) : r;
    else if (r < 0.01 && M_min >= 10E-10) s[0] = 0;
    else if (M_min < 10E-10) s[0] = 1.0;
    else *tie = 1;
    s[1] = 0.5 * (1 + pow(s[0], 0.2));
    r = sqrt(Re_max * Re_max - Re_min * Re_min) / (Re_min + eps);
    s[2] = 0.0;
    if (r > 0.01) s[2] = (1.0 < r) ? S(1.0) : r;
    else if (r < 0.01 && Re_min >= 10E-10) s[2] = 0.0;
    else if (Re_min <= 10E-10) s[2] = 1.0;
    else *tie = 1;
    s[3] = 0.5 * (1 + pow(s[2], 0.2));
}

// NodeUnpack (src/FDV/Function/NodeUnpack.h:9-29) for one node
template <int NDIM, typename TH, typename S>
CFDV_HD void unpack(const TH& th, const S* U, S& rho, S v[NDIM], S& E, S& T) {
    rho = U[0];
    S vsqr = 0;
#pragma unroll
    for (int i = 0; i < NDIM; i++) {
        v[i] = U[1 + i] / U[0];
        vsqr += v[i] * v[i];
    }
    E = U[NDIM + 1] / U[0];
    const S e = E - 0.5 * vsqr;
    T = e / th.Cv;
}

// ---------------------------------------------------------------------------
// eval2dboundary (src/FDV/Function/eval2dboundary.h:14-371): bc edits of
// the face Gauss-point state.  Returns false for an unknown code.
// ---------------------------------------------------------------------------
template <typename TH, typename S>
CFDV_HD bool eval2dboundary(const TH& th, int bc, GPState<3, S>& f, const double en[3]) {
    switch (bc) {
    case 2: case 12:
        for (int i = 0; i < 3; i++) f.dV[i][2] = 0;
        f.dT[2] = 0;
        return true;                                  // tau stays 0
    case 22:
        for (int i = 0; i < 3; i++) f.dV[i][1] = 0;
        f.dT[1] = 0;
        calc_tau<3>(f);
        f.tau[0][1] = 0.0; f.tau[1][1] = 0.0; f.tau[2][1] = 0.0;
        f.tau[1][0] = 0.0; f.tau[1][2] = 0.0;
        return true;
    case 1: case 51:
        if (th.adiabatic) f.dT[2] = 0.0;
        calc_tau<3>(f);
        return true;
    case 21: case 31:
        if (th.adiabatic) f.dT[1] = 0.0;
        calc_tau<3>(f);
        return true;
    case 11:
        if (th.adiabatic) f.dT[0] = 0.0;
        calc_tau<3>(f);
        return true;
    case -1:
        return true;                                  // returns before tau (Q6)
    case 6:
        f.dT[0] = 0.0;
        calc_tau<3>(f);
        for (int i = 0; i < 3; i++) f.tau[i][0] = 0;
        return true;
    case 7:
        f.dT[2] = 0.0;
        calc_tau<3>(f);
        for (int i = 0; i < 3; i++) f.tau[i][2] = 0;
        // no break in the source: continues into 77/32 (Q6)
        f.dT[1] = 0.0;
        calc_tau<3>(f);
        for (int i = 0; i < 3; i++) f.tau[i][1] = 0;
        return true;
    case 77: case 32:
        f.dT[1] = 0.0;
        calc_tau<3>(f);
        for (int i = 0; i < 3; i++) f.tau[i][1] = 0;
        return true;
    case 4: {
        // projection with the unnormalised en; tau not recomputed; only the
        // first three equations' gradients are projected (loop bound jmax)
        f.v[0] = f.v[1] = f.v[2] = 0;
        const S dudot = f.dV[0][0] * en[0] + f.dV[0][1] * en[1] + f.dV[0][2] * en[2];
        const S dvdot = f.dV[1][0] * en[0] + f.dV[1][1] * en[1] + f.dV[1][2] * en[2];
        const S dtdot = f.dT[0] * en[0] + f.dT[1] * en[1] + f.dT[2] * en[2];
        for (int j = 0; j < 3; j++) {
            f.dV[0][j] -= dudot * en[j];
            f.dV[1][j] -= dvdot * en[j];
            f.dT[j] -= dtdot * en[j];
        }
        S dUdot[3];
        for (int i = 0; i < 3; i++) dUdot[i] = f.dU[i][0] * en[0] + f.dU[i][1] * en[1] + f.dU[i][2] * en[2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) f.dU[i][j] -= dUdot[i] * en[j];
        return true;
    }
    case 8:
        calc_tau<3>(f);
        return true;
    case -12:
        f.dV[0][2] = 0.0; f.dV[1][2] = 0.0; f.dV[2][2] = 0.0;
        f.dT[2] = 0.0;
        f.v[2] = 0.0;
        return true;
    default:
        return false;
    }
}

// ---------------------------------------------------------------------------
// hex8 build kernel
// ---------------------------------------------------------------------------
// per-wave node data in LDS: x(3), U(5), v(3), E, T, rho
#define ND_X 0
#define ND_U 3
#define ND_V 8
#define ND_E 11
#define ND_T 12
#define ND_RHO 13
#define ND_STRIDE 14

// node row from coordinates x[3] and state U[5] (of scalar type S)
template <typename TH, typename S>
CFDV_HD void load_node3d_t(const TH& th, const double* x, const S* Ul, S* row) {
    for (int i = 0; i < 3; i++) row[ND_X + i] = x[i];
    for (int q = 0; q < 5; q++) row[ND_U + q] = Ul[q];
    S rho, v[3], E, T;
    unpack<3>(th, Ul, rho, v, E, T);
    for (int i = 0; i < 3; i++) row[ND_V + i] = v[i];
    row[ND_E] = E;
    row[ND_T] = T;
    row[ND_RHO] = rho;
}

CFDV_HD void load_node3d(const BuildArgs& A, int g, double* row) {
    double Ul[5];
    for (int q = 0; q < 5; q++) Ul[q] = A.U[(size_t)g * 5 + q];
    load_node3d_t(A.th, A.x + (size_t)g * 3, Ul, row);
}

// Accumulates, for one Gauss point of the domain or of a face, the
// direction-streamed factorised contraction into R and the rhs pieces Q.
// wa/ws/wg: node-pair weights; A_i = a_i + b_i is also returned through
// AQ[r] += wr_i * sum_s A_i(r,s) D(s) for the rhs.
template <typename TH, typename S>
CFDV_HD void contract3(const TH& th, const GPState<3, S>& f, S R[5][5], double cI, const double wg[3],
                       const double wa[3], const double ws[3], double cA, const double wr[3], const S D[5],
                       S AQ[5], S s1, S s2, S s3, S s4) {
    S Ah[5][5], Sh[5][5];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        AQ[r] = 0.0;
#pragma unroll
        for (int s = 0; s < 5; s++) {
            Ah[r][s] = 0.0;
            Sh[r][s] = 0.0;
        }
        R[r][r] += cI;
    }
    S a[5][5], b[5][5];
#define DIR(I)                                                                      \
    {                                                                               \
        calc_a3_dir<I>(th, f, a);                                                   \
        calc_b3_dir<I>(th, f, b);                                                   \
        _Pragma("unroll") for (int r = 0; r < 5; r++) {                             \
            S q = 0.0;                                                              \
            _Pragma("unroll") for (int s = 0; s < 5; s++) {                         \
                const S A_ = a[r][s] + b[r][s];                                     \
                R[r][s] += wg[I] * (s1 * a[r][s] + s3 * b[r][s]);                   \
                Ah[r][s] += wa[I] * A_;                                             \
                Sh[r][s] += ws[I] * (s2 * a[r][s] + s4 * b[r][s]);                  \
                q += A_ * D[s];                                                     \
            }                                                                       \
            AQ[r] += wr[I] * q;                                                     \
        }                                                                           \
    }
    DIR(0)
    DIR(1)
    DIR(2)
#undef DIR
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int s = 0; s < 5; s++) {
            S p = 0.0;
#pragma unroll
            for (int t = 0; t < 5; t++) p += Ah[r][t] * Sh[t][s];
            R[r][s] += cA * p;
        }
}

// Face integrals (FDVGalerkin.h:455-533, 643-756) of every face with
// 0 < bc < 100, added for lane (N, M) to R and (for M == 0) rhs.
template <typename TH, typename S>
CFDV_HD void build3d_faces(const BuildArgs& A, const TH& th, const GaussTables& G, const S (*nd)[ND_STRIDE],
                           long e, int lane, const S sp[4], S R[5][5], S rhs[5], int* status) {
    constexpr int NEQ = 5;
    const int N = lane >> 3, M = lane & 7;
    const S s1 = sp[0], s2 = sp[1], s3 = sp[2], s4 = sp[3];
    const double dt = A.dt, dtsq_2 = (dt * dt) / 2.0;
    const double reinv = 1.0 / th.creyn;
    (void)s1; (void)s2; (void)s4;
    const int* bcs = A.bc + e * 6;
    const double* Hn = (A.adap && A.adap[e]) ? A.hnm + e * 64 : nullptr;
    int anyface = 0;
#pragma unroll
    for (int fc = 0; fc < 6; fc++) anyface |= (bcs[fc] != 0 && bcs[fc] < 100);
    if (anyface) {
        const int FN[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {6, 7, 4, 5}};
        for (int fgp = 0; fgp < 4; fgp++) {
            const double w0 = G.wf[fgp][0], w1 = G.wf[fgp][1];
            const double z0 = G.zf[fgp][0], z1 = G.zf[fgp][1];
            for (int fc = 0; fc < 6; fc++) {
                const int bc = bcs[fc];
                if (!(bc != 0 && bc < 100)) continue;
                int fl[4];
                double nl[4][3];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    fl[t] = FN[fc][t];
#pragma unroll
                    for (int i = 0; i < 3; i++) nl[t][i] = value_of(nd[fl[t]][ND_X + i]);
                }
                if (Hn) {
                    // face corners from the Hnm-mapped element nodes
                    // (calc_TF_norm3D.h:35-66)
                    for (int t = 0; t < 4; t++)
#pragma unroll
                        for (int i = 0; i < 3; i++) {
                            double acc = 0.0;
                            for (int m = 0; m < 8; m++) acc += Hn[fl[t] + 8 * m] * value_of(nd[m][ND_X + i]);
                            nl[t][i] = acc;
                        }
                }
                // TFNorm3D (calc_TF_norm3D.h:9-167), including the Lambda/Mu
                // magnitude recomputed after each component (:208-212)
                double phib[4];
                phib[0] = 0.25 * (1 - z0) * (1 - z1);
                phib[1] = 0.25 * (1 + z0) * (1 - z1);
                phib[2] = 0.25 * (1 + z0) * (1 + z1);
                phib[3] = 0.25 * (1 - z0) * (1 + z1);
                if (Hn) {
                    // adap_face (MeshRefine2D.h:903-952): only phibou is
                    // carried over, dpidx stays untransformed
                    double pt[4] = {0.0, 0.0, 0.0, 0.0};
                    for (int n = 0; n < 4; n++)
                        for (int m = 0; m < 4; m++) pt[m] += phib[n] * Hn[fl[n] + 8 * fl[m]];
#pragma unroll
                    for (int t = 0; t < 4; t++) phib[t] = pt[t];
                }
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
                double en[3];
#pragma unroll
                for (int i = 0; i < 3; i++) en[i] = en_[i] * w0 * w1 * detjac;

                // face Gauss-point state; raw (unedited) sums kept for G
                GPState<3, S> f;
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
                int fN = -1, fM = -1;
                double phN = 0, phM = 0, dpM[3] = {0, 0, 0};
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int n = fl[t];
                    double dpx[3];
                    const double d0 = (dtr[1][1] * dp[t][0] - dtr[1][0] * dp[t][1]) * djinv;
                    const double d1 = (-dtr[0][1] * dp[t][0] + dtr[0][0] * dp[t][1]) * djinv;
#pragma unroll
                    for (int i = 0; i < 3; i++) dpx[i] = Lm[i] * d0 + ga[i] * d1;
                    const double ph = phib[t];
                    if (n == N) { fN = t; phN = ph; }
                    if (n == M) { fM = t; phM = ph; dpM[0] = dpx[0]; dpM[1] = dpx[1]; dpM[2] = dpx[2]; }
                    Sph += ph;
                    f.E += ph * nd[n][ND_E];
                    const S Tt = nd[n][ND_T];
                    f.T += ph * Tt;
#pragma unroll
                    for (int q = 0; q < NEQ; q++) {
                        const S u_ = nd[n][ND_U + q];
                        f.U[q] += ph * u_;
#pragma unroll
                        for (int j = 0; j < 3; j++) f.dU[q][j] += dpx[j] * u_;
                    }
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        const S v_ = nd[n][ND_V + i];
                        Sdp[i] += dpx[i];
                        f.v[i] += ph * v_;
                        f.dT[i] += dpx[i] * Tt;
#pragma unroll
                        for (int j = 0; j < 3; j++) f.dV[i][j] += dpx[j] * v_;
                    }
                }
                S vph[3], dVp[3][3];
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
                    continue;
                }
                if (fN < 0) continue;   // this lane's row is not on the face
                S D[5];
                D[0] = 0.0;
#pragma unroll
                for (int s = 1; s < 4; s++) {
                    S acc = 0.0;
#pragma unroll
                    for (int j = 0; j < 3; j++) acc += (-f.tau[j][s - 1] * reinv) * Sdp[j];
                    D[s] = acc;
                }
                {
                    S acc = 0.0;
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        S td = 0.0;
#pragma unroll
                        for (int k = 0; k < 3; k++) td += f.tau[j][k] * dVp[k][j];
                        acc += (-td - f.k * f.dT[j] * Sdp[j]) * reinv;
                    }
                    D[4] = acc;
                }
                const double onM = (fM >= 0) ? 1.0 : 0.0;
                double wg[3], wr[3], wa[3];
                S W[3][3];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    wg[i] = dt * phN * phM * en[i] * onM;
                    wr[i] = phN * en[i];
                    wa[i] = en[i] * onM;
#pragma unroll
                    for (int j = 0; j < 3; j++) W[i][j] = dt * s3 * phN * en[i] * dpM[j] * onM;
                }
                S AQ[5];
                contract3(th, f, R, 0.0, wg, wa, dpM, -dtsq_2 * phN, wr, D, AQ, s1, s2, s3, s4);
                accum_c3(th, f, W, R);
                if (M == 0) {
#pragma unroll
                    for (int r = 0; r < NEQ; r++) {
                        S acc = 0.0;
                        if (r > 0) {
#pragma unroll
                            for (int i = 0; i < 3; i++) {
                                S P;
                                if (r < 4) P = dt * (-f.tau[i][r - 1] * reinv) * Sph;
                                else {
                                    S tv = 0.0;
#pragma unroll
                                    for (int k = 0; k < 3; k++) tv += f.tau[i][k] * vph[k];
                                    P = dt * ((-tv - f.k * f.dT[i] * Sph) * reinv);
                                }
                                acc += wr[i] * P;
                            }
                        }
                        rhs[r] += -acc + dtsq_2 * AQ[r];
                    }
                }
            }
        }
    }

}

// ApplyBC (ApplyBC.h:14-133): constrained rows of lane (N, M) from the
// element's dof mask; row -> 0.5 on the diagonal, rhs -> 0.
template <typename S>
CFDV_HD void apply_bc3(const BuildArgs& A, long e, int lane, S R[5][5], S rhs[5]) {
    constexpr int NEQ = 5;
    const int N = lane >> 3, M = lane & 7;
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

// One hex8 element seen from lane (N, M) = (lane/8, lane%8): FDVParam,
// the 8-point domain integral, the face integrals and ApplyBC for the 5x5
// sub-block R(5N+r, 5M+s) and (for M == 0) rhs(5N+r).  nd holds the element's
// node data (x, U, v, E, T, rho per node).
template <typename TH, typename S>
CFDV_HD void build3d_elem(const BuildArgs& A, const TH& th, const GaussTables& G, const S (*nd)[ND_STRIDE],
                          long e, int lane, S R[5][5], S rhs[5], S sp[4], int* tie, int* status) {
    constexpr int NNOD = 8, NEQ = 5;
    const int N = lane >> 3, M = lane & 7;
    {
        S vn[NNOD][3], Tn[NNOD], rhon[NNOD];
#pragma unroll
        for (int n = 0; n < NNOD; n++) {
#pragma unroll
            for (int i = 0; i < 3; i++) vn[n][i] = nd[n][ND_V + i];
            Tn[n] = nd[n][ND_T];
            rhon[n] = nd[n][ND_RHO];
        }
        fdvparam<3, 8>(th, vn, Tn, rhon, sp, tie);
    }
    const S s1 = sp[0], s2 = sp[1], s3 = sp[2], s4 = sp[3];
    const double dt = A.dt, dtsq_2 = (dt * dt) / 2.0;
    const double reinv = 1.0 / th.creyn;

#pragma unroll
    for (int r = 0; r < NEQ; r++) {
        rhs[r] = 0.0;
#pragma unroll
        for (int s = 0; s < NEQ; s++) R[r][s] = 0.0;
    }

    // hanging-node element: test functions Phi Hnm (adap(), MeshRefine2D.h:877-900)
    const double* Hn = (A.adap && A.adap[e]) ? A.hnm + e * 64 : nullptr;

    // ---------------- domain integral: 8 Gauss points ----------------
    for (int gp = 0; gp < 8; gp++) {
        double phg[8], dpg[8][3];
        if (Hn) {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                phg[m] = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) dpg[m][j] = 0.0;
            }
            for (int n = 0; n < 8; n++)
                for (int m = 0; m < 8; m++) {
                    const double h = Hn[n + 8 * m];
                    phg[m] += G.phi3[gp][n] * h;
#pragma unroll
                    for (int j = 0; j < 3; j++) dpg[m][j] += G.dpxi3[gp][n][j] * h;
                }
        } else {
#pragma unroll
            for (int m = 0; m < 8; m++) {
                phg[m] = G.phi3[gp][m];
#pragma unroll
                for (int j = 0; j < 3; j++) dpg[m][j] = G.dpxi3[gp][m][j];
            }
        }
        double jac[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < 8; n++) acc += dpg[n][j] * value_of(nd[n][ND_X + i]);
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
        // Gauss-point state (FDVGalerkin::sum, :761-797) and my node pair
        GPState<3, S> f;
        f.E = 0; f.T = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            f.v[i] = 0; f.dT[i] = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) f.dV[i][j] = 0;
        }
#pragma unroll
        for (int q = 0; q < NEQ; q++) {
            f.U[q] = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) f.dU[q][j] = 0;
        }
        double SPhi = 0.0, SD[3] = {0, 0, 0};
        double PN = 0, PM = 0, DN[3] = {0, 0, 0}, DM[3] = {0, 0, 0};
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double dphi[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < 3; j++) acc += Jinv * (ji[i][j] * dpg[n][j]);
                dphi[i] = acc;
            }
            const double ph = phg[n];
            if (n == N) { PN = ph; DN[0] = dphi[0]; DN[1] = dphi[1]; DN[2] = dphi[2]; }
            if (n == M) { PM = ph; DM[0] = dphi[0]; DM[1] = dphi[1]; DM[2] = dphi[2]; }
            SPhi += ph;
            f.E += ph * nd[n][ND_E];
            const S Tnn = nd[n][ND_T];
            f.T += ph * Tnn;
#pragma unroll
            for (int q = 0; q < NEQ; q++) {
                const S u_ = nd[n][ND_U + q];
                f.U[q] += ph * u_;
#pragma unroll
                for (int j = 0; j < 3; j++) f.dU[q][j] += dphi[j] * u_;
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const S v_ = nd[n][ND_V + i];
                SD[i] += dphi[i];
                f.v[i] += ph * v_;
                f.dT[i] += dphi[i] * Tnn;
#pragma unroll
                for (int j = 0; j < 3; j++) f.dV[i][j] += dphi[j] * v_;
            }
        }
        f.mu = calc_mu(th, f.T);
        f.k = calc_k(th, f.mu);
        calc_tau<3>(f);
        const double wJ = G.w3[gp] * J;

        // rhs pieces: P(i,r) = dt sum_b Phi_b G_b(i,r) and D(s) = sum_jb dPhi_bj
        // G_b(j,s), G_b from the Gauss-point tau/dT/k with node velocities
        // (sumFGB :801-811, calc_G NavierStokes3D.h:41-64; F = 0, quirk Q1)
        S D[5];
        D[0] = 0.0;
#pragma unroll
        for (int s = 1; s < 4; s++) {
            S acc = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) acc += (-f.tau[j][s - 1] * reinv) * SD[j];
            D[s] = acc;
        }
        {
            S acc = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                S td = 0.0;
#pragma unroll
                for (int k = 0; k < 3; k++) td += f.tau[j][k] * f.dV[k][j];
                acc += (-td - f.k * f.dT[j] * SD[j]) * reinv;
            }
            D[4] = acc;
        }
        double wg[3], wr[3];
        S W[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            wg[i] = -dt * wJ * PM * DN[i];
            wr[i] = wJ * DN[i];
#pragma unroll
            for (int j = 0; j < 3; j++) W[i][j] = -dt * s3 * wJ * DN[i] * DM[j];
        }
        S AQ[5];
        contract3(th, f, R, wJ * PN * PM, wg, DN, DM, dtsq_2 * wJ, wr, D, AQ, s1, s2, s3, s4);
        accum_c3(th, f, W, R);
#pragma unroll
        for (int r = 0; r < NEQ; r++) {
            S acc = 0.0;
            if (r > 0) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    S P;
                    if (r < 4) P = dt * (-f.tau[i][r - 1] * reinv) * SPhi;
                    else {
                        S tv = 0.0;
#pragma unroll
                        for (int k = 0; k < 3; k++) tv += f.tau[i][k] * f.v[k];
                        P = dt * ((-tv - f.k * f.dT[i] * SPhi) * reinv);
                    }
                    acc += wr[i] * P;
                }
            }
            rhs[r] += acc - dtsq_2 * AQ[r];
        }
    }

    build3d_faces(A, th, G, nd, e, lane, sp, R, rhs, status);
    apply_bc3(A, e, lane, R, rhs);
}

}  // namespace cfdv
