// Device-side physics of the FDV element build (gfx950, FP64).
//
// Each function restates a piece of the reference element build; the
// citations name the reference lines.  The element matrix is assembled in
// the factorised form
//
//   R_NM += cI*I + sum_i wg_i (s1 a_i + s3 b_i) + sum_ij wc_ij c_ij
//           + cA (sum_i wa_i A_i)(sum_j ws_j S_j),
//   A_i = a_i + b_i,  S_j = s2 a_j + s4 b_j,
//
// which is the reference's sumDomainMatrix/sumBoundaryMatrix (FDVGalerkin.h
// :405-451, :718-756) with its g/h/k/m intermediates (:536-716) expanded
// using d = 0, B = 0 and s5 = s6 = 0 (FDVGalerkin.h:128-129,
// NavierStokes3D.h:67-70,480-483): h_ij = -dt s3 c_ij + dt^2/2 A_i S_j,
// g_i = -dt (s1 a_i + s3 b_i), k = -g, m = -h, eta = I.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dual.h"

// Element-level physics is callable from host code too, so that the exact
// device arithmetic can be exercised by a host-side debugging harness.
#define CFDV_HD __host__ __device__ __forceinline__

namespace cfdv {

struct DevThermo {
    double gamma, gamm1, gm1d2, gamm3, gm3d2;
    double Pr, csuth, Cv, cgas, cmach, creyn, Twall;
    int adiabatic;
};

// Gauss-point tables, filled from the host (Newton-iterated 2-point rule).
struct GaussTables {
    // 3D domain: 8 points, 3D faces: 4 points
    double w3[8];               // w0*w1*w2
    double phi3[8][8];          // [gp][N]
    double dpxi3[8][8][3];      // [gp][N][j]
    double wf[4][2];            // face point weights (w0, w1)
    double zf[4][2];            // face point coordinates
    // 2D domain: 4 points, 2D edges: 2 points
    double w2[4];
    double phi2[4][4];
    double dpxi2[4][4][2];
    double z2[4][2];
    double we[2];
    double ze[2];
};

// Element blocks in HBM: chunks of 64 elements; inside a chunk the k*k
// entries are grouped V = 16 bytes / sizeof(T) at a time (V = 2 doubles or
// 4 floats), group-major, so that a wavefront of 64 threads (one element
// each) loads one 16-byte vector per lane from 1 KB of contiguous memory.
template <typename T>
CFDV_HD size_t blk_offset(long e, int q, int KK) {
    constexpr int V = 16 / (int)sizeof(T);
    const int r = q % V;
    return (size_t)(e >> 6) * 64 * KK + (size_t)(q - r) * 64 + (size_t)(e & 63) * V + r;
}

// FP64 element blocks of contexts with assembled node rows: chunks of 64
// element slots, node-pair major inside a chunk.  Pair (N, M) of slot e is
// one NEQ x NEQ block, row-major (R(i, j), i = N NEQ + r, j = M NEQ + s),
// and the 64 slots of a chunk hold pair (N, M) in 64 consecutive blocks: the
// row assembly reads a node-pair block as one contiguous run, and the same
// pair of consecutive slots of a color as one longer run.
template <int NNOD, int NEQ>
CFDV_HD size_t pair_base(long e, int pair) {
    constexpr int BB = NEQ * NEQ, KK = NNOD * NNOD * BB;
    return (size_t)(e >> 6) * 64 * KK + (size_t)pair * 64 * BB + (size_t)(e & 63) * BB;
}
template <int NNOD, int NEQ>
CFDV_HD size_t pair_offset(long e, int q) {
    constexpr int K = NNOD * NEQ;
    const int i = q / K, j = q % K;
    return pair_base<NNOD, NEQ>(e, (i / NEQ) * NNOD + j / NEQ) + (size_t)((i % NEQ) * NEQ + j % NEQ);
}

template <int NDIM> struct Dims;
template <> struct Dims<3> { static constexpr int NNOD = 8, NEQ = 5, NB = 4, NFACE = 6, NGP = 8, NFGP = 4; };
template <> struct Dims<2> { static constexpr int NNOD = 4, NEQ = 4, NB = 2, NFACE = 4, NGP = 4, NFGP = 2; };

// Thermo with the Mach-dependent constants of scalar type S (double, or a
// dual number for tangents with respect to the Mach number)
template <typename S>
struct ThermoT {
    double gamma, gamm1, gm1d2, gamm3, gm3d2;
    double Pr, csuth;
    S Cv, cgas, cmach;
    double creyn, Twall;
    int adiabatic;
};

template <typename TH, typename S>
CFDV_HD S calc_mu(const TH& th, const S& T) {
    // Thermo::calc_mu, Thermo.h:53-70
    if (T < 1.0E-12) return S(1.0);
    return ((1 + th.csuth) / (T + th.csuth)) * T * sqrt(T);
}
template <typename TH, typename S>
CFDV_HD S calc_k(const TH& th, const S& mu) {
    // Thermo::calc_k, Thermo.h:72-75
    return mu / ((th.gamma - 1.0) * (th.cmach * th.cmach) * th.Pr);
}

// coth(x) = cosh(x)/sinh(x) as the reference's host libm (glibc 2.35)
// evaluates it, for the quad4 numerical diffusion: alpha = coth(R/2) - 2/R
// (FDVGalerkin.h:46-50, 308-316) cancels as R -> 0, so the reference's
// blocks carry the last bits of its cosh and sinh.  The method is fdlibm's
// (Sun Microsystems, 1993): expm1 by reduction to |r| <= ln2/2 with
// k = round(x/ln2) and a rational approximation in r^2 whose polynomial is
// evaluated in glibc's grouping R1 + h2 R2 + h4 R3; then
// sinh(x) = sign(x)/2 (2t - t^2/(t+1)) for |x| < 1, sign(x)/2 (t + t/(t+1))
// for |x| < 22, and cosh(x) = 1 + t^2/(2(1+t)) for |x| < ln2/2, with
// t = expm1(|x|).  Measured bit-identical to glibc 2.35's expm1, sinh and
// cosh on 25 million random arguments in those ranges.  Elsewhere the device
// libm is used: there |R| > ln 2 and the subtraction no longer cancels.
CFDV_HD double fdlibm_expm1(double x) {
#pragma clang fp contract(off)
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double invln2 = 1.44269504088896338700e+00;
    const double Q1 = -3.33333333333331316428e-02, Q2 = 1.58730158725481460165e-03,
                 Q3 = -7.93650757867487942473e-05, Q4 = 4.00821782732936239552e-06,
                 Q5 = -2.01099218183624371326e-07;
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
    const unsigned hx = (unsigned)(bits >> 32) & 0x7fffffffu;
    const bool neg = (bits >> 63) != 0;
    if (hx >= 0x4043687Au) return exp(x) - 1.0;   // |x| >= 38.8: not reached by the callers
    int k = 0;
    double c = 0.0;
    if (hx > 0x3fd62e42u) {   // |x| > ln2/2: reduce
        double hi, lo;
        if (hx < 0x3FF0A2B2u) {   // |x| < 3 ln2 / 2
            if (!neg) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
            else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
        } else {
            k = (int)(invln2 * x + (neg ? -0.5 : 0.5));
            const double t = k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (hx < 0x3c900000u) {   // |x| < 2^-54
        return x;
    }
    const double hfx = 0.5 * x;
    const double hxs = x * hfx;
    const double R1 = 1.0 + hxs * Q1, h2 = hxs * hxs;
    const double R2 = Q2 + hxs * Q3, h4 = h2 * h2;
    const double R3 = Q4 + hxs * Q5;
    const double r1 = R1 + h2 * R2 + h4 * R3;
    const double t = 3.0 - r1 * hfx;
    double e = hxs * ((r1 - t) / (6.0 - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e -= hxs;
    if (k == -1) return 0.5 * (x - e) - 0.5;
    if (k == 1) {
        if (x < -0.25) return -2.0 * (e - (x + 0.5));
        return 1.0 + 2.0 * (x - e);
    }
    const double twopk = __builtin_bit_cast(double, (unsigned long long)(0x3ff + k) << 52);
    if (k <= -2 || k > 56) {
        const double y = 1.0 - (e - x);
        return y * twopk - 1.0;
    }
    if (k < 20) {
        const double o = __builtin_bit_cast(double, (unsigned long long)(0x3ff00000u - (0x200000u >> k)) << 32);
        return (o - (e - x)) * twopk;
    }
    const double o = __builtin_bit_cast(double, (unsigned long long)(0x3ff - k) << 52);
    return ((x - (e + o)) + 1.0) * twopk;
}
CFDV_HD double libm_sinh(double x) {
#pragma clang fp contract(off)
    const double ax = fabs(x), h = x < 0 ? -0.5 : 0.5;
    if (!(ax < 22.0)) return sinh(x);
    if (ax < 0x1p-28) return x;
    const double t = fdlibm_expm1(ax);
    if (ax < 1.0) return h * (2.0 * t - t * t / (t + 1.0));
    return h * (t + t / (t + 1.0));
}
CFDV_HD double libm_cosh(double x) {
#pragma clang fp contract(off)
    const double ax = fabs(x);
    if (!(ax < __builtin_bit_cast(double, 0x3fd62e4300000000ull))) return cosh(x);   // ~ln2/2
    if (ax < 0x1p-55) return 1.0;
    const double t = fdlibm_expm1(ax), w = 1.0 + t;
    return 1.0 + (t * t) / (w + w);
}

// Gauss-point state (Flow.h:9-62 with Tau)
template <int NDIM, typename S = double> struct GPState {
    static constexpr int NEQ = Dims<NDIM>::NEQ;
    S E, T;
    S v[NDIM];
    S U[NEQ];
    S dU[NEQ][NDIM];
    S dV[NDIM][NDIM];
    S dT[NDIM];
    S k, mu;
    S tau[NDIM][NDIM];
};

template <int NDIM, typename S>
CFDV_HD void calc_tau(GPState<NDIM, S>& f) {
    // Tau::calculate, tau.h:14-42
    S vkk = 0;
#pragma unroll
    for (int k = 0; k < NDIM; k++) vkk += f.dV[k][k];
#pragma unroll
    for (int i = 0; i < NDIM; i++)
#pragma unroll
        for (int j = 0; j < NDIM; j++) {
            f.tau[i][j] = f.mu * (f.dV[i][j] + f.dV[j][i]);
            if (i == j) f.tau[i][j] -= (2.0 / 3.0) * f.mu * vkk;
        }
}

// ---------------------------------------------------------------------------
// Flux Jacobians.  3D: NavierStokes3D.h:74-477, 2D: NavierStokes2D.h:59-222.
// a[i][r][s] = dF_i/dU, b = dG_i/dU, c[i][j] = dG_i/dU_,j.
// ---------------------------------------------------------------------------
// Per-direction forms of calc_a/calc_b (3D) so that only one direction's
// 2 x 25 coefficients are live at a time.  Entries shared between
// directions in the reference (b(1,1,.) = b(0,2,.), b(2,1,.) = b(0,3,.),
// b(2,2,.) = b(1,3,.), NavierStokes3D.h:275-277, 307-312) are recomputed
// from the same expressions.
template <int I, typename TH, typename S>
CFDV_HD void calc_a3_dir(const TH& th, const GPState<3, S>& f, S a[5][5]) {
    const S gamma = th.gamma, gamma1 = th.gamm1, gamma3 = th.gamm3, gm3d2 = th.gm3d2, gm1d2 = th.gm1d2;
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int s = 0; s < 5; s++) a[r][s] = 0.0;
    const S u = f.v[0], v = f.v[1], w = f.v[2];
    const S uu = u * u, uv = u * v, vv = v * v, ww = w * w, uw = u * w, vw = v * w;
    const S E = f.E;
    const S q2 = uu + vv + ww;
    if (I == 0) {
        a[0][1] = 1.0;
        a[1][0] = gm3d2 * uu + gm1d2 * (vv + ww);
        a[1][1] = -gamma3 * u; a[1][2] = -gamma1 * v; a[1][3] = -gamma1 * w; a[1][4] = gamma1;
        a[2][0] = -uv; a[2][1] = v; a[2][2] = u;
        a[3][0] = -uw; a[3][1] = w; a[3][3] = u;
        a[4][0] = -gamma * E * u + gamma1 * u * q2;
        a[4][1] = gamma * E - gm1d2 * (3. * uu + (vv + ww));
        a[4][2] = -gamma1 * uv; a[4][3] = -gamma1 * uw; a[4][4] = gamma * u;
    } else if (I == 1) {
        a[0][2] = 1.0;
        a[1][0] = -uv; a[1][1] = v; a[1][2] = u;
        a[2][0] = gm1d2 * (uu + ww) + gm3d2 * vv;
        a[2][1] = -gamma1 * u; a[2][2] = -gamma3 * v; a[2][3] = -gamma1 * w; a[2][4] = gamma1;
        a[3][0] = -vw; a[3][2] = w; a[3][3] = v;
        a[4][0] = -gamma * E * v + gamma1 * v * q2;
        a[4][1] = -gamma1 * uv;
        a[4][2] = gamma * E - gm1d2 * ((uu + ww) + 3. * vv);
        a[4][3] = -gamma1 * vw; a[4][4] = gamma * v;
    } else {
        a[0][3] = 1.0;
        a[1][0] = -uw; a[1][1] = w; a[1][3] = u;
        a[2][0] = -vw; a[2][2] = w; a[2][3] = v;
        a[3][0] = gm1d2 * (uu + vv) + gm3d2 * ww;
        a[3][1] = -gamma1 * u; a[3][2] = -gamma1 * v; a[3][3] = -gamma3 * w; a[3][4] = gamma1;
        a[4][0] = -gamma * E * w + gamma1 * w * q2;
        a[4][1] = -gamma1 * uw; a[4][2] = -gamma1 * vw;
        a[4][3] = gamma * E - gm1d2 * ((uu + vv) + 3. * ww);
        a[4][4] = gamma * w;
    }
}

template <int I, typename TH, typename S>
CFDV_HD void calc_b3_dir(const TH& th, const GPState<3, S>& f, S b[5][5]) {
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int s = 0; s < 5; s++) b[r][s] = 0.0;
    const S mu = f.mu, E = f.E;
    const S lambda = -2. * mu / 3.;
    const S mu_R = 2. * mu + lambda;
    const S denom = 1. / (f.U[0] * f.U[0]);
    const S TERM1 = f.k * denom / th.Cv;
    const S rinv = 1.0 / f.U[0];
    const S u = f.v[0], v = f.v[1], w = f.v[2];
    const S q2 = u * u + v * v + w * w;
    const S drdx = f.dU[0][0], dldx = f.dU[1][0], dmdx = f.dU[2][0], dndx = f.dU[3][0];
    const S drdy = f.dU[0][1], dldy = f.dU[1][1], dmdy = f.dU[2][1], dndy = f.dU[3][1];
    const S drdz = f.dU[0][2], dldz = f.dU[1][2], dmdz = f.dU[2][2], dndz = f.dU[3][2];
    // rows 1..3 of direction I (b(I, 1..3, .))
    if (I == 0) {
        b[1][0] = denom * (mu_R * dldx + lambda * (dmdy + dndz) + mu_R * (-2. * u * drdx + v * drdy + w * drdz));
        b[1][1] = mu_R * drdx * denom; b[1][2] = lambda * drdy * denom; b[1][3] = lambda * drdz * denom;
        b[2][0] = mu * denom * (dldy + dmdx - 2. * (u * drdy + v * drdx));
        b[2][1] = mu * drdy * denom; b[2][2] = mu * drdx * denom;
        b[3][0] = mu * denom * (dldz + dndx - 2. * (u * drdz + w * drdx));
        b[3][1] = mu * drdz * denom; b[3][3] = mu * drdx * denom;
    } else if (I == 1) {
        b[1][0] = mu * denom * (dldy + dmdx - 2. * (u * drdy + v * drdx));
        b[1][1] = mu * drdy * denom; b[1][2] = mu * drdx * denom;
        b[2][0] = denom * (lambda * (dldx + dndz) + mu_R * (dmdy + u * drdx - 2. * v * drdy + w * drdz));
        b[2][1] = lambda * drdx * denom; b[2][2] = mu_R * drdy * denom; b[2][3] = lambda * drdz * denom;
        b[3][0] = mu * denom * (dmdz + dndy - 2. * (v * drdz + w * drdy));
        b[3][2] = mu * denom * drdz; b[3][3] = mu * denom * drdy;
    } else {
        b[1][0] = mu * denom * (dldz + dndx - 2. * (u * drdz + w * drdx));
        b[1][1] = mu * drdz * denom; b[1][3] = mu * drdx * denom;
        b[2][0] = mu * denom * (dmdz + dndy - 2. * (v * drdz + w * drdy));
        b[2][2] = mu * denom * drdz; b[2][3] = mu * denom * drdy;
        b[3][0] = denom * (mu_R * (dndz + u * drdx + v * drdy - 2. * w * drdz) + lambda * (dldx + dmdy));
        b[3][1] = lambda * drdx * denom; b[3][2] = lambda * drdy * denom; b[3][3] = mu_R * drdz * denom;
    }
    // row 4 of direction I, derivative index X = I
    const S drdX = f.dU[0][I], dldX = f.dU[1][I], dmdX = f.dU[2][I], dndX = f.dU[3][I], dedX = f.dU[4][I];
    b[4][0] = u * b[1][0] + v * b[2][0] + w * b[3][0] +
              (denom * (f.U[1] * f.tau[I][0] + f.U[2] * f.tau[I][1] + f.U[3] * f.tau[I][2]) -
               TERM1 * (-dedX + (2. * E - 3. * q2) * drdX + 2. * (u * dldX + v * dmdX + w * dndX)));
    b[4][1] = u * b[1][1] + v * b[2][1] + w * b[3][1] + (-f.tau[I][0] * rinv - TERM1 * (2. * u * drdX - dldX));
    b[4][2] = u * b[1][2] + v * b[2][2] + w * b[3][2] + (-f.tau[I][1] * rinv - TERM1 * (2. * v * drdX - dmdX));
    b[4][3] = u * b[1][3] + v * b[2][3] + w * b[3][3] + (-f.tau[I][2] * rinv - TERM1 * (2. * w * drdX - dndX));
    b[4][4] = TERM1 * drdX;
    const S reinv = 1.0 / th.creyn;
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int s = 0; s < 5; s++) b[r][s] *= reinv;
}

// c_ij (NavierStokes3D::calc_c, NavierStokes3D.h:347-477), contracted on
// the fly:  out[r][s] += sum_ij W[i][j] c_ij[r][s].
template <typename TH, typename S>
CFDV_HD void accum_c3(const TH& th, const GPState<3, S>& f, const S W[3][3], S out[5][5]) {
    const S mu = f.mu;
    const S lambda = -2. * mu / 3.;
    const S mu_R = 2. * mu + lambda;
    const S rho = f.U[0];
    const S rinv = 1.0 / rho;
    const S denom = 1. / (rho * rho);
    const S l = f.U[1], m = f.U[2], n = f.U[3], re = f.U[4];
    const S ll = l * l, mm = m * m, nn = n * n;
    const S hkocv = f.k / th.Cv;
    const S TERM1 = f.k / th.Cv * denom;
    const S reinv = 1.0 / th.creyn;
    const S lr = l * rinv, mr = m * rinv, nr = n * rinv;
    const S en = TERM1 * (-re + (ll + mm + nn) * rinv);

    // diagonal blocks c_ii
    S c[5][5];
#define CSET_DIAG(I, MUX, MUY, MUZ, LX, MX, NX)                                   \
    {                                                                             \
        for (int r = 0; r < 5; r++)                                               \
            for (int s = 0; s < 5; s++) c[r][s] = 0.0;                            \
        c[1][0] = MUX * l * denom; c[1][1] = (-MUX * rinv);                       \
        c[2][0] = MUY * m * denom; c[2][2] = (-MUY * rinv);                       \
        c[3][0] = MUZ * n * denom; c[3][3] = (-MUZ * rinv);                       \
        c[4][0] = (l * c[1][0] + m * c[2][0] + n * c[3][0]) * rinv - en;          \
        c[4][1] = lr * c[1][1] + TERM1 * l;                                       \
        c[4][2] = mr * c[2][2] + TERM1 * m;                                       \
        c[4][3] = nr * c[3][3] + TERM1 * n;                                       \
        c[4][4] = (-hkocv * rinv);                                                \
        const S wgt = W[I][I] * reinv;                                       \
        for (int r = 0; r < 5; r++)                                               \
            for (int s = 0; s < 5; s++) out[r][s] += wgt * c[r][s];               \
    }
    CSET_DIAG(0, mu_R, mu, mu, 0, 0, 0)
    CSET_DIAG(1, mu, mu_R, mu, 0, 0, 0)
    CSET_DIAG(2, mu, mu, mu_R, 0, 0, 0)
#undef CSET_DIAG
    // off-diagonal blocks c_ij, i != j: nonzeros at (1+i,0), (1+i,1+j),
    // (1+j,0), (1+j,1+i), (4,0), (4,1+i)... restated per reference entry
    // c(i,j,1+i,0) = lambda * U(1+j) * denom, c(i,j,1+i,1+j) = -lambda/rho,
    // c(i,j,1+j,0) = mu * U(1+i) * denom,     c(i,j,1+j,1+i) = -mu/rho,
    // c(i,j,4,0)   = U(1+i)/rho c(i,j,1+i,0) + U(1+j)/rho c(i,j,1+j,0),
    // c(i,j,4,1+i) = U(1+j)/rho c(i,j,1+j,1+i), c(i,j,4,1+j) = U(1+i)/rho c(i,j,1+i,1+j).
    const S Uc[3] = {l, m, n};
    const S Ur[3] = {lr, mr, nr};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (i == j) continue;
            const S wgt = W[i][j] * reinv;
            const S c_i0 = lambda * Uc[j] * denom;
            const S c_ij = (-lambda * rinv);
            const S c_j0 = mu * Uc[i] * denom;
            const S c_ji = (-mu * rinv);
            out[1 + i][0] += wgt * c_i0;
            out[1 + i][1 + j] += wgt * c_ij;
            out[1 + j][0] += wgt * c_j0;
            out[1 + j][1 + i] += wgt * c_ji;
            out[4][0] += wgt * (Ur[i] * c_i0 + Ur[j] * c_j0);
            out[4][1 + i] += wgt * (Ur[j] * c_ji);
            out[4][1 + j] += wgt * (Ur[i] * c_ij);
        }
}

// 2D Jacobians (NavierStokes2D.h:59-222)
CFDV_HD void calc_a2(const DevThermo& th, const GPState<2>& f, double a[2][4][4]) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int s = 0; s < 4; s++) a[i][r][s] = 0.0;
    const double u = f.v[0], v = f.v[1];
    const double uu = u * u, uv = u * v, vv = v * v;
    const double E = f.E;
    a[0][0][1] = 1.0;
    a[0][1][0] = th.gm3d2 * uu + th.gm1d2 * vv;
    a[0][1][1] = -th.gamm3 * u;
    a[0][1][2] = -th.gamm1 * v;
    a[0][1][3] = th.gamm1;
    a[0][2][0] = -uv; a[0][2][1] = v; a[0][2][2] = u;
    a[0][3][0] = -th.gamma * E * u + th.gamm1 * u * (uu + vv);
    a[0][3][1] = th.gamma * E - th.gm1d2 * (3. * uu + vv);
    a[0][3][2] = -th.gamm1 * uv;
    a[0][3][3] = th.gamma * u;
    a[1][0][2] = 1.0;
    a[1][1][0] = -uv; a[1][1][1] = v; a[1][1][2] = u;
    a[1][2][0] = th.gm1d2 * uu + th.gm3d2 * vv;
    a[1][2][1] = -th.gamm1 * u;
    a[1][2][2] = -th.gamm3 * v;
    a[1][2][3] = th.gamm1;
    a[1][3][0] = -th.gamma * E * v + th.gamm1 * v * (uu + vv);
    a[1][3][1] = -th.gamm1 * uv;
    a[1][3][2] = th.gamma * E - th.gm1d2 * (uu + 3. * vv);
    a[1][3][3] = th.gamma * v;
}

CFDV_HD void calc_b2(const DevThermo& th, const GPState<2>& f, double b[2][4][4]) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int s = 0; s < 4; s++) b[i][r][s] = 0.0;
    const double reinv = 1.0 / th.creyn;
    const double mu = f.mu, E = f.E;
    const double lambda = -2. * mu / 3.;
    const double mu_R = 2. * mu + lambda;
    const double DENOM = 1. / (f.U[0] * f.U[0]);
    const double TERM1 = f.k * DENOM / th.Cv;
    const double rho = f.U[0];
    const double u = f.v[0], v = f.v[1];
    const double uu = u * u, vv = v * v;
    const double drdx = f.dU[0][0], dldx = f.dU[1][0], dmdx = f.dU[2][0], dedx = f.dU[3][0];
    const double drdy = f.dU[0][1], dldy = f.dU[1][1], dmdy = f.dU[2][1], dedy = f.dU[3][1];
    b[0][1][0] = (DENOM * (mu_R * dldx + lambda * dmdy - 2. * (mu_R * u * drdx + lambda * v * drdy))) * reinv;
    b[0][1][1] = mu_R * drdx * DENOM * reinv;
    b[0][1][2] = lambda * drdy * DENOM * reinv;
    b[0][2][0] = (mu * DENOM * (dldy + dmdx - 2. * (u * drdy + v * drdx))) * reinv;
    b[0][2][1] = mu * drdy * DENOM * reinv;
    b[0][2][2] = mu * drdx * DENOM * reinv;
    b[0][3][0] = u * b[0][1][0] + v * b[0][2][0] +
                 ((u * f.tau[0][0] + v * f.tau[1][0]) / rho -
                  TERM1 * (-dedx + (2. * E - 3. * uu - 3. * vv) * drdx + 2. * u * dldx + 2. * v * dmdx)) * reinv;
    b[0][3][1] = u * b[0][1][1] + v * b[0][2][1] + (-f.tau[0][0] / rho - TERM1 * (2. * u * drdx - dldx)) * reinv;
    b[0][3][2] = u * b[0][1][2] + v * b[0][2][2] + (-f.tau[1][0] / rho - TERM1 * (-dmdx + 2. * v * drdx)) * reinv;
    b[0][3][3] = TERM1 * drdx * reinv;
    b[1][1][0] = b[0][2][0];
    b[1][1][1] = b[0][2][1];
    b[1][1][2] = b[0][2][2];
    b[1][2][0] = (DENOM * (lambda * (dldx + 4. * v * drdy) + mu_R * (u * drdx + dmdy))) * reinv;
    b[1][2][1] = lambda * drdx * DENOM * reinv;
    b[1][2][2] = mu_R * drdy * DENOM * reinv;
    b[1][3][0] = u * b[1][1][0] + v * b[1][2][0] +
                 ((u * f.tau[0][1] + v * f.tau[1][1]) / rho -
                  TERM1 * (-dedy + (2. * E - 3. * uu - 3. * vv) * drdy + 2. * u * dldy + 2. * v * dmdy)) * reinv;
    b[1][3][1] = u * b[1][1][1] + v * b[1][2][1] + (-f.tau[1][0] / rho - TERM1 * (-dldy + 2. * u * drdy)) * reinv;
    b[1][3][2] = u * b[1][1][2] + v * b[1][2][2] + (-f.tau[1][1] / rho - TERM1 * (2. * v * drdy - dmdy)) * reinv;
    b[1][3][3] = TERM1 * drdy * reinv;
}

CFDV_HD void accum_c2(const DevThermo& th, const GPState<2>& f, const double W[2][2],
                                         double out[4][4]) {
    const double reinv = 1.0 / th.creyn;
    const double mu = f.mu;
    const double lambda = -2. * mu / 3.;
    const double mu_R = 2. * mu + lambda;
    const double rho = f.U[0];
    const double denom = 1. / (rho * rho);
    const double denom2 = denom / rho;
    const double l = f.U[1], m = f.U[2], re = f.U[3];
    const double ll = l * l, lm = l * m, mm = m * m;
    const double rv = ll + mm;
    const double hkocv = f.k / th.Cv;
    double c[2][2][4][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) c[i][j][r][s] = 0.0;
    c[0][0][1][0] = mu_R * l * denom * reinv;
    c[0][0][1][1] = -mu_R / rho * reinv;
    c[0][0][2][0] = mu * m * denom * reinv;
    c[0][0][2][2] = -mu / rho * reinv;
    c[0][0][3][0] = ((mu_R * ll * denom2 + mu * mm * denom2) - hkocv * (-re * denom + rv * denom2)) * reinv;
    c[0][0][3][1] = (-mu_R + hkocv) * l * denom * reinv;
    c[0][0][3][2] = (-mu + hkocv) * m * denom * reinv;
    c[0][0][3][3] = -hkocv / rho * reinv;
    c[0][1][1][0] = lambda * m * denom * reinv;
    c[0][1][1][2] = -lambda / rho * reinv;
    c[0][1][2][0] = mu * l * denom * reinv;
    c[0][1][2][1] = -mu / rho * reinv;
    c[0][1][3][0] = (lambda + mu) * lm * denom2 * reinv;
    c[0][1][3][1] = -mu * m * denom * reinv;
    c[0][1][3][2] = -lambda * l * denom * reinv;
    c[1][0][1][0] = mu * m * denom * reinv;
    c[1][0][1][2] = -mu / rho * reinv;
    c[1][0][2][0] = lambda * l * denom * reinv;
    c[1][0][2][1] = -lambda / rho * reinv;
    c[1][0][3][0] = (mu + lambda) * lm * denom2 * reinv;
    c[1][0][3][1] = -lambda * m * denom * reinv;
    c[1][0][3][2] = -mu * l * denom * reinv;
    c[1][1][1][0] = mu * l * denom * reinv;
    c[1][1][1][1] = -mu / rho * reinv;
    c[1][1][2][0] = mu_R * m * denom * reinv;
    c[1][1][2][2] = -mu_R / rho * reinv;
    c[1][1][3][0] = ((mu * ll * denom2 + mu_R * mm * denom2) - hkocv * (-re * denom + rv * denom2)) * reinv;
    c[1][1][3][1] = (-mu + hkocv) * l * denom * reinv;
    c[1][1][3][2] = (-mu_R + hkocv) * m * denom * reinv;
    c[1][1][3][3] = -hkocv / rho * reinv;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int s = 0; s < 4; s++) out[r][s] += W[i][j] * c[i][j][r][s];
}

}  // namespace cfdv
