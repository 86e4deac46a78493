// Host-side check of the interior build's per-point tables (build.hip),
// compiled with hipcc and run on the CPU (the functions are __host__
// __device__): for random Gauss-point states,
//   * gp_tab_dir_rot(I) -- direction 0's formulas on the axis-swapped
//     state -- equals calc_a3_dir<I> / calc_b3_dir<I> entry by entry, its
//     rotated M1 slots cover exactly direction I's 17-entry pattern (slot 0
//     is s1), and PQ = P - dt^2/2 (a + b) D;
//   * accum_c3_ctab over the coefficient table that the three direction
//     calls fill between them (each its share, gp_tab_c_dir) equals
//     accum_c3 (NavierStokes3D.h:347-477) for random weights W; the table
//     starts as NaNs, and one that no direction writes fails.
// Prints the worst relative difference of each and exits 1 above 1e-14.
#include "../../cfdv_amd/csrc/build.hip"
#include <cmath>
#include <cstdio>
#include <random>
using namespace cfdv;

int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1, 1);
    DevThermo th{};
    th.gamma = 1.4; th.gamm1 = 0.4; th.gm1d2 = 0.2; th.gamm3 = 1.4 - 3; th.gm3d2 = (1.4 - 3) / 2;
    th.Pr = 0.72; th.csuth = 110.0 / 300.0; th.cmach = 6; th.creyn = 5e5;
    th.Cv = 1. / 1.4 / 0.4 / 36.; th.cgas = 1. / 1.4 / 36.;
    double worst_dir = 0, worst_c = 0;
    for (int trial = 0; trial < 500; trial++) {
        GPShared g;
        GPState<3>& f = g.f;
        f.E = 2 + U(rng); f.T = 1 + 0.3 * U(rng); f.mu = 1 + 0.2 * U(rng); f.k = 0.5 + 0.1 * U(rng);
        for (int i = 0; i < 3; i++) {
            f.v[i] = U(rng); f.dT[i] = U(rng);
            for (int j = 0; j < 3; j++) f.dV[i][j] = U(rng);
        }
        for (int q = 0; q < 5; q++) {
            f.U[q] = (q == 0 ? 1.5 : 0) + U(rng);
            for (int j = 0; j < 3; j++) f.dU[q][j] = U(rng);
        }
        calc_tau<3>(f);
        for (int c = 0; c < 5; c++) g.D[c] = U(rng);
        for (int i = 0; i < 3; i++) {
            g.P[i][0] = 0;
            for (int r = 1; r < 5; r++) g.P[i][r] = U(rng);
        }
        const double sp[4] = {0.5 + 0.5 * U(rng), 0.7 + 0.2 * U(rng), 0.5 + 0.5 * U(rng), 0.7 + 0.2 * U(rng)};
        const double dt = 0.05 * (1.5 + U(rng));
        GPTab t;
        for (int c = 0; c < CT_N; c++) t.C[c] = NAN;
        for (int I = 0; I < 3; I++) gp_tab_dir_rot(th, g, sp, dt, I, t);
        for (int c = 0; c < CT_N; c++)
            if (std::isnan(t.C[c])) {
                printf("c_ij table entry %d written by no direction\n", c);
                return 1;
            }
        for (int I = 0; I < 3; I++) {
            double a[5][5], b[5][5];
            if (I == 0) { calc_a3_dir<0>(th, f, a); calc_b3_dir<0>(th, f, b); }
            if (I == 1) { calc_a3_dir<1>(th, f, a); calc_b3_dir<1>(th, f, b); }
            if (I == 2) { calc_a3_dir<2>(th, f, a); calc_b3_dir<2>(th, f, b); }
            double scale = 0;
            for (int r = 0; r < 5; r++)
                for (int c = 0; c < 5; c++) scale = fmax(scale, fabs(a[r][c]) + fabs(b[r][c]));
            bool seen[5][5] = {};
            for (int p = 0; p < 17; p++) {
                const int r = sig_eq(I, JPAT_R[0][p]), c = sig_eq(I, JPAT_S[0][p]);
                seen[r][c] = true;
                const double m1 = p == 0 ? sp[0] : t.M1[I][p - 1];
                worst_dir = fmax(worst_dir, fabs(m1 - (sp[0] * a[r][c] + sp[2] * b[r][c])) / scale);
            }
            for (int p = 0; p < 17; p++)
                if (!seen[JPAT_R[I][p]][JPAT_S[I][p]]) {
                    printf("direction %d: pattern entry %d not covered\n", I, p);
                    return 1;
                }
            for (int r = 0; r < 5; r++)
                for (int c = 0; c < 5; c++) {
                    if (!seen[r][c] && (a[r][c] != 0 || b[r][c] != 0)) {
                        printf("direction %d: nonzero (%d, %d) outside the pattern\n", I, r, c);
                        return 1;
                    }
                    worst_dir = fmax(worst_dir, fabs(t.Af[I][r * 5 + c] - (a[r][c] + b[r][c])) / scale);
                    worst_dir = fmax(worst_dir, fabs(t.Sf[I][r * 5 + c] - (sp[1] * a[r][c] + sp[3] * b[r][c])) / scale);
                }
            for (int r = 0; r < 5; r++) {
                double q = 0;
                for (int c = 0; c < 5; c++) q += (a[r][c] + b[r][c]) * g.D[c];
                const double pq = g.P[I][r] - dt * dt / 2 * q;
                worst_dir = fmax(worst_dir, fabs(t.PQ[I][sig_eq(I, r)] - pq) / (fabs(g.P[I][r]) + scale * 5));
            }
        }
        double W[3][3], o1[5][5] = {}, o2[5][5] = {};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) W[i][j] = U(rng);
        accum_c3(th, f, W, o1);
        accum_c3_ctab(t.C, W, o2);
        double sc = 0;
        for (int r = 0; r < 5; r++)
            for (int c = 0; c < 5; c++) sc = fmax(sc, fabs(o1[r][c]));
        for (int r = 0; r < 5; r++)
            for (int c = 0; c < 5; c++) worst_c = fmax(worst_c, fabs(o1[r][c] - o2[r][c]) / sc);
    }
    printf("directions %.3e c_ij table %.3e\n", worst_dir, worst_c);
    return (worst_dir < 1e-14 && worst_c < 1e-14) ? 0 : 1;
}
