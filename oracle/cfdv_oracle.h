/*
 * CPU oracle for the FDV time step of hahnpv/CFDV.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithm, used as the checker by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in
 * the product library (cfdv_amd/) links, loads or calls it.
 *
 * Every routine follows the reference's own loop and operand order (so that
 * it reproduces the reference's rounding as closely as plain C allows; build
 * with -ffp-contract=off) and cites the reference file:line it restates.
 * Reference quirks (numbered Q1..Q16 where they are reproduced) are kept, not fixed.
 *
 * Pinning: each function is checked in tests/test_oracle_*.py against closed
 * forms, invariants or finite differences (see the header comment of each
 * test module).  Functions that no such check reaches are marked
 * "parity unpinned" below.
 */
#ifndef CFDV_ORACLE_H
#define CFDV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Thermo singleton, src/FDV/Thermo.h:12-33 */
typedef struct {
    double gamma, gamm1, gm1d2, gamm3, gm3d2;
    double Pr, csuth, Cv, cgas, cmach, creyn, Twall;
    int adiabatic;
} orc_thermo;

/* Node primitives written by NodeUnpack, src/FDV/Function/NodeUnpack.h:9-29 */
typedef struct {
    double rho, v[3], E, e, p, T;
} orc_prim;

/* Gauss-point state, src/FDV/Flow.h:9-62 (+ Tau, src/FDV/tau.h) */
typedef struct {
    double E, T;
    double v[3], x[3];
    double U[5];
    double dU[5][3];   /* dU(i, j): eqn i, direction j */
    double dV[3][3];   /* dV(i, j) = d v_i / d x_j      */
    double dT[3];
    double k, mu;
    double tau[3][3];
} orc_flow;

/* Mesh as read by LoadBinary (src/Utility/IO/ASCIItoBinary/LoadBinary.h) */
typedef struct {
    int ndim, nnod, neqn, nbnod, nface;
    int n_elem, n_node;
    const int *conn;                 /* [n_elem][nnod] global node ids      */
    const double *x;                 /* [n_node][ndim]                      */
    const int *bc;                   /* [n_elem][nface], 0 = no face        */
    const unsigned char *dirichlet;  /* [n_node][neqn]                      */
    const unsigned char *adap;       /* [n_elem] or NULL                    */
    const double *hnm;               /* [n_elem][nnod*nnod], Hnm(N,M) at
                                        N + nnod*M, or NULL                 */
} orc_mesh;

/* GMRES statistics */
typedef struct {
    double tol;          /* 1e-9 * ||b|| (quirk) or rtol * ||b|| (true) */
    double bnorm;
    double last_resid;   /* |g(m)| after the last restart (Givens estimate) */
    double true_resid;   /* ||b - A du|| recomputed at exit (true mode)   */
    int matvecs;
    int iters;           /* Arnoldi iterations over all restarts          */
    int restarts;
    int breakdown;
} orc_gmres_stats;

/* ---- thermodynamics ---------------------------------------------------- */
void   orc_thermo_set(orc_thermo *th, double gamma, double M, double Re,
                      double Pr, double Tinf, double Tw, int adiabatic);
double orc_calc_mu(const orc_thermo *th, double T);
double orc_calc_k(const orc_thermo *th, double mu);

/* ---- finite-element basis --------------------------------------------- */
void orc_gauss(int N, double *w, double *zeta);
void orc_shape2d(const double *zeta, double *phi, double *dpxi /*[4][2]*/);
void orc_shape3d(const double *zeta, double *phi, double *dpxi /*[8][3]*/);
/* J (row-major i,j) = sum_N dPxi(N,j) x_N(i); returns det; dphi [nnod][ndim] */
double orc_jacobian3d(const double *dpxi, const double *xn /*[8][3]*/,
                      double *jac, double *dphi);
double orc_jacobian2d(const double *dpxi, const double *xn /*[4][2]*/,
                      double *jac, double *dphi);
/* Face test functions/normal (calc_TF_norm3D.h:9-167); nodeloc [4][3] */
double orc_tfnorm3d(const double *nodeloc, const double *w, const double *zeta,
                    double *phibou, double *en, double *dpidx /*[4][3]*/);

/* ---- node / element glue ---------------------------------------------- */
void orc_node_unpack(const orc_thermo *th, int ndim, const double *U,
                     orc_prim *out);
void orc_node_unpack_all(const orc_thermo *th, int ndim, int neqn, int n_node,
                         const double *U, orc_prim *out);
void orc_calc_length(const orc_mesh *m, double *area, double *cleng);
int  orc_cfl(const orc_thermo *th, const orc_mesh *m, const orc_prim *prim,
             const double *cleng, double cfl, double *dt);
/* FDVParam; s is in/out (s1..s4), carried across calls like the functor
   member (quirk Q11).  Returns 1 when r == 0.01 exactly left s1 or s3
   unassigned, else 0. */
int  orc_fdvparam(const orc_thermo *th, const orc_mesh *m, const orc_prim *prim,
                  int e, double *s);

/* ---- Navier-Stokes pieces (3D: NavierStokes3D.h, 2D: NavierStokes2D.h) - */
void orc_tau(int ndim, double mu, const double dV[3][3], double tau[3][3]);
void orc_calc_a(const orc_thermo *th, int ndim, const orc_flow *f,
                double *a /*[ndim][neqn][neqn]*/);
void orc_calc_b(const orc_thermo *th, int ndim, const orc_flow *f,
                double *b /*[ndim][neqn][neqn]*/);
void orc_calc_c(const orc_thermo *th, int ndim, const orc_flow *f,
                double *c /*[ndim][ndim][neqn][neqn]*/);
void orc_calc_G(int ndim, const double tau[3][3], const double *v,
                const double *dT, double HK, double Re, double *G /*[ndim][neqn]*/);
void orc_calc_F2d(const double *U, double p, double *F /*[2][4]*/);
/* numerical diffusion + discontinuity capturing of one Gauss point
   (FDVGalerkin.h:225-387): Psi [nnod]; xn [nnod][nd] element nodes, zeta
   the point, dPhi [nnod][nd] at the point (Psi = 0 in 3D) */
void orc_psi(int nd, int nn, const orc_flow *f, const double *xn, const double *zeta, const double *dPhi,
             double *Psi);
/* Boundary edits of the Gauss-point state; returns 0 or -1 (unknown bc) */
int  orc_eval2dboundary(const orc_thermo *th, int bc, orc_flow *f,
                        const double *en);
int  orc_eval1dboundary(int bc, orc_flow *f, const double *xn0,
                        const double *xn1, const double *en);

/* ---- element build ---------------------------------------------------- */
/* One FDVGalerkin::operator() on element e (FDVGalerkin.h:121-154).
   R is k*k column-major (R(i,j) at i + k*j), rhs k; both accumulate
   (caller clears, as ClearElement does).  s is the carried s1..s4. */
int orc_build_element(const orc_thermo *th, const orc_mesh *m,
                      const double *U, const orc_prim *prim, int e, double dt,
                      double *s, double *R, double *rhs);
/* All elements in order: R [n_elem][k*k], rhs [n_elem][k], s_out [n_elem][4]
   (may be NULL).  Returns number of Q11 ties seen. */
int orc_build(const orc_thermo *th, const orc_mesh *m, const double *U,
              const orc_prim *prim, double dt, double *R, double *rhs,
              double *s_out);
/* ApplyBC on every element (ApplyBC.h:14-133) */
void orc_apply_bc(const orc_thermo *th, const orc_mesh *m, double *R,
                  double *rhs);

/* ---- EBE solve -------------------------------------------------------- */
void orc_compose_rhs(const orc_mesh *m, const double *rhs, double *b);
/* y = sum_e P_e^T R_e P_e x (sum_Ax order, MPI_GMRES.h:570-693); y is
   overwritten */
void orc_matvec(const orc_mesh *m, const double *R, const double *x,
                double *y);
/* y = sum_e P_e^T R_e^T P_e x (transpose operator) */
void orc_matvec_t(const orc_mesh *m, const double *R, const double *x,
                  double *y);
/* MPI_GMRES::iterate at one rank (MPI_GMRES.h:125-220), quirks Q2-Q4 */
int orc_gmres_quirk(const orc_mesh *m, const double *R, const double *rhs,
                    int restart, int mdim, double *du, orc_gmres_stats *st);
/* The same, recording per restart the Hessenberg matrix as the Arnoldi
   process leaves it (before the rotations, Htr [restart][(mdim+1)*mdim],
   H(i,j) at i + (mdim+1) j), the rotated right-hand side g (gtr
   [restart][mdim+1]) and the number of Arnoldi steps (mmtr [restart]);
   any of the three may be NULL. */
int orc_gmres_quirk_trace(const orc_mesh *m, const double *R, const double *rhs,
                          int restart, int mdim, double *du, orc_gmres_stats *st,
                          double *Htr, double *gtr, int *mmtr);
/* Textbook restarted GMRES(m): Givens, correct back substitution, exit at
   ||b - A du|| <= rtol ||b|| or maxrestart restarts.  transpose != 0 solves
   with the transpose operator. */
int orc_gmres_true(const orc_mesh *m, const double *R, const double *b,
                   int mdim, int maxrestart, double rtol, int transpose,
                   double *du, orc_gmres_stats *st);
/* MPI_GMRES::update (MPI_GMRES.h:224-301): U0 = U; U += du */
void orc_update(int n, double *U, double *U0, const double *du);
/* MPI_RMSError (MPI_RMSError.h:19-74): Urms_i = sqrt(sum U^2/sum U0^2) - 1 */
void orc_rms(int n_node, int neqn, const double *U, const double *U0,
             double *urms);

/* ---- refinement flags and coloring ------------------------------------ */
/* AdaptiveRefineList flag pass (RefineListAdaptive.h:28-38): flag = s1 > 0.8
   with s carried across elements.  Returns number flagged. */
int orc_refine_flags(const orc_thermo *th, const orc_mesh *m,
                     const orc_prim *prim, unsigned char *flags);
/* Greedy first-fit element coloring in element order; two elements conflict
   when they share a node.  Returns number of colors. */
int orc_color_greedy(const orc_mesh *m, int *color);
/* Same, visiting elements in increasing order of keys[e]. */
int orc_color_greedy_keys(const orc_mesh *m, const unsigned long long *keys, int *color);

/* ---- one full time step (main.cpp:162-206 without Save/Output) ---------- */
typedef struct {
    double cfl;
    int restart, mdim;   /* gmresrestart, gmresiter */
    int mode;            /* 0 = quirk (reference), 1 = true GMRES */
    double rtol;         /* true mode */
    int maxrestart;      /* true mode */
} orc_step_opts;
typedef struct {
    double dt;
    double urms[5];
    orc_gmres_stats gm;
} orc_step_stats;
int orc_step(const orc_thermo *th, const orc_mesh *m, double *U, double *U0,
             const orc_step_opts *o, orc_step_stats *st);

/* Multithreaded timing variant of orc_step with mode 1 (cfdv_oracle_par.c):
   the build over elements, the product one element color at a time (cptr
   [ncolor + 1] offsets into celem, the elements of each color), GMRES vector
   work over rows, on nthreads OpenMP threads.  t_phase [4] (may be NULL)
   receives seconds for set-up, build, ApplyBC + rhs, GMRES + update;
   verbose != 0 prints build and restart progress to stderr. */
int orc_step_par(const orc_thermo *th, const orc_mesh *m, double *U, double *U0, const orc_step_opts *o,
                 const int *cptr, const int *celem, int ncolor, int nthreads, orc_step_stats *st,
                 double *t_phase, int verbose);

#ifdef __cplusplus
}
#endif
#endif
