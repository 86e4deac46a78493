/*
 * C ABI of libcfdv_amd: the MI355X (gfx950) FDV time step of hahnpv/CFDV.
 *
 * Plain pointers and sizes only.  The caller owns every host array; the
 * context owns every device buffer.  No pointer is retained past a call.
 * Every function returns CFDV_OK (0) or a negative error code and never
 * exits, throws across the ABI or reads stdin; cfdv_last_error() gives the
 * message.  A context is not thread-safe; it runs on one HIP stream and every
 * call is synchronous on return unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to the reference src/):
 *   cfdv_create/destroy  Thermo::set (FDV/Thermo.h:22-33) + the mesh held by
 *                        LoadBinaryData (Utility/IO/ASCIItoBinary/LoadBinary.h:
 *                        156-318: nodes, elements, faces, Dirichlet flags)
 *                        and CalcLength (FDV/Function/CalcLength.h:34-110,
 *                        main.cpp:120-121)
 *   cfdv_set_state       node U after LoadBinary::loadNodes (LoadBinary.h:
 *                        174-200) or any later state
 *   cfdv_cfl             CFL (FDV/Function/MPI_CFL.h:9-56, main.cpp:179)
 *   cfdv_build           for_each(ClearElement), for_each(NodeUnpack),
 *                        for_each(FDVGalerkin), for_each(ApplyBC)
 *                        (main.cpp:169-194; FDV/FDVGalerkin.h:121-154,
 *                        FDV/Function/ApplyBC.h:14-133)
 *   cfdv_get_blocks      Element::R, Element::rhs (FDV/Element.h:33-34)
 *   cfdv_solve           GMRESbase::iterate (Utility/Solvers/GMRESbase.h:
 *                        16-17) as implemented by MPI_GMRES::iterate
 *                        (Utility/Solvers/MPI_GMRES.h:125-220)
 *   cfdv_update          GMRESbase::update / MPI_GMRES::update
 *                        (MPI_GMRES.h:224-301)
 *   cfdv_rms             MPI_RMSError (FDV/Function/MPI_RMSError.h:19-74)
 *   cfdv_refine_flags    the flag pass of AdaptiveRefineList
 *                        (adap/RefineListAdaptive.h:28-38)
 *   cfdv_step            one iteration of the time loop (main.cpp:162-206)
 *                        without Save/Output/Sensor
 *   cfdv_color, cfdv_matvec, cfdv_profile_matvec, cfdv_matvec_multi,
 *   cfdv_profile_matvec_multi: no reference counterpart
 *                        (element coloring for the atomic-free EBE matvec,
 *                        and test/measurement hooks).
 *   cfdv_refine_*        host-side refinement, no GPU: cfdv_refine_list is
 *                        AdaptiveRefineList (adap/RefineListAdaptive.h:8-126)
 *                        after its flag pass, with ElementAssociation
 *                        (adap/ElementAssociation.h:42-249); cfdv_refine_apply
 *                        is MeshRefine2D (adap/MeshRefine2D.h:39-872)
 *                        (main.cpp:143-151); cfdv_refine_create/_get stand for
 *                        the element and node vectors those mutate
 *   cfdv_set_thermo      Thermo::set (FDV/Thermo.h:22-33) after create
 *   cfdv_adjoint_solve, cfdv_adjoint_solve_multi,
 *   cfdv_build_vjp:      no reference counterpart (discrete
 *                        adjoint of one step: transpose solve, and the
 *                        forward-mode tangents of the element build).
 */
#ifndef CFDV_H
#define CFDV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFDV_OK 0
#define CFDV_ERR_ARG (-1)        /* bad argument or call order               */
#define CFDV_ERR_HIP (-2)        /* HIP runtime error (incl. out of memory)  */
#define CFDV_ERR_BC (-3)         /* unknown boundary-condition code          */
#define CFDV_ERR_GEOM (-4)       /* non-positive element or face Jacobian    */
#define CFDV_ERR_NAN (-5)        /* NaN/Inf in the solve                     */
#define CFDV_ERR_BREAKDOWN (-6)  /* Krylov breakdown (zero Arnoldi vector)   */
#define CFDV_ERR_CFL (-7)        /* no element set dt                        */
#define CFDV_ERR_NOCONV (-8)     /* tolerance not met within the limits      */
#define CFDV_ERR_REFINE (-9)     /* refinement reached a state where the
                                    reference blocks on stdin or recurses
                                    without end                             */

typedef struct cfdv_ctx cfdv_ctx;

typedef struct {
    int ndim;              /* 3 (hex8) or 2 (quad4)                               */
    int nnod;              /* nodes per element: 8 or 4                           */
    int neqn;              /* equations per node: 5 or 4                          */
    int64_t n_elem;
    int64_t n_node;
    const int32_t *conn;   /* [n_elem][nnod] node ids (0-based)                   */
    const double *x;       /* [n_node][ndim] coordinates                          */
    const int32_t *bc;     /* [n_elem][nface] face bc codes, 0 = interior; face
                              numbering of get_nodes_from_face (FiniteElement/
                              get_face.h:80-165)                                  */
    const uint8_t *dirichlet; /* [n_node][neqn] Node::dirichlet (LoadBinary.h:
                              277-316)                                            */
    const uint8_t *adap;   /* [n_elem] Element::adap (hanging-node elements of a
                              refined mesh, FDV/Element.h:39-41) or NULL         */
    const double *hnm;     /* [n_elem][nnod*nnod] Element::Hnm, Hnm(N, M) at
                              N + nnod*M (Tensor column-major), read where adap
                              is set; NULL when adap is NULL.  The adjoint
                              entries (cfdv_build_vjp) take hex8 only          */
} cfdv_mesh;

typedef struct {          /* the aero deck keys (MPIBinaryConfiguration.h:54-63) */
    double gamma, mach, reynolds, prandtl, tinf, twall;
    int adiabatic;
} cfdv_thermo;

/* block storage requested at create time */
#define CFDV_BLOCKS_FP64 1
#define CFDV_BLOCKS_FP32 2
/* also sum the FP64 element blocks into assembled node rows (FP64 and FP32
   copies) after every build (no reference counterpart: the reference
   multiplies element by element, MPI_GMRES.h:442-566).  Needs
   CFDV_BLOCKS_FP64.  Products K x of the FP64 and mixed solves and of
   cfdv_matvec then use the rows (27 blocks per hex8 node instead of 64 per
   element).  K^T x (transposed solves, the adjoint) uses them too: the
   first K^T x after a K x re-sums the rows as those of K^T (one assembly)
   and the first K x after a K^T x back.  The reference solve stays element
   by element.
   CFDV_BLOCKS_ROWS alone (hex8 meshes): the rows without element blocks.
   The build adds each element block into a staged copy of the FP64 rows,
   one color after the other (each slot sums its blocks in color order), and
   lays the FP64 and FP32 rows out from it; K^T's rows are laid out from the
   same copy through the mirrored slots.  The FP64 and mixed solves, the
   adjoint solves, the build tangents and the products work as above;
   cfdv_get_blocks(R), cfdv_get_blocks32 and the reference solve, which need
   element blocks, return CFDV_ERR_ARG.  CFDV_BLOCKS_ROWS with
   CFDV_BLOCKS_FP32 needs CFDV_BLOCKS_FP64. */
#define CFDV_BLOCKS_ROWS 4

/* solver modes */
#define CFDV_SOLVE_REFERENCE 0   /* MPI_GMRES::iterate as written: MGS Arnoldi,
                                    fixed number of restarts, early-break
                                    estimate and back substitution exactly as
                                    the reference (quirks Q2-Q4)               */
#define CFDV_SOLVE_FP64 1        /* restarted GMRES(m), FP64 blocks and vectors,
                                    exit at ||b - K du|| <= rtol ||b||         */
#define CFDV_SOLVE_MIXED 2       /* FP64 iterative refinement around an FP32
                                    GMRES(inner_m) on FP32 blocks; outer
                                    residual, Hessenberg/Givens and update in
                                    FP64                                        */

typedef struct {
    int mode;          /* CFDV_SOLVE_*                                            */
    int restarts;      /* reference: gmresrestart; FP64: max restarts; mixed: max
                          outer refinement sweeps                                */
    int m;             /* Krylov dimension (gmresiter)                            */
    double rtol;       /* FP64/mixed: relative residual target                   */
    int inner_m;       /* mixed: inner Krylov dimension                           */
    double inner_rtol; /* mixed: inner relative residual target (a sweep that
                          needs less to reach rtol ||b|| stops at what it
                          needs, at most 0.5)                                  */
    int inner_restarts;/* mixed: max inner restarts per sweep                    */
    int transpose;     /* solve K^T du = b (adjoint); FP64 and mixed modes only   */
    int precond;       /* FP64 and mixed modes: CFDV_PRECOND_NONE or
                          CFDV_PRECOND_NODE_BLOCK (right preconditioner; the
                          stopping test stays on the true residual)            */
    int warm_start;    /* FP64 and mixed modes, K du = b: start from the last
                          update's du scaled by dt / dt_prev (1), or from the
                          increment rate r_k = du_k / dt_k of the last updates
                          extrapolated: dt (2 r_1 - r_2) (2) or
                          dt (3 r_1 - 3 r_2 + r_3) (3), with one or two more
                          du kept on the device; instead of 0 (the stopping
                          test is unchanged, ||b - K du|| <= rtol ||b||).
                          Ignored for the first solve after create,
                          cfdv_set_state or cfdv_set_thermo; the next ones use
                          the lower modes until enough updates are kept     */
} cfdv_solve_opts;

#define CFDV_PRECOND_NONE 0
/* node-block Jacobi: P_n = sum over the elements at node n of their
   neqn x neqn diagonal block R_e(n, n); applied as K P^-1 z = b, du = P^-1 z */
#define CFDV_PRECOND_NODE_BLOCK 1

typedef struct {
    double bnorm;        /* ||b||                                                 */
    double resid;        /* final ||b - K du|| recomputed in FP64                  */
    double rel_resid;    /* resid / bnorm                                          */
    double est_resid;    /* last Givens residual estimate                          */
    int matvecs64;       /* products with FP64 blocks (rows or EBE)                */
    int matvecs32;       /* products with FP32 blocks (rows or EBE)                */
    int iters;           /* Arnoldi iterations (all restarts / sweeps)             */
    int restarts;        /* restarts (FP64, reference) or refinement sweeps        */
    int breakdown;
} cfdv_solve_stats;

typedef struct {
    double dt;
    double urms[5];
    double ms_cfl, ms_build, ms_solve, ms_update;   /* HIP-event timings */
    double ms_matvec64, ms_matvec32;   /* summed matvec time (events around
                                          each full product: the row kernel,
                                          or zero-fill + one EBE launch per
                                          color)                             */
    int n_matvec64, n_matvec32;
    cfdv_solve_stats solve;
    double ms_assemble;   /* node rows from the element blocks (CFDV_BLOCKS_ROWS),
                             or laid out from the staged rows (CFDV_BLOCKS_ROWS
                             alone), not part of ms_build                   */
} cfdv_step_stats;

int cfdv_create(const cfdv_mesh *mesh, const cfdv_thermo *thermo, int block_flags,
                cfdv_ctx **out);
void cfdv_destroy(cfdv_ctx *ctx);
const char *cfdv_last_error(const cfdv_ctx *ctx);
int cfdv_mesh_info(const cfdv_ctx *ctx, int64_t *n_elem, int64_t *n_node, int *ncolors,
                   int64_t *n_slots);

/* the assembled node rows (CFDV_BLOCKS_ROWS; zeros without): n_blocks node
   pairs that share an element (the neqn x neqn blocks a product reads),
   n_slots stored block slots including the padding of 64-row chunks (no
   reference counterpart; measurement hook) */
int cfdv_rows_info(const cfdv_ctx *ctx, int64_t *n_blocks, int64_t *n_slots);
int cfdv_set_state(cfdv_ctx *ctx, const double *U);            /* [n_node][neqn] */
int cfdv_get_state(cfdv_ctx *ctx, double *U, double *U0);      /* either may be NULL */
int cfdv_cfl(cfdv_ctx *ctx, double cfl, double *dt);
int cfdv_build(cfdv_ctx *ctx, double dt);
/* R: [n_elem][k][k] with R(i,j) at [e][j][i] (the reference Tensor's
   column-major order), rhs: [n_elem][k], s: [n_elem][4] (s1..s4).  Any may
   be NULL. */
int cfdv_get_blocks(cfdv_ctx *ctx, double *R, double *rhs, double *s);
int cfdv_get_blocks32(cfdv_ctx *ctx, float *R);
int cfdv_get_rhs_vector(cfdv_ctx *ctx, double *b);             /* assembled b */
int cfdv_solve(cfdv_ctx *ctx, const cfdv_solve_opts *opts, cfdv_solve_stats *st);
int cfdv_get_du(cfdv_ctx *ctx, double *du);
/* The last CFDV_SOLVE_REFERENCE solve of the context, per restart r <
   *n_restarts, for parity checks upstream of its back substitution (no
   reference counterpart; the values are those of MPI_GMRES::iterate,
   MPI_GMRES.h:125-220): H [r][m][m+1] the Hessenberg matrix as the Arnoldi
   process left it, before the rotations (H(i,j) at [r][j][i]); g [r][m+1] the
   right-hand side after rotate (:733-752); mm [r] the Arnoldi steps taken.
   Any output may be NULL (call with NULLs first for m and n_restarts). */
int cfdv_reference_trace(cfdv_ctx *ctx, int32_t *m, int32_t *n_restarts, double *H, double *g, int32_t *mm);
int cfdv_update(cfdv_ctx *ctx);
int cfdv_rms(cfdv_ctx *ctx, double *urms);
/* flags [n_elem] = s1 > 0.8; ties = elements where r == 0.01 exactly left
   s1/s3 unassigned in the reference (quirk Q11) */
int cfdv_refine_flags(cfdv_ctx *ctx, uint8_t *flags, double *s1, int32_t *nflagged,
                      int32_t *nties);
/* cfdv_refine_flags with no device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The flags, s1 and the tie counter live in a fixed 32 MiB
   device array that the library holds on each GPU until the process ends;
   meshes above CFDV_FLAGS_NOALLOC_MAX_ELEM elements are rejected with
   CFDV_ERR_ARG.  Every call on one GPU uses that GPU's array, so calls that
   overlap on one GPU (several contexts or threads) race. */
#define CFDV_FLAGS_NOALLOC_MAX_ELEM 3700000
int cfdv_refine_flags_noalloc(cfdv_ctx *ctx, uint8_t *flags, double *s1, int32_t *nflagged,
                              int32_t *nties);
int cfdv_color(cfdv_ctx *ctx, int32_t *color, int32_t *ncolors);
/* y = K x (or K^T x) with the built blocks; precision 64, 32 (FP32 blocks and
   vectors) or 3 (FP32 blocks, FP64 vectors); host arrays [n_node*neqn].
   K x and K^T x use the assembled rows when the context holds them */
int cfdv_matvec(cfdv_ctx *ctx, const double *x, double *y, int precision, int transpose);
/* average time of one full K x (the rows, or EBE over all colors),
   measured with HIP events on the context stream; precision 64 or 32 */
int cfdv_profile_matvec(cfdv_ctx *ctx, int precision, int reps, double *ms);
int cfdv_step(cfdv_ctx *ctx, double cfl, const cfdv_solve_opts *opts, cfdv_step_stats *st);
/* Y = K X (or K^T X) for 8 vectors in one pass over the FP32 blocks (no
   reference counterpart): X, Y host arrays [n_node*neqn][8] (the 8 values
   of a dof contiguous); nvec must be 8; hex8 meshes.  K X and K^T X run
   over the FP32 node rows when the context holds them, else element by
   element with CFDV_BLOCKS_FP32 */
int cfdv_matvec_multi(cfdv_ctx *ctx, const float *X, float *Y, int nvec, int transpose);
/* average time of one 8-vector product K X (node rows, or EBE over all
   colors), HIP events */
int cfdv_profile_matvec_multi(cfdv_ctx *ctx, int reps, double *ms);

/* ---- discrete adjoint of one step (no reference counterpart) ----------
   For J = g^T U1 with U1 = U0 + du, K(U0, M) du = b(U0, M) at a fixed dt:
     dJ/dU0 = g + gU,  dJ/dM = gM,  where K^T lambda = g and (gU, gM) is the
   gradient of Phi = lambda^T (b - K du) from cfdv_build_vjp. */
/* replace the Thermo constants (Thermo::set, FDV/Thermo.h:22-33); the
   adiabatic flag must stay as at create; invalidates the build */
int cfdv_set_thermo(cfdv_ctx *ctx, const cfdv_thermo *thermo);
/* K^T lambda = g with the FP64 or mixed solver (opts->transpose ignored);
   host arrays [n_node*neqn] */
int cfdv_adjoint_solve(cfdv_ctx *ctx, const cfdv_solve_opts *opts, const double *g, double *lambda,
                       cfdv_solve_stats *st);
/* K^T lambda_s = g_s for 8 right-hand sides at once (no reference
   counterpart; the adjoints of 8 functionals of one step).  Mixed
   precision: per right-hand side FP64 refinement r_s = g_s - K64^T lambda_s,
   lambda_s += ||r_s|| d_s, around one FP32 global GMRES cycle per sweep
   (Jbilou, Messaoudi and Sadok, 1999) whose Krylov vectors hold the 8
   normalised residuals side by side, so that each Arnoldi step is one
   8-vector product K^T X (cfdv_matvec_multi).  G, Lambda: host arrays
   [8][n_node*neqn] (right-hand side s contiguous at s*n_node*neqn);
   rel_resid[8] receives ||g_s - K^T lambda_s|| / ||g_s||.  opts: mode
   CFDV_SOLVE_MIXED, rtol per right-hand side, inner_m (2 to 500),
   inner_rtol for the Frobenius norm of the inner residual, restarts = the
   most refinement sweeps, precond; st: summed counts, rel_resid the
   largest.  CFDV_ERR_NOCONV when a right-hand side stays above rtol.
   hex8 meshes. */
int cfdv_adjoint_solve_multi(cfdv_ctx *ctx, const cfdv_solve_opts *opts, const double *G, double *Lambda,
                             int nvec, double *rel_resid, cfdv_solve_stats *st);
/* gradient of Phi = lambda^T (b - K du) with respect to the node state U
   (gU [n_node][neqn]) and the Thermo Mach number (*gmach), at the state and
   dt of the last cfdv_build, lambda and du held fixed (du NULL: the last
   solved du).  hex8 meshes. */
int cfdv_build_vjp(cfdv_ctx *ctx, const double *lambda, const double *du, double *gU, double *gmach);

/* ---- host-side adaptive refinement (no GPU needed) ---------------------
   A cfdv_refine object holds a copy of the mesh and node states that the
   reference's AdaptiveRefineList and MeshRefine2D read and append to:
   element and node numbers are append order, a refined element keeps its
   number and appends nnod-1 children, a hanging local node keeps the node of
   the element before refinement and gets an Hnm row (Element::Hnm, adap).
   Elements without adap carry the identity as Hnm. */
typedef struct cfdv_refine cfdv_refine;
#define CFDV_REFINE_PARITY 0   /* repeated-node test on x and y only, |d| < tol
                                  (MeshRefine2D.h:337-361, quirk Q14: in 3D,
                                  nodes differing only in z merge)            */
#define CFDV_REFINE_3D 1       /* repeated-node test on every coordinate      */
/* U0 may be NULL (U0 = U); refine_level [n_elem] may be NULL (all 0); adap
   and hnm of the mesh are read when adap is set (hex8 and quad4) */
int cfdv_refine_create(const cfdv_mesh *mesh, const double *U, const double *U0,
                       const int32_t *refine_level, cfdv_refine **out);
void cfdv_refine_destroy(cfdv_refine *r);
const char *cfdv_refine_last_error(const cfdv_refine *r);
/* AdaptiveRefineList from the flags [n_elem] (s1 > 0.8, cfdv_refine_flags):
   dependency closure over face and edge neighbours, first-occurrence dedupe,
   buckets by refine level.  elelist capacity n_elem, level_ptr capacity
   (max refine level + 2); *n_level_ptr = max refine level + 2 */
int cfdv_refine_list(cfdv_refine *r, const uint8_t *flags, int32_t *elelist, int32_t *n_list,
                     int32_t *level_ptr, int32_t *n_level_ptr);
/* MeshRefine2D: refine elelist level bucket by level bucket (level_ptr as
   returned by cfdv_refine_list); mode CFDV_REFINE_*; tol <= 0 selects the
   reference's 1e-3 */
int cfdv_refine_apply(cfdv_refine *r, const int32_t *elelist, int32_t n_list, const int32_t *level_ptr,
                      int32_t n_level_ptr, int mode, double tol);
int cfdv_refine_size(const cfdv_refine *r, int64_t *n_elem, int64_t *n_node, int32_t *max_level);
/* copy the current mesh out (any pointer may be NULL): conn [n_elem][nnod],
   bc [n_elem][nface] (SaveBinary.h:59-69), x, U, U0, dirichlet, adap,
   hnm [n_elem][nnod*nnod] with Hnm(N, M) at N + nnod*M, refine_level */
int cfdv_refine_get(const cfdv_refine *r, int32_t *conn, int32_t *bc, double *x, double *U, double *U0,
                    uint8_t *dirichlet, uint8_t *adap, double *hnm, int32_t *refine_level);

#ifdef __cplusplus
}
#endif
#endif
