/*
 * Multithreaded timing variant of the oracle's full FDV step.
 *
 * TEST INFRASTRUCTURE ONLY: this is what the benchmark's cpu_baseline leg
 * times on the host cores; nothing checks results against it and nothing in
 * the product library loads it.  The per-element arithmetic is the oracle's
 * own (orc_build_element, the orc_matvec loop), spread over OpenMP threads:
 *
 *   build      over elements, static chunks.  Each thread carries its own s
 *              (FDVParam, FDVParam.h:15-111); at an r == 0.01 tie the serial
 *              build keeps the previous element's s, here the previous
 *              element of the thread's chunk (0 at a chunk's start).
 *   product    sum_Ax (MPI_GMRES.h:602-641) one element color at a time:
 *              elements of one color share no node, so their scatters into
 *              Ax never meet.
 *   GMRES      orc_gmres_true's restarted GMRES(m) with modified
 *              Gram-Schmidt, dot products and updates over rows.
 *
 * The results equal orc_step's (mode 1) up to the order of the sums.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "cfdv_oracle.h"

static void matvec_colored(const orc_mesh *m, const double *R, const double *xn, double *Ax,
                           const int *cptr, const int *celem, int ncolor)
{
    const int nq = m->neqn, nn = m->nnod, k = nn * nq;
    const size_t nd = (size_t)m->n_node * nq;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (size_t i = 0; i < nd; i++) Ax[i] = 0.0;
        for (int c = 0; c < ncolor; c++) {
#pragma omp for schedule(static)
            for (int p = cptr[c]; p < cptr[c + 1]; p++) {
                const int j = celem[p];
                const double *Re = R + (size_t)j * k * k;
                const int *cn = m->conn + (size_t)j * nn;
                for (int y = 0; y < nn; y++) {
                    const int gy = cn[y];
                    for (int x = 0; x < nn; x++) {
                        const int gx = cn[x];
                        for (int xx = 0; xx < nq; xx++)
                            for (int yy = 0; yy < nq; yy++)
                                Ax[(size_t)gx * nq + xx] +=
                                    Re[(x * nq + xx) + k * (y * nq + yy)] * xn[(size_t)gy * nq + yy];
                    }
                }
            }
        }
    }
}

static double pdot(size_t n, const double *a, const double *b)
{
    double r = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : r)
    for (size_t i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

static int gmres_par(const orc_mesh *m, const double *R, const double *b, int mdim, int maxrestart,
                     double rtol, const int *cptr, const int *celem, int ncolor, double *du,
                     orc_gmres_stats *st, int verbose)
{
    const size_t size = (size_t)m->n_node * m->neqn;
    double *r = (double *)malloc(size * sizeof(double));
    double *Ax = (double *)malloc(size * sizeof(double));
    double *v = (double *)malloc(size * (mdim + 1) * sizeof(double));
    double *h = (double *)calloc((size_t)(mdim + 1) * mdim, sizeof(double));
    double *g = (double *)calloc(mdim + 1, sizeof(double));
    double *cs = (double *)calloc(mdim, sizeof(double));
    double *sn = (double *)calloc(mdim, sizeof(double));
    double *y = (double *)calloc(mdim, sizeof(double));
    if (!r || !Ax || !v || !h || !g || !cs || !sn || !y) return -4;
#define H(i, j) h[(i) + (mdim + 1) * (j)]
    memset(st, 0, sizeof(*st));
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < size; n++) du[n] = 0.0;
    const double bnorm = sqrt(pdot(size, b, b));
    const double tol = rtol * bnorm;
    st->tol = tol;
    st->bnorm = bnorm;
    for (int rs = 0; rs < maxrestart; rs++) {
        matvec_colored(m, R, du, Ax, cptr, celem, ncolor);
        st->matvecs++;
#pragma omp parallel for schedule(static)
        for (size_t n = 0; n < size; n++) r[n] = b[n] - Ax[n];
        const double beta = sqrt(pdot(size, r, r));
        if (beta <= tol) break;
        memset(g, 0, sizeof(double) * (mdim + 1));
        memset(h, 0, sizeof(double) * (mdim + 1) * mdim);
        g[0] = beta;
#pragma omp parallel for schedule(static)
        for (size_t n = 0; n < size; n++) v[n] = r[n] / beta;
        int jdone = 0;
        for (int j = 0; j < mdim; j++) {
            double *w = v + size * (j + 1);
            matvec_colored(m, R, v + size * j, w, cptr, celem, ncolor);
            st->matvecs++;
            st->iters++;
            for (int i = 0; i <= j; i++) {   /* modified Gram-Schmidt */
                const double hij = pdot(size, w, v + size * i);
                H(i, j) = hij;
                const double *vi = v + size * i;
#pragma omp parallel for schedule(static)
                for (size_t n = 0; n < size; n++) w[n] -= hij * vi[n];
            }
            const double hn = sqrt(pdot(size, w, w));
            H(j + 1, j) = hn;
            if (hn != 0) {
#pragma omp parallel for schedule(static)
                for (size_t n = 0; n < size; n++) w[n] /= hn;
            } else {
                st->breakdown = 1;
            }
            for (int i = 0; i < j; i++) {
                const double h0 = H(i, j), h1 = H(i + 1, j);
                H(i, j) = cs[i] * h0 + sn[i] * h1;
                H(i + 1, j) = -sn[i] * h0 + cs[i] * h1;
            }
            const double mag = sqrt(H(j, j) * H(j, j) + H(j + 1, j) * H(j + 1, j));
            cs[j] = H(j, j) / mag;
            sn[j] = H(j + 1, j) / mag;
            H(j, j) = mag;
            H(j + 1, j) = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jdone = j + 1;
            if (fabs(g[j + 1]) <= tol || st->breakdown) break;
        }
        for (int j = jdone - 1; j >= 0; j--) {
            y[j] = g[j];
            for (int i = j + 1; i < jdone; i++) y[j] -= H(j, i) * y[i];
            y[j] /= H(j, j);
        }
#pragma omp parallel for schedule(static)
        for (size_t n = 0; n < size; n++) {
            double acc = du[n];
            for (int i = 0; i < jdone; i++) acc += y[i] * v[size * i + n];
            du[n] = acc;
        }
        st->restarts++;
        st->last_resid = fabs(g[jdone]);
        if (verbose) {
            fprintf(stderr, "oracle step: GMRES restart %d, %d products, estimated rel resid %.2e\n",
                    st->restarts, st->matvecs, st->last_resid / bnorm);
            fflush(stderr);
        }
        if (st->breakdown) break;
    }
    matvec_colored(m, R, du, Ax, cptr, celem, ncolor);
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < size; n++) r[n] = b[n] - Ax[n];
    st->true_resid = sqrt(pdot(size, r, r));
#undef H
    free(r); free(Ax); free(v); free(h); free(g); free(cs); free(sn); free(y);
    return 0;
}

int orc_step_par(const orc_thermo *th, const orc_mesh *m, double *U, double *U0, const orc_step_opts *o,
                 const int *cptr, const int *celem, int ncolor, int nthreads, orc_step_stats *st,
                 double *t_phase, int verbose)
{
    const int k = m->nnod * m->neqn;
    const size_t nd = (size_t)m->n_node * m->neqn;
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
    double t0 = omp_get_wtime();
    orc_prim *prim = (orc_prim *)malloc(sizeof(orc_prim) * m->n_node);
    double *area = (double *)malloc(sizeof(double) * m->n_elem);
    double *cleng = (double *)malloc(sizeof(double) * m->n_elem);
    double *R = (double *)malloc(sizeof(double) * (size_t)m->n_elem * k * k);
    double *rhs = (double *)malloc(sizeof(double) * (size_t)m->n_elem * k);
    double *du = (double *)malloc(sizeof(double) * nd);
    double *b = (double *)malloc(sizeof(double) * nd);
    if (!prim || !area || !cleng || !R || !rhs || !du || !b) return -4;
    memset(st, 0, sizeof(*st));
    orc_calc_length(m, area, cleng);
    orc_node_unpack_all(th, m->ndim, m->neqn, m->n_node, U, prim);
    int rc = orc_cfl(th, m, prim, cleng, o->cfl, &st->dt);
    double t1 = omp_get_wtime(), t2 = t1, t3 = t1;
    if (rc == 0) {
        int bad = 0;
#pragma omp parallel
        {
            double s[4] = {0, 0, 0, 0};
            const int tid = omp_get_thread_num(), nth = omp_get_num_threads();
            const int chunk = (m->n_elem + nth - 1) / nth, step = chunk / 8 > 0 ? chunk / 8 : 1;
#pragma omp for schedule(static) reduction(| : bad)
            for (int e = 0; e < m->n_elem; e++) {
                double *Re = R + (size_t)e * k * k, *be = rhs + (size_t)e * k;
                memset(Re, 0, sizeof(double) * (size_t)k * k);
                memset(be, 0, sizeof(double) * k);
                if (orc_build_element(th, m, U, prim, e, st->dt, s, Re, be) < 0) bad |= 1;
                if (verbose && tid == 0 && (e + 1) % step == 0) {
                    fprintf(stderr, "oracle step: build, thread 0 at %d of its %d elements\n", e + 1, chunk);
                    fflush(stderr);
                }
            }
        }
        if (bad) rc = -1;
        t2 = omp_get_wtime();
        if (rc == 0) {
            orc_apply_bc(th, m, R, rhs);
            orc_compose_rhs(m, rhs, b);
            t3 = omp_get_wtime();
            rc = gmres_par(m, R, b, o->mdim, o->maxrestart, o->rtol, cptr, celem, ncolor, du, &st->gm, verbose);
            if (rc == 0) {
                orc_update((int)nd, U, U0, du);
                orc_rms(m->n_node, m->neqn, U, U0, st->urms);
            }
        }
    }
    double t4 = omp_get_wtime();
    if (t_phase) {
        t_phase[0] = t1 - t0;   /* lengths, unpack, CFL */
        t_phase[1] = t2 - t1;   /* element build */
        t_phase[2] = t3 - t2;   /* ApplyBC, compose rhs */
        t_phase[3] = t4 - t3;   /* GMRES, update */
    }
    free(prim); free(area); free(cleng); free(R); free(rhs); free(du); free(b);
    return rc < 0 ? rc : 0;
}
