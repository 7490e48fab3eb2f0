/*
 * nr_oracle.c -- CPU FP64 Newton-Raphson AC power flow: the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against (tests/, the smoke check and bench.py's CPU-baseline leg).  Nothing
 * in the gpusim2grid_amd package links, loads or calls it.
 *
 * It restates, independently of the GPU code, the algorithm of gpusim2grid's
 * bare [pvpq | pq] Newton-Raphson:
 *   - Jacobian layout: rows/cols = [P of sorted pvpq | Q of pq in input order],
 *     each row's columns ascending, one J coefficient per (Ybus entry, block)
 *     (src/_cpp/acpf_nr.cu:145-236, build_J_structure);
 *   - mismatch F = -[Re(V conj(Ybus V) - Sbus)](pvpq) ; -[Im(...)](pq)
 *     (src/_cpp/acpf_nr_kernels.cu:493-554, fill_FP/fill_FQ);
 *   - Jacobian values dS/dVa, dS/dVm per Ybus entry
 *     (src/_cpp/acpf_nr_kernels.cu:559-654, fill_J_kernel);
 *   - polar update, theta first then |V| (acpf_nr_kernels.cu:728-788), with
 *     the optional per-row step scaling (acpf_nr_kernels.cu:677-723);
 *   - convergence: strict ||F||inf < tol in per unit, NaN-sticky
 *     (acpf_nr_kernels.cu:1076-1128; contingency/row_status.hpp:28-50);
 *     a row's iteration count is the number of Newton updates applied before
 *     ||F(V)||inf < tol held (lightsim2grid's count, the reference's continuous
 *     schedule at k=1, contingency/batch_pf_driver.cu:2895-2913).
 *
 * The linear solver is its own: a minimum-degree column order on A+A^T and a
 * left-looking (Gilbert-Peierls) sparse LU with threshold partial pivoting
 * that prefers the diagonal, KLU's method; optionally (the CPU baseline's
 * setting) a row's later iterations refactorize numerically on its first
 * factorization's pivots and pattern, as KLU's klu_refactor does.  Rows of a
 * batch run on POSIX threads.
 *
 * Status codes: 0 Converged, 1 MaxIter, 2 Diverged (row_status.hpp:33-43).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* Jacobian pattern                                                           */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n, nY, npvpq, npq, d, nJ;
    const int *Yp, *Yi;          /* Ybus CSR (sorted columns) */
    const double *Yx;            /* interleaved re,im [2 nY] */
    int *pvpq;                   /* sorted */
    const int *pq;
    int *row_p, *row_q;          /* per bus: J row of its P / Q equation, -1 if none */
    int *Jp, *Ji;                /* J CSR */
    int *map11, *map12, *map21, *map22; /* per Ybus entry: J value position or -1 */
} orc_sys;

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Builds the J CSR for the [P(pvpq sorted) | Q(pq)] layout. Returns nJ. */
static int orc_sys_init(orc_sys *s, int n, const int *Yp, const int *Yi, const double *Yx,
                        int npv, const int *pv, int npq, const int *pq)
{
    memset(s, 0, sizeof(*s));
    s->n = n; s->Yp = Yp; s->Yi = Yi; s->Yx = Yx; s->nY = Yp[n];
    s->npvpq = npv + npq; s->npq = npq; s->pq = pq;
    s->d = s->npvpq + npq;
    s->pvpq = (int *)malloc(sizeof(int) * (s->npvpq + 1));
    memcpy(s->pvpq, pv, sizeof(int) * npv);
    memcpy(s->pvpq + npv, pq, sizeof(int) * npq);
    qsort(s->pvpq, s->npvpq, sizeof(int), cmp_int);
    s->row_p = (int *)malloc(sizeof(int) * n);
    s->row_q = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) s->row_p[i] = s->row_q[i] = -1;
    for (int k = 0; k < s->npvpq; ++k) s->row_p[s->pvpq[k]] = k;
    for (int k = 0; k < npq; ++k) s->row_q[pq[k]] = s->npvpq + k;
    /* the variable layout equals the equation layout: theta col of bus = row_p,
     * |V| col of bus = row_q */
    const int d = s->d;
    int *cnt = (int *)calloc(d + 1, sizeof(int));
    for (int i = 0; i < n; ++i) {
        int rows[2] = {s->row_p[i], s->row_q[i]};
        for (int a = 0; a < 2; ++a) {
            if (rows[a] < 0) continue;
            for (int k = Yp[i]; k < Yp[i + 1]; ++k) {
                int j = Yi[k];
                cnt[rows[a] + 1] += (s->row_p[j] >= 0) + (s->row_q[j] >= 0);
            }
        }
    }
    for (int r = 0; r < d; ++r) cnt[r + 1] += cnt[r];
    s->nJ = cnt[d];
    s->Jp = (int *)malloc(sizeof(int) * (d + 1));
    memcpy(s->Jp, cnt, sizeof(int) * (d + 1));
    s->Ji = (int *)malloc(sizeof(int) * (s->nJ + 1));
    s->map11 = (int *)malloc(sizeof(int) * s->nY);
    s->map12 = (int *)malloc(sizeof(int) * s->nY);
    s->map21 = (int *)malloc(sizeof(int) * s->nY);
    s->map22 = (int *)malloc(sizeof(int) * s->nY);
    for (int k = 0; k < s->nY; ++k) s->map11[k] = s->map12[k] = s->map21[k] = s->map22[k] = -1;
    /* fill each row with its (column, ybus k, block) triples, then sort by column */
    int *head = (int *)malloc(sizeof(int) * d);
    memcpy(head, cnt, sizeof(int) * d);
    int *tag = (int *)malloc(sizeof(int) * (s->nJ + 1)); /* ybus k * 4 + block */
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 2; ++a) {
            int r = a == 0 ? s->row_p[i] : s->row_q[i];
            if (r < 0) continue;
            for (int k = Yp[i]; k < Yp[i + 1]; ++k) {
                int j = Yi[k];
                if (s->row_p[j] >= 0) { s->Ji[head[r]] = s->row_p[j]; tag[head[r]++] = 4 * k + (a == 0 ? 0 : 2); }
                if (s->row_q[j] >= 0) { s->Ji[head[r]] = s->row_q[j]; tag[head[r]++] = 4 * k + (a == 0 ? 1 : 3); }
            }
        }
    }
    /* insertion sort per row (rows are short) */
    for (int r = 0; r < d; ++r) {
        for (int p = s->Jp[r] + 1; p < s->Jp[r + 1]; ++p) {
            int c = s->Ji[p], t = tag[p], q = p - 1;
            while (q >= s->Jp[r] && s->Ji[q] > c) { s->Ji[q + 1] = s->Ji[q]; tag[q + 1] = tag[q]; --q; }
            s->Ji[q + 1] = c; tag[q + 1] = t;
        }
        for (int p = s->Jp[r]; p < s->Jp[r + 1]; ++p) {
            int k = tag[p] >> 2, blk = tag[p] & 3;
            int *m = blk == 0 ? s->map11 : blk == 1 ? s->map12 : blk == 2 ? s->map21 : s->map22;
            m[k] = p;
        }
    }
    free(cnt); free(head); free(tag);
    return s->nJ;
}

static void orc_sys_free(orc_sys *s)
{
    free(s->pvpq); free(s->row_p); free(s->row_q); free(s->Jp); free(s->Ji);
    free(s->map11); free(s->map12); free(s->map21); free(s->map22);
}

/* Ibus = Ybus V ; F ; returns ||F||inf (NaN-sticky) */
static double orc_mismatch(const orc_sys *s, const double *V, const double *S, double *I, double *F)
{
    const int n = s->n;
    for (int i = 0; i < n; ++i) {
        double re = 0.0, im = 0.0;
        for (int k = s->Yp[i]; k < s->Yp[i + 1]; ++k) {
            const int j = s->Yi[k];
            const double yr = s->Yx[2 * k], yi = s->Yx[2 * k + 1];
            const double vr = V[2 * j], vi = V[2 * j + 1];
            re += yr * vr - yi * vi;
            im += yr * vi + yi * vr;
        }
        I[2 * i] = re; I[2 * i + 1] = im;
    }
    double nrm = 0.0;
    int nan = 0;
    for (int i = 0; i < n; ++i) {
        const double vr = V[2 * i], vi = V[2 * i + 1], ir = I[2 * i], ii = I[2 * i + 1];
        /* S_calc = V conj(I) */
        const double sr = vr * ir + vi * ii, si = vi * ir - vr * ii;
        if (s->row_p[i] >= 0) {
            double f = -(sr - S[2 * i]);
            F[s->row_p[i]] = f;
            if (isnan(f)) nan = 1; else if (fabs(f) > nrm) nrm = fabs(f);
        }
        if (s->row_q[i] >= 0) {
            double f = -(si - S[2 * i + 1]);
            F[s->row_q[i]] = f;
            if (isnan(f)) nan = 1; else if (fabs(f) > nrm) nrm = fabs(f);
        }
    }
    return nan ? NAN : nrm;
}

/* J values in the CSR order of s->Jp/Ji */
static void orc_jacobian(const orc_sys *s, const double *V, const double *I, double *Jx)
{
    for (int i = 0; i < s->n; ++i) {
        const double vir = V[2 * i], vii = V[2 * i + 1];
        for (int k = s->Yp[i]; k < s->Yp[i + 1]; ++k) {
            const int j = s->Yi[k];
            const double yr = s->Yx[2 * k], yi = s->Yx[2 * k + 1];
            double va_r, va_i, vm_r, vm_i; /* dS/dVa, dS/dVm */
            if (i == j) {
                /* dS/dVa = j V_i conj(I_i - Y_ii V_i) */
                const double tr = s->Yx[2 * k] * vir - s->Yx[2 * k + 1] * vii;
                const double ti = s->Yx[2 * k] * vii + s->Yx[2 * k + 1] * vir;
                const double cr = I[2 * i] - tr, ci = -(I[2 * i + 1] - ti); /* conj */
                const double pr = vir * cr - vii * ci, pi = vir * ci + vii * cr;
                va_r = -pi; va_i = pr;
                /* dS/dVm = V_i conj(Y_ii V_i/|V_i|) + conj(I_i) V_i/|V_i| */
                const double m = hypot(vir, vii);
                const double nr = vir / m, ni = vii / m;
                const double ur = yr * nr - yi * ni, ui = yr * ni + yi * nr;
                const double t1r = vir * ur + vii * ui, t1i = vii * ur - vir * ui;
                const double t2r = I[2 * i] * nr + I[2 * i + 1] * ni;
                const double t2i = I[2 * i] * ni - I[2 * i + 1] * nr;
                vm_r = t1r + t2r; vm_i = t1i + t2i;
            } else {
                const double vjr = V[2 * j], vji = V[2 * j + 1];
                /* dS/dVa = -j V_i conj(Y_ij V_j) */
                const double ur = yr * vjr - yi * vji, ui = yr * vji + yi * vjr;
                const double pr = vir * ur + vii * ui, pi = vii * ur - vir * ui; /* V_i conj(u) */
                va_r = pi; va_i = -pr;
                /* dS/dVm = V_i conj(Y_ij V_j/|V_j|) */
                const double m = hypot(vjr, vji);
                vm_r = pr / m; vm_i = pi / m;
            }
            if (s->map11[k] >= 0) Jx[s->map11[k]] = va_r;
            if (s->map12[k] >= 0) Jx[s->map12[k]] = vm_r;
            if (s->map21[k] >= 0) Jx[s->map21[k]] = va_i;
            if (s->map22[k] >= 0) Jx[s->map22[k]] = vm_i;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Sparse LU: minimum-degree order + left-looking LU with partial pivoting     */
/* ------------------------------------------------------------------------- */

/* Minimum degree on the graph of A + A^T (A given by CSR rows = columns of the
 * transposed problem; the pattern is symmetrised first).  Plain elimination
 * graph with explicit adjacency sets; fine for the low-fill power-flow graphs. */
static void orc_min_degree(int d, const int *Ap, const int *Ai, int *perm)
{
    int *deg = (int *)calloc(d, sizeof(int));
    int **adj = (int **)malloc(sizeof(int *) * d);
    int *cap = (int *)malloc(sizeof(int) * d);
    for (int i = 0; i < d; ++i) { cap[i] = 8; adj[i] = (int *)malloc(sizeof(int) * cap[i]); }
    char *mark = (char *)calloc(d, 1);
#define ADD_EDGE(a, b)                                                                 \
    do {                                                                               \
        int _a = (a), _b = (b), _f = 0;                                                \
        for (int _t = 0; _t < deg[_a]; ++_t) if (adj[_a][_t] == _b) { _f = 1; break; } \
        if (!_f) {                                                                     \
            if (deg[_a] == cap[_a]) { cap[_a] *= 2; adj[_a] = (int *)realloc(adj[_a], sizeof(int) * cap[_a]); } \
            adj[_a][deg[_a]++] = _b;                                                   \
        }                                                                              \
    } while (0)
    for (int i = 0; i < d; ++i)
        for (int p = Ap[i]; p < Ap[i + 1]; ++p) {
            int j = Ai[p];
            if (j == i) continue;
            ADD_EDGE(i, j);
            ADD_EDGE(j, i);
        }
    char *done = (char *)calloc(d, 1);
    for (int step = 0; step < d; ++step) {
        int best = -1;
        for (int i = 0; i < d; ++i)
            if (!done[i] && (best < 0 || deg[i] < deg[best])) best = i;
        perm[step] = best;
        done[best] = 1;
        /* neighbours of best form a clique */
        int nb = 0;
        int *nbr = adj[best];
        for (int t = 0; t < deg[best]; ++t) if (!done[nbr[t]]) nbr[nb++] = nbr[t];
        for (int a = 0; a < nb; ++a) {
            int u = nbr[a];
            /* drop best from u */
            int w = 0;
            for (int t = 0; t < deg[u]; ++t) if (adj[u][t] != best) adj[u][w++] = adj[u][t];
            deg[u] = w;
            for (int t = 0; t < deg[u]; ++t) mark[adj[u][t]] = 1;
            for (int b = 0; b < nb; ++b) {
                int v = nbr[b];
                if (v != u && !mark[v]) {
                    if (deg[u] == cap[u]) { cap[u] *= 2; adj[u] = (int *)realloc(adj[u], sizeof(int) * cap[u]); }
                    adj[u][deg[u]++] = v;
                    mark[v] = 1;
                }
            }
            for (int t = 0; t < deg[u]; ++t) mark[adj[u][t]] = 0;
        }
        deg[best] = 0;
    }
#undef ADD_EDGE
    for (int i = 0; i < d; ++i) free(adj[i]);
    free(adj); free(cap); free(deg); free(mark); free(done);
}

typedef struct {
    int d;
    int *q;          /* column order */
    int *pinv;       /* row i of A is pivot row pinv[i] */
    /* L: unit lower, columns; U: upper, columns (diag last per column) */
    int *Lp, *Li; double *Lx; int Lcap;
    int *Up, *Ui; double *Ux; int Ucap;
    /* workspace */
    double *x; int *xi; int *stack; int *pstack; char *visited;
    /* A in CSC (transposed from the CSR input) */
    int *Ap, *Ai; int *Amap; /* Amap[csc pos] = csr pos */
} orc_lu;

static void orc_lu_init(orc_lu *f, int d, const int *Jp, const int *Ji)
{
    memset(f, 0, sizeof(*f));
    f->d = d;
    const int nnz = Jp[d];
    f->Ap = (int *)calloc(d + 1, sizeof(int));
    f->Ai = (int *)malloc(sizeof(int) * (nnz + 1));
    f->Amap = (int *)malloc(sizeof(int) * (nnz + 1));
    for (int p = 0; p < nnz; ++p) f->Ap[Ji[p] + 1]++;
    for (int c = 0; c < d; ++c) f->Ap[c + 1] += f->Ap[c];
    int *head = (int *)malloc(sizeof(int) * d);
    memcpy(head, f->Ap, sizeof(int) * d);
    for (int r = 0; r < d; ++r)
        for (int p = Jp[r]; p < Jp[r + 1]; ++p) {
            int c = Ji[p];
            f->Ai[head[c]] = r;
            f->Amap[head[c]++] = p;
        }
    free(head);
    f->q = (int *)malloc(sizeof(int) * d);
    orc_min_degree(d, Jp, Ji, f->q);
    f->pinv = (int *)malloc(sizeof(int) * d);
    f->Lcap = f->Ucap = 4 * nnz + d;
    f->Lp = (int *)malloc(sizeof(int) * (d + 1));
    f->Li = (int *)malloc(sizeof(int) * f->Lcap);
    f->Lx = (double *)malloc(sizeof(double) * f->Lcap);
    f->Up = (int *)malloc(sizeof(int) * (d + 1));
    f->Ui = (int *)malloc(sizeof(int) * f->Ucap);
    f->Ux = (double *)malloc(sizeof(double) * f->Ucap);
    f->x = (double *)calloc(d, sizeof(double));
    f->xi = (int *)malloc(sizeof(int) * d);
    f->stack = (int *)malloc(sizeof(int) * d);
    f->pstack = (int *)malloc(sizeof(int) * d);
    f->visited = (char *)calloc(d, 1);
}

static void orc_lu_free(orc_lu *f)
{
    free(f->Ap); free(f->Ai); free(f->Amap); free(f->q); free(f->pinv);
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux);
    free(f->x); free(f->xi); free(f->stack); free(f->pstack); free(f->visited);
}

/* reach of column j's pattern through L (rows already pivoted), topological order
 * into xi[top..d-1]; returns top */
static int orc_reach(orc_lu *f, int col)
{
    const int d = f->d;
    int top = d;
    for (int p = f->Ap[col]; p < f->Ap[col + 1]; ++p) {
        int i = f->Ai[p];
        if (f->visited[i]) continue;
        /* DFS from row i */
        int head = 0;
        f->stack[0] = i;
        f->pstack[0] = -2;
        while (head >= 0) {
            int r = f->stack[head];
            int k = f->pinv[r];          /* pivot column of row r, -1 if not yet pivoted */
            if (f->pstack[head] == -2) {
                f->visited[r] = 1;
                f->pstack[head] = k < 0 ? 0 : f->Lp[k] + 1; /* skip L diag */
            }
            int done = 1;
            if (k >= 0) {
                int end = f->Lp[k + 1];
                while (f->pstack[head] < end) {
                    int r2 = f->Li[f->pstack[head]++];
                    if (!f->visited[r2]) {
                        f->stack[++head] = r2;
                        f->pstack[head] = -2;
                        done = 0;
                        break;
                    }
                }
            }
            if (done) {
                --head;
                f->xi[--top] = r;
            }
        }
    }
    for (int p = top; p < d; ++p) f->visited[f->xi[p]] = 0;
    return top;
}

/* Factorizes A (CSR values Jx in the pattern given at init). 0 on success. */
static int orc_lu_factor(orc_lu *f, const double *Jx)
{
    const int d = f->d;
    int lnz = 0, unz = 0;
    for (int i = 0; i < d; ++i) f->pinv[i] = -1;
    for (int k = 0; k < d; ++k) {
        f->Lp[k] = lnz;
        f->Up[k] = unz;
        const int col = f->q[k];
        if (lnz + d > f->Lcap) { f->Lcap = 2 * f->Lcap + d; f->Li = (int *)realloc(f->Li, sizeof(int) * f->Lcap); f->Lx = (double *)realloc(f->Lx, sizeof(double) * f->Lcap); }
        if (unz + d > f->Ucap) { f->Ucap = 2 * f->Ucap + d; f->Ui = (int *)realloc(f->Ui, sizeof(int) * f->Ucap); f->Ux = (double *)realloc(f->Ux, sizeof(double) * f->Ucap); }
        int top = orc_reach(f, col);
        for (int p = top; p < d; ++p) f->x[f->xi[p]] = 0.0;
        for (int p = f->Ap[col]; p < f->Ap[col + 1]; ++p) f->x[f->Ai[p]] = Jx[f->Amap[p]];
        /* triangular solve with the pivoted part of L */
        for (int p = top; p < d; ++p) {
            int r = f->xi[p];
            int kk = f->pinv[r];
            if (kk < 0) continue;
            double xr = f->x[r];
            f->Ui[unz] = kk; f->Ux[unz++] = xr;
            for (int t = f->Lp[kk] + 1; t < f->Lp[kk + 1]; ++t) f->x[f->Li[t]] -= f->Lx[t] * xr;
        }
        /* pivot: prefer the diagonal row (row index == col) when within threshold */
        int ipiv = -1;
        double amax = -1.0;
        for (int p = top; p < d; ++p) {
            int r = f->xi[p];
            if (f->pinv[r] < 0 && fabs(f->x[r]) > amax) { amax = fabs(f->x[r]); ipiv = r; }
        }
        if (ipiv < 0 || amax == 0.0 || !isfinite(amax)) return -1;
        if (f->pinv[col] < 0 && fabs(f->x[col]) >= 1e-3 * amax) ipiv = col;
        double piv = f->x[ipiv];
        f->Ui[unz] = k; f->Ux[unz++] = piv;
        f->pinv[ipiv] = k;
        f->Li[lnz] = ipiv; f->Lx[lnz++] = 1.0;
        for (int p = top; p < d; ++p) {
            int r = f->xi[p];
            if (f->pinv[r] < 0) { f->Li[lnz] = r; f->Lx[lnz++] = f->x[r] / piv; }
            f->x[r] = 0.0;
        }
    }
    f->Lp[d] = lnz;
    f->Up[d] = unz;
    /* L row indices -> pivot indices */
    for (int p = 0; p < lnz; ++p) f->Li[p] = f->pinv[f->Li[p]];
    return 0;
}

/* Numeric refactorization on the last factorization's pivots and L / U
 * pattern (KLU's klu_refactor: no reach, no pivot search).  The U entries
 * of a column are stored in the reach's topological order, so applying
 * them in stored order is the left-looking triangular solve.  Returns -1
 * (the caller factorizes afresh) when a pivot is zero, non-finite, or under
 * 1e-3 of its column's largest L entry, the full factorization's threshold. */
static int orc_lu_refactor(orc_lu *f, const double *Jx)
{
    const int d = f->d;
    double *x = f->x;  /* indexed by pivot position, all zero between columns */
    for (int k = 0; k < d; ++k) {
        const int col = f->q[k];
        for (int p = f->Ap[col]; p < f->Ap[col + 1]; ++p) x[f->pinv[f->Ai[p]]] = Jx[f->Amap[p]];
        const int pd = f->Up[k + 1] - 1;  /* the diagonal, last */
        for (int t = f->Up[k]; t < pd; ++t) {
            const int kk = f->Ui[t];
            const double xr = x[kk];
            x[kk] = 0.0;
            f->Ux[t] = xr;
            for (int t2 = f->Lp[kk] + 1; t2 < f->Lp[kk + 1]; ++t2) x[f->Li[t2]] -= f->Lx[t2] * xr;
        }
        const double piv = x[k];
        x[k] = 0.0;
        double lmax = 0.0;
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; ++t) lmax = fmax(lmax, fabs(x[f->Li[t]]));
        if (!(fabs(piv) > 0.0) || !isfinite(piv) || fabs(piv) < 1e-3 * lmax) {
            for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; ++t) x[f->Li[t]] = 0.0;
            for (int kk = k + 1; kk < d; ++kk) x[kk] = 0.0;
            return -1;
        }
        f->Ux[pd] = piv;
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; ++t) {
            const int r = f->Li[t];
            f->Lx[t] = x[r] / piv;
            x[r] = 0.0;
        }
    }
    return 0;
}

/* x = A^-1 b (x, b may alias) */
static void orc_lu_solve(const orc_lu *f, const double *b, double *x, double *w)
{
    const int d = f->d;
    for (int i = 0; i < d; ++i) w[f->pinv[i]] = b[i];
    for (int k = 0; k < d; ++k) {
        double v = w[k];
        for (int t = f->Lp[k] + 1; t < f->Lp[k + 1]; ++t) w[f->Li[t]] -= f->Lx[t] * v;
    }
    for (int k = d - 1; k >= 0; --k) {
        int pd = f->Up[k + 1] - 1;       /* diagonal is last */
        w[k] /= f->Ux[pd];
        double v = w[k];
        for (int t = f->Up[k]; t < pd; ++t) w[f->Ui[t]] -= f->Ux[t] * v;
    }
    for (int k = 0; k < d; ++k) x[f->q[k]] = w[k];
}

/* ------------------------------------------------------------------------- */
/* Newton-Raphson                                                             */
/* ------------------------------------------------------------------------- */

typedef struct {
    orc_sys s;
    orc_lu lu;
    double *I, *F, *Jx, *dx, *w;
    int refactor;  /* iterations after a row's first refactor on its pivots */
} orc_solver;

static void orc_solver_init(orc_solver *o, int n, const int *Yp, const int *Yi, const double *Yx,
                            int npv, const int *pv, int npq, const int *pq)
{
    orc_sys_init(&o->s, n, Yp, Yi, Yx, npv, pv, npq, pq);
    orc_lu_init(&o->lu, o->s.d, o->s.Jp, o->s.Ji);
    o->I = (double *)malloc(sizeof(double) * 2 * n);
    o->F = (double *)malloc(sizeof(double) * (o->s.d + 1));
    o->Jx = (double *)malloc(sizeof(double) * (o->s.nJ + 1));
    o->dx = (double *)malloc(sizeof(double) * (o->s.d + 1));
    o->w = (double *)malloc(sizeof(double) * (o->s.d + 1));
}

static void orc_solver_free(orc_solver *o)
{
    orc_lu_free(&o->lu);
    orc_sys_free(&o->s);
    free(o->I); free(o->F); free(o->Jx); free(o->dx); free(o->w);
}

/* Step scaling (acpf_nr_kernels.cu:677-723): dx *= alpha, alpha = min(1,
 * max_dva / max|dtheta|, max_dvm / max|dVm|), maxima NaN-sticky (a NaN
 * maximum leaves alpha at 1).  max_dva <= 0: off. */
static void orc_scale_step(const orc_sys *s, double *dx, double max_dva, double max_dvm)
{
    if (!(max_dva > 0.0)) return;
    double mt = 0.0, mv = 0.0;
    for (int k = 0; k < s->npvpq; ++k) {
        const double a = fabs(dx[k]);
        if (!(mt >= a)) mt = a;
    }
    for (int k = 0; k < s->npq; ++k) {
        const double a = fabs(dx[s->npvpq + k]);
        if (!(mv >= a)) mv = a;
    }
    double alpha = 1.0;
    if (mt > max_dva && max_dva / mt < alpha) alpha = max_dva / mt;
    if (mv > max_dvm && max_dvm / mv < alpha) alpha = max_dvm / mv;
    if (alpha == 1.0) return;
    for (int k = 0; k < s->d; ++k) dx[k] *= alpha;
}

/* One row: V (in/out, interleaved complex), S. Returns status. */
static int orc_solver_run(orc_solver *o, const double *Yx, double *V, const double *S, int max_iter,
                          double tol, double max_dva, double max_dvm, int *iters, double *res)
{
    orc_sys *s = &o->s;
    s->Yx = Yx;
    int it = 0;
    for (;;) {
        double r = orc_mismatch(s, V, S, o->I, o->F);
        *res = r;
        *iters = it;
        if (!isfinite(r)) return 2;
        if (r < tol) return 0;
        if (it >= max_iter) return 1;
        orc_jacobian(s, V, o->I, o->Jx);
        if (!(o->refactor && it > 0 && orc_lu_refactor(&o->lu, o->Jx) == 0))
            if (orc_lu_factor(&o->lu, o->Jx) != 0) { *res = NAN; return 2; }
        orc_lu_solve(&o->lu, o->F, o->dx, o->w);
        orc_scale_step(s, o->dx, max_dva, max_dvm);
        /* theta of pvpq buses, then |V| of pq buses */
        for (int k = 0; k < s->npvpq; ++k) {
            const int b = s->pvpq[k];
            const double vm = hypot(V[2 * b], V[2 * b + 1]);
            const double va = atan2(V[2 * b + 1], V[2 * b]) + o->dx[k];
            V[2 * b] = vm * cos(va);
            V[2 * b + 1] = vm * sin(va);
        }
        for (int k = 0; k < s->npq; ++k) {
            const int b = s->pq[k];
            const double vm = hypot(V[2 * b], V[2 * b + 1]) + o->dx[s->npvpq + k];
            const double va = atan2(V[2 * b + 1], V[2 * b]);
            V[2 * b] = vm * cos(va);
            V[2 * b + 1] = vm * sin(va);
        }
        ++it;
    }
}

/* ------------------------------------------------------------------------- */
/* Exported entry points (ctypes)                                             */
/* ------------------------------------------------------------------------- */

/* J pattern in the reference layout. Jp [d+1], Ji [nJ], maps [4 nY] (11,12,21,22). */
int orc_jpattern_size(int n, const int *Yp, const int *Yi, int npv, const int *pv, int npq,
                      const int *pq, int *d_out)
{
    orc_sys s;
    int nJ = orc_sys_init(&s, n, Yp, Yi, NULL, npv, pv, npq, pq);
    *d_out = s.d;
    orc_sys_free(&s);
    return nJ;
}

void orc_jpattern(int n, const int *Yp, const int *Yi, int npv, const int *pv, int npq,
                  const int *pq, int *Jp, int *Ji, int *maps)
{
    orc_sys s;
    orc_sys_init(&s, n, Yp, Yi, NULL, npv, pv, npq, pq);
    memcpy(Jp, s.Jp, sizeof(int) * (s.d + 1));
    memcpy(Ji, s.Ji, sizeof(int) * s.nJ);
    memcpy(maps, s.map11, sizeof(int) * s.nY);
    memcpy(maps + s.nY, s.map12, sizeof(int) * s.nY);
    memcpy(maps + 2 * s.nY, s.map21, sizeof(int) * s.nY);
    memcpy(maps + 3 * s.nY, s.map22, sizeof(int) * s.nY);
    orc_sys_free(&s);
}

/* F (reference layout) and J values (CSR order) at V; returns ||F||inf */
double orc_f_and_j(int n, const int *Yp, const int *Yi, const double *Yx, int npv, const int *pv,
                   int npq, const int *pq, const double *V, const double *S, double *F, double *Jx)
{
    orc_sys s;
    orc_sys_init(&s, n, Yp, Yi, Yx, npv, pv, npq, pq);
    double *I = (double *)malloc(sizeof(double) * 2 * n);
    double r = orc_mismatch(&s, V, S, I, F);
    if (Jx) orc_jacobian(&s, V, I, Jx);
    free(I);
    orc_sys_free(&s);
    return r;
}

/* x = A^-1 b for a CSR matrix (oracle LU). Returns 0 on success. */
int orc_sparse_solve(int d, const int *Ap, const int *Ai, const double *Ax, const double *b, double *x)
{
    orc_lu f;
    orc_lu_init(&f, d, Ap, Ai);
    int rc = orc_lu_factor(&f, Ax);
    if (rc == 0) {
        double *w = (double *)malloc(sizeof(double) * (d + 1));
        orc_lu_solve(&f, b, x, w);
        free(w);
    }
    orc_lu_free(&f);
    return rc;
}

/* Factorizes A1 (CSR Ap/Ai/Ax1), refactorizes on its pivots with A2's
 * values (same pattern) and solves A2 x = b.  Returns the refactorization's
 * code (-1: a weak pivot, x untouched) or the first factorization's. */
int orc_sparse_refactor_solve(int d, const int *Ap, const int *Ai, const double *Ax1, const double *Ax2,
                              const double *b, double *x)
{
    orc_lu f;
    orc_lu_init(&f, d, Ap, Ai);
    int rc = orc_lu_factor(&f, Ax1);
    if (rc == 0) rc = orc_lu_refactor(&f, Ax2);
    if (rc == 0) {
        double *w = (double *)malloc(sizeof(double) * (d + 1));
        orc_lu_solve(&f, b, x, w);
        free(w);
    }
    orc_lu_free(&f);
    return rc;
}

typedef struct {
    int n; const int *Yp, *Yi; const double *Yx; long long y_stride; /* 0: shared */
    int npv; const int *pv; int npq; const int *pq;
    double *V; long long v_stride; const double *S; long long s_stride;
    int max_iter; double tol; double max_dva, max_dvm;
    int *status; int *iters; double *res;
    int row0, row1;
    int refactor;
} orc_job;

static void *orc_worker(void *arg)
{
    orc_job *j = (orc_job *)arg;
    orc_solver o;
    orc_solver_init(&o, j->n, j->Yp, j->Yi, j->Yx, j->npv, j->pv, j->npq, j->pq);
    o.refactor = j->refactor;
    for (int r = j->row0; r < j->row1; ++r) {
        const double *Yx = j->Yx + (size_t)r * (size_t)j->y_stride;
        j->status[r] = orc_solver_run(&o, Yx, j->V + (size_t)r * (size_t)j->v_stride,
                                      j->S + (size_t)r * (size_t)j->s_stride, j->max_iter, j->tol,
                                      j->max_dva, j->max_dvm, &j->iters[r], &j->res[r]);
    }
    orc_solver_free(&o);
    return NULL;
}

/* Solves n_rows power flows. V: [rows, n] complex in/out (interleaved re,im);
 * S: [rows, n] complex, or shared when s_stride == 0; Yx: [rows, nY] complex,
 * or shared when y_stride == 0 (strides counted in complex numbers).
 * Runs on n_threads POSIX threads.  max_dva > 0 turns on step scaling
 * with bounds max_dva (angles, rad) and max_dvm (magnitudes, pu).  refactor:
 * each row's iterations after its first refactorize on the first one's
 * pivots (KLU-style), falling back to a fresh factorization on a weak pivot. */
void orc_nr_batch(int n, const int *Yp, const int *Yi, const double *Yx, long long y_stride, int npv,
                  const int *pv, int npq, const int *pq, int n_rows, double *V, const double *S,
                  long long s_stride, int max_iter, double tol, int *status, int *iters, double *res,
                  int n_threads, double max_dva, double max_dvm, int refactor)
{
    if (n_threads < 1) n_threads = 1;
    if (n_threads > n_rows) n_threads = n_rows > 0 ? n_rows : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
    orc_job *jobs = (orc_job *)malloc(sizeof(orc_job) * n_threads);
    for (int t = 0; t < n_threads; ++t) {
        orc_job *j = &jobs[t];
        j->n = n; j->Yp = Yp; j->Yi = Yi; j->Yx = Yx; j->y_stride = 2 * y_stride;
        j->npv = npv; j->pv = pv; j->npq = npq; j->pq = pq;
        j->V = V; j->v_stride = 2LL * n; j->S = S; j->s_stride = 2 * s_stride;
        j->max_iter = max_iter; j->tol = tol; j->max_dva = max_dva; j->max_dvm = max_dvm;
        j->status = status; j->iters = iters; j->res = res; j->refactor = refactor;
        j->row0 = (int)((long long)n_rows * t / n_threads);
        j->row1 = (int)((long long)n_rows * (t + 1) / n_threads);
        if (n_threads == 1) orc_worker(j);
        else pthread_create(&th[t], NULL, orc_worker, j);
    }
    if (n_threads > 1)
        for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}
