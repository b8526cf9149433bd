/*
 * cexec.c — FP64 CPU executor of a flattened contraction plan, one syndrome
 * at a time.  TEST INFRASTRUCTURE ONLY: used by tests/ as a fast oracle at
 * d >= 5 and by bench.py as the timed host-CPU baseline.
 *
 * It restates what the reference does per call of
 * `decode(ct::CompiledTNMMAP, syn)` (src/decoding/tndecoder.jl:160-165):
 *   - update_syndrome! (:148-158): the evidence leaves become one-hot
 *     vectors [1,0] / [0,1] (leaves sliced by a syndrome bit take the slice);
 *   - ct.code(ct.tensors...) (:162): every pairwise step of the tree is
 *     evaluated, syndrome-independent ones included, as OMEinsum does;
 *   - findmax (:163): first maximum in column-major order.
 * Each step is C[c] = sum A[a] * B[b] over the union of its labels, with the
 * (a, b, c) index triples precomputed once per plan.
 *
 * Parallelism: OpenMP over syndromes (one private scratch per thread).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int rank;
    int64_t off;       /* scratch offset of the node's values */
} ox_node;

typedef struct {
    int a, b, c;       /* node ids */
    int64_t nu;        /* 2^|union| */
    uint32_t *ai, *bi, *ci;
} ox_step;

typedef struct {
    int n_leaves, n_steps, n_syn, n_out;
    ox_node *nodes;
    ox_step *steps;
    int64_t scratch;   /* doubles per thread */
    /* leaves */
    double **leaf_data;    /* data in leaf label order (effective labels) for constants */
    int *leaf_ev;          /* one-hot evidence bit or -1 */
    int *leaf_nsl;         /* number of sliced labels */
    int **leaf_slbit;      /* syndrome bits of the sliced labels */
    double **leaf_var;     /* per-variant sliced data: 2^nsl x 2^rank */
    int *root_perm;        /* final label j -> position in root labels */
} ox_plan;

static int find(const int *v, int n, int x) {
    for (int i = 0; i < n; ++i)
        if (v[i] == x) return i;
    return -1;
}

void ox_plan_free(ox_plan *p);

ox_plan *ox_plan_new(int n_leaves, const int *leaf_rank, const int *leaf_labels, const double *leaf_data,
                     const int *leaf_ev, int n_steps, const int *lhs, const int *rhs, const int *out_rank,
                     const int *out_labels, int n_out, const int *final_labels, int n_syn, int n_sliced,
                     const int *sliced_label, const int *sliced_bit) {
    ox_plan *p = (ox_plan *)calloc(1, sizeof(ox_plan));
    if (!p) return NULL;
    p->n_leaves = n_leaves;
    p->n_steps = n_steps;
    p->n_syn = n_syn;
    p->n_out = n_out;
    const int NN = n_leaves + n_steps;
    p->nodes = (ox_node *)calloc(NN, sizeof(ox_node));
    p->steps = (ox_step *)calloc(n_steps > 0 ? n_steps : 1, sizeof(ox_step));
    p->leaf_data = (double **)calloc(n_leaves, sizeof(double *));
    p->leaf_ev = (int *)calloc(n_leaves, sizeof(int));
    p->leaf_nsl = (int *)calloc(n_leaves, sizeof(int));
    p->leaf_slbit = (int **)calloc(n_leaves, sizeof(int *));
    p->leaf_var = (double **)calloc(n_leaves, sizeof(double *));
    p->root_perm = (int *)calloc(n_out > 0 ? n_out : 1, sizeof(int));
    int **labels = (int **)calloc(NN, sizeof(int *));
    int *nlab = (int *)calloc(NN, sizeof(int));
    if (!p->nodes || !p->steps || !labels || !nlab) { free(labels); free(nlab); ox_plan_free(p); return NULL; }
    int64_t loff = 0, doff = 0, scratch = 0;
    for (int t = 0; t < n_leaves; ++t) {
        const int r = leaf_rank[t];
        const int *L = leaf_labels + loff;
        labels[t] = (int *)malloc(sizeof(int) * (r > 0 ? r : 1));
        int eff = 0, nsl = 0;
        int slpos[32], slb[32];
        for (int q = 0; q < r; ++q) {
            int j = find(sliced_label, n_sliced, L[q]);
            if (j >= 0) { slpos[nsl] = q; slb[nsl] = sliced_bit[j]; ++nsl; }
            else labels[t][eff++] = L[q];
        }
        nlab[t] = eff;
        p->leaf_ev[t] = leaf_ev ? leaf_ev[t] : -1;
        p->leaf_nsl[t] = nsl;
        const int64_t n_eff = (int64_t)1 << eff;
        if (p->leaf_ev[t] < 0) {
            const double *D = leaf_data + doff;
            if (nsl == 0) {
                p->leaf_data[t] = (double *)malloc(sizeof(double) * n_eff);
                memcpy(p->leaf_data[t], D, sizeof(double) * n_eff);
            } else {
                p->leaf_slbit[t] = (int *)malloc(sizeof(int) * nsl);
                memcpy(p->leaf_slbit[t], slb, sizeof(int) * nsl);
                p->leaf_var[t] = (double *)calloc((size_t)n_eff << nsl, sizeof(double));
                for (int64_t x = 0; x < ((int64_t)1 << r); ++x) {
                    int v = 0;
                    for (int j = 0; j < nsl; ++j) v |= (int)((x >> slpos[j]) & 1) << j;
                    int64_t ki = 0;
                    int kp = 0;
                    for (int q = 0; q < r; ++q) {
                        int sl = 0;
                        for (int j = 0; j < nsl; ++j) sl |= (slpos[j] == q);
                        if (!sl) ki |= ((x >> q) & 1) << kp++;
                    }
                    p->leaf_var[t][((int64_t)v << eff) + ki] = D[x];
                }
            }
        }
        p->nodes[t].rank = eff;
        p->nodes[t].off = scratch;
        scratch += n_eff;
        loff += r;
        doff += (int64_t)1 << r;
    }
    int64_t ooff = 0;
    for (int i = 0; i < n_steps; ++i) {
        const int node = n_leaves + i, r = out_rank[i];
        labels[node] = (int *)malloc(sizeof(int) * (r > 0 ? r : 1));
        memcpy(labels[node], out_labels + ooff, sizeof(int) * r);
        nlab[node] = r;
        ooff += r;
        p->nodes[node].rank = r;
        p->nodes[node].off = scratch;
        scratch += (int64_t)1 << r;
        ox_step *s = &p->steps[i];
        s->a = lhs[i];
        s->b = rhs[i];
        s->c = node;
        int U[64], nu = 0;
        for (int q = 0; q < r; ++q) U[nu++] = labels[node][q];
        for (int q = 0; q < nlab[s->a]; ++q) if (find(U, nu, labels[s->a][q]) < 0) U[nu++] = labels[s->a][q];
        for (int q = 0; q < nlab[s->b]; ++q) if (find(U, nu, labels[s->b][q]) < 0) U[nu++] = labels[s->b][q];
        s->nu = (int64_t)1 << nu;
        s->ai = (uint32_t *)malloc(sizeof(uint32_t) * s->nu);
        s->bi = (uint32_t *)malloc(sizeof(uint32_t) * s->nu);
        s->ci = (uint32_t *)malloc(sizeof(uint32_t) * s->nu);
        int pa[64], pb[64];
        for (int q = 0; q < nlab[s->a]; ++q) pa[q] = find(U, nu, labels[s->a][q]);
        for (int q = 0; q < nlab[s->b]; ++q) pb[q] = find(U, nu, labels[s->b][q]);
        for (int64_t u = 0; u < s->nu; ++u) {
            uint32_t ia = 0, ib = 0;
            for (int q = 0; q < nlab[s->a]; ++q) ia |= (uint32_t)((u >> pa[q]) & 1) << q;
            for (int q = 0; q < nlab[s->b]; ++q) ib |= (uint32_t)((u >> pb[q]) & 1) << q;
            s->ai[u] = ia;
            s->bi[u] = ib;
            s->ci[u] = (uint32_t)(u & ((((int64_t)1) << r) - 1));
        }
    }
    const int root = NN - 1;
    for (int j = 0; j < n_out; ++j) p->root_perm[j] = find(labels[root], nlab[root], final_labels[j]);
    p->scratch = scratch;
    for (int t = 0; t < NN; ++t) free(labels[t]);
    free(labels);
    free(nlab);
    return p;
}

void ox_plan_free(ox_plan *p) {
    if (!p) return;
    if (p->steps)
        for (int i = 0; i < p->n_steps; ++i) { free(p->steps[i].ai); free(p->steps[i].bi); free(p->steps[i].ci); }
    for (int t = 0; p->leaf_data && t < p->n_leaves; ++t) {
        free(p->leaf_data[t]);
        if (p->leaf_slbit) free(p->leaf_slbit[t]);
        if (p->leaf_var) free(p->leaf_var[t]);
    }
    free(p->leaf_data); free(p->leaf_ev); free(p->leaf_nsl); free(p->leaf_slbit); free(p->leaf_var);
    free(p->nodes); free(p->steps); free(p->root_perm);
    free(p);
}

int64_t ox_plan_scratch(const ox_plan *p) { return p->scratch; }

/* Multiply-adds per syndrome (every step, as OMEinsum evaluates them). */
double ox_plan_macs(const ox_plan *p) {
    double s = 0;
    for (int i = 0; i < p->n_steps; ++i) s += (double)p->steps[i].nu;
    return s;
}

static void decode_one(const ox_plan *p, const uint8_t *syn, double *W, double *mar, int32_t *cls, uint8_t *tie,
                       double tie_tol) {
    for (int t = 0; t < p->n_leaves; ++t) {
        double *dst = W + p->nodes[t].off;
        const int64_t n = (int64_t)1 << p->nodes[t].rank;
        if (p->leaf_ev[t] >= 0) {
            const int s = syn[p->leaf_ev[t]] & 1;
            dst[0] = s ? 0.0 : 1.0;
            dst[1] = s ? 1.0 : 0.0;
        } else if (p->leaf_nsl[t] > 0) {
            int v = 0;
            for (int j = 0; j < p->leaf_nsl[t]; ++j) v |= (syn[p->leaf_slbit[t][j]] & 1) << j;
            memcpy(dst, p->leaf_var[t] + ((int64_t)v << p->nodes[t].rank), sizeof(double) * n);
        } else {
            memcpy(dst, p->leaf_data[t], sizeof(double) * n);
        }
    }
    for (int i = 0; i < p->n_steps; ++i) {
        const ox_step *s = &p->steps[i];
        const double *A = W + p->nodes[s->a].off;
        const double *B = W + p->nodes[s->b].off;
        double *C = W + p->nodes[s->c].off;
        memset(C, 0, sizeof(double) * ((int64_t)1 << p->nodes[s->c].rank));
        const uint32_t *ai = s->ai, *bi = s->bi, *ci = s->ci;
        for (int64_t u = 0; u < s->nu; ++u) C[ci[u]] += A[ai[u]] * B[bi[u]];
    }
    const double *R = W + p->nodes[p->n_leaves + p->n_steps - 1].off;
    const int nout = 1 << p->n_out;
    double best = -1, second = -1;
    int bi = 0;
    for (int j = 0; j < nout; ++j) {
        int64_t idx = 0;
        for (int q = 0; q < p->n_out; ++q) idx |= (int64_t)((j >> q) & 1) << p->root_perm[q];
        const double x = R[idx];
        if (mar) mar[j] = x;
        if (j == 0 || x > best) {
            if (j > 0) second = best;
            best = x;
            bi = j;
        } else if (x > second) {
            second = x;
        }
    }
    if (cls) *cls = bi;
    if (tie) *tie = (nout > 1 && second >= 0 && (best - second) <= tie_tol * best) ? 1 : 0;
}

int ox_decode(const ox_plan *p, const uint8_t *syn, int64_t n, double *marginals, int32_t *classes,
              uint8_t *ties, double tie_tol, int nthreads) {
    if (!p) return -1;
    const int nout = 1 << p->n_out;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *W = (double *)malloc(sizeof(double) * (size_t)p->scratch);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t b = 0; b < n; ++b) {
            decode_one(p, syn + b * p->n_syn, W, marginals ? marginals + b * nout : NULL,
                       classes ? classes + b : NULL, ties ? ties + b : NULL, tie_tol);
        }
        free(W);
    }
    return 0;
}
