/*
 * blocks_gravity.c — self-gravity on block meshes (TEST INFRASTRUCTURE ONLY;
 * included by blocks_oracle.c after the block mesh helpers).
 *
 * The reference solves ∇²φ = 4πGρ on the composite leaf mesh with one operator
 * built from its face walk (src/gravity.jl:1-21): _build_poisson_ka_struct /
 * _assemble_poisson_struct (src/gravity.jl:488-600) push one edge per interior
 * face of for_each_face — conforming faces once from the hi side, coarse↔fine
 * faces once from the fine side with the fine area — with coefficient
 * area/d, d = ½(w_left + w_right) along the axis (_face_distance,
 * src/gravity.jl:86-93), and turn the edge list into CSR rows by a counting
 * sort (each row keeps its edges in walk order). The same walk gives the
 * gravity stencil (glo, ghi, d_lo + d_hi per axis): a side that no face reached
 * is a domain boundary (homogeneous Neumann: the cell itself at half width),
 * and the coarse cell of every level jump takes its representative neighbours
 * instead (face_neighbor_handles). CG, the right-hand side, g = −∇φ and the
 * source follow src/gravity.jl:230-447 as in gravity_impl.h. φ's boundaries
 * are the mesh's own: periodic block sides wrap, every other uncovered side is
 * homogeneous Neumann.
 */

typedef struct {
    int64_t n, nnz;
    int64_t* rowptr;   /* n + 1 */
    int64_t* colidx;   /* nnz */
    double* coef;      /* nnz: area / face distance */
    double* vol;       /* n */
    int64_t* glo;      /* 3 * n: lo neighbour per axis (self at a Neumann side) */
    int64_t* ghi;      /* 3 * n */
    double* gsum;      /* 3 * n: d_lo + d_hi */
} bo_poisson;

void bo_poisson_destroy(bo_poisson* P) {
    if (!P) return;
    free(P->rowptr);
    free(P->colidx);
    free(P->coef);
    free(P->vol);
    free(P->glo);
    free(P->ghi);
    free(P->gsum);
    free(P);
}

static double bo_face_distance(const bo_mesh* m, int64_t i, int64_t j, int d) {
    int64_t t[3];
    int bi = bo_locate(m, i, t), bj = bo_locate(m, j, t);
    return 0.5 * (bo_width(m, bi, d) + bo_width(m, bj, d));
}

bo_poisson* bo_poisson_create(const bo_mesh* m) {
    int64_t n = m->ncell;
    int rank = m->rank;
    bo_poisson* P = (bo_poisson*)calloc(1, sizeof(bo_poisson));
    P->n = n;
    P->vol = (double*)malloc(sizeof(double) * n);
    P->glo = (int64_t*)malloc(sizeof(int64_t) * 3 * n);
    P->ghi = (int64_t*)malloc(sizeof(int64_t) * 3 * n);
    P->gsum = (double*)calloc(3 * n, sizeof(double));
    double* dlo = (double*)calloc(3 * n, sizeof(double));
    double* dhi = (double*)calloc(3 * n, sizeof(double));
    uint8_t* marked = (uint8_t*)calloc(n, 1);
    int32_t* lev = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int b = 0; b < m->nblocks; ++b)
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            P->vol[c] = bo_volume(m, b);
            lev[c] = m->level[b];
        }
    for (int64_t c = 0; c < n; ++c)
        for (int d = 0; d < 3; ++d) P->glo[3 * c + d] = P->ghi[3 * c + d] = c;
    /* edges in for_each_face order: per leaf, per axis, the hi face when the
     * neighbour is at the same level or coarser, the lo face when it is coarser */
    int64_t cap = 3 * n + 16, ne = 0;
    int64_t* ei = (int64_t*)malloc(sizeof(int64_t) * cap);
    int64_t* ej = (int64_t*)malloc(sizeof(int64_t) * cap);
    double* ec = (double*)malloc(sizeof(double) * cap);
    for (int64_t g = 0; g < n; ++g) {
        int64_t ijk[3];
        int b = bo_locate(m, g, ijk);
        for (int d = 0; d < rank; ++d) {
            double area = bo_area(m, b, d);
            for (int side = 1; side >= 0; --side) {
                int64_t rep;
                int k = bo_neighbor(m, g, d, side, &rep, NULL, NULL);
                if (!(k == NB_COARSER || (side == 1 && k == NB_SAME))) continue;
                int64_t i = side ? g : rep, j = side ? rep : g;
                double fd = bo_face_distance(m, i, j, d);
                if (ne == cap) {
                    cap *= 2;
                    ei = (int64_t*)realloc(ei, sizeof(int64_t) * cap);
                    ej = (int64_t*)realloc(ej, sizeof(int64_t) * cap);
                    ec = (double*)realloc(ec, sizeof(double) * cap);
                }
                ei[ne] = i;
                ej[ne] = j;
                ec[ne] = area / fd;
                ++ne;
                P->ghi[3 * i + d] = j;
                dhi[3 * i + d] = fd;
                P->glo[3 * j + d] = i;
                dlo[3 * j + d] = fd;
                if (lev[i] != lev[j]) marked[lev[i] < lev[j] ? i : j] = 1;
            }
        }
    }
    /* CSR by counting sort: row i then row j per edge, rows keep walk order */
    P->rowptr = (int64_t*)calloc(n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < ne; ++k) {
        P->rowptr[ei[k] + 1]++;
        P->rowptr[ej[k] + 1]++;
    }
    for (int64_t c = 0; c < n; ++c) P->rowptr[c + 1] += P->rowptr[c];
    P->nnz = P->rowptr[n];
    P->colidx = (int64_t*)malloc(sizeof(int64_t) * (P->nnz + 1));
    P->coef = (double*)malloc(sizeof(double) * (P->nnz + 1));
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    memcpy(cur, P->rowptr, sizeof(int64_t) * (n + 1));
    for (int64_t k = 0; k < ne; ++k) {
        int64_t p = cur[ei[k]]++;
        P->colidx[p] = ej[k];
        P->coef[p] = ec[k];
        int64_t q = cur[ej[k]]++;
        P->colidx[q] = ei[k];
        P->coef[q] = ec[k];
    }
    /* gravity stencil: the face walk's neighbours (a side no face reached is a
     * Neumann boundary), the representatives for the coarse cell of a level jump */
    for (int64_t c = 0; c < n; ++c) {
        int64_t ijk[3];
        int b = bo_locate(m, c, ijk);
        for (int d = 0; d < rank; ++d) {
            double hw = 0.5 * bo_width(m, b, d);
            double dl, dh;
            if (!marked[c]) {
                dl = dlo[3 * c + d];
                dh = dhi[3 * c + d];
                if (dl == 0.0) dl = hw;
                if (dh == 0.0) dh = hw;
            } else {
                int64_t rep;
                if (bo_neighbor(m, c, d, 0, &rep, NULL, NULL) == NB_BOUNDARY) {
                    P->glo[3 * c + d] = c;
                    dl = hw;
                } else {
                    P->glo[3 * c + d] = rep;
                    dl = bo_face_distance(m, c, rep, d);
                }
                if (bo_neighbor(m, c, d, 1, &rep, NULL, NULL) == NB_BOUNDARY) {
                    P->ghi[3 * c + d] = c;
                    dh = hw;
                } else {
                    P->ghi[3 * c + d] = rep;
                    dh = bo_face_distance(m, c, rep, d);
                }
            }
            P->gsum[3 * c + d] = dl + dh;
        }
    }
    free(cur);
    free(ei);
    free(ej);
    free(ec);
    free(dlo);
    free(dhi);
    free(marked);
    free(lev);
    return P;
}

int64_t bo_poisson_nnz(const bo_poisson* P) { return P->nnz; }

/* CSR and stencil copies for the tests: rowptr (n+1), colidx/coef (nnz),
 * glo/ghi/gsum (3n, [cell][axis]). Any pointer may be NULL. */
void bo_poisson_export(const bo_poisson* P, int64_t* rowptr, int64_t* colidx, double* coef, int64_t* glo,
                       int64_t* ghi, double* gsum) {
    if (rowptr) memcpy(rowptr, P->rowptr, sizeof(int64_t) * (P->n + 1));
    if (colidx) memcpy(colidx, P->colidx, sizeof(int64_t) * P->nnz);
    if (coef) memcpy(coef, P->coef, sizeof(double) * P->nnz);
    if (glo) memcpy(glo, P->glo, sizeof(int64_t) * 3 * P->n);
    if (ghi) memcpy(ghi, P->ghi, sizeof(int64_t) * 3 * P->n);
    if (gsum) memcpy(gsum, P->gsum, sizeof(double) * 3 * P->n);
}

#define BT double
#define BSUF f64
#define BSQRT sqrt
#define BFABS fabs
#include "blocks_gravity_impl.h"
#undef BT
#undef BSUF
#undef BSQRT
#undef BFABS

#define BT float
#define BSUF f32
#define BSQRT sqrtf
#define BFABS fabsf
#include "blocks_gravity_impl.h"
#undef BT
#undef BSUF
#undef BSQRT
#undef BFABS
