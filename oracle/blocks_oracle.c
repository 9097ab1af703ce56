/*
 * blocks_oracle.c — CPU oracle for block-structured AMR meshes (TEST INFRASTRUCTURE ONLY).
 *
 * Restates Vespa.jl's flux-divergence step on a composite (leaf-only) mesh with
 * hanging nodes, the semantics HGBackend gives it:
 *   - face enumeration (lib/HGBackend/src/HGBackend.jl:202-276): per leaf, per
 *     axis, the hi face is emitted when the neighbour is at the same level or
 *     coarser (or a domain boundary); the lo face only when it is a domain
 *     boundary or the lo neighbour is coarser. A coarse↔fine face is therefore
 *     emitted once from the fine side, carrying the fine face area;
 *   - reconstruction uses one representative neighbour per face
 *     (face_neighbor_handles, lib/HGBackend/src/HGBackend.jl:178-200);
 *   - the flux, accumulation and RK2 update are src/driver.jl:209-359; level
 *     restricted stepping, flux registers and subcycling are src/driver.jl:367-385,
 *     523-605 and src/reflux.jl:51-103.
 *
 * Leaves are the cells of blocks. Block b has a level L_b, cells per axis
 * dims_b, and an integer origin on the lattice of the finest cells (unit = one
 * cell of level maxlevel); a level-L cell spans 2^(maxlevel-L) lattice units
 * and h0/2^L physical units. Leaves are ordered block by block, column-major
 * inside a block (the leaf order of for_each_cell). Periodic boundaries wrap a
 * block onto itself (RefMesh semantics); between blocks only geometric contact
 * counts. HierarchicalGrids (absent here) chooses the representative neighbour
 * across a coarse↔fine face; this restatement uses the fine cell that contains
 * the lo corner of the coarse face (parity with HG on that choice is unpinned).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { BO_PER = 0, BO_OUT = 1, BO_REF = 2 };
enum { NB_SAME = 0, NB_COARSER = 1, NB_FINER = 2, NB_BOUNDARY = 3 };

typedef struct {
    int32_t rank, nblocks, maxlevel;
    double h0[3];
    double gamma;
    int32_t* level;
    int32_t* dims;
    int64_t* origin;
    int32_t* bc;
    int64_t* offset; /* nblocks + 1 */
    int64_t ncell;
    /* candidate neighbour blocks per block face */
    int32_t** cand;
    int32_t* ncand;
} bo_mesh;

static int64_t bo_span(const bo_mesh* m, int b) { return (int64_t)1 << (m->maxlevel - m->level[b]); }

static double bo_width(const bo_mesh* m, int b, int d) {
    return m->h0[d] / (double)((int64_t)1 << m->level[b]);
}
static double bo_volume(const bo_mesh* m, int b) {
    double v = bo_width(m, b, 0);
    for (int d = 1; d < m->rank; ++d) v = v * bo_width(m, b, d);
    return v;
}
static double bo_area(const bo_mesh* m, int b, int axis) {
    double a = 1.0;
    for (int d = 0; d < m->rank; ++d)
        if (d != axis) a *= bo_width(m, b, d);
    return a;
}

/* block containing global cell g and its local coordinates */
static int bo_locate(const bo_mesh* m, int64_t g, int64_t ijk[3]) {
    int lo = 0, hi = m->nblocks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (m->offset[mid] <= g) lo = mid; else hi = mid - 1;
    }
    int b = lo;
    int64_t l = g - m->offset[b];
    const int32_t* n = m->dims + 3 * b;
    ijk[0] = l % n[0];
    ijk[1] = (l / n[0]) % n[1];
    ijk[2] = l / ((int64_t)n[0] * n[1]);
    return b;
}

static int64_t bo_index(const bo_mesh* m, int b, const int64_t ijk[3]) {
    const int32_t* n = m->dims + 3 * b;
    return m->offset[b] + (ijk[2] * n[1] + ijk[1]) * n[0] + ijk[0];
}

/* does block b contain lattice point q (rank axes) */
static int bo_contains(const bo_mesh* m, int b, const int64_t q[3]) {
    int64_t s = bo_span(m, b);
    for (int d = 0; d < m->rank; ++d) {
        int64_t o = m->origin[3 * b + d];
        if (q[d] < o || q[d] >= o + (int64_t)m->dims[3 * b + d] * s) return 0;
    }
    return 1;
}

/* Neighbour of cell g across face (d, side): kind, representative cell. For
 * NB_FINER, fine[] receives the 2^(rank-1) fine cells touching the face in the
 * fine block's leaf order (count in *nfine). */
static int bo_neighbor(const bo_mesh* m, int64_t g, int d, int side, int64_t* rep, int64_t* fine, int* nfine) {
    int64_t ijk[3];
    int b = bo_locate(m, g, ijk);
    const int32_t* n = m->dims + 3 * b;
    int64_t j = ijk[d] + (side ? 1 : -1);
    if (j >= 0 && j < n[d]) {
        int64_t t[3] = {ijk[0], ijk[1], ijk[2]};
        t[d] = j;
        *rep = bo_index(m, b, t);
        return NB_SAME;
    }
    if (m->bc[6 * b + 2 * d + side] == BO_PER) {
        int64_t t[3] = {ijk[0], ijk[1], ijk[2]};
        t[d] = j < 0 ? n[d] - 1 : 0;
        *rep = bo_index(m, b, t);
        return NB_SAME;
    }
    int64_t s = bo_span(m, b);
    int64_t q[3] = {0, 0, 0};
    for (int e = 0; e < m->rank; ++e) q[e] = m->origin[3 * b + e] + ijk[e] * s;
    q[d] = side ? q[d] + s : q[d] - 1;
    int fb = 2 * d + side;
    for (int k = 0; k < m->ncand[6 * b + fb]; ++k) {
        int c = m->cand[6 * b + fb][k];
        if (!bo_contains(m, c, q)) continue;
        int64_t sc = bo_span(m, c);
        int64_t t[3] = {0, 0, 0};
        for (int e = 0; e < m->rank; ++e) t[e] = (q[e] - m->origin[3 * c + e]) / sc;
        *rep = bo_index(m, c, t);
        int dl = m->level[c] - m->level[b];
        if (dl == 0) return NB_SAME;
        if (dl == -1) return NB_COARSER;
        /* finer: the 2^(rank-1) fine cells on the face, in leaf order */
        if (fine) {
            int cnt = 0;
            int e1 = -1, e2 = -1;
            for (int e = 0; e < m->rank; ++e) {
                if (e == d) continue;
                if (e1 < 0) e1 = e; else e2 = e;
            }
            int n2 = e2 >= 0 ? 2 : 1, n1 = e1 >= 0 ? 2 : 1;
            for (int a2 = 0; a2 < n2; ++a2)
                for (int a1 = 0; a1 < n1; ++a1) {
                    int64_t u[3] = {t[0], t[1], t[2]};
                    if (e1 >= 0) u[e1] += a1;
                    if (e2 >= 0) u[e2] += a2;
                    fine[cnt++] = bo_index(m, c, u);
                }
            *nfine = cnt;
        }
        return NB_FINER;
    }
    return NB_BOUNDARY;
}

void bo_destroy(bo_mesh* m) {
    if (!m) return;
    if (m->cand)
        for (int i = 0; i < 6 * m->nblocks; ++i) free(m->cand[i]);
    free(m->cand);
    free(m->ncand);
    free(m->level);
    free(m->dims);
    free(m->origin);
    free(m->bc);
    free(m->offset);
    free(m);
}

/* Returns NULL on an invalid mesh (overlap is not checked; level jumps > 1 across
 * a face and misaligned fine blocks are rejected by bo_validate). */
bo_mesh* bo_create(int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0, double gamma,
                   const int32_t* level, const int32_t* dims, const int64_t* origin, const int32_t* bc) {
    if (rank < 1 || rank > 3 || nblocks < 1) return NULL;
    bo_mesh* m = (bo_mesh*)calloc(1, sizeof(bo_mesh));
    m->rank = rank;
    m->nblocks = nblocks;
    m->maxlevel = maxlevel;
    for (int d = 0; d < 3; ++d) m->h0[d] = h0[d];
    m->gamma = gamma;
    m->level = (int32_t*)malloc(sizeof(int32_t) * nblocks);
    m->dims = (int32_t*)malloc(sizeof(int32_t) * 3 * nblocks);
    m->origin = (int64_t*)malloc(sizeof(int64_t) * 3 * nblocks);
    m->bc = (int32_t*)malloc(sizeof(int32_t) * 6 * nblocks);
    m->offset = (int64_t*)malloc(sizeof(int64_t) * (nblocks + 1));
    memcpy(m->level, level, sizeof(int32_t) * nblocks);
    memcpy(m->dims, dims, sizeof(int32_t) * 3 * nblocks);
    memcpy(m->origin, origin, sizeof(int64_t) * 3 * nblocks);
    memcpy(m->bc, bc, sizeof(int32_t) * 6 * nblocks);
    m->offset[0] = 0;
    for (int b = 0; b < nblocks; ++b) {
        int64_t c = 1;
        for (int d = 0; d < 3; ++d) c *= (d < rank ? dims[3 * b + d] : 1);
        m->offset[b + 1] = m->offset[b] + c;
    }
    m->ncell = m->offset[nblocks];
    /* candidate neighbours: blocks whose extent touches the face plane */
    m->cand = (int32_t**)calloc(6 * nblocks, sizeof(int32_t*));
    m->ncand = (int32_t*)calloc(6 * nblocks, sizeof(int32_t));
    for (int b = 0; b < nblocks; ++b) {
        int64_t sb = bo_span(m, b);
        for (int d = 0; d < rank; ++d)
            for (int side = 0; side < 2; ++side) {
                int fb = 6 * b + 2 * d + side;
                m->cand[fb] = (int32_t*)malloc(sizeof(int32_t) * nblocks);
                int64_t plane = side ? m->origin[3 * b + d] + dims[3 * b + d] * sb : m->origin[3 * b + d];
                for (int c = 0; c < nblocks; ++c) {
                    if (c == b) continue;
                    int64_t sc = bo_span(m, c);
                    int64_t clo = m->origin[3 * c + d], chi = clo + dims[3 * c + d] * sc;
                    if (side ? clo != plane : chi != plane) continue;
                    int overlap = 1;
                    for (int e = 0; e < rank; ++e) {
                        if (e == d) continue;
                        int64_t alo = m->origin[3 * b + e], ahi = alo + dims[3 * b + e] * sb;
                        int64_t blo = m->origin[3 * c + e], bhi = blo + dims[3 * c + e] * sc;
                        if (bhi <= alo || ahi <= blo) overlap = 0;
                    }
                    if (overlap) m->cand[fb][m->ncand[fb]++] = c;
                }
            }
    }
    return m;
}

int64_t bo_ncell(const bo_mesh* m) { return m->ncell; }

/* 0 if every face sees a level jump of at most one and every coarse face maps
 * onto one fine block; >0 otherwise. */
int bo_validate(const bo_mesh* m) {
    int64_t fine[4];
    int nf;
    for (int64_t g = 0; g < m->ncell; ++g) {
        int64_t ijk[3];
        int b = bo_locate(m, g, ijk);
        for (int d = 0; d < m->rank; ++d)
            for (int side = 0; side < 2; ++side) {
                int64_t rep;
                int k = bo_neighbor(m, g, d, side, &rep, fine, &nf);
                if (k == NB_BOUNDARY && m->bc[6 * b + 2 * d + side] == BO_PER) return 1;
                if (k == NB_SAME || k == NB_COARSER || k == NB_FINER) {
                    int64_t t[3];
                    int c = bo_locate(m, rep, t);
                    int dl = m->level[c] - m->level[b];
                    if (dl > 1 || dl < -1) return 2;
                    if (k == NB_FINER) {
                        for (int i = 0; i < nf; ++i) {
                            int64_t u[3];
                            if (bo_locate(m, fine[i], u) != c) return 3;
                        }
                    }
                }
            }
    }
    return 0;
}

#define BT double
#define BSUF f64
#define BSQRT sqrt
#define BFABS fabs
#include "blocks_impl.h"
#undef BT
#undef BSUF
#undef BSQRT
#undef BFABS

#define BT float
#define BSUF f32
#define BSQRT sqrtf
#define BFABS fabsf
#include "blocks_impl.h"
#undef BT
#undef BSUF
#undef BSQRT
#undef BFABS

/* Self-gravity on block meshes (composite Poisson operator, CG, g, source). */
#include "blocks_gravity.c"
