/* Uniform-mesh description and geometry helpers shared by the C oracle
 * (vespa_oracle.c) and its dual-number instantiation (jvp_oracle.cc).
 * TEST INFRASTRUCTURE ONLY: see vespa_oracle.c.
 * Follows lib/RefMesh/src/RefMesh.jl:89-161 (mesh, widths, areas, volume,
 * neighbours) and Julia's IEEE min/max. */
#ifndef VO_GRID_IMPL_H
#define VO_GRID_IMPL_H

#include <math.h>
#include <stdint.h>

enum { VO_BC_PERIODIC = 0, VO_BC_OUTFLOW = 1, VO_BC_REFLECTING = 2, VO_BC_PARENT = 3 };

/* Uniform mesh description (lib/RefMesh/src/RefMesh.jl:89-100). Axes beyond
 * rank have dims = 1 and take no faces. bc[2*axis + side], side 0 = lo, 1 = hi.
 * VO_BC_PARENT is ParentGhost (lib/MeshInterface/src/MeshInterface.jl:86-99,
 * src/driver.jl:127-134): the boundary ghost of a face is the primitive state
 * the closure returns for (axis, side, cell); pg holds those states at the
 * state's precision, 5 per boundary face in the slots of vo_bface_slot. */
typedef struct {
    int32_t rank;
    int32_t bc[6];
    int64_t dims[3];
    double lo[3];
    double hi[3];
    double gamma;
    const void* pg;
} vo_grid;

static int vo_check(const vo_grid* m) {
    if (m->rank < 1 || m->rank > 3) return 1;
    for (int d = 0; d < 3; ++d) {
        if (m->dims[d] < 1) return 1;
        if (d >= m->rank && m->dims[d] != 1) return 1;
    }
    for (int d = 0; d < m->rank; ++d) {
        int lo = m->bc[2 * d], hi = m->bc[2 * d + 1];
        if (lo < 0 || lo > 3 || hi < 0 || hi > 3) return 1;
        if ((lo == VO_BC_PERIODIC) != (hi == VO_BC_PERIODIC)) return 1;
        if ((lo == VO_BC_PARENT || hi == VO_BC_PARENT) && !m->pg) return 1;
    }
    return 0;
}

/* The adjoint restatement covers the synthesizing boundary rules only. */
static int vo_has_parent(const vo_grid* m) {
    for (int s = 0; s < 2 * m->rank; ++s)
        if (m->bc[s] == VO_BC_PARENT) return 1;
    return 0;
}

static inline int64_t vo_ncell(const vo_grid* m) { return m->dims[0] * m->dims[1] * m->dims[2]; }

/* lib/RefMesh/src/RefMesh.jl:122-123 _width = (hi - lo)/dims in Tg = Float64. */
static inline double vo_cell_width(const vo_grid* m, int d) {
    return (m->hi[d] - m->lo[d]) / (double)m->dims[d];
}
/* lib/RefMesh/src/RefMesh.jl:127 cell_volume = prod(widths), left fold. */
static inline double vo_cell_volume(const vo_grid* m) {
    double v = vo_cell_width(m, 0);
    for (int d = 1; d < m->rank; ++d) v = v * vo_cell_width(m, d);
    return v;
}
/* lib/RefMesh/src/RefMesh.jl:135-142 face_area: a = one; a *= w[d] for d != axis. */
static inline double vo_face_area(const vo_grid* m, int axis) {
    double a = 1.0;
    for (int d = 0; d < m->rank; ++d)
        if (d != axis) a *= vo_cell_width(m, d);
    return a;
}

static inline int64_t vo_stride(const vo_grid* m, int d) {
    return d == 0 ? 1 : (d == 1 ? m->dims[0] : m->dims[0] * m->dims[1]);
}
static inline int64_t vo_coord(const vo_grid* m, int64_t c, int d) {
    return (c / vo_stride(m, d)) % m->dims[d];
}

/* lib/RefMesh/src/RefMesh.jl:147-161 neighbor: returns 1 and the neighbor index for
 * an interior (or periodic-wrapped) neighbor, 0 for a domain boundary. */
static inline int vo_neighbor(const vo_grid* m, int64_t c, int d, int side, int64_t* nb) {
    int64_t n = m->dims[d], s = vo_stride(m, d);
    int64_t i = vo_coord(m, c, d);
    int64_t j = i + (side ? 1 : -1);
    if (j >= 0 && j < n) { *nb = c + (j - i) * s; return 1; }
    if (m->bc[2 * d + side] == VO_BC_PERIODIC) {
        int64_t jw = j < 0 ? n - 1 : 0;
        *nb = c + (jw - i) * s;
        return 1;
    }
    return 0;
}

/* Slot of a boundary face in the flat boundary-flux array: planes ordered
 * (axis 0 lo, axis 0 hi, axis 1 lo, ...), each column-major over the other axes. */
static int64_t vo_bface_slot(const vo_grid* m, int64_t c, int d, int side) {
    int64_t n = vo_ncell(m);
    int64_t base = 0;
    for (int e = 0; e < d; ++e) base += 2 * (n / m->dims[e]);
    if (side) base += n / m->dims[d];
    int64_t idx = 0, mul = 1;
    for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        idx += vo_coord(m, c, e) * mul;
        mul *= m->dims[e];
    }
    return base + idx;
}


static inline double vo_jmax_f64_(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y > x) return y;
    if (x > y) return x;
    return signbit(y) ? x : y;
}
static inline double vo_jmin_f64_(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y < x) return y;
    if (x < y) return x;
    return signbit(y) ? y : x;
}

#endif /* VO_GRID_IMPL_H */
