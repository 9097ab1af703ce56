// Block-structured AMR context: host topology builder, flux registers and the
// vh_blocks_* C-ABI entry points.
//
// The mesh is a list of blocks (level, cells per axis, integer origin on the
// lattice of the finest cells, boundary codes), leaves ordered block by block
// and column-major inside a block. From it the builder derives, once per mesh:
//   - per block face: the representative neighbour of every face cell
//     (repmap), used by the tiled stage kernel for slopes across the face;
//   - link faces: every face between two different blocks, emitted once as
//     HGBackend emits it (same level from the hi side, coarse-fine from the fine
//     side with the fine area; lib/HGBackend/src/HGBackend.jl:202-276), with the
//     representative neighbours of both cells for the reconstruction
//     (lib/HGBackend/src/HGBackend.jl:178-200);
//   - link-cell records: for each cell touching link faces, the order in which
//     the face walk adds them (own walk position, or the block order of the
//     emitting cells);
//   - flux-register lists: per level, the coarse cells with finer neighbours
//     (src/reflux.jl:51-103);
//   - tile lists for one persistent launch per stage (all blocks, per level,
//     and per block for the launch-per-block comparison).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vespahip.h"
#include "host.h"
#include "profiles.h"

namespace vh {

enum { NB_SAME = 0, NB_COARSER = 1, NB_FINER = 2, NB_BOUNDARY = 3 };

// Map from non-negative 64-bit keys (cell indices, cell·6 + slot) to
// non-negative ints, open addressing with linear probing: the table build does
// a few lookups per cell on every linked block face, where node-based maps
// spend most of the build's time.
class FlatMap {
 public:
  void reserve(size_t n) {
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    keys_.assign(cap, -1);
    vals_.assign(cap, 0);
    mask_ = cap - 1;
    n_ = 0;
  }
  int find(long long k) const {
    if (keys_.empty()) return -1;
    for (size_t i = slot(k);; i = (i + 1) & mask_) {
      if (keys_[i] == k) return vals_[i];
      if (keys_[i] < 0) return -1;
    }
  }
  // Sets key k to v; returns false (and leaves the value) when k is present and keep is set.
  bool put(long long k, int v, bool keep = false) {
    if (2 * (n_ + 1) > keys_.size()) grow();
    for (size_t i = slot(k);; i = (i + 1) & mask_) {
      if (keys_[i] == k) {
        if (keep) return false;
        vals_[i] = v;
        return true;
      }
      if (keys_[i] < 0) {
        keys_[i] = k;
        vals_[i] = v;
        ++n_;
        return true;
      }
    }
  }
  template <class F>
  void for_each(F f) const {
    for (size_t i = 0; i < keys_.size(); ++i)
      if (keys_[i] >= 0) f(keys_[i], vals_[i]);
  }
  size_t size() const { return n_; }

 private:
  size_t slot(long long k) const { return (size_t)(((unsigned long long)k * 0x9E3779B97F4A7C15ULL) >> 20) & mask_; }
  void grow() {
    std::vector<long long> k0;
    std::vector<int> v0;
    k0.swap(keys_);
    v0.swap(vals_);
    reserve(k0.size());
    for (size_t i = 0; i < k0.size(); ++i)
      if (k0[i] >= 0) put(k0[i], v0[i]);
  }
  std::vector<long long> keys_;
  std::vector<int> vals_;
  size_t mask_ = 0, n_ = 0;
};

// Tile limits of block_stage_kernel (BNT lanes, LDS sized for these bounds).
static const int kLanes = 256, kMaxW = 432, kMaxFS = 592, kMaxE = 32;

struct Topo {
  int rank = 0, nb = 0, maxlevel = 0;
  double h0[3] = {1, 1, 1};
  std::vector<int> level, dims, bc;      // dims: 3 per block, bc: 6 per block
  std::vector<long long> origin, off;    // origin: 3 per block, off: nb + 1
  std::vector<std::vector<int>> cand;    // 6 per block: blocks touching the face plane

  long long span(int b) const { return 1LL << (maxlevel - level[b]); }
  double width(int b, int d) const { return h0[d] / (double)(1LL << level[b]); }
  double volume(int b) const {
    double v = width(b, 0);
    for (int d = 1; d < rank; ++d) v = v * width(b, d);
    return v;
  }
  double area(int b, int axis) const {
    double a = 1.0;
    for (int d = 0; d < rank; ++d)
      if (d != axis) a *= width(b, d);
    return a;
  }
  int locate(long long g, long long ijk[3]) const {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (off[mid] <= g) lo = mid;
      else hi = mid - 1;
    }
    const long long l = g - off[lo];
    const int* n = &dims[3 * lo];
    ijk[0] = l % n[0];
    ijk[1] = (l / n[0]) % n[1];
    ijk[2] = l / ((long long)n[0] * n[1]);
    return lo;
  }
  long long index(int b, const long long ijk[3]) const {
    const int* n = &dims[3 * b];
    return off[b] + (ijk[2] * n[1] + ijk[1]) * n[0] + ijk[0];
  }
  bool contains(int b, const long long q[3]) const {
    const long long s = span(b);
    for (int d = 0; d < rank; ++d) {
      const long long o = origin[3 * b + d];
      if (q[d] < o || q[d] >= o + (long long)dims[3 * b + d] * s) return false;
    }
    return true;
  }
  // Neighbour of cell g across face (d, side): kind and representative cell;
  // for NB_FINER the 2^(rank-1) finer cells on the face in their leaf order.
  int neighbor(long long g, int d, int side, long long* rep, long long* fine, int* nfine, int* nbblock) const {
    long long ijk[3];
    const int b = locate(g, ijk);
    return neighbor_at(b, ijk, d, side, rep, fine, nfine, nbblock);
  }
  // The same for cell ijk of block b (no search for the block).
  int neighbor_at(int b, const long long ijk[3], int d, int side, long long* rep, long long* fine, int* nfine,
                  int* nbblock) const {
    const int* n = &dims[3 * b];
    const long long j = ijk[d] + (side ? 1 : -1);
    if (j >= 0 && j < n[d]) {
      long long t[3] = {ijk[0], ijk[1], ijk[2]};
      t[d] = j;
      *rep = index(b, t);
      if (nbblock) *nbblock = b;
      return NB_SAME;
    }
    if (bc[6 * b + 2 * d + side] == BC_PERIODIC) {
      long long t[3] = {ijk[0], ijk[1], ijk[2]};
      t[d] = j < 0 ? n[d] - 1 : 0;
      *rep = index(b, t);
      if (nbblock) *nbblock = b;
      return NB_SAME;
    }
    const long long s = span(b);
    long long q[3] = {0, 0, 0};
    for (int e = 0; e < rank; ++e) q[e] = origin[3 * b + e] + ijk[e] * s;
    q[d] = side ? q[d] + s : q[d] - 1;
    for (int c : cand[6 * b + 2 * d + side]) {
      if (!contains(c, q)) continue;
      const long long sc = span(c);
      long long t[3] = {0, 0, 0};
      for (int e = 0; e < rank; ++e) t[e] = (q[e] - origin[3 * c + e]) / sc;
      *rep = index(c, t);
      if (nbblock) *nbblock = c;
      const int dl = level[c] - level[b];
      if (dl == 0) return NB_SAME;
      if (dl == -1) return NB_COARSER;
      if (dl != 1) return -1;
      if (fine) {
        int e1 = -1, e2 = -1;
        for (int e = 0; e < rank; ++e) {
          if (e == d) continue;
          if (e1 < 0) e1 = e;
          else e2 = e;
        }
        const int n2 = e2 >= 0 ? 2 : 1, n1 = e1 >= 0 ? 2 : 1;
        int cnt = 0;
        for (int a2 = 0; a2 < n2; ++a2)
          for (int a1 = 0; a1 < n1; ++a1) {
            long long u[3] = {t[0], t[1], t[2]};
            if (e1 >= 0) u[e1] += a1;
            if (e2 >= 0) u[e2] += a2;
            for (int e = 0; e < rank; ++e)
              if (u[e] >= dims[3 * c + e]) return -2;   // the face is not covered by one finer block
            fine[cnt++] = index(c, u);
          }
        *nfine = cnt;
      }
      return NB_FINER;
    }
    return NB_BOUNDARY;
  }
};

// Face-plane index of a face cell (column-major over the other two axes); the
// device helper plane_index uses the same rule.
static int plane_idx(const int* n, int d, long long i, long long j, long long k) {
  return (int)(d == 0 ? k * n[1] + j : (d == 1 ? k * n[0] + i : j * n[0] + i));
}
static int plane_size(const int* n, int d) { return d == 0 ? n[1] * n[2] : (d == 1 ? n[0] * n[2] : n[0] * n[1]); }

struct RegList {
  std::vector<int> cell, lc, block;
};

struct HostTables {
  std::vector<BlockDev> blocks;
  std::vector<int> repmap, lcmap;
  std::vector<LinkFaceDev> faces;
  std::vector<int> slot_faces;
  std::vector<LinkCellDev> linkcells;
  std::vector<RegList> regs;             // per coarse level
  // tiles: [block order][all, cost-sorted][level 0 .. maxlevel, cost-sorted]
  std::vector<TileDev> tiles;
  std::vector<int> tile_list_off;        // list l: [tile_list_off[l], tile_list_off[l+1])
  std::vector<int> block_tile_start;     // per block, in the block-order list (nb + 1)
  // CFL block lists: [all][level 0 .. maxlevel]
  std::vector<int> cfl_blocks;
  std::vector<long long> cfl_prefix;
  std::vector<int> cfl_list_off;         // into cfl_blocks; prefix list l starts at cfl_list_off[l] + l
};

static int balanced(int n, int maxT) {
  const int nt = (n + maxT - 1) / maxT;
  return (n + nt - 1) / nt;
}

// Builds every table of a mesh. Returns 0 or a negative code (message set).
static int build_tables(Topo& T, int zchunk, HostTables& H, int max_tx = 32) {
  const int nb = T.nb, rank = T.rank;
  // ---- candidate neighbours by face plane ----
  T.cand.assign(6 * nb, {});
  for (int d = 0; d < rank; ++d) {
    std::unordered_map<long long, std::vector<int>> lo_at;
    for (int c = 0; c < nb; ++c) lo_at[T.origin[3 * c + d]].push_back(c);
    std::unordered_map<long long, std::vector<int>> hi_at;
    for (int c = 0; c < nb; ++c) hi_at[T.origin[3 * c + d] + (long long)T.dims[3 * c + d] * T.span(c)].push_back(c);
    for (int b = 0; b < nb; ++b) {
      const long long sb = T.span(b);
      for (int side = 0; side < 2; ++side) {
        const long long plane = side ? T.origin[3 * b + d] + (long long)T.dims[3 * b + d] * sb : T.origin[3 * b + d];
        auto& m = side ? lo_at : hi_at;
        auto it = m.find(plane);
        if (it == m.end()) continue;
        for (int c : it->second) {
          if (c == b) continue;
          const long long sc = T.span(c);
          bool overlap = true;
          for (int e = 0; e < rank; ++e) {
            if (e == d) continue;
            const long long alo = T.origin[3 * b + e], ahi = alo + T.dims[3 * b + e] * sb;
            const long long blo = T.origin[3 * c + e], bhi = blo + T.dims[3 * c + e] * sc;
            if (bhi <= alo || ahi <= blo) overlap = false;
          }
          if (overlap) T.cand[6 * b + 2 * d + side].push_back(c);
        }
        if (!T.cand[6 * b + 2 * d + side].empty() && T.bc[6 * b + 2 * d] == BC_PERIODIC)
          return fail(-1, "block %s%lld touches another block along a periodic axis", "", b);
      }
    }
  }

  // ---- blocks, representative maps, link faces ----
  H.blocks.assign(nb, BlockDev{});
  FlatMap facemap;   // emitting cell * 6 + slot -> face
  long long cells_on_linked_planes = 0;
  for (int b = 0; b < nb; ++b)
    for (int d = 0; d < rank; ++d)
      for (int side = 0; side < 2; ++side)
        if (!T.cand[6 * b + 2 * d + side].empty()) cells_on_linked_planes += plane_size(&T.dims[3 * b], d);
  facemap.reserve((size_t)cells_on_linked_planes + 16);
  long long fine[4];
  int nfine = 0;
  for (int b = 0; b < nb; ++b) {
    BlockDev& bk = H.blocks[b];
    for (int d = 0; d < 3; ++d) bk.n[d] = T.dims[3 * b + d];
    for (int s = 0; s < 6; ++s) bk.bc[s] = T.bc[6 * b + s];
    bk.level = T.level[b];
    bk.off = T.off[b];
    for (int d = 0; d < 3; ++d) {
      bk.w[d] = d < rank ? T.width(b, d) : 1.0;
      bk.area[d] = d < rank ? T.area(b, d) : 1.0;
    }
    bk.V = T.volume(b);
    for (int s = 0; s < 6; ++s) bk.repoff[s] = bk.lcoff[s] = -1;
    for (int d = 0; d < rank; ++d)
      for (int side = 0; side < 2; ++side) {
        const int slot = 2 * d + side;
        if (T.cand[6 * b + slot].empty()) continue;
        const int ps = plane_size(bk.n, d);
        if ((long long)H.repmap.size() + ps > 0x7fffffffLL) return fail(-1, "mesh too large%s", "");
        bk.repoff[slot] = (int)H.repmap.size();
        H.repmap.resize(H.repmap.size() + ps, -1);
        const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
        for (int v = 0; v < bk.n[e2]; ++v)
          for (int u = 0; u < bk.n[e1]; ++u) {
            long long ijk[3];
            ijk[d] = side ? bk.n[d] - 1 : 0;
            ijk[e1] = u;
            ijk[e2] = v;
            const long long g = T.index(b, ijk);
            long long rep = -1;
            const int kind = T.neighbor_at(b, ijk, d, side, &rep, fine, &nfine, nullptr);
            if (kind == -1) return fail(-1, "block %s%lld touches a block more than one level apart", "", b);
            if (kind < 0)
              return fail(-1, "block %s%lld: a coarse face is not covered by one finer block (fine blocks must be "
                          "aligned to coarse cells)", "", b);
            H.repmap[bk.repoff[slot] + plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2])] =
                kind == NB_BOUNDARY ? -1 : (int)rep;
            const bool emit = (side == 1 && (kind == NB_SAME || kind == NB_COARSER)) || (side == 0 && kind == NB_COARSER);
            if (!emit) continue;
            LinkFaceDev f{};
            f.iL = (int)(side ? g : rep);
            f.iR = (int)(side ? rep : g);
            f.axis = d;
            f.area = T.area(b, d);
            facemap.put(g * 6 + slot, (int)H.faces.size());
            H.faces.push_back(f);
          }
      }
  }
  // representative neighbours of both cells of every link face
  for (LinkFaceDev& f : H.faces) {
    const int d = f.axis;
    for (int which = 0; which < 2; ++which) {
      const long long c = which ? f.iR : f.iL;
      long long ijk[3];
      const int b = T.locate(c, ijk);
      int* nbr = which ? f.nR : f.nL;
      for (int s = 0; s < 2; ++s) {
        long long rep = -1;
        const int kind = T.neighbor_at(b, ijk, d, s, &rep, nullptr, nullptr, nullptr);
        if (kind < 0) return fail(-1, "invalid neighbour across a link face%s", "");
        nbr[s] = kind == NB_BOUNDARY ? -1 - T.bc[6 * b + 2 * d + s] : (int)rep;
      }
    }
  }

  // ---- link-cell records ----
  // A slot lists at most 2^(rank-1) <= 4 faces: [0] count, [1..4] faces.
  using SlotList = std::array<int, 5>;
  FlatMap recof;   // cell -> record
  recof.reserve(2 * H.faces.size() + 16);
  std::vector<int> recblock;
  std::vector<long long> reccell;             // record -> cell
  std::vector<int> recgrpblock;               // record*6 + slot -> emitting block of a type-2 slot
  std::vector<LinkCellDev> recs;
  std::vector<SlotList> slotlists;            // record*6 + slot
  {
    // at most two records per link face, and one per cell on a linked face plane
    const size_t nrec = std::min((size_t)cells_on_linked_planes, 2 * H.faces.size());
    recs.reserve(nrec);
    recblock.reserve(nrec);
    reccell.reserve(nrec);
    recgrpblock.reserve(6 * nrec);
    slotlists.reserve(6 * nrec);
  }
  auto record_for = [&](long long g, int b) -> int {
    const int found = recof.find(g);
    if (found >= 0) return found;
    const int r = (int)recs.size();
    recof.put(g, r);
    reccell.push_back(g);
    LinkCellDev R{};
    recs.push_back(R);
    recblock.push_back(b);
    for (int q = 0; q < 6; ++q) {
      slotlists.push_back(SlotList{0, -1, -1, -1, -1});
      recgrpblock.push_back(-1);
    }
    return r;
  };
  for (int b = 0; b < nb; ++b) {
    const BlockDev& bk = H.blocks[b];
    for (int d = 0; d < rank; ++d)
      for (int side = 0; side < 2; ++side) {
        const int slot = 2 * d + side;
        if (bk.repoff[slot] < 0) continue;
        const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
        for (int v = 0; v < bk.n[e2]; ++v)
          for (int u = 0; u < bk.n[e1]; ++u) {
            long long ijk[3];
            ijk[d] = side ? bk.n[d] - 1 : 0;
            ijk[e1] = u;
            ijk[e2] = v;
            const long long g = T.index(b, ijk);
            long long rep = -1;
            int nbb = -1;
            const int kind = T.neighbor_at(b, ijk, d, side, &rep, fine, &nfine, &nbb);
            if (kind == NB_BOUNDARY) continue;
            int type = 0, emitter_block = -1;
            SlotList fl{0, -1, -1, -1, -1};
            auto face_of = [&](long long cell, int sl) -> int {
              return facemap.find(cell * 6 + sl);
            };
            if ((side == 1 && (kind == NB_SAME || kind == NB_COARSER)) || (side == 0 && kind == NB_COARSER)) {
              type = 1;
              fl[++fl[0]] = face_of(g, slot);
            } else if (side == 0 && kind == NB_SAME) {
              type = 2;
              emitter_block = nbb;
              fl[++fl[0]] = face_of(rep, 2 * d + 1);
            } else if (kind == NB_FINER) {
              type = 2;
              emitter_block = nbb;
              for (int i = 0; i < nfine; ++i) fl[++fl[0]] = face_of(fine[i], side ? 2 * d : 2 * d + 1);
            }
            for (int i = 1; i <= fl[0]; ++i)
              if (fl[i] < 0)
                return fail(-1, "block %s%lld: a link face has no emitting cell (mesh not 2:1 aligned)", "", b);
            const int r = record_for(g, b);
            recs[r].type[slot] = type;
            slotlists[6 * r + slot] = fl;
            recgrpblock[6 * r + slot] = emitter_block;
            if (kind == NB_FINER) recs[r].finer |= 1 << slot;
            for (int i = 1; i <= fl[0]; ++i) {   // each listed face must touch this cell
              const LinkFaceDev& F = H.faces[fl[i]];
              if (F.iL != g && F.iR != g) return fail(-1, "link face bookkeeping error at block %s%lld", "", b);
            }
          }
      }
  }
  H.linkcells.resize(recs.size());
  H.slot_faces.reserve(2 * H.faces.size());   // each link face is listed by both of its cells
  for (size_t r = 0; r < recs.size(); ++r) {
    LinkCellDev R = recs[r];
    std::pair<int, int> grp[6];   // (emitting block, slot), at most one per slot
    int ng = 0;
    for (int s = 0; s < 6; ++s) {
      const SlotList& sl = slotlists[6 * r + s];
      R.start[s] = (int)H.slot_faces.size();
      R.count[s] = sl[0];
      for (int i = 1; i <= sl[0]; ++i) H.slot_faces.push_back(sl[i]);
      if (R.type[s] == 2) grp[ng++] = {recgrpblock[6 * r + s], s};
    }
    std::sort(grp, grp + ng);
    R.ngrp = ng;
    R.npre = 0;
    for (int q = 0; q < 6; ++q) R.grp[q] = -1;
    for (int q = 0; q < ng; ++q) {
      if (q > 0 && grp[q].first == grp[q - 1].first)
        return fail(-1, "a cell touches block %s%lld through two faces", "", grp[q].first);
      R.grp[q] = grp[q].second;
      if (grp[q].first < recblock[r]) R.npre++;
    }
    H.linkcells[r] = R;
  }
  // every link face must be gathered by both of its cells (fails when a fine
  // block is not aligned to the coarse cells it touches)
  {
    std::vector<int> seen(H.faces.size(), 0);
    for (int f : H.slot_faces) seen[f]++;
    for (size_t f = 0; f < seen.size(); ++f)
      if (seen[f] != 2)
        return fail(-1, "link face %s%lld is not shared by two cells: fine blocks must be aligned to coarse cells", "",
                    (long long)f);
  }
  // link-cell maps per block face
  for (int b = 0; b < nb; ++b) {
    BlockDev& bk = H.blocks[b];
    bk.haslinks = 0;
    for (int d = 0; d < rank; ++d)
      for (int side = 0; side < 2; ++side) {
        const int slot = 2 * d + side;
        if (bk.repoff[slot] < 0) continue;
        const int ps = plane_size(bk.n, d);
        std::vector<int> pm(ps, -1);
        bool any = false;
        const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
        for (int v = 0; v < bk.n[e2]; ++v)
          for (int u = 0; u < bk.n[e1]; ++u) {
            long long ijk[3];
            ijk[d] = side ? bk.n[d] - 1 : 0;
            ijk[e1] = u;
            ijk[e2] = v;
            const int r = recof.find(T.index(b, ijk));
            if (r < 0) continue;
            pm[plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2])] = r;
            any = true;
          }
        if (!any) continue;
        if ((long long)H.lcmap.size() + ps > 0x7fffffffLL) return fail(-1, "mesh too large%s", "");
        bk.lcoff[slot] = (int)H.lcmap.size();
        H.lcmap.insert(H.lcmap.end(), pm.begin(), pm.end());
        bk.haslinks = 1;
      }
  }
  // a record may belong to a cell that is on a linked face only through
  // another slot (edge/corner cells): make sure every record is reachable
  for (int rec = 0; rec < (int)reccell.size(); ++rec) {
    long long ijk[3];
    const int b = T.locate(reccell[rec], ijk);
    const BlockDev& bk = H.blocks[b];
    bool found = false;
    for (int d = 0; d < rank && !found; ++d)
      for (int side = 0; side < 2 && !found; ++side) {
        const int slot = 2 * d + side;
        if (ijk[d] != (side ? bk.n[d] - 1 : 0) || bk.lcoff[slot] < 0) continue;
        found = H.lcmap[bk.lcoff[slot] + plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2])] == rec;
      }
    if (!found) return fail(-1, "unreachable link-cell record in block %s%lld", "", b);
  }

  // ---- flux-register lists (coarse cells with finer neighbours) ----
  H.regs.assign(T.maxlevel + 1, RegList{});
  {
    std::vector<std::pair<long long, int>> order(reccell.size());
    for (size_t r = 0; r < reccell.size(); ++r) order[r] = {reccell[r], (int)r};
    std::sort(order.begin(), order.end());
    for (auto& kv : order) {
      const LinkCellDev& R = H.linkcells[kv.second];
      if (!R.finer) continue;
      const int b = recblock[kv.second];
      RegList& L = H.regs[T.level[b]];
      L.cell.push_back((int)kv.first);
      L.lc.push_back(kv.second);
      L.block.push_back(b);
    }
  }

  // ---- tiles ----
  std::vector<TileDev> base;
  std::vector<long long> cost;
  H.block_tile_start.assign(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    const BlockDev& bk = H.blocks[b];
    H.block_tile_start[b] = (int)base.size();
    int tx, ty;
    if (rank == 1) {
      tx = balanced(bk.n[0], 64);
      ty = 1;
    } else {
      tx = balanced(bk.n[0], std::min(kMaxE, std::max(4, max_tx)));
      int maxty = kLanes / tx;
      if (maxty > kMaxE) maxty = kMaxE;
      ty = balanced(bk.n[1], maxty);
    }
    if (tx * ty > kLanes || (tx + 4) * (ty + 4) > kMaxW || ty * (tx + 2) + (rank > 1 ? (ty + 2) * tx : 0) > kMaxFS ||
        (rank > 1 && tx > kMaxE) || ty > kMaxE)
      return fail(-1, "no tile shape fits block %s%lld", "", b);
    int zc = rank == 3 ? (zchunk > 0 ? zchunk : bk.n[2]) : 1;
    if (zc > bk.n[2]) zc = bk.n[2];
    if (rank == 3) zc = balanced(bk.n[2], zc);
    for (int z0 = 0; z0 < bk.n[2]; z0 += zc)
      for (int y0 = 0; y0 < bk.n[1]; y0 += ty)
        for (int x0 = 0; x0 < bk.n[0]; x0 += tx) {
          TileDev t{};
          t.block = b;
          t.x0 = x0;
          t.y0 = y0;
          t.z0 = z0;
          t.z1 = std::min(bk.n[2], z0 + zc);
          t.tx = tx;
          t.ty = ty;
          base.push_back(t);
          const int ex = std::min(tx, bk.n[0] - x0), ey = std::min(ty, bk.n[1] - y0);
          cost.push_back((long long)(t.z1 - t.z0 + (rank == 3 ? 4 : 0)) * (ex + 4) * (ey + (rank > 1 ? 4 : 0)));
        }
  }
  H.block_tile_start[nb] = (int)base.size();
  const int nt = (int)base.size();
  auto sorted_list = [&](int lev) {
    std::vector<int> idx;
    for (int i = 0; i < nt; ++i)
      if (lev < 0 || T.level[base[i].block] == lev) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    return idx;
  };
  H.tiles = base;
  H.tile_list_off.clear();
  H.tile_list_off.push_back(nt);   // list 0: all, starts after the block-order copy
  for (int l = -1; l <= T.maxlevel; ++l) {
    for (int i : sorted_list(l)) H.tiles.push_back(base[i]);
    H.tile_list_off.push_back((int)H.tiles.size());
  }

  // ---- CFL block lists ----
  H.cfl_blocks.clear();
  H.cfl_prefix.clear();
  H.cfl_list_off.clear();
  for (int l = -1; l <= T.maxlevel; ++l) {
    H.cfl_list_off.push_back((int)H.cfl_blocks.size());
    long long acc = 0;
    H.cfl_prefix.push_back(0);
    for (int b = 0; b < nb; ++b) {
      if (l >= 0 && T.level[b] != l) continue;
      H.cfl_blocks.push_back(b);
      acc += T.off[b + 1] - T.off[b];
      H.cfl_prefix.push_back(acc);
    }
  }
  H.cfl_list_off.push_back((int)H.cfl_blocks.size());
  return 0;
}

static int make_topo(Topo& T, int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0, const int32_t* level,
                     const int32_t* dims, const int64_t* origin, const int32_t* bc) {
  if (rank < 1 || rank > 3) return fail(-1, "rank must be 1..3%s (got %lld)", "", rank);
  if (nblocks < 1) return fail(-1, "need at least one block%s (got %lld)", "", nblocks);
  if (maxlevel < 0 || maxlevel > 20) return fail(-1, "maxlevel out of range%s (%lld)", "", maxlevel);
  if (!h0 || !level || !dims || !origin || !bc) return fail(-1, "null argument%s", "");
  T.rank = rank;
  T.nb = nblocks;
  T.maxlevel = maxlevel;
  for (int d = 0; d < 3; ++d) {
    T.h0[d] = d < rank ? h0[d] : 1.0;
    if (d < rank && !(h0[d] > 0.0)) return fail(-1, "h0 must be positive%s (axis %lld)", "", d);
  }
  T.level.assign(level, level + nblocks);
  T.dims.assign(dims, dims + 3 * nblocks);
  T.origin.assign(origin, origin + 3 * nblocks);
  T.bc.assign(bc, bc + 6 * nblocks);
  T.off.assign(nblocks + 1, 0);
  for (int b = 0; b < nblocks; ++b) {
    if (T.level[b] < 0 || T.level[b] > maxlevel) return fail(-1, "block %s%lld: level out of range", "", b);
    long long c = 1;
    for (int d = 0; d < 3; ++d) {
      if (d >= rank) {
        T.dims[3 * b + d] = 1;
        T.origin[3 * b + d] = 0;
        T.bc[6 * b + 2 * d] = T.bc[6 * b + 2 * d + 1] = BC_OUTFLOW;
        continue;
      }
      const int n = T.dims[3 * b + d];
      if (n < 1 || n > (1 << 20)) return fail(-1, "block %s%lld: cells per axis out of range", "", b);
      if (T.origin[3 * b + d] % T.span(b)) return fail(-1, "block %s%lld: origin not aligned to its cells", "", b);
      for (int s = 0; s < 2; ++s)
        if (T.bc[6 * b + 2 * d + s] < 0 || T.bc[6 * b + 2 * d + s] > 2)
          return fail(-1, "block %s%lld: unknown boundary code", "", b);
      if ((T.bc[6 * b + 2 * d] == BC_PERIODIC) != (T.bc[6 * b + 2 * d + 1] == BC_PERIODIC))
        return fail(-1, "block %s%lld: periodic must be set on both sides of an axis", "", b);
      if (T.bc[6 * b + 2 * d] == BC_PERIODIC && n < 2)
        return fail(-1, "block %s%lld: a periodic axis needs at least 2 cells", "", b);
      c *= n;
    }
    T.off[b + 1] = T.off[b] + c;
  }
  if (T.off[nblocks] > 0x7fffffffLL) return fail(-1, "mesh too large%s (%lld cells)", "", T.off[nblocks]);
  return 0;
}

// ----------------------------------------------------------------------------
// profile-independent kernels (compiled with -ffp-contract=off)

// _reflux_capture! (src/reflux.jl:70-89) for every coarse cell of one level:
// delta += (s·F)·area when the coarse cell is the face's left cell, else
// delta += ((−s)·F)·area, over its finer groups in the walk order.
template <typename T>
__global__ void __launch_bounds__(256) capture_kernel(const int* __restrict__ cells, const int* __restrict__ lcs,
                                                      int nreg, BlocksDev B, const T* __restrict__ F, double sd,
                                                      T* __restrict__ delta) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nreg) return;
  const int c = cells[e];
  const LinkCellDev R = B.linkcells[lcs[e]];
  const T s = T(sd);
  T acc[5];
  for (int f = 0; f < 5; ++f) acc[f] = delta[5 * (long long)e + f];
  for (int q = 0; q < R.ngrp; ++q) {
    const int slot = R.grp[q];
    if (!((R.finer >> slot) & 1)) continue;
    for (int i = 0; i < R.count[slot]; ++i) {
      const int fi = B.slot_faces[R.start[slot] + i];
      const LinkFaceDev L = B.faces[fi];
      const T aT = T(L.area);
      const T* Fv = F + 5 * (long long)fi;
      if (c == L.iL) {
        for (int f = 0; f < 5; ++f) acc[f] = acc[f] + (s * Fv[f]) * aT;
      } else {
        for (int f = 0; f < 5; ++f) acc[f] = acc[f] + ((-s) * Fv[f]) * aT;
      }
    }
  }
  for (int f = 0; f < 5; ++f) delta[5 * (long long)e + f] = acc[f];
}

// _reflux_apply! (src/reflux.jl:94-103): U −= (1/V)·delta, then clear delta.
template <typename T>
__global__ void __launch_bounds__(256) reflux_kernel(const int* __restrict__ cells, const int* __restrict__ blk,
                                                     int nreg, BlocksDev B, T* __restrict__ U,
                                                     T* __restrict__ delta) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nreg) return;
  const long long c = cells[e];
  const T invV = T(1) / T(B.blocks[blk[e]].V);
  for (int f = 0; f < 5; ++f) {
    const long long i = f * B.ncell + c;
    U[i] = U[i] - invV * delta[5 * (long long)e + f];
    delta[5 * (long long)e + f] = T(0);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) zero_kernel(T* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = T(0);
}

// Σ U·V per workgroup over all cells (src/diagnostics.jl:16-31), in double.
template <typename T>
__global__ void __launch_bounds__(256) block_totals_kernel(const T* __restrict__ U, BlocksDev B,
                                                           const int* __restrict__ blist,
                                                           const long long* __restrict__ prefix, int nb,
                                                           double* __restrict__ partials) {
  __shared__ double red[16];
  const long long total = prefix[nb];
  double s[5] = {0, 0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const BlockDev& bk = B.blocks[blist[lo]];
    const long long c = bk.off + (i - prefix[lo]);
    for (int f = 0; f < 5; ++f) s[f] += double(U[f * B.ncell + c]) * bk.V;
  }
  for (int f = 0; f < 5; ++f) {
    const double t = block_sum<256>(s[f], red);
    if (threadIdx.x == 0) partials[5 * blockIdx.x + f] = t;
  }
}

}  // namespace vh

using namespace vh;

// ----------------------------------------------------------------------------
// context

struct DevReg {
  int n = 0;
  int* cells = nullptr;
  int* lcs = nullptr;
  int* blk = nullptr;
  void* delta = nullptr;
};

struct vh_blocks {
  Topo T;
  HostTables H;
  int dtype = VH_F64, mode = VH_EXACT, device = 0;
  size_t esize = 8;
  int zchunk = 8;
  int max_tx = 32;        // widest tile along x
  int max_wgs = 0;        // persistent workgroups (0: 3 per CU)
  int ncu = 256;
  int stage_wgs = 0;      // workgroups of the last all-blocks stage-2 launch
  int use_graph = 1;
  double gamma = 1.4;
  // device tables
  BlockDev* d_blocks = nullptr;
  TileDev* d_tiles = nullptr;
  int* d_repmap = nullptr;
  int* d_lcmap = nullptr;
  LinkCellDev* d_linkcells = nullptr;
  int* d_slot_faces = nullptr;
  LinkFaceDev* d_faces = nullptr;
  void* d_flux = nullptr;
  int* d_cfl_blocks = nullptr;
  long long* d_cfl_prefix = nullptr;
  double* d_ind = nullptr;   // per-block refinement indicator maximum
  std::vector<DevReg> regs;
  // state
  void* U = nullptr;
  void* U1 = nullptr;
  double* partials = nullptr;
  int npartials = 0;
  double* clock = nullptr;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_np = -1;
};

static const int kCflWgs = 2048;

static BlocksDev blocks_dev(const vh_blocks* m) {
  BlocksDev B{};
  B.blocks = m->d_blocks;
  B.tiles = m->d_tiles;
  B.repmap = m->d_repmap;
  B.lcmap = m->d_lcmap;
  B.linkcells = m->d_linkcells;
  B.slot_faces = m->d_slot_faces;
  B.faces = m->d_faces;
  B.fluxbuf = m->d_flux;
  B.ncell = m->T.off[m->T.nb];
  B.gamma = m->gamma;
  B.igm1 = 1.0 / (m->gamma - 1.0);
  B.ntiles = 0;
  B.nfaces = (int)m->H.faces.size();
  return B;
}

static void drop_graph(vh_blocks* m) {
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  m->graph = nullptr;
  m->graph_np = -1;
}

template <typename V>
static int upload_vec(V** dst, const std::vector<V>& src) {
  *dst = nullptr;
  const size_t bytes = sizeof(V) * (src.empty() ? 1 : src.size());
  VH_HIP(hipMalloc((void**)dst, bytes));
  if (!src.empty()) VH_HIP(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return 0;
}

// Tile list l: 0 all blocks, 1 + L level L.
static int list_wgs(const vh_blocks* m, int l) {
  const int n = m->H.tile_list_off[l + 1] - m->H.tile_list_off[l];
  const int cap = m->max_wgs > 0 ? m->max_wgs : 3 * m->ncu;
  return n < cap ? n : cap;
}

static void enqueue_link(vh_blocks* m, const void* Us) {
  const BlocksDev B = blocks_dev(m);
  if (m->mode == VH_FAST) fast::launch_link_flux(m->dtype, B.nfaces, m->stream, Us, B, m->d_flux);
  else exact::launch_link_flux(m->dtype, B.nfaces, m->stream, Us, B, m->d_flux);
}

static void enqueue_tiles(vh_blocks* m, bool stage2, int l, const void* Us, const void* U0, void* Uout) {
  BlocksDev B = blocks_dev(m);
  B.tiles = m->d_tiles + m->H.tile_list_off[l];
  B.ntiles = m->H.tile_list_off[l + 1] - m->H.tile_list_off[l];
  const int wgs = list_wgs(m, l);
  const double* dtp = m->clock + 1;
  if (m->mode == VH_FAST) fast::launch_block_stage(m->dtype, m->T.rank, stage2, wgs, m->stream, Us, U0, Uout, dtp, m->partials, B);
  else exact::launch_block_stage(m->dtype, m->T.rank, stage2, wgs, m->stream, Us, U0, Uout, dtp, m->partials, B);
  if (l == 0 && stage2) m->stage_wgs = wgs;
}

// One launch per block (the comparison loop): tiles of block b only.
static void enqueue_tiles_per_block(vh_blocks* m, bool stage2, const void* Us, const void* U0, void* Uout) {
  const double* dtp = m->clock + 1;
  for (int b = 0; b < m->T.nb; ++b) {
    BlocksDev B = blocks_dev(m);
    const int t0 = m->H.block_tile_start[b], t1 = m->H.block_tile_start[b + 1];
    B.tiles = m->d_tiles + t0;
    B.ntiles = t1 - t0;
    if (m->mode == VH_FAST)
      fast::launch_block_stage(m->dtype, m->T.rank, stage2, t1 - t0, m->stream, Us, U0, Uout, dtp, m->partials + t0, B);
    else
      exact::launch_block_stage(m->dtype, m->T.rank, stage2, t1 - t0, m->stream, Us, U0, Uout, dtp, m->partials + t0, B);
  }
  if (stage2) m->stage_wgs = m->H.block_tile_start[m->T.nb];
}

static void enqueue_cfl(vh_blocks* m, int level) {
  const int l = level < 0 ? 0 : level + 1;
  const int b0 = m->H.cfl_list_off[l], b1 = m->H.cfl_list_off[l + 1];
  const BlocksDev B = blocks_dev(m);
  const int* bl = m->d_cfl_blocks + b0;
  const long long* pf = m->d_cfl_prefix + b0 + l;
  if (b1 == b0) {   // no cell on this level: max = 0, dt = +Inf
    hipLaunchKernelGGL(zero_kernel<double>, dim3((kCflWgs + 255) / 256), dim3(256), 0, m->stream, m->partials,
                       (long long)kCflWgs);
    return;
  }
  if (m->mode == VH_FAST) fast::launch_block_cfl(m->dtype, m->T.rank, kCflWgs, m->stream, m->U, B, bl, pf, b1 - b0, m->partials);
  else exact::launch_block_cfl(m->dtype, m->T.rank, kCflWgs, m->stream, m->U, B, bl, pf, b1 - b0, m->partials);
}

static void enqueue_capture(vh_blocks* m, int cl, double scale) {
  const DevReg& R = m->regs[cl];
  if (R.n == 0) return;
  const BlocksDev B = blocks_dev(m);
  const unsigned g = (unsigned)((R.n + 255) / 256);
  if (m->dtype == VH_F64)
    hipLaunchKernelGGL(capture_kernel<double>, dim3(g), dim3(256), 0, m->stream, R.cells, R.lcs, R.n, B,
                       (const double*)m->d_flux, scale, (double*)R.delta);
  else
    hipLaunchKernelGGL(capture_kernel<float>, dim3(g), dim3(256), 0, m->stream, R.cells, R.lcs, R.n, B,
                       (const float*)m->d_flux, scale, (float*)R.delta);
}

static void enqueue_reflux(vh_blocks* m, int cl) {
  const DevReg& R = m->regs[cl];
  if (R.n == 0) return;
  const BlocksDev B = blocks_dev(m);
  const unsigned g = (unsigned)((R.n + 255) / 256);
  if (m->dtype == VH_F64)
    hipLaunchKernelGGL(reflux_kernel<double>, dim3(g), dim3(256), 0, m->stream, R.cells, R.blk, R.n, B,
                       (double*)m->U, (double*)R.delta);
  else
    hipLaunchKernelGGL(reflux_kernel<float>, dim3(g), dim3(256), 0, m->stream, R.cells, R.blk, R.n, B,
                       (float*)m->U, (float*)R.delta);
}

// Register set of a level step: coarse level and sign (src/driver.jl:523-551).
struct ActiveRegs {
  int n = 0;
  int cl[2];
  double sign[2];
};

// One SSP-RK2 step reading dt from clock[1] (src/driver.jl:333-378). level < 0:
// every leaf, fused CFL partials of the new state; level >= 0: that level's
// leaves only (the other leaves are read, not written).
static void enqueue_step(vh_blocks* m, int level, double dt_for_regs, const ActiveRegs& regs, bool per_block = false,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  const size_t bytes = (size_t)5 * m->T.off[m->T.nb] * m->esize;
  const int l = level < 0 ? 0 : level + 1;
  if (level >= 0) (void)hipMemcpyAsync(m->U1, m->U, bytes, hipMemcpyDeviceToDevice, m->stream);
  for (int stage = 0; stage < 2; ++stage) {
    const void* Us = stage ? m->U1 : m->U;
    enqueue_link(m, Us);
    for (int r = 0; r < regs.n; ++r) enqueue_capture(m, regs.cl[r], regs.sign[r] * (0.5 * dt_for_regs));
    if (per_block) {
      if (stage == 0) enqueue_tiles_per_block(m, false, m->U, m->U, m->U1);
      else enqueue_tiles_per_block(m, true, m->U1, m->U, m->U);
    } else {
      if (stage == 0) {
        enqueue_tiles(m, false, l, m->U, m->U, m->U1);
      } else {
        if (ev0) (void)hipEventRecord(ev0, m->stream);
        enqueue_tiles(m, true, l, m->U1, m->U, m->U);
        if (ev1) (void)hipEventRecord(ev1, m->stream);
      }
    }
  }
}

static int get_clock(vh_blocks* m, double clk[4]) {
  VH_HIP(hipMemcpyAsync(clk, m->clock, 4 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

static int compute_dt_sync(vh_blocks* m, double cfl, int level, double* dt) {
  enqueue_cfl(m, level);
  enqueue_dt_finalize(m->stream, m->partials, kCflWgs, cfl, 0.0, 0, m->clock);
  VH_HIP(hipGetLastError());
  double clk[4];
  int rc = get_clock(m, clk);
  if (rc) return rc;
  *dt = clk[1];
  return 0;
}

// Single-rate step body (dt finalize from the partials of the previous stage
// 2, step, clock advance), optionally replayed from a captured graph.
static int single_rate_step(vh_blocks* m, double cfl, double tfinal, int clip, int np) {
  const ActiveRegs none{};
  if (m->use_graph) {
    if (!m->graph || m->graph_np != np) {
      drop_graph(m);
      hipGraph_t gr = nullptr;
      VH_HIP(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
      enqueue_dt_finalize(m->stream, m->partials, np, cfl, tfinal, clip, m->clock);
      enqueue_step(m, -1, 0.0, none);
      enqueue_advance_clock(m->stream, m->clock);
      VH_HIP(hipStreamEndCapture(m->stream, &gr));
      hipError_t e = hipGraphInstantiate(&m->graph, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) { m->graph = nullptr; return fail(-2, "graph instantiation failed: %s", hipGetErrorString(e)); }
      m->graph_np = np;
    }
    VH_HIP(hipGraphLaunch(m->graph, m->stream));
  } else {
    enqueue_dt_finalize(m->stream, m->partials, np, cfl, tfinal, clip, m->clock);
    enqueue_step(m, -1, 0.0, none);
    enqueue_advance_clock(m->stream, m->clock);
  }
  VH_HIP(hipGetLastError());
  return 0;
}

static int level_present_max(const vh_blocks* m) {
  int l = 0;
  for (int b = 0; b < m->T.nb; ++b) l = std::max(l, m->T.level[b]);
  return l;
}

// evolve_level! (src/driver.jl:553-578): n substeps of the level, each followed
// by the finer level's subcycles and the reflux of this level's register.
static int evolve_level(vh_blocks* m, int level, double dt_target, double cfl, bool parent) {
  double dt_stable = 0.0;
  int rc = compute_dt_sync(m, cfl, level, &dt_stable);
  if (rc) return rc;
  if (!std::isfinite(dt_stable)) return 0;
  const double q = dt_target / dt_stable * (1 - 1e-12);
  long long ns = (long long)std::ceil(q);
  if (ns < 1) ns = 1;
  const double dt_sub = dt_target / (double)ns;
  const bool has_finer = level < level_present_max(m);
  ActiveRegs regs;
  if (parent) { regs.cl[regs.n] = level - 1; regs.sign[regs.n++] = 1.0; }
  if (has_finer) { regs.cl[regs.n] = level; regs.sign[regs.n++] = -1.0; }
  for (long long s = 0; s < ns; ++s) {
    enqueue_set_dt(m->stream, m->clock, dt_sub);
    enqueue_step(m, level, dt_sub, regs);
    VH_HIP(hipGetLastError());
    if (has_finer) {
      rc = evolve_level(m, level + 1, dt_sub, cfl, true);
      if (rc) return rc;
      enqueue_reflux(m, level);
      VH_HIP(hipGetLastError());
    }
  }
  return 0;
}

static int alloc_partials(vh_blocks* m) {
  int need = kCflWgs;
  need = std::max(need, m->H.block_tile_start[m->T.nb]);
  need = std::max(need, m->max_wgs > 0 ? m->max_wgs : 3 * m->ncu);
  if (need > m->npartials) {
    if (m->partials) VH_HIP(hipFree(m->partials));
    m->partials = nullptr;
    VH_HIP(hipMalloc(&m->partials, sizeof(double) * need * 5));
    m->npartials = need;
  }
  return 0;
}

static void free_tables(vh_blocks* m) {
  (void)hipFree(m->d_blocks);
  (void)hipFree(m->d_tiles);
  (void)hipFree(m->d_repmap);
  (void)hipFree(m->d_lcmap);
  (void)hipFree(m->d_linkcells);
  (void)hipFree(m->d_slot_faces);
  (void)hipFree(m->d_faces);
  (void)hipFree(m->d_flux);
  (void)hipFree(m->d_cfl_blocks);
  (void)hipFree(m->d_cfl_prefix);
  (void)hipFree(m->d_ind);
  for (DevReg& r : m->regs) {
    (void)hipFree(r.cells);
    (void)hipFree(r.lcs);
    (void)hipFree(r.blk);
    (void)hipFree(r.delta);
  }
  m->regs.clear();
  m->d_blocks = nullptr;
  m->d_tiles = nullptr;
  m->d_repmap = m->d_lcmap = m->d_slot_faces = m->d_cfl_blocks = nullptr;
  m->d_linkcells = nullptr;
  m->d_faces = nullptr;
  m->d_flux = nullptr;
  m->d_cfl_prefix = nullptr;
  m->d_ind = nullptr;
}

static int upload_tables(vh_blocks* m) {
  const HostTables& H = m->H;
  int rc = 0;
  if ((rc = upload_vec(&m->d_blocks, H.blocks))) return rc;
  if ((rc = upload_vec(&m->d_tiles, H.tiles))) return rc;
  if ((rc = upload_vec(&m->d_repmap, H.repmap))) return rc;
  if ((rc = upload_vec(&m->d_lcmap, H.lcmap))) return rc;
  if ((rc = upload_vec(&m->d_linkcells, H.linkcells))) return rc;
  if ((rc = upload_vec(&m->d_slot_faces, H.slot_faces))) return rc;
  if ((rc = upload_vec(&m->d_faces, H.faces))) return rc;
  if ((rc = upload_vec(&m->d_cfl_blocks, H.cfl_blocks))) return rc;
  if ((rc = upload_vec(&m->d_cfl_prefix, H.cfl_prefix))) return rc;
  VH_HIP(hipMalloc(&m->d_flux, m->esize * 5 * std::max<size_t>(1, H.faces.size())));
  VH_HIP(hipMalloc(&m->d_ind, sizeof(double) * m->T.nb));
  m->regs.assign(H.regs.size(), DevReg{});
  for (size_t l = 0; l < H.regs.size(); ++l) {
    DevReg& R = m->regs[l];
    R.n = (int)H.regs[l].cell.size();
    if ((rc = upload_vec(&R.cells, H.regs[l].cell))) return rc;
    if ((rc = upload_vec(&R.lcs, H.regs[l].lc))) return rc;
    if ((rc = upload_vec(&R.blk, H.regs[l].block))) return rc;
    const long long nv = 5LL * std::max(1, R.n);
    VH_HIP(hipMalloc(&R.delta, m->esize * nv));
    const unsigned g = (unsigned)((nv + 255) / 256);
    if (m->dtype == VH_F64) hipLaunchKernelGGL(zero_kernel<double>, dim3(g), dim3(256), 0, m->stream, (double*)R.delta, nv);
    else hipLaunchKernelGGL(zero_kernel<float>, dim3(g), dim3(256), 0, m->stream, (float*)R.delta, nv);
    VH_HIP(hipGetLastError());
  }
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

extern "C" {

int vh_blocks_describe(int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0, const int32_t* level,
                       const int32_t* dims, const int64_t* origin, const int32_t* bc, int32_t zchunk,
                       int64_t* counts6) {
  if (!counts6) return fail(-1, "null argument%s", "");
  Topo T;
  int rc = make_topo(T, rank, nblocks, maxlevel, h0, level, dims, origin, bc);
  if (rc) return rc;
  HostTables H;
  rc = build_tables(T, zchunk, H);
  if (rc) return rc;
  long long nreg = 0;
  for (const RegList& r : H.regs) nreg += (long long)r.cell.size();
  counts6[0] = T.off[T.nb];
  counts6[1] = H.block_tile_start[T.nb];
  counts6[2] = (int64_t)H.faces.size();
  counts6[3] = (int64_t)H.linkcells.size();
  counts6[4] = nreg;
  counts6[5] = (int64_t)H.slot_faces.size();
  return 0;
}

int vh_blocks_create(vh_blocks** out, int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0,
                     double gamma, const int32_t* level, const int32_t* dims, const int64_t* origin,
                     const int32_t* bc, int32_t dtype) {
  if (!out) return fail(-1, "null argument%s", "");
  *out = nullptr;
  if (dtype != VH_F64 && dtype != VH_F32) return fail(-1, "dtype must be VH_F64 or VH_F32%s (got %lld)", "", dtype);
  if (!(gamma > 1.0)) return fail(-1, "gamma must exceed 1%s", "");
  vh_blocks* m = new vh_blocks();
  int rc = make_topo(m->T, rank, nblocks, maxlevel, h0, level, dims, origin, bc);
  if (!rc) rc = build_tables(m->T, m->zchunk, m->H, m->max_tx);
  if (rc) { delete m; return rc; }
  m->dtype = dtype;
  m->esize = dtype == VH_F64 ? 8 : 4;
  m->gamma = gamma;
  if (hipGetDevice(&m->device) != hipSuccess) { delete m; return fail(-2, "no HIP device%s", ""); }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, m->device) == hipSuccess && ncu > 0)
    m->ncu = ncu;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return fail(-2, "stream creation failed%s", "");
  }
  const long long n = m->T.off[m->T.nb];
  const size_t bytes = (size_t)5 * n * m->esize;
  if (hipMalloc(&m->U, bytes) != hipSuccess || hipMalloc(&m->U1, bytes) != hipSuccess ||
      hipMalloc(&m->clock, 4 * sizeof(double)) != hipSuccess) {
    vh_blocks_destroy(m);
    return fail(-2, "device allocation of %s%lld bytes failed", "", (long long)(2 * bytes));
  }
  rc = upload_tables(m);
  if (!rc) rc = alloc_partials(m);
  if (rc) { vh_blocks_destroy(m); return rc; }
  VH_HIP(hipMemsetAsync(m->clock, 0, 4 * sizeof(double), m->stream));
  VH_HIP(hipMemsetAsync(m->U, 0, bytes, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  *out = m;
  return 0;
}

int vh_blocks_destroy(vh_blocks* m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  drop_graph(m);
  free_tables(m);
  (void)hipFree(m->U);
  (void)hipFree(m->U1);
  (void)hipFree(m->partials);
  (void)hipFree(m->clock);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int64_t vh_blocks_ncell(const vh_blocks* m) { return m ? m->T.off[m->T.nb] : -1; }

int vh_blocks_set_mode(vh_blocks* m, int32_t mode) {
  if (!m) return fail(-1, "null argument%s", "");
  if (mode != VH_EXACT && mode != VH_FAST) return fail(-1, "mode must be VH_EXACT or VH_FAST%s (got %lld)", "", mode);
  VH_HIP(hipStreamSynchronize(m->stream));
  if (mode != m->mode) drop_graph(m);
  m->mode = mode;
  return 0;
}

int vh_blocks_upload(vh_blocks* m, const void* U) {
  if (!m || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipMemcpyAsync(m->U, U, (size_t)5 * m->T.off[m->T.nb] * m->esize, hipMemcpyHostToDevice, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_download(vh_blocks* m, void* U) {
  if (!m || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipMemcpyAsync(U, m->U, (size_t)5 * m->T.off[m->T.nb] * m->esize, hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_compute_dt(vh_blocks* m, double cfl, int32_t level, double* dt) {
  if (!m || !dt) return fail(-1, "null argument%s", "");
  if (level > m->T.maxlevel) return fail(-1, "level out of range%s (%lld)", "", level);
  VH_HIP(hipSetDevice(m->device));
  return compute_dt_sync(m, cfl, level, dt);
}

int vh_blocks_step(vh_blocks* m, double dt, int32_t level) {
  if (!m) return fail(-1, "null argument%s", "");
  if (level > m->T.maxlevel) return fail(-1, "level out of range%s (%lld)", "", level);
  VH_HIP(hipSetDevice(m->device));
  enqueue_set_dt(m->stream, m->clock, dt);
  enqueue_step(m, level < 0 ? -1 : level, dt, ActiveRegs());
  VH_HIP(hipGetLastError());
  return 0;
}

int vh_blocks_run(vh_blocks* m, double cfl, double tfinal, int64_t max_steps, int32_t subcycle, double* t_io,
                  int64_t* steps) {
  if (!m || !t_io) return fail(-1, "null argument%s", "");
  const bool clip = tfinal > 0.0;
  if (!clip && max_steps <= 0) return fail(-1, "need tfinal > 0 or max_steps > 0%s", "");
  if (subcycle && !clip) return fail(-1, "subcycling needs tfinal > 0%s", "");
  VH_HIP(hipSetDevice(m->device));
  double clk[4] = {*t_io, 0.0, 0.0, 0.0};
  VH_HIP(hipMemcpyAsync(m->clock, clk, sizeof clk, hipMemcpyHostToDevice, m->stream));
  int64_t k = 0;
  double t = *t_io;
  int rc = 0;
  if (subcycle) {
    // evolve! with subcycling (src/driver.jl:581-605): root dt from level 0.
    while (t < tfinal * (1 - 1e-12) && (max_steps <= 0 || k < max_steps)) {
      double dt = 0.0;
      if ((rc = compute_dt_sync(m, cfl, 0, &dt))) return rc;
      dt = std::fmin(dt, tfinal - t);
      if ((rc = evolve_level(m, 0, dt, cfl, false))) return rc;
      t += dt;
      ++k;
    }
    clk[0] = t;
    clk[3] = (double)k;
    VH_HIP(hipMemcpyAsync(m->clock, clk, sizeof clk, hipMemcpyHostToDevice, m->stream));
    VH_HIP(hipStreamSynchronize(m->stream));
  } else {
    enqueue_cfl(m, -1);
    int np = kCflWgs;
    while (true) {
      if (max_steps > 0 && k >= max_steps) break;
      if (clip && !(t < tfinal * (1 - 1e-12))) break;
      if ((rc = single_rate_step(m, cfl, tfinal, clip ? 1 : 0, np))) return rc;
      np = list_wgs(m, 0);
      ++k;
      if (clip) {
        if ((rc = get_clock(m, clk))) return rc;
        t = clk[0];
      }
    }
    if ((rc = get_clock(m, clk))) return rc;
    t = clk[0];
  }
  *t_io = t;
  if (steps) *steps = k;
  return 0;
}

int vh_blocks_totals(vh_blocks* m, double* out5) {
  if (!m || !out5) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  const BlocksDev B = blocks_dev(m);
  const int nb = m->H.cfl_list_off[1] - m->H.cfl_list_off[0];
  if (m->dtype == VH_F64)
    hipLaunchKernelGGL(block_totals_kernel<double>, dim3(kCflWgs), dim3(256), 0, m->stream, (const double*)m->U, B,
                       m->d_cfl_blocks, m->d_cfl_prefix, nb, m->partials);
  else
    hipLaunchKernelGGL(block_totals_kernel<float>, dim3(kCflWgs), dim3(256), 0, m->stream, (const float*)m->U, B,
                       m->d_cfl_blocks, m->d_cfl_prefix, nb, m->partials);
  VH_HIP(hipGetLastError());
  std::vector<double> h(5 * kCflWgs);
  VH_HIP(hipMemcpyAsync(h.data(), m->partials, sizeof(double) * h.size(), hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  for (int f = 0; f < 5; ++f) out5[f] = 0.0;
  for (int b = 0; b < kCflWgs; ++b)
    for (int f = 0; f < 5; ++f) out5[f] += h[5 * b + f];
  return 0;
}

int vh_blocks_set_launch(vh_blocks* m, int32_t zchunk, int32_t max_workgroups, int32_t use_graph) {
  if (!m) return fail(-1, "null argument%s", "");
  if (zchunk < 0 || max_workgroups < 0) return fail(-1, "bad launch argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipStreamSynchronize(m->stream));
  drop_graph(m);
  m->use_graph = use_graph ? 1 : 0;
  m->max_wgs = max_workgroups;
  if (zchunk != m->zchunk) {
    HostTables H;
    int rc = build_tables(m->T, zchunk, H, m->max_tx);
    if (rc) return rc;
    free_tables(m);
    m->H = std::move(H);
    m->zchunk = zchunk;
    rc = upload_tables(m);
    if (rc) return rc;
  }
  return alloc_partials(m);
}

int vh_blocks_set_tiling(vh_blocks* m, int32_t max_tile_x) {
  if (!m) return fail(-1, "null argument%s", "");
  if (max_tile_x < 4 || max_tile_x > 32) return fail(-1, "max_tile_x must be 4..32%s (got %lld)", "", max_tile_x);
  if (max_tile_x == m->max_tx) return 0;
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipStreamSynchronize(m->stream));
  HostTables H;
  int rc = build_tables(m->T, m->zchunk, H, max_tile_x);
  if (rc) return rc;
  drop_graph(m);
  free_tables(m);
  m->H = std::move(H);
  m->max_tx = max_tile_x;
  rc = upload_tables(m);
  if (rc) return rc;
  return alloc_partials(m);
}

int vh_blocks_info(const vh_blocks* m, int64_t* counts6, int32_t* zchunk, int32_t* workgroups) {
  if (!m) return fail(-1, "null argument%s", "");
  if (counts6) {
    long long nreg = 0;
    for (const RegList& r : m->H.regs) nreg += (long long)r.cell.size();
    counts6[0] = m->T.off[m->T.nb];
    counts6[1] = m->H.block_tile_start[m->T.nb];
    counts6[2] = (int64_t)m->H.faces.size();
    counts6[3] = (int64_t)m->H.linkcells.size();
    counts6[4] = nreg;
    counts6[5] = (int64_t)m->H.slot_faces.size();
  }
  if (zchunk) *zchunk = m->zchunk;
  if (workgroups) *workgroups = list_wgs(m, 0);
  return 0;
}

int vh_blocks_sync(vh_blocks* m) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_indicator(vh_blocks* m, double* block_max, double* cell_g) {
  if (!m || !block_max) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  const long long n = m->T.off[m->T.nb];
  double* dcell = nullptr;
  if (cell_g) VH_HIP(hipMalloc(&dcell, sizeof(double) * n));
  regrid_indicator(m->stream, m->dtype, m->T.nb, m->d_blocks, m->d_repmap, m->T.rank, m->U, dcell, m->d_ind);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(block_max, m->d_ind, sizeof(double) * m->T.nb, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess && cell_g)
    e = hipMemcpyAsync(cell_g, dcell, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (dcell) (void)hipFree(dcell);
  if (e != hipSuccess) return fail(-2, "HIP error %s in the refinement indicator%s", hipGetErrorString(e));
  return 0;
}

int vh_blocks_remap(vh_blocks* dst, const vh_blocks* src, const int32_t* src_block) {
  if (!dst || !src || !src_block) return fail(-1, "null argument%s", "");
  if (dst == src) return fail(-1, "remap needs two different contexts%s", "");
  const Topo& D = dst->T;
  const Topo& S = src->T;
  if (dst->dtype != src->dtype) return fail(-1, "remap between contexts of different dtypes%s", "");
  if (dst->device != src->device) return fail(-1, "remap between contexts on different devices%s", "");
  if (D.rank != S.rank || D.maxlevel != S.maxlevel) return fail(-1, "remap between meshes of different rank or maxlevel%s", "");
  for (int d = 0; d < D.rank; ++d)
    if (D.h0[d] != S.h0[d]) return fail(-1, "remap between meshes of different base cell widths%s (axis %lld)", "", d);
  std::vector<RemapDev> map(D.nb);
  for (int b = 0; b < D.nb; ++b) {
    const int p = src_block[b];
    if (p < 0 || p >= S.nb) return fail(-1, "block %s%lld: source block out of range", "", b);
    const int shift = D.level[b] - S.level[p];
    if (shift < 0)
      return fail(-1, "block %s%lld is coarser than its source block (the remap prolongs by injection only)", "", b);
    map[b].src = p;
    map[b].shift = shift;
    for (int d = 0; d < 3; ++d) {
      map[b].o[d] = 0;
      if (d >= D.rank) continue;
      const long long ss = S.span(p);
      const long long rel = D.origin[3 * b + d] - S.origin[3 * p + d];
      if (rel < 0 || rel % ss) return fail(-1, "block %s%lld does not start on a cell of its source block", "", b);
      const long long o = rel / ss;
      if (o + ((long long)(D.dims[3 * b + d] - 1) >> shift) >= S.dims[3 * p + d])
        return fail(-1, "block %s%lld is not contained in its source block", "", b);
      map[b].o[d] = (int)o;
    }
  }
  VH_HIP(hipSetDevice(dst->device));
  VH_HIP(hipStreamSynchronize(src->stream));
  RemapDev* dmap = nullptr;
  int rc = upload_vec(&dmap, map);
  if (rc) {
    (void)hipFree(dmap);
    return rc;
  }
  regrid_remap(dst->stream, dst->dtype, D.nb, dst->d_blocks, src->d_blocks, dmap, src->U, S.off[S.nb], dst->U,
               D.off[D.nb]);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(dst->stream);
  (void)hipFree(dmap);
  if (e != hipSuccess) return fail(-2, "HIP error %s in the remap%s", hipGetErrorString(e));
  return 0;
}

int vh_blocks_time_steps(vh_blocks* m, double cfl, int64_t nsteps, int32_t method, double* ms_step,
                         double* ms_stage2) {
  if (!m || nsteps <= 0 || nsteps > 100000) return fail(-1, "bad argument%s (nsteps %lld)", "", nsteps);
  if (method < 0 || method > 2) return fail(-1, "method must be 0, 1 or 2%s (got %lld)", "", method);
  VH_HIP(hipSetDevice(m->device));
  const int saved_graph = m->use_graph;
  m->use_graph = method == 2;
  const ActiveRegs none{};
  const int np_steady = method == 1 ? m->H.block_tile_start[m->T.nb] : list_wgs(m, 0);
  auto one_step = [&](int np) -> int {
    if (method == 1) {
      enqueue_dt_finalize(m->stream, m->partials, np, cfl, 0.0, 0, m->clock);
      enqueue_step(m, -1, 0.0, none, true);
      enqueue_advance_clock(m->stream, m->clock);
      VH_HIP(hipGetLastError());
      return 0;
    }
    return single_rate_step(m, cfl, 0.0, 0, np);
  };
  // two untimed steps: the first reduces the CFL partials of the initial state,
  // the second builds the steady-state graph when graphs are used
  enqueue_cfl(m, -1);
  int rc = one_step(kCflWgs);
  if (!rc) rc = one_step(np_steady);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float tot = 0;
  if (!rc && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    (void)hipEventRecord(e0, m->stream);
    for (int64_t k = 0; k < nsteps && !rc; ++k) rc = one_step(np_steady);
    (void)hipEventRecord(e1, m->stream);
    hipError_t e = hipStreamSynchronize(m->stream);
    if (!rc && e != hipSuccess) rc = fail(-2, "HIP error %s while timing%s", hipGetErrorString(e));
    if (!rc) (void)hipEventElapsedTime(&tot, e0, e1);
  } else if (!rc) {
    rc = fail(-2, "hipEventCreate failed%s", "");
  }
  // Stage-2 launches of the persistent kernel timed one by one (eager steps).
  double st2 = 0.0;
  const int k2 = (int)(nsteps < 20 ? nsteps : 20);
  std::vector<hipEvent_t> ev(2 * k2, nullptr);
  if (!rc && ms_stage2) {
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) rc = fail(-2, "hipEventCreate failed%s", "");
    for (int k = 0; k < k2 && !rc; ++k) {
      enqueue_dt_finalize(m->stream, m->partials, list_wgs(m, 0), cfl, 0.0, 0, m->clock);
      enqueue_step(m, -1, 0.0, none, false, ev[2 * k], ev[2 * k + 1]);
      enqueue_advance_clock(m->stream, m->clock);
    }
    hipError_t e = hipStreamSynchronize(m->stream);
    if (!rc && e != hipSuccess) rc = fail(-2, "HIP error %s while timing%s", hipGetErrorString(e));
    for (int k = 0; k < k2 && !rc; ++k) {
      float a = 0;
      (void)hipEventElapsedTime(&a, ev[2 * k], ev[2 * k + 1]);
      st2 += a;
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  m->use_graph = saved_graph;
  if (rc) return rc;
  if (ms_step) *ms_step = (double)tot / nsteps;
  if (ms_stage2) *ms_stage2 = st2 / k2;
  return 0;
}

}  // extern "C"
