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
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/vespahip.h"
#include "host.h"
#include "profiles.h"

namespace vh {

enum { NB_SAME = 0, NB_COARSER = 1, NB_FINER = 2, NB_BOUNDARY = 3 };

// Tile limits of block_stage_kernel (BNT lanes, LDS sized for these bounds).
static const int kLanes = 256, kMaxW = 432, kMaxFS = 592, kMaxE = 32;

struct Topo {
  int rank = 0, nb = 0, maxlevel = 0;
  double h0[3] = {1, 1, 1};
  std::vector<int> level, dims, bc;      // dims: 3 per block, bc: 6 per block
  std::vector<long long> origin, off;    // origin: 3 per block, off: nb + 1
  // blocks touching face plane k = 6·block + slot: cand_ids[cand_off[k] .. cand_off[k+1])
  std::vector<int> cand_off, cand_ids;

  int shift(int b) const { return maxlevel - level[b]; }
  long long span(int b) const { return 1LL << shift(b); }
  bool has_cand(int b, int slot) const { return cand_off[6 * b + slot + 1] > cand_off[6 * b + slot]; }
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
    cell_ijk(lo, g, ijk);
    return lo;
  }
  void cell_ijk(int b, long long g, long long ijk[3]) const {
    const int l = (int)(g - off[b]);   // blocks hold fewer than 2^31 cells
    const int* n = &dims[3 * b];
    ijk[0] = l % n[0];
    ijk[1] = (l / n[0]) % n[1];
    ijk[2] = l / (n[0] * n[1]);
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
  // Neighbour of cell ijk of block b across face (d, side): kind and
  // representative cell; for NB_FINER the 2^(rank-1) finer cells on the face in
  // their leaf order. nbblock / nijk (optional): the neighbour's block and the
  // representative's cell coordinates in it.
  int neighbor_at(int b, const long long ijk[3], int d, int side, long long* rep, long long* fine, int* nfine,
                  int* nbblock, long long* nijk = nullptr) const {
    const int* n = &dims[3 * b];
    const long long j = ijk[d] + (side ? 1 : -1);
    if (j >= 0 && j < n[d]) {
      long long t[3] = {ijk[0], ijk[1], ijk[2]};
      t[d] = j;
      *rep = index(b, t);
      if (nbblock) *nbblock = b;
      if (nijk) for (int e = 0; e < 3; ++e) nijk[e] = t[e];
      return NB_SAME;
    }
    if (bc[6 * b + 2 * d + side] == BC_PERIODIC) {
      long long t[3] = {ijk[0], ijk[1], ijk[2]};
      t[d] = j < 0 ? n[d] - 1 : 0;
      *rep = index(b, t);
      if (nbblock) *nbblock = b;
      if (nijk) for (int e = 0; e < 3; ++e) nijk[e] = t[e];
      return NB_SAME;
    }
    const int sb = shift(b);
    long long q[3] = {0, 0, 0};
    for (int e = 0; e < rank; ++e) q[e] = origin[3 * b + e] + (ijk[e] << sb);
    q[d] = side ? q[d] + (1LL << sb) : q[d] - 1;
    const int k = 6 * b + 2 * d + side;
    for (int ci = cand_off[k]; ci < cand_off[k + 1]; ++ci) {
      const int c = cand_ids[ci];
      if (!contains(c, q)) continue;
      const int sc = shift(c);
      long long t[3] = {0, 0, 0};
      for (int e = 0; e < rank; ++e) t[e] = (q[e] - origin[3 * c + e]) >> sc;   // q lies inside c
      *rep = index(c, t);
      if (nbblock) *nbblock = c;
      if (nijk) for (int e = 0; e < 3; ++e) nijk[e] = t[e];
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

struct BuildScratch;
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
  // list l, shape class c (block_shape_class): tiles [tile_list_off[l] + tile_class_off[l*(K+1) + c],
  // tile_list_off[l] + tile_class_off[l*(K+1) + c + 1]), K = kBlockClasses
  std::vector<int> tile_class_off;
  std::vector<int> block_tile_start;     // per block, in the block-order list (nb + 1)
  // CFL block lists: [all][level 0 .. maxlevel]
  std::vector<int> cfl_blocks;
  std::vector<long long> cfl_prefix;
  std::vector<int> cfl_list_off;         // into cfl_blocks; prefix list l starts at cfl_list_off[l] + l
  // the builder's working buffers, kept with the tables so that a rebuild
  // (every regrid) reuses their memory instead of allocating and faulting it in
  std::shared_ptr<BuildScratch> scratch;
};

static int balanced(int n, int maxT) {
  const int nt = (n + maxT - 1) / maxT;
  return (n + nt - 1) / nt;
}

// ----------------------------------------------------------------------------
// host thread pool for the table builder: run(n, f) calls f(i) for i in [0, n)
// on the pool's workers and the caller, returning when all are done. Created on
// first use with min(hardware threads, 16) threads (VESPAHIP_HOST_THREADS
// overrides, 1 = serial); a child process after fork() builds its own (it
// inherits no worker threads). The pool is never destroyed: its workers sleep
// until the process ends.
class HostPool {
 public:
  static int threads() {
    static const int n = [] {
      const char* e = getenv("VESPAHIP_HOST_THREADS");
      int v = e ? atoi(e) : (int)std::thread::hardware_concurrency();
      return std::max(1, std::min(v, e ? 256 : 16));
    }();
    return n;
  }
  static void run(int n, const std::function<void(int)>& f) {
    if (n <= 1 || threads() <= 1) {
      for (int i = 0; i < n; ++i) f(i);
      return;
    }
    static std::mutex create_mu;
    static HostPool* pool = nullptr;
    HostPool* p;
    {
      std::lock_guard<std::mutex> lk(create_mu);
      if (!pool || pool->pid_ != getpid()) pool = new HostPool(threads() - 1);   // the old one stays (forked child)
      p = pool;
    }
    p->go(n, f);
  }

 private:
  // A build issues a few jobs back to back: threads spin this many pauses
  // (tens of microseconds) for the next job or the end of the current one
  // before they sleep, which saves a wake-up per job.
  static constexpr int kSpin = 1 << 14;
  static void pause() { __builtin_ia32_pause(); }

  explicit HostPool(int workers) : pid_(getpid()) {
    for (int i = 0; i < workers; ++i) std::thread([this] { loop(); }).detach();
    nworkers_ = workers;
  }
  void loop() {
    long long seen = 0;
    for (;;) {
      long long g = gen_.load(std::memory_order_acquire);
      for (int i = 0; g == seen && i < kSpin; ++i) {
        pause();
        g = gen_.load(std::memory_order_acquire);
      }
      if (g == seen) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
        g = gen_.load(std::memory_order_acquire);
      }
      seen = g;
      const std::function<void(int)>* f = job_;
      const int n = njob_;
      for (int i; (i = next_.fetch_add(1)) < n;) (*f)(i);
      if (active_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
        std::lock_guard<std::mutex> lk(mu_);
        done_.notify_one();
      }
    }
  }
  void go(int n, const std::function<void(int)>& f) {
    std::lock_guard<std::mutex> run_lk(run_mu_);   // one job at a time
    job_ = &f;
    njob_ = n;
    next_.store(0);
    active_.store(nworkers_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      gen_.fetch_add(1, std::memory_order_release);   // publishes the job to spinning and sleeping workers
    }
    cv_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) f(i);
    for (int i = 0; active_.load(std::memory_order_acquire) != 0 && i < kSpin; ++i) pause();
    if (active_.load(std::memory_order_acquire) != 0) {
      std::unique_lock<std::mutex> lk(mu_);
      done_.wait(lk, [&] { return active_.load(std::memory_order_acquire) == 0; });
    }
  }
  pid_t pid_;
  int nworkers_ = 0;
  std::mutex run_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int njob_ = 0;
  std::atomic<long long> gen_{0};
  std::atomic<int> next_{0}, active_{0};
};

// Output of one chunk (a contiguous run of blocks) of the parallel table
// phases; indices inside are chunk-local until the chunks are joined in order,
// which reproduces the serial walk's order exactly.
struct ChunkOut {
  int b0 = 0, b1 = 0;
  int rc = 0;
  std::string err;
  std::vector<LinkFaceDev> faces;
  std::vector<LinkCellDev> recs;
  std::vector<int> slot_faces, lcmap;
  std::vector<RegList> regs;   // per level
  // working buffers of the record phase
  std::vector<int> loc;
  std::vector<long long> bcell;
  std::vector<LinkCellDev> brec;
  std::vector<std::array<int, 5>> bslots;
  std::vector<int> bgrp;
  std::vector<std::pair<long long, int>> bfiner;
  void reset(int nlev) {
    rc = 0;
    err.clear();
    faces.clear();
    recs.clear();
    slot_faces.clear();
    lcmap.clear();
    if ((int)regs.size() < nlev) regs.resize(nlev);
    for (RegList& r : regs) {
      r.cell.clear();
      r.lc.clear();
      r.block.clear();
    }
  }
  int set_error(int code) {
    if (!rc) {
      rc = code;
      err = vh_last_error();
    }
    return code;
  }
};

// Candidate-search key: a block's lo (or hi) plane along an axis and its lo
// corner along the first other axis.
struct CandKey {
  long long plane, e1lo;
  int b;
  bool operator<(const CandKey& o) const {
    return plane != o.plane ? plane < o.plane : (e1lo != o.e1lo ? e1lo < o.e1lo : b < o.b);
  }
};

struct BuildScratch {
  std::vector<CandKey> at[3][2];
  std::vector<std::vector<int>> cids;
  std::vector<int> counts, perr, fmap, nbk, seen, cls, idx, chunk_of;
  std::vector<ChunkOut> ch;
  std::vector<TileDev> base;
  std::vector<long long> cost;
};

// Builds every table of a mesh. Returns 0 or a negative code (message set).
// nthreads: 0 = the pool's size, 1 = serial; the tables do not depend on it.
static int build_tables(Topo& T, int zchunk, HostTables& H, int max_tx = 32, int nthreads = 0) {
  const int nb = T.nb, rank = T.rank;
  // ---- candidate neighbours by face plane ----
  // Blocks grouped by the coordinate of their lo (hi) plane along each axis and
  // sorted inside a group by their lo corner along the first other axis; a
  // block's side meets the blocks of the opposite group at its own plane whose
  // extents overlap its own along the other axes. Blocks are disjoint, so the
  // search along that axis starts one largest extent before the side's lo end.
  // Blocks are searched in parallel chunks; each side lists its candidates in
  // ascending block order.
  const int nthr = nthreads > 0 ? nthreads : HostPool::threads();
  if (!H.scratch) H.scratch = std::make_shared<BuildScratch>();
  BuildScratch& X = *H.scratch;
  {
    using Key = CandKey;
    auto& at = X.at;   // [axis][0: lo planes, 1: hi planes]
    long long emax[3] = {1, 1, 1};
    for (int d = 0; d < rank; ++d) {
      const int e1 = d == 0 ? 1 : 0;
      for (int c = 0; c < nb; ++c)
        if (e1 < rank) emax[d] = std::max(emax[d], (long long)T.dims[3 * c + e1] * T.span(c));
    }
    // the 2·rank key lists are filled and sorted side by side
    HostPool::run(nb < 256 ? 1 : 2 * rank, [&](int q) {
      for (int j = (nb < 256 ? 0 : q); j < 2 * rank; j += (nb < 256 ? 1 : 2 * rank)) {
        const int d = j >> 1, side = j & 1, e1 = d == 0 ? 1 : 0;
        std::vector<Key>& v = at[d][side];
        v.resize(nb);
        for (int c = 0; c < nb; ++c) {
          const long long o1 = e1 < rank ? T.origin[3 * c + e1] : 0;
          const long long pl = T.origin[3 * c + d] + (side ? (long long)T.dims[3 * c + d] * T.span(c) : 0);
          v[c] = {pl, o1, c};
        }
        std::sort(v.begin(), v.end());
      }
    });
    const int ncc = nb < 256 ? 1 : std::min(4 * nthr, nb);
    auto& cids = X.cids;                               // candidates of the chunk's sides in (block, slot) order
    if ((int)cids.size() < ncc) cids.resize(ncc);
    for (int q = 0; q < ncc; ++q) cids[q].clear();
    auto& counts = X.counts;
    counts.assign(6 * (size_t)nb, 0);
    auto& perr = X.perr;
    perr.assign(ncc, -1);                              // block with a periodic-axis contact
    HostPool::run(ncc, [&](int q) {
      const int b0 = (int)((long long)nb * q / ncc), b1 = (int)((long long)nb * (q + 1) / ncc);
      int tmp[64];
      for (int b = b0; b < b1; ++b) {
        const long long sb = T.span(b);
        for (int d = 0; d < rank; ++d) {
          const int e1 = d == 0 ? 1 : 0;
          const long long a1 = e1 < rank ? T.origin[3 * b + e1] : 0;
          const long long a1hi = e1 < rank ? a1 + T.dims[3 * b + e1] * sb : 1;
          for (int side = 0; side < 2; ++side) {
            const long long plane =
                side ? T.origin[3 * b + d] + (long long)T.dims[3 * b + d] * sb : T.origin[3 * b + d];
            const auto& v = at[d][side ? 0 : 1];
            int k = 0;
            for (auto it = std::lower_bound(v.begin(), v.end(), Key{plane, a1 - emax[d] + 1, -1});
                 it != v.end() && it->plane == plane && it->e1lo < a1hi; ++it) {
              const int c = it->b;
              if (c == b) continue;
              const long long sc = T.span(c);
              bool overlap = true;
              for (int e = 0; e < rank; ++e) {
                if (e == d) continue;
                const long long alo = T.origin[3 * b + e], ahi = alo + T.dims[3 * b + e] * sb;
                const long long blo = T.origin[3 * c + e], bhi = blo + T.dims[3 * c + e] * sc;
                if (bhi <= alo || ahi <= blo) overlap = false;
              }
              if (!overlap) continue;
              if (k < 64) tmp[k] = c;
              ++k;
            }
            if (k > 64) {   // more than 64 blocks on one side: collect again without the cap
              std::vector<int> big;
              for (auto it = std::lower_bound(v.begin(), v.end(), Key{plane, a1 - emax[d] + 1, -1});
                   it != v.end() && it->plane == plane && it->e1lo < a1hi; ++it) {
                const int c = it->b;
                if (c == b) continue;
                const long long sc = T.span(c);
                bool overlap = true;
                for (int e = 0; e < rank; ++e) {
                  if (e == d) continue;
                  const long long alo = T.origin[3 * b + e], ahi = alo + T.dims[3 * b + e] * sb;
                  const long long blo = T.origin[3 * c + e], bhi = blo + T.dims[3 * c + e] * sc;
                  if (bhi <= alo || ahi <= blo) overlap = false;
                }
                if (overlap) big.push_back(c);
              }
              std::sort(big.begin(), big.end());
              cids[q].insert(cids[q].end(), big.begin(), big.end());
            } else {
              std::sort(tmp, tmp + k);
              cids[q].insert(cids[q].end(), tmp, tmp + k);
            }
            counts[6 * (size_t)b + 2 * d + side] = k;
            if (k > 0 && T.bc[6 * b + 2 * d] == BC_PERIODIC && perr[q] < 0) perr[q] = b;
          }
        }
      }
    });
    for (int q = 0; q < ncc; ++q)
      if (perr[q] >= 0) return fail(-1, "block %s%lld touches another block along a periodic axis", "", perr[q]);
    T.cand_off.assign(6 * (size_t)nb + 1, 0);
    for (size_t k = 0; k < 6 * (size_t)nb; ++k) T.cand_off[k + 1] = T.cand_off[k] + counts[k];
    T.cand_ids.clear();
    T.cand_ids.reserve(T.cand_off[6 * (size_t)nb]);
    for (int q = 0; q < ncc; ++q) T.cand_ids.insert(T.cand_ids.end(), cids[q].begin(), cids[q].end());
  }

  // ---- blocks and the offsets of their face-plane maps (serial) ----
  H.blocks.assign(nb, BlockDev{});
  long long nplane = 0;
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
      for (int side = 0; side < 2; ++side)
        if (T.has_cand(b, 2 * d + side)) {
          bk.repoff[2 * d + side] = (int)nplane;
          nplane += plane_size(bk.n, d);
        }
    if (nplane > 0x7fffffffLL) return fail(-1, "mesh too large%s", "");
  }
  H.repmap.assign((size_t)nplane, -1);
  // fmap runs beside repmap: the link face a face cell emits through that side, or -1
  std::vector<int>& fmap = X.fmap;
  fmap.assign((size_t)nplane, -1);
  // nbk beside them: the neighbour block across that side and the kind, 8·block + kind
  // (set by the face walk, read back by the record walk)
  std::vector<int>& nbk = X.nbk;
  nbk.resize((size_t)nplane);

  // chunks: contiguous block ranges of similar face-plane work
  int nchunk = std::min(4 * nthr, nb);
  if (nb < 64 || nplane < 4096 || nthreads == 1) nchunk = 1;
  // chunk outputs keep their capacity between builds: the vector only grows
  std::vector<ChunkOut>& ch = X.ch;
  if ((int)ch.size() < nchunk) ch.resize(nchunk);
  for (int c = 0; c < nchunk; ++c) ch[c].reset(T.maxlevel + 1);
  {
    long long acc = 0;
    int c = 0;
    ch[0].b0 = 0;
    for (int b = 0; b < nb; ++b) {
      const BlockDev& bk = H.blocks[b];
      long long w = 1;
      for (int s = 0; s < 6; ++s)
        if (bk.repoff[s] >= 0) w += plane_size(bk.n, s >> 1);
      acc += w;
      if (c + 1 < nchunk && acc * nchunk >= (long long)(c + 1) * (nplane + nb)) {
        ch[c].b1 = b + 1;
        ch[++c].b0 = b + 1;
      }
    }
    ch[c].b1 = nb;
    nchunk = c + 1;
  }
  std::vector<int>& chunk_of = X.chunk_of;   // block -> its chunk
  chunk_of.resize(nb);
  for (int c = 0; c < nchunk; ++c)
    for (int b = ch[c].b0; b < ch[c].b1; ++b) chunk_of[b] = c;
  auto for_chunks = [&](const std::function<void(ChunkOut&)>& f) {
    HostPool::run(nchunk, [&](int i) { f(ch[i]); });
    for (int c = 0; c < nchunk; ++c)
      if (ch[c].rc) return fail(ch[c].rc, "%s", ch[c].err.c_str());
    return 0;
  };

  // ---- representative maps and link faces (per chunk) ----
  // A face is emitted once, as the face walk emits it: same level from the hi
  // side, coarse↔fine from the fine side (lib/HGBackend/src/HGBackend.jl:202-276),
  // with the representative neighbours of both cells for the reconstruction.
  int rc = for_chunks([&](ChunkOut& o) {
    long long fine[4];
    int nfine = 0;
    for (int b = o.b0; b < o.b1 && !o.rc; ++b) {
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
              long long rep = -1, nijk[3];
              int nbb = -1;
              const int kind = T.neighbor_at(b, ijk, d, side, &rep, fine, &nfine, &nbb, nijk);
              if (kind == -1)
                return (void)o.set_error(fail(-1, "block %s%lld touches a block more than one level apart", "", b));
              if (kind < 0)
                return (void)o.set_error(fail(-1, "block %s%lld: a coarse face is not covered by one finer block "
                                              "(fine blocks must be aligned to coarse cells)", "", b));
              const int pi = bk.repoff[slot] + plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2]);
              H.repmap[pi] = kind == NB_BOUNDARY ? -1 : (int)rep;
              nbk[pi] = 8 * nbb + kind;
              const bool emit = (side == 1 && (kind == NB_SAME || kind == NB_COARSER)) || (side == 0 && kind == NB_COARSER);
              if (!emit) continue;
              LinkFaceDev f{};
              f.iL = (int)(side ? g : rep);
              f.iR = (int)(side ? rep : g);
              f.axis = d;
              f.area = T.area(b, d);
              for (int which = 0; which < 2; ++which) {
                const bool own = (which == 0) == (side == 1);   // iL is this cell on the hi side
                const int cb = own ? b : nbb;
                const long long* q = own ? ijk : nijk;
                int* nbr = which ? f.nR : f.nL;
                for (int s = 0; s < 2; ++s) {
                  long long r2 = -1;
                  const int k2 = T.neighbor_at(cb, q, d, s, &r2, nullptr, nullptr, nullptr);
                  if (k2 < 0) return (void)o.set_error(fail(-1, "invalid neighbour across a link face%s", ""));
                  nbr[s] = k2 == NB_BOUNDARY ? -1 - T.bc[6 * cb + 2 * d + s] : (int)r2;
                }
              }
              fmap[pi] = (int)o.faces.size();   // chunk-local: + face_base[chunk_of[b]]
              o.faces.push_back(f);
            }
        }
    }
  });
  if (rc) return rc;
  // faces are numbered in chunk order: global = chunk-local + face_base[chunk]
  // (the chunks' faces are copied into H.faces with the records below)
  std::vector<int> face_base(nchunk + 1, 0);
  for (int c = 0; c < nchunk; ++c) face_base[c + 1] = face_base[c] + (int)ch[c].faces.size();
  H.faces.resize(face_base[nchunk]);
  auto face_ref = [&](int f) -> const LinkFaceDev& {
    const int c = (int)(std::upper_bound(face_base.begin(), face_base.begin() + nchunk + 1, f) - face_base.begin()) - 1;
    return ch[c].faces[f - face_base[c]];
  };

  // the link face that cell q of block blk emits through side sl, or -1
  auto face_at = [&](int blk, const long long q[3], int sl) -> int {
    const BlockDev& ob = H.blocks[blk];
    if (ob.repoff[sl] < 0) return -1;
    const int d = sl >> 1;
    if (q[d] != ((sl & 1) ? ob.n[d] - 1 : 0)) return -1;
    const int v = fmap[ob.repoff[sl] + plane_idx(ob.n, d, q[0], q[1], q[2])];
    return v < 0 ? -1 : v + face_base[chunk_of[blk]];
  };

  // ---- link-cell records, their per-face maps and the flux-register lists (per chunk) ----
  // One record per cell touching link faces, created block by block in the
  // order the block's face cells are visited (sides in slot order); a slot
  // lists at most 2^(rank-1) <= 4 faces: [0] count, [1..4] faces.
  using SlotList = std::array<int, 5>;
  rc = for_chunks([&](ChunkOut& o) {
    long long cmax = 1;
    for (int b = o.b0; b < o.b1; ++b) cmax = std::max(cmax, T.off[b + 1] - T.off[b]);
    std::vector<int>& loc = o.loc;            // cell of the current block -> its record (block-local)
    loc.assign((size_t)cmax, -1);
    std::vector<long long>& bcell = o.bcell;  // block-local record -> cell
    std::vector<LinkCellDev>& brec = o.brec;
    std::vector<SlotList>& bslots = o.bslots; // block-local record * 6 + slot
    std::vector<int>& bgrp = o.bgrp;          // block-local record * 6 + slot -> emitting block of a type-2 slot
    std::vector<std::pair<long long, int>>& bfiner = o.bfiner;
    for (int b = o.b0; b < o.b1 && !o.rc; ++b) {
      BlockDev& bk = H.blocks[b];
      bcell.clear();
      brec.clear();
      bslots.clear();
      bgrp.clear();
      const int r0 = (int)o.recs.size();
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
              const int pi = bk.repoff[slot] + plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2]);
              if (H.repmap[pi] < 0) continue;   // a domain boundary
              const long long g = T.index(b, ijk);
              // the face walk's neighbour: block, kind and (from the representative) its cell
              const int kind = nbk[pi] & 7, nbb = nbk[pi] >> 3;
              long long nijk[3];
              {
                const int* nn = &T.dims[3 * nbb];
                const long long t = H.repmap[pi] - T.off[nbb];
                nijk[0] = t % nn[0];
                nijk[1] = (t / nn[0]) % nn[1];
                nijk[2] = t / ((long long)nn[0] * nn[1]);
              }
              int type = 0, emitter_block = -1;
              SlotList fl{0, -1, -1, -1, -1};
              if ((side == 1 && (kind == NB_SAME || kind == NB_COARSER)) || (side == 0 && kind == NB_COARSER)) {
                type = 1;
                fl[++fl[0]] = fmap[pi] + face_base[chunk_of[b]];
              } else if (side == 0 && kind == NB_SAME) {
                type = 2;
                emitter_block = nbb;
                fl[++fl[0]] = face_at(nbb, nijk, 2 * d + 1);
              } else if (kind == NB_FINER) {
                // the finer cells on the face in leaf order: the representative
                // (lo corner) and its +1 steps along the face axes
                type = 2;
                emitter_block = nbb;
                const int n2 = e2 < rank ? 2 : 1, n1 = e1 < rank ? 2 : 1;
                for (int a2 = 0; a2 < n2; ++a2)
                  for (int a1 = 0; a1 < n1; ++a1) {
                    long long q[3] = {nijk[0], nijk[1], nijk[2]};
                    q[e1] += a1;
                    q[e2] += a2;
                    fl[++fl[0]] = face_at(nbb, q, side ? 2 * d : 2 * d + 1);
                  }
              }
              for (int i = 1; i <= fl[0]; ++i)
                if (fl[i] < 0)
                  return (void)o.set_error(
                      fail(-1, "block %s%lld: a link face has no emitting cell (mesh not 2:1 aligned)", "", b));
              int& r = loc[g - T.off[b]];
              if (r < 0) {
                r = (int)brec.size();
                bcell.push_back(g);
                brec.push_back(LinkCellDev{});
                for (int q = 0; q < 6; ++q) {
                  bslots.push_back(SlotList{0, -1, -1, -1, -1});
                  bgrp.push_back(-1);
                }
              }
              brec[r].tc |= (unsigned)type << (2 * slot);
              bslots[6 * r + slot] = fl;
              bgrp[6 * r + slot] = emitter_block;
              if (kind == NB_FINER) brec[r].gr |= 1u << (24 + slot);
              for (int i = 1; i <= fl[0]; ++i) {   // each listed face must touch this cell
                const LinkFaceDev& F = face_ref(fl[i]);
                if (F.iL != g && F.iR != g)
                  return (void)o.set_error(fail(-1, "link face bookkeeping error at block %s%lld", "", b));
              }
            }
        }
      // the block's records: face lists and groups
      bfiner.clear();
      for (size_t r = 0; r < brec.size(); ++r) {
        LinkCellDev R = brec[r];
        std::pair<int, int> grp[6];   // (emitting block, slot), at most one per slot
        int ng = 0;
        R.start = (int)o.slot_faces.size();   // chunk-local until joined
        for (int s = 0; s < 6; ++s) {
          const SlotList& sl = bslots[6 * r + s];
          R.tc |= (unsigned)sl[0] << (12 + 3 * s);
          for (int i = 1; i <= sl[0]; ++i) o.slot_faces.push_back(sl[i]);
          if (lc_type(R, s) == 2) grp[ng++] = {bgrp[6 * r + s], s};
        }
        std::sort(grp, grp + ng);
        int npre = 0;
        unsigned gbits = 0;
        for (int q = 0; q < 6; ++q) {
          if (q > 0 && q < ng && grp[q].first == grp[q - 1].first)
            return (void)o.set_error(fail(-1, "a cell touches block %s%lld through two faces", "", grp[q].first));
          gbits |= (q < ng ? (unsigned)grp[q].second : 7u) << (3 * q);
          if (q < ng && grp[q].first < b) npre++;
        }
        R.gr |= gbits | ((unsigned)npre << 18) | ((unsigned)ng << 21);
        o.recs.push_back(R);
        if (lc_finer(R)) bfiner.push_back({bcell[r], r0 + (int)r});
      }
      // link-cell maps of the block's faces (every record sits on one of them)
      bk.haslinks = 0;
      for (int d = 0; d < rank && !brec.empty(); ++d)
        for (int side = 0; side < 2; ++side) {
          const int slot = 2 * d + side;
          if (bk.repoff[slot] < 0) continue;
          const int ps = plane_size(bk.n, d);
          const size_t base = o.lcmap.size();
          bool any = false;
          o.lcmap.resize(base + ps, -1);
          const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
          for (int v = 0; v < bk.n[e2]; ++v)
            for (int u = 0; u < bk.n[e1]; ++u) {
              long long ijk[3];
              ijk[d] = side ? bk.n[d] - 1 : 0;
              ijk[e1] = u;
              ijk[e2] = v;
              const int r = loc[T.index(b, ijk) - T.off[b]];
              if (r < 0) continue;
              o.lcmap[base + plane_idx(bk.n, d, ijk[0], ijk[1], ijk[2])] = r0 + r;   // chunk-local
              any = true;
            }
          if (!any) {
            o.lcmap.resize(base);
            continue;
          }
          bk.lcoff[slot] = (int)base;   // chunk-local
          bk.haslinks = 1;
        }
      // flux-register lists: the coarse cells with finer neighbours, in cell order
      std::sort(bfiner.begin(), bfiner.end());
      RegList& L = o.regs[T.level[b]];
      for (const auto& kv : bfiner) {
        L.cell.push_back((int)kv.first);
        L.lc.push_back(kv.second);   // chunk-local
        L.block.push_back(b);
      }
      for (long long g : bcell) loc[g - T.off[b]] = -1;
    }
  });
  if (rc) return rc;
  std::vector<long long> rec_base(nchunk + 1, 0), sf_base(nchunk + 1, 0), lc_base(nchunk + 1, 0);
  for (int c = 0; c < nchunk; ++c) {
    rec_base[c + 1] = rec_base[c] + (long long)ch[c].recs.size();
    sf_base[c + 1] = sf_base[c] + (long long)ch[c].slot_faces.size();
    lc_base[c + 1] = lc_base[c] + (long long)ch[c].lcmap.size();
  }
  if (lc_base[nchunk] > 0x7fffffffLL || sf_base[nchunk] > 0x7fffffffLL) return fail(-1, "mesh too large%s", "");
  H.linkcells.resize(rec_base[nchunk]);
  H.slot_faces.resize(sf_base[nchunk]);
  H.lcmap.resize(lc_base[nchunk]);
  HostPool::run(nchunk, [&](int c) {
    const ChunkOut& o = ch[c];
    std::copy(o.faces.begin(), o.faces.end(), H.faces.begin() + face_base[c]);
    const int rb = (int)rec_base[c], sb = (int)sf_base[c], lb = (int)lc_base[c];
    for (size_t r = 0; r < o.recs.size(); ++r) {
      LinkCellDev R = o.recs[r];
      R.start += sb;
      H.linkcells[rb + r] = R;
    }
    std::copy(o.slot_faces.begin(), o.slot_faces.end(), H.slot_faces.begin() + sb);
    for (size_t i = 0; i < o.lcmap.size(); ++i) H.lcmap[lb + i] = o.lcmap[i] >= 0 ? o.lcmap[i] + rb : -1;
    for (int b = o.b0; b < o.b1; ++b)
      for (int s = 0; s < 6; ++s)
        if (H.blocks[b].lcoff[s] >= 0) H.blocks[b].lcoff[s] += lb;
  });
  if ((int)H.regs.size() != T.maxlevel + 1) H.regs.resize(T.maxlevel + 1);
  for (RegList& r : H.regs) {
    r.cell.clear();
    r.lc.clear();
    r.block.clear();
  }
  for (int c = 0; c < nchunk; ++c)
    for (int l = 0; l <= T.maxlevel; ++l) {
      const RegList& s = ch[c].regs[l];
      RegList& L = H.regs[l];
      L.cell.insert(L.cell.end(), s.cell.begin(), s.cell.end());
      L.block.insert(L.block.end(), s.block.begin(), s.block.end());
      for (int v : s.lc) L.lc.push_back(v + (int)rec_base[c]);
    }
  // every link face must be gathered by both of its cells (fails when a fine
  // block is not aligned to the coarse cells it touches)
  {
    std::vector<int>& seen = X.seen;
    seen.assign(H.faces.size(), 0);
    for (int f : H.slot_faces) seen[f]++;
    for (size_t f = 0; f < seen.size(); ++f)
      if (seen[f] != 2)
        return fail(-1, "link face %s%lld is not shared by two cells: fine blocks must be aligned to coarse cells", "",
                    (long long)f);
  }

  // ---- tiles ----
  std::vector<TileDev>& base = X.base;
  std::vector<long long>& cost = X.cost;
  base.clear();
  cost.clear();
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
  // A list holds its tiles grouped by shape class (one launch per class), each
  // group in decreasing cost.
  std::vector<int>& cls = X.cls;
  cls.resize(nt);
  for (int i = 0; i < nt; ++i) cls[i] = block_shape_class(rank, base[i].tx, base[i].ty);
  auto sorted_list = [&](int lev) -> const std::vector<int>& {
    std::vector<int>& idx = X.idx;
    idx.clear();
    for (int i = 0; i < nt; ++i)
      if (lev < 0 || T.level[base[i].block] == lev) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return cls[a] != cls[b] ? cls[a] < cls[b] : cost[a] > cost[b]; });
    return idx;
  };
  H.tiles = base;
  H.tile_list_off.clear();
  H.tile_class_off.clear();
  H.tile_list_off.push_back(nt);   // list 0: all, starts after the block-order copy
  for (int l = -1; l <= T.maxlevel; ++l) {
    const std::vector<int>& idx = sorted_list(l);
    for (int c = 0, q = 0; c <= kBlockClasses; ++c) {
      while (c < kBlockClasses && q < (int)idx.size() && cls[idx[q]] < c) ++q;
      H.tile_class_off.push_back(c < kBlockClasses ? q : (int)idx.size());
    }
    for (int i : idx) H.tiles.push_back(base[i]);
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

// FNV-1a over 8-byte words (a trailing partial word zero-padded) of each table.
template <class V>
static void digest_vec(uint64_t& h, const std::vector<V>& v) {
  const size_t n = v.size() * sizeof(V);
  h = (h ^ (uint64_t)n) * 1099511628211ULL;
  const unsigned char* p = (const unsigned char*)v.data();
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    memcpy(&w, p + i, 8);
    h = (h ^ w) * 1099511628211ULL;
  }
  if (i < n) {
    uint64_t w = 0;
    memcpy(&w, p + i, n - i);
    h = (h ^ w) * 1099511628211ULL;
  }
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
  const int finer = lc_finer(R), ngrp = lc_ngrp(R);
  for (int q = 0; q < ngrp; ++q) {
    const int slot = lc_grp(R, q);
    if (!((finer >> slot) & 1)) continue;
    const int st = lc_start(R, slot), cnt = lc_count(R, slot);
    for (int i = 0; i < cnt; ++i) {
      const int fi = B.slot_faces[st + i];
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
  // Every table above (and the flux-register lists) lives in one device arena;
  // an in-place regrid uploads the new mesh's tables into the spare arena while
  // the remap still reads the current one. Both keep their capacity, as do the
  // pinned staging of the upload and the spare host tables.
  void* arena[2] = {nullptr, nullptr};
  size_t arena_cap[2] = {0, 0};
  int arena_cur = 0;
  void* stage = nullptr;
  size_t stage_cap = 0;
  size_t state_cap = 0;      // bytes of U and of U1
  Topo T2;
  HostTables H2;
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
  // A regrid does not wait for the stream: the graph it replaces is kept here
  // until the next regrid has waited for stage_ev (recorded after the upload
  // that followed the graph's last launch), and stage_ev guards the pinned
  // staging against a copy still reading it.
  hipGraphExec_t retired = nullptr;
  hipEvent_t stage_ev = nullptr;
  struct BlockGrav* grav = nullptr;   // self-gravity (vh_blocks_enable_gravity)
};

static const int kCflWgs = 2048;

// Self-gravity on a block mesh: the composite operator as CSR (built on the
// host per mesh, as the reference builds its PoissonKAStruct), φ and the CG
// vectors at the state's precision, and the step's g field.
struct BlockGrav {
  GravParams P;
  GravBuffers B;
  GravCsr A;
  int* rowptr = nullptr;
  int* colidx = nullptr;
  double* coef = nullptr;
  double* vol = nullptr;
  int* glo = nullptr;
  int* ghi = nullptr;
  double* gsum = nullptr;
  void* g = nullptr;       // 3·ncell values: g_x, g_y, g_z of the current φ
};

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
  if (m->grav) {
    B.gravg = m->grav->g;
    B.gravG = m->grav->P.G;
  }
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

// Tile list l: 0 all blocks, 1 + L level L. A stage launches one persistent
// kernel per tile shape class of the list; max_wgs (default 3 per CU) caps the
// workgroups of 256 lanes, a class with smaller workgroups gets proportionally
// more.
static int class_tiles(const vh_blocks* m, int l, int c) {
  const int* o = m->H.tile_class_off.data() + (size_t)l * (kBlockClasses + 1);
  return o[c + 1] - o[c];
}
static int class_wgs(const vh_blocks* m, int l, int c) {
  const int n = class_tiles(m, l, c);
  const int cap = (m->max_wgs > 0 ? m->max_wgs : 3 * m->ncu) * (256 / block_class_lanes(c));
  return n < cap ? n : cap;
}
// Workgroups of a stage launch over list l (= the CFL partials its stage 2 writes).
static int list_wgs(const vh_blocks* m, int l) {
  int np = 0;
  for (int c = 0; c < kBlockClasses; ++c) np += class_wgs(m, l, c);
  return np;
}

static void enqueue_link(vh_blocks* m, const void* Us) {
  const BlocksDev B = blocks_dev(m);
  if (m->mode == VH_FAST) fast::launch_link_flux(m->dtype, B.nfaces, m->stream, Us, B, m->d_flux);
  else exact::launch_link_flux(m->dtype, B.nfaces, m->stream, Us, B, m->d_flux);
}

static void enqueue_tiles(vh_blocks* m, bool stage2, int l, const void* Us, const void* U0, void* Uout) {
  const double* dtp = m->clock + 1;
  const int* co = m->H.tile_class_off.data() + (size_t)l * (kBlockClasses + 1);
  int np = 0;
  for (int c = 0; c < kBlockClasses; ++c) {
    const int wgs = class_wgs(m, l, c);
    if (wgs == 0) continue;
    BlocksDev B = blocks_dev(m);
    B.tiles = m->d_tiles + m->H.tile_list_off[l] + co[c];
    B.ntiles = co[c + 1] - co[c];
    if (m->mode == VH_FAST)
      fast::launch_block_stage(m->dtype, m->T.rank, c, stage2, wgs, m->stream, Us, U0, Uout, dtp, m->partials + np, B);
    else
      exact::launch_block_stage(m->dtype, m->T.rank, c, stage2, wgs, m->stream, Us, U0, Uout, dtp, m->partials + np, B);
    np += wgs;
  }
  if (l == 0 && stage2) m->stage_wgs = np;
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
      fast::launch_block_stage(m->dtype, m->T.rank, 0, stage2, t1 - t0, m->stream, Us, U0, Uout, dtp, m->partials + t0, B);
    else
      exact::launch_block_stage(m->dtype, m->T.rank, 0, stage2, t1 - t0, m->stream, Us, U0, Uout, dtp, m->partials + t0, B);
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
  need = std::max(need, list_wgs(m, 0));
  if (need > m->npartials) {
    VH_HIP(hipStreamSynchronize(m->stream));   // queued stages may still write the old partials
    if (m->partials) VH_HIP(hipFree(m->partials));
    m->partials = nullptr;
    VH_HIP(hipMalloc(&m->partials, sizeof(double) * need * 5));
    m->npartials = need;
  }
  return 0;
}

static void forget_tables(vh_blocks* m) {
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

static void free_tables(vh_blocks* m) {
  for (int a = 0; a < 2; ++a) {
    (void)hipFree(m->arena[a]);
    m->arena[a] = nullptr;
    m->arena_cap[a] = 0;
  }
  if (m->stage) (void)hipHostFree(m->stage);
  m->stage = nullptr;
  m->stage_cap = 0;
  forget_tables(m);
}

// Uploads every device table of H into arena slot a (allocated at twice the
// need when too small) through the pinned staging buffer in one copy, zeroes the flux-
// register accumulators and points the context's tables there. extra/extra_bytes
// (optional) is packed behind the tables; *extra_dev receives its address.
// Stream-ordered; the staging is rewritten only after stage_ev, recorded behind
// the last copy out of it, has completed.
static int upload_tables(vh_blocks* m, const HostTables& H, int nblocks, int a, const void* extra = nullptr,
                         size_t extra_bytes = 0, void** extra_dev = nullptr) {
  struct Part {
    const void* src;
    size_t bytes;
    size_t off;
  };
  std::vector<Part> parts;
  size_t total = 0;
  auto add = [&](const void* src, size_t bytes) -> size_t {
    const size_t off = total;
    parts.push_back({src, bytes, off});
    total += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    return off;
  };
  const size_t o_blocks = add(H.blocks.data(), sizeof(BlockDev) * H.blocks.size());
  const size_t o_tiles = add(H.tiles.data(), sizeof(TileDev) * H.tiles.size());
  const size_t o_repmap = add(H.repmap.data(), sizeof(int) * H.repmap.size());
  const size_t o_lcmap = add(H.lcmap.data(), sizeof(int) * H.lcmap.size());
  const size_t o_lc = add(H.linkcells.data(), sizeof(LinkCellDev) * H.linkcells.size());
  const size_t o_sf = add(H.slot_faces.data(), sizeof(int) * H.slot_faces.size());
  const size_t o_faces = add(H.faces.data(), sizeof(LinkFaceDev) * H.faces.size());
  const size_t o_cflb = add(H.cfl_blocks.data(), sizeof(int) * H.cfl_blocks.size());
  const size_t o_cflp = add(H.cfl_prefix.data(), sizeof(long long) * H.cfl_prefix.size());
  std::vector<size_t> o_reg(3 * H.regs.size());
  for (size_t l = 0; l < H.regs.size(); ++l) {
    o_reg[3 * l] = add(H.regs[l].cell.data(), sizeof(int) * H.regs[l].cell.size());
    o_reg[3 * l + 1] = add(H.regs[l].lc.data(), sizeof(int) * H.regs[l].lc.size());
    o_reg[3 * l + 2] = add(H.regs[l].block.data(), sizeof(int) * H.regs[l].block.size());
  }
  const size_t o_extra = extra_bytes ? add(extra, extra_bytes) : 0;
  const size_t staged = total;   // everything above is copied; below are device-only buffers
  const size_t o_flux = add(nullptr, m->esize * 5 * std::max<size_t>(1, H.faces.size()));
  const size_t o_ind = add(nullptr, sizeof(double) * (size_t)nblocks);
  std::vector<size_t> o_delta(H.regs.size());
  const size_t zero_from = total;
  for (size_t l = 0; l < H.regs.size(); ++l)
    o_delta[l] = add(nullptr, m->esize * 5 * std::max<size_t>(1, H.regs[l].cell.size()));
  const size_t zero_to = total;
  if (total > m->arena_cap[a]) {
    VH_HIP(hipStreamSynchronize(m->stream));   // queued work may still read the old slot
    (void)hipFree(m->arena[a]);
    m->arena[a] = nullptr;
    m->arena_cap[a] = 0;
    // regrids grow the mesh: keep room for the next ones (at least 8 MiB, so a
    // mesh growing from a few blocks does not reallocate at every regrid)
    const size_t cap = std::max<size_t>(2 * total, (size_t)8 << 20);
    VH_HIP(hipMalloc(&m->arena[a], cap));
    m->arena_cap[a] = cap;
  }
  if (m->stage_ev) VH_HIP(hipEventSynchronize(m->stage_ev));   // the last copy out of the staging is done
  if (staged > m->stage_cap) {
    if (m->stage) (void)hipHostFree(m->stage);
    m->stage = nullptr;
    m->stage_cap = 0;
    const size_t cap = std::max<size_t>(2 * staged, (size_t)4 << 20);
    VH_HIP(hipHostMalloc(&m->stage, cap, 0));
    m->stage_cap = cap;
  }
  // the tables are packed into the staging by the builder's threads, in pieces
  // of at most 256 KiB
  char* st = (char*)m->stage;
  struct Piece {
    const char* src;
    char* dst;
    size_t bytes;
  };
  std::vector<Piece> pieces;
  for (const Part& p : parts)
    if (p.src && p.bytes && p.off < staged)
      for (size_t o = 0; o < p.bytes; o += (size_t)1 << 18)
        pieces.push_back({(const char*)p.src + o, st + p.off + o, std::min(p.bytes - o, (size_t)1 << 18)});
  HostPool::run((int)pieces.size(), [&](int i) { memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes); });
  char* base = (char*)m->arena[a];
  VH_HIP(hipMemcpyAsync(base, st, staged, hipMemcpyHostToDevice, m->stream));
  if (!m->stage_ev) VH_HIP(hipEventCreateWithFlags(&m->stage_ev, hipEventDisableTiming));
  VH_HIP(hipEventRecord(m->stage_ev, m->stream));
  const long long nz = (long long)((zero_to - zero_from) / sizeof(float));
  if (nz > 0)
    hipLaunchKernelGGL(zero_kernel<float>, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, m->stream,
                       (float*)(base + zero_from), nz);
  VH_HIP(hipGetLastError());
  m->d_blocks = (BlockDev*)(base + o_blocks);
  m->d_tiles = (TileDev*)(base + o_tiles);
  m->d_repmap = (int*)(base + o_repmap);
  m->d_lcmap = (int*)(base + o_lcmap);
  m->d_linkcells = (LinkCellDev*)(base + o_lc);
  m->d_slot_faces = (int*)(base + o_sf);
  m->d_faces = (LinkFaceDev*)(base + o_faces);
  m->d_cfl_blocks = (int*)(base + o_cflb);
  m->d_cfl_prefix = (long long*)(base + o_cflp);
  m->d_flux = base + o_flux;
  m->d_ind = (double*)(base + o_ind);
  m->regs.assign(H.regs.size(), DevReg{});
  for (size_t l = 0; l < H.regs.size(); ++l) {
    DevReg& R = m->regs[l];
    R.n = (int)H.regs[l].cell.size();
    R.cells = (int*)(base + o_reg[3 * l]);
    R.lcs = (int*)(base + o_reg[3 * l + 1]);
    R.blk = (int*)(base + o_reg[3 * l + 2]);
    R.delta = base + o_delta[l];
  }
  if (extra_dev) *extra_dev = extra_bytes ? base + o_extra : nullptr;
  m->arena_cur = a;
  return 0;
}

// ---- self-gravity (src/gravity.jl) ----

// The composite operator A = −∇²·V of the mesh as CSR, from the face walk of
// for_each_face (src/gravity.jl:512-562, _assemble_poisson_struct): per leaf in
// leaf order, per axis the hi face when the neighbour is at the same level or
// coarser and the lo face when it is coarser; each face (left i, right j, area
// of the emitting leaf) is one edge with coef = area/d, d = ½(w_i + w_j),
// pushed to row i then row j. The gravity stencil takes the face walk's
// neighbours (a side no face reached is a Neumann boundary: the cell itself at
// half width); the coarse cell of a level jump takes its representative
// neighbours (src/gravity.jl:576-595).
struct HostCsr {
  std::vector<int> rowptr, colidx, glo, ghi;
  std::vector<double> coef, vol, gsum;
};

static int build_poisson(const Topo& T, HostCsr& C) {
  const long long n = T.off[T.nb];
  const int rank = T.rank;
  std::vector<int> ei, ej, lev(n);
  std::vector<double> ec, dlo(3 * n, 0.0), dhi(3 * n, 0.0);
  std::vector<char> marked(n, 0);
  ei.reserve(rank * n);
  ej.reserve(rank * n);
  ec.reserve(rank * n);
  C.vol.resize(n);
  C.glo.resize(3 * n);
  C.ghi.resize(3 * n);
  C.gsum.assign(3 * n, 0.0);
  for (int b = 0; b < T.nb; ++b)
    for (long long c = T.off[b]; c < T.off[b + 1]; ++c) {
      C.vol[c] = T.volume(b);
      lev[c] = T.level[b];
    }
  for (long long c = 0; c < n; ++c)
    for (int d = 0; d < 3; ++d) C.glo[d * n + c] = C.ghi[d * n + c] = (int)c;
  for (int b = 0; b < T.nb; ++b) {
    const int* nn = &T.dims[3 * b];
    for (long long k = 0; k < nn[2]; ++k)
      for (long long j = 0; j < nn[1]; ++j)
        for (long long i = 0; i < nn[0]; ++i) {
          const long long ijk[3] = {i, j, k};
          const long long g = T.index(b, ijk);
          for (int d = 0; d < rank; ++d) {
            const double area = T.area(b, d);
            for (int side = 1; side >= 0; --side) {
              long long rep = -1;
              int nbb = -1;
              const int kind = T.neighbor_at(b, ijk, d, side, &rep, nullptr, nullptr, &nbb);
              if (kind < 0) return fail(-1, "invalid neighbour while building the Poisson operator%s", "");
              if (!(kind == NB_COARSER || (side == 1 && kind == NB_SAME))) continue;
              const long long li = side ? g : rep, rj = side ? rep : g;
              const int bi = side ? b : nbb, bj = side ? nbb : b;
              const double fd = 0.5 * (T.width(bi, d) + T.width(bj, d));
              ei.push_back((int)li);
              ej.push_back((int)rj);
              ec.push_back(area / fd);
              C.ghi[d * n + li] = (int)rj;
              dhi[3 * li + d] = fd;
              C.glo[d * n + rj] = (int)li;
              dlo[3 * rj + d] = fd;
              if (lev[li] != lev[rj]) marked[lev[li] < lev[rj] ? li : rj] = 1;
            }
          }
        }
  }
  C.rowptr.assign(n + 1, 0);
  for (size_t e = 0; e < ei.size(); ++e) {
    C.rowptr[ei[e] + 1]++;
    C.rowptr[ej[e] + 1]++;
  }
  for (long long c = 0; c < n; ++c) C.rowptr[c + 1] += C.rowptr[c];
  C.colidx.resize(C.rowptr[n]);
  C.coef.resize(C.rowptr[n]);
  std::vector<int> cur(C.rowptr.begin(), C.rowptr.end());
  for (size_t e = 0; e < ei.size(); ++e) {
    const int p = cur[ei[e]]++;
    C.colidx[p] = ej[e];
    C.coef[p] = ec[e];
    const int q = cur[ej[e]]++;
    C.colidx[q] = ei[e];
    C.coef[q] = ec[e];
  }
  for (int b = 0; b < T.nb; ++b) {
    const int* nn = &T.dims[3 * b];
    for (long long k = 0; k < nn[2]; ++k)
      for (long long j = 0; j < nn[1]; ++j)
        for (long long i = 0; i < nn[0]; ++i) {
          const long long ijk[3] = {i, j, k};
          const long long c = T.index(b, ijk);
          for (int d = 0; d < rank; ++d) {
            const double hw = 0.5 * T.width(b, d);
            double dl, dh;
            if (!marked[c]) {
              dl = dlo[3 * c + d] == 0.0 ? hw : dlo[3 * c + d];
              dh = dhi[3 * c + d] == 0.0 ? hw : dhi[3 * c + d];
            } else {
              long long rep = -1;
              int nbb = -1;
              if (T.neighbor_at(b, ijk, d, 0, &rep, nullptr, nullptr, &nbb) == NB_BOUNDARY) {
                C.glo[d * n + c] = (int)c;
                dl = hw;
              } else {
                C.glo[d * n + c] = (int)rep;
                dl = 0.5 * (T.width(b, d) + T.width(nbb, d));
              }
              if (T.neighbor_at(b, ijk, d, 1, &rep, nullptr, nullptr, &nbb) == NB_BOUNDARY) {
                C.ghi[d * n + c] = (int)c;
                dh = hw;
              } else {
                C.ghi[d * n + c] = (int)rep;
                dh = 0.5 * (T.width(b, d) + T.width(nbb, d));
              }
            }
            C.gsum[d * n + c] = dl + dh;
          }
        }
  }
  return 0;
}

static void free_grav(vh_blocks* m) {
  BlockGrav* G = m->grav;
  if (!G) return;
  for (void* p : {(void*)G->rowptr, (void*)G->colidx, (void*)G->coef, (void*)G->vol, (void*)G->glo, (void*)G->ghi,
                  (void*)G->gsum, G->g, G->B.phi, G->B.work, G->B.state, (void*)G->B.scal, (void*)G->B.partials})
    (void)hipFree(p);
  if (G->B.host_state) (void)hipHostFree(G->B.host_state);
  delete G;
  m->grav = nullptr;
}

// g of the current φ into the step's field (before a step with self-gravity).
static void enqueue_grav_accel(vh_blocks* m) {
  grav_csr_accel(m->stream, m->dtype, m->grav->A, m->grav->B.phi, m->grav->g);
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

int vh_blocks_table_digest(int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0, const int32_t* level,
                           const int32_t* dims, const int64_t* origin, const int32_t* bc, int32_t zchunk,
                           int32_t nthreads, uint64_t* digest) {
  if (!digest) return fail(-1, "null argument%s", "");
  if (nthreads < 0) return fail(-1, "nthreads must be >= 0%s", "");
  Topo T;
  int rc = make_topo(T, rank, nblocks, maxlevel, h0, level, dims, origin, bc);
  if (rc) return rc;
  HostTables H;
  rc = build_tables(T, zchunk, H, 32, nthreads);
  if (rc) return rc;
  uint64_t h = 1469598103934665603ULL;
  digest_vec(h, H.blocks);
  digest_vec(h, H.repmap);
  digest_vec(h, H.lcmap);
  digest_vec(h, H.faces);
  digest_vec(h, H.slot_faces);
  {
    // the records in their unpacked form (type, start, count and grp per slot,
    // npre, ngrp, finer, 0), the layout the digests were first recorded with
    std::vector<std::array<int, 28>> wide(H.linkcells.size());
    for (size_t i = 0; i < wide.size(); ++i) {
      const LinkCellDev& R = H.linkcells[i];
      std::array<int, 28>& w = wide[i];
      for (int s = 0; s < 6; ++s) {
        w[s] = lc_type(R, s);
        w[6 + s] = lc_start(R, s);
        w[12 + s] = lc_count(R, s);
        w[18 + s] = lc_grp(R, s);
      }
      w[24] = lc_npre(R);
      w[25] = lc_ngrp(R);
      w[26] = lc_finer(R);
      w[27] = 0;
    }
    digest_vec(h, wide);
  }
  for (const RegList& r : H.regs) {
    digest_vec(h, r.cell);
    digest_vec(h, r.lc);
    digest_vec(h, r.block);
  }
  digest_vec(h, H.tiles);
  digest_vec(h, H.tile_list_off);
  digest_vec(h, H.block_tile_start);
  digest_vec(h, H.cfl_blocks);
  digest_vec(h, H.cfl_prefix);
  digest_vec(h, H.cfl_list_off);
  *digest = h;
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
  m->state_cap = bytes;
  rc = upload_tables(m, m->H, m->T.nb, 0);
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
  if (m->retired) (void)hipGraphExecDestroy(m->retired);
  if (m->stage_ev) (void)hipEventDestroy(m->stage_ev);
  free_grav(m);
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
  if (m->grav) enqueue_grav_accel(m);   // φ (hence g) held over the step
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
    // evolve! with subcycling (src/driver.jl:581-605): root dt from level 0;
    // with self-gravity one Poisson solve per root step, g held over every
    // substep of every level (src/driver.jl:589-592)
    while (t < tfinal * (1 - 1e-12) && (max_steps <= 0 || k < max_steps)) {
      if (m->grav) {
        long long it = 0;
        double rr = 0.0;
        m->grav->B.rho = m->U;
        if ((rc = grav_csr_solve(m->stream, m->dtype, m->grav->A, m->grav->P, m->grav->B, &it, &rr))) return rc;
        enqueue_grav_accel(m);
      }
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
      if (m->grav) {
        // solve_poisson! before each step (src/driver.jl:498), g held over the step;
        // the CG loop synchronises, so these steps are enqueued eagerly
        long long it = 0;
        double rr = 0.0;
        m->grav->B.rho = m->U;
        if ((rc = grav_csr_solve(m->stream, m->dtype, m->grav->A, m->grav->P, m->grav->B, &it, &rr))) return rc;
        enqueue_grav_accel(m);
        const int saved = m->use_graph;
        m->use_graph = 0;
        rc = single_rate_step(m, cfl, tfinal, clip ? 1 : 0, np);
        m->use_graph = saved;
        if (rc) return rc;
      } else if ((rc = single_rate_step(m, cfl, tfinal, clip ? 1 : 0, np))) {
        return rc;
      }
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

int vh_blocks_enable_gravity(vh_blocks* m, double G, int64_t maxiter, double tol, int32_t project_mean) {
  if (!m) return fail(-1, "null argument%s", "");
  if (!(G >= 0.0) || maxiter < 0 || !(tol > 0.0)) return fail(-1, "bad gravity parameters%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipStreamSynchronize(m->stream));
  drop_graph(m);    // captured steps hold the old kernel arguments
  free_grav(m);
  HostCsr C;
  int rc = build_poisson(m->T, C);
  if (rc) return rc;
  BlockGrav* Gv = new BlockGrav();
  m->grav = Gv;
  const long long n = m->T.off[m->T.nb];
  Gv->P.G = G;
  Gv->P.maxiter = maxiter;
  Gv->P.tol = tol;
  Gv->P.project_mean = project_mean ? 1 : 0;
  if ((rc = upload_vec(&Gv->rowptr, C.rowptr)) || (rc = upload_vec(&Gv->colidx, C.colidx)) ||
      (rc = upload_vec(&Gv->coef, C.coef)) || (rc = upload_vec(&Gv->vol, C.vol)) ||
      (rc = upload_vec(&Gv->glo, C.glo)) || (rc = upload_vec(&Gv->ghi, C.ghi)) || (rc = upload_vec(&Gv->gsum, C.gsum))) {
    free_grav(m);
    return rc;
  }
  hipError_t e = hipMalloc(&Gv->B.phi, m->esize * n);
  if (e == hipSuccess) e = hipMalloc(&Gv->B.work, 4 * m->esize * n);
  if (e == hipSuccess) e = hipMalloc(&Gv->g, 3 * m->esize * n);
  if (e == hipSuccess) e = hipMalloc(&Gv->B.state, kGravStateBytes);
  if (e == hipSuccess) e = hipMalloc(&Gv->B.scal, 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&Gv->B.partials, 4096 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&Gv->B.host_state, kGravStateBytes, 0);
  if (e == hipSuccess) e = hipMemsetAsync(Gv->B.phi, 0, m->esize * n, m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(Gv->g, 0, 3 * m->esize * n, m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(Gv->B.scal, 0, 8 * sizeof(double), m->stream);
  if (e == hipSuccess) e = hipMemsetAsync(Gv->B.state, 0, kGravStateBytes, m->stream);
  // Σ V accumulated cell by cell in leaf order in Float64 (the mean projections'
  // denominator, src/gravity.jl:230-246, 310-320)
  double sumv = 0.0;
  for (long long i = 0; i < n; ++i) sumv += C.vol[i];
  if (e == hipSuccess) e = hipMemcpyAsync(Gv->B.scal + 1, &sumv, sizeof(double), hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    free_grav(m);
    return fail(-2, "HIP error %s while enabling gravity%s", hipGetErrorString(e));
  }
  Gv->A.n = n;
  Gv->A.rowptr = Gv->rowptr;
  Gv->A.colidx = Gv->colidx;
  Gv->A.coef = Gv->coef;
  Gv->A.vol = Gv->vol;
  Gv->A.glo = Gv->glo;
  Gv->A.ghi = Gv->ghi;
  Gv->A.gsum = Gv->gsum;
  Gv->A.rank = m->T.rank;
  Gv->B.rho = m->U;
  return 0;
}

int vh_blocks_disable_gravity(vh_blocks* m) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipStreamSynchronize(m->stream));
  drop_graph(m);
  free_grav(m);
  return 0;
}

static int need_grav(vh_blocks* m) {
  if (!m) return fail(-1, "null argument%s", "");
  if (!m->grav) return fail(-1, "gravity is not enabled on this mesh%s", "");
  return 0;
}

int vh_blocks_solve_poisson(vh_blocks* m, int64_t* iters, double* relres) {
  if (int rc = need_grav(m)) return rc;
  VH_HIP(hipSetDevice(m->device));
  long long it = 0;
  double rr = 0.0;
  m->grav->B.rho = m->U;
  if (int rc = grav_csr_solve(m->stream, m->dtype, m->grav->A, m->grav->P, m->grav->B, &it, &rr)) return rc;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return 0;
}

int vh_blocks_phi_upload(vh_blocks* m, const void* host) {
  if (int rc = need_grav(m)) return rc;
  if (!host) return fail(-1, "null argument%s", "");
  VH_HIP(hipMemcpyAsync(m->grav->B.phi, host, m->esize * m->T.off[m->T.nb], hipMemcpyHostToDevice, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_phi_download(vh_blocks* m, void* host) {
  if (int rc = need_grav(m)) return rc;
  if (!host) return fail(-1, "null argument%s", "");
  VH_HIP(hipMemcpyAsync(host, m->grav->B.phi, m->esize * m->T.off[m->T.nb], hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_apply_laplacian(vh_blocks* m, const void* dx, void* dout) {
  if (int rc = need_grav(m)) return rc;
  if (!dx || !dout) return fail(-1, "null argument%s", "");
  grav_csr_apply(m->stream, m->dtype, m->grav->A, dx, dout);
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_blocks_gravity_accel(vh_blocks* m, void* dg) {
  if (int rc = need_grav(m)) return rc;
  if (!dg) return fail(-1, "null argument%s", "");
  grav_csr_accel(m->stream, m->dtype, m->grav->A, m->grav->B.phi, dg);
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(m->stream));
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
    HostTables& H = m->H2;
    int rc = build_tables(m->T, zchunk, H, m->max_tx);
    if (rc) return rc;
    std::swap(m->H, H);
    m->zchunk = zchunk;
    rc = upload_tables(m, m->H, m->T.nb, m->arena_cur);   // the stream is idle: the arena can be rewritten
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
  HostTables& H = m->H2;
  int rc = build_tables(m->T, m->zchunk, H, max_tile_x);
  if (rc) return rc;
  drop_graph(m);
  std::swap(m->H, H);
  m->max_tx = max_tile_x;
  rc = upload_tables(m, m->H, m->T.nb, m->arena_cur);   // the stream is idle: the arena can be rewritten
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

// The injection map of a remap from mesh S onto mesh D (src_block[b] = the
// block of S containing block b of D, at the same or a coarser level).
static int build_remap(const Topo& D, const Topo& S, const int32_t* src_block, std::vector<RemapDev>& map) {
  if (D.rank != S.rank || D.maxlevel != S.maxlevel) return fail(-1, "remap between meshes of different rank or maxlevel%s", "");
  for (int d = 0; d < D.rank; ++d)
    if (D.h0[d] != S.h0[d]) return fail(-1, "remap between meshes of different base cell widths%s (axis %lld)", "", d);
  map.resize(D.nb);
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
  return 0;
}

int vh_blocks_remap(vh_blocks* dst, const vh_blocks* src, const int32_t* src_block) {
  if (!dst || !src || !src_block) return fail(-1, "null argument%s", "");
  if (dst == src) return fail(-1, "remap needs two different contexts%s", "");
  const Topo& D = dst->T;
  const Topo& S = src->T;
  if (dst->dtype != src->dtype) return fail(-1, "remap between contexts of different dtypes%s", "");
  if (dst->device != src->device) return fail(-1, "remap between contexts on different devices%s", "");
  std::vector<RemapDev> map;
  int rc = build_remap(D, S, src_block, map);
  if (rc) return rc;
  VH_HIP(hipSetDevice(dst->device));
  VH_HIP(hipStreamSynchronize(src->stream));
  RemapDev* dmap = nullptr;
  rc = upload_vec(&dmap, map);
  if (rc) {
    (void)hipFree(dmap);
    return rc;
  }
  regrid_remap(dst->stream, dst->dtype, D.nb, dst->d_blocks, src->d_blocks, dmap, src->U, S.off[S.nb], dst->U,
               D.off[D.nb], 5);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(dst->stream);
  (void)hipFree(dmap);
  if (e != hipSuccess) return fail(-2, "HIP error %s in the remap%s", hipGetErrorString(e));
  return 0;
}

int vh_blocks_regrid(vh_blocks* m, int32_t nblocks, const int32_t* level, const int32_t* dims, const int64_t* origin,
                     const int32_t* bc, const int32_t* src_block) {
  if (!m || !src_block) return fail(-1, "null argument%s", "");
  Topo& N = m->T2;
  HostTables& H = m->H2;
  const double h0[3] = {m->T.h0[0], m->T.h0[1], m->T.h0[2]};
  int rc = make_topo(N, m->T.rank, nblocks, m->T.maxlevel, h0, level, dims, origin, bc);
  if (!rc) rc = build_tables(N, m->zchunk, H, m->max_tx);
  std::vector<RemapDev> map;
  if (!rc) rc = build_remap(N, m->T, src_block, map);
  if (rc) return rc;
  VH_HIP(hipSetDevice(m->device));
  // No wait for the stream: the new tables go to the spare arena, which the
  // steps since the last regrid do not read, and everything below is enqueued
  // behind them. The graph of the old tables is retired (destroyed once the
  // next regrid has seen the upload after its last launch complete).
  if (m->stage_ev) VH_HIP(hipEventSynchronize(m->stage_ev));
  if (m->retired) (void)hipGraphExecDestroy(m->retired);
  m->retired = m->graph;
  m->graph = nullptr;
  m->graph_np = -1;
  const long long n_old = m->T.off[m->T.nb], n_new = N.off[N.nb];
  const BlockDev* old_blocks = m->d_blocks;
  const int old_arena = m->arena_cur;
  // gravity: φ carried over by the same injection (the next solve's warm start)
  void* phi_new = nullptr;
  GravParams gp{};
  const bool grav = m->grav != nullptr;
  if (grav) {
    gp = m->grav->P;
    VH_HIP(hipMalloc(&phi_new, m->esize * std::max<long long>(1, n_new)));
  }
  void* dmap = nullptr;
  rc = upload_tables(m, H, N.nb, 1 - old_arena, map.data(), sizeof(RemapDev) * map.size(), &dmap);
  if (rc) {
    (void)hipFree(phi_new);
    return rc;   // the context keeps no valid tables: destroy it
  }
  // the state: remapped into U1 when it fits (then U and U1 swap), else into a
  // new allocation with headroom
  const size_t need = (size_t)5 * n_new * m->esize;
  void* dst = m->U1;
  bool fresh = false;
  if (need > m->state_cap) {
    const size_t cap = std::max<size_t>(2 * need, (size_t)4 << 20);
    VH_HIP(hipMalloc(&dst, cap));
    fresh = true;
    m->state_cap = cap;
  }
  regrid_remap(m->stream, m->dtype, N.nb, m->d_blocks, old_blocks, (const RemapDev*)dmap, m->U, n_old, dst, n_new, 5);
  if (grav)
    regrid_remap(m->stream, m->dtype, N.nb, m->d_blocks, old_blocks, (const RemapDev*)dmap, m->grav->B.phi, n_old,
                 phi_new, n_new, 1);
  hipError_t e = hipGetLastError();
  // the old state is freed (growth) and φ carried over (gravity) only after the remap
  if (e == hipSuccess && (fresh || grav)) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    if (fresh) (void)hipFree(dst);
    (void)hipFree(phi_new);
    return fail(-2, "HIP error %s in the regrid remap%s", hipGetErrorString(e));
  }
  if (fresh) {
    (void)hipFree(m->U);
    (void)hipFree(m->U1);
    m->U = dst;
    m->U1 = nullptr;
    if (hipMalloc(&m->U1, m->state_cap) != hipSuccess) return fail(-2, "device allocation of %s%lld bytes failed", "",
                                                                     (long long)m->state_cap);
  } else {
    std::swap(m->U, m->U1);
  }
  std::swap(m->T, N);
  std::swap(m->H, H);
  rc = alloc_partials(m);
  if (rc) return rc;
  if (grav) {
    rc = vh_blocks_enable_gravity(m, gp.G, gp.maxiter, gp.tol, gp.project_mean);
    if (!rc) {
      e = hipMemcpyAsync(m->grav->B.phi, phi_new, m->esize * n_new, hipMemcpyDeviceToDevice, m->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
      if (e != hipSuccess) rc = fail(-2, "HIP error %s carrying φ over%s", hipGetErrorString(e));
    }
    (void)hipFree(phi_new);
    if (rc) return rc;
  }
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
