// Host-side interface between the C ABI (sdf_abi.cpp) and the kernel
// translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>

#include "sdf_common.h"

namespace sdf {

// Thrown inside the library only; the ABI layer converts it to a status.
struct Error {
  int code;
  std::string msg;
};

#define SDF_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      throw ::sdf::Error{-2, std::string("HIP error in ") + #call + ": " + \
                                 hipGetErrorString(e_)};                   \
  } while (0)

// Grow-only named device buffers owned by a context.
class Arena {
 public:
  ~Arena() { release(); }
  void *get(const std::string &name, size_t bytes) {
    Buf &b = bufs_[name];
    if (b.bytes < bytes) {
      if (b.p) SDF_HIP(hipFree(b.p));
      b.p = nullptr;
      b.bytes = 0;
      size_t want = bytes < 256 ? 256 : bytes;
      SDF_HIP(hipMalloc(&b.p, want));
      b.bytes = want;
    }
    return b.p;
  }
  template <typename T>
  T *get(const std::string &name, size_t count) {
    return static_cast<T *>(get(name, count * sizeof(T)));
  }
  void release() {
    for (auto &kv : bufs_)
      if (kv.second.p) (void)hipFree(kv.second.p);
    bufs_.clear();
  }
  size_t bytes() const {
    size_t s = 0;
    for (auto &kv : bufs_) s += kv.second.bytes;
    return s;
  }
  // Prefix for scratch names (scan / sort temporaries) so that work issued
  // on concurrent streams uses distinct buffers.
  std::string scope;
  std::string tmp(const std::string &name) const { return scope + name; }

 private:
  struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Buf> bufs_;
};

// Per-point state that the fused kernel could not settle with the six axis
// rays; finished by the adaptive-vote kernel.  bits: parities of the rays
// the current vote phase traces (bit r for ray r), zero when the item is
// written, set by the ray stage.
struct Pending {
  int64_t idx;
  double dist;
  int32_t inside;
  int32_t used;
  unsigned long long bits;
};

// The sign kernel's pending lists: [0] points whose axis-ray decision fell
// inside a rounding margin (they vote from ray 1), [1] six votes split 5:1,
// [2] six votes split 4:2 or 3:3, which cannot reach the 0.85 confidence by
// ray 10 and take all their rays in the vote's first phase.
constexpr int kPendLists = 3;
// (no member initialisers: the sign kernel keeps a copy in LDS; value-
// initialise it, PendOut{}, on the host)
struct PendOut {
  Pending *p[kPendLists];
  uint32_t *n[kPendLists];
  uint32_t cap[kPendLists];
};

// An item of the winding-number fallback: the point, its distance and the
// row of its chunk sums when they were computed ahead (else -1: computed
// after the vote, at the item's position in the list).
struct WindItem {
  int64_t idx;
  double dist;
  int64_t row;
};

// Error flags written by kernels (bit mask).
enum : uint32_t {
  kErrBadIndex = 1u,       // IEN entry outside 1..nnp
  kErrPendingOverflow = 2u,
  kErrHitOverflow = 4u,    // one ray crossed more triangles than the buffer
  kErrStackOverflow = 8u,
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
};

// Sources of a status block: word i = *p[i] (32- or 64-bit), 0 when null.
struct StatusSrc {
  const void *p[16] = {};
  int is64[16] = {};
  int n = 0;
  void set(int i, const void *ptr, bool wide) {
    p[i] = ptr;
    is64[i] = wide ? 1 : 0;
  }
};

// ---------------- kernel launchers (implemented in .hip files) -----------
void launch_status(hipStream_t s, const StatusSrc &src, uint64_t *out);
// build side
void launch_fill_u32(hipStream_t s, uint32_t *p, int64_t n, uint32_t v);
void launch_fill_u64(hipStream_t s, uint64_t *p, int64_t n, uint64_t v);
void launch_fill_f64(hipStream_t s, double *p, int64_t n, double v);
void launch_gather_tris(hipStream_t s, const double *X, int64_t nnp,
                        const int64_t *IEN, int64_t nel, Tri64 *tri,
                        uint32_t *err);
// exclusive scan of n uint32 counts into n+1 uint64 offsets (out[n] = total)
void scan_counts(hipStream_t s, Arena &ar, const uint32_t *in, uint64_t *out,
                 int64_t n);

// Header of one distance work item (a chunk of <= clen of a tile's
// triangles), one 16-byte scalar load: the tile, the tile's first work item
// and, packed, the chunk's first index into the tile's triangle list
// (bits 0-47), its triangle count (bits 48-54) and whether it is the tile's
// only chunk (bit 55).
struct DistItem {
  uint32_t tile;
  uint32_t first;
  uint64_t range;
};
constexpr uint64_t kItemBeginMask = (1ull << 48) - 1;

// Brick rows' byte-1 / byte-7 segment masks (4 words each: m1 lo, hi, m7 lo,
// hi), the AND over a row's lines of theirs: x rows (by, bz) over bx, y rows
// (bx, bz) over by, z rows (bx, by) over bz; nbx, nby the sign-brick counts
// along x and y.
struct SameRows {
  uint64_t *x = nullptr, *y = nullptr, *z = nullptr;
  int32_t nbx = 0, nby = 0;
};

struct Bins {
  // tile binning (distance)
  uint64_t *tile_off = nullptr;  // per tile: first pair of its run (0 when empty)
  uint64_t *tile_end = nullptr;  // per tile: end of its run (0 when empty)
  uint32_t *tile_tris = nullptr;
  const uint32_t *tile_keys = nullptr;  // the sorted pairs' tile ids
  // distances by tile (d2buf slot = tile id, spare slot ntiles; the pair
  // windows of launch_dist_pairs) rather than by work item
  bool d2_by_tile = false;
  // tile_off / tile_end are left to tile_bins_bounds (d2's slots are set)
  bool bounds_deferred = false;
  // axis-line hits (signs)
  uint64_t *line_off = nullptr;  // nlines + 1
  uint32_t *line_hit_tri = nullptr;  // per hit: its triangle (hits grouped by line)
  int64_t line_base[4] = {0, 0, 0, 0};
  LineFast *line_fast = nullptr;   // per hit, sorted by c within each line
  uint32_t *line_meta = nullptr;   // per line: cluster count | kLineExact
  LineInline *line_inline = nullptr;  // per line: inline summary (<= 4 hits)
  // per brick row of each axis: the AND over its lines of their byte-1 and
  // byte-7 segment masks (null when a line has more than 128 segments)
  SameRows same_rows;
  uint32_t *seg_words = nullptr;      // sign-brick segment bytes, 4 per word (SegLayout),
                                      // then the sentinel word 0x01010101
  int64_t ntile_pairs = 0, nline_hits = 0;
  // distance work items: each non-empty tile's list in chunks of 64
  uint64_t *chunk_off = nullptr;  // ntiles + 1 (first work item of a tile)
  uint32_t *chunk_tile = nullptr;  // work item -> tile
  DistItem *chunk_item = nullptr;  // work item -> its header (distance kernel)
  int64_t nchunks = 0;
  // Capacities above; the actual counts live on the device (written by the
  // scans).  A count above its capacity means the lists were truncated and
  // the call is repeated with larger capacities.
  const uint64_t *npairs_dev = nullptr, *nhits_dev = nullptr, *nchunks_dev = nullptr;
  int32_t chunk_len = kTilePts;  // triangles per distance work item (<= 64)
};

struct FusedOut {
  double *sdf = nullptr;    // dist * sign (nullable)
  double *dist = nullptr;   // unsigned distance (nullable)
  double *signs = nullptr;  // +-1 (nullable)
  double *conf = nullptr;   // confidence (nullable)
};

// Most jobs (meshes) one queue launch serves.

// Distance work of one lattice: items (tile, chunk of <= clen triangles).
struct DistJob {
  GridP g;
  TileDims td;
  const Tri64 *tri;
  const uint64_t *tile_off;
  const uint32_t *tile_tris;
  const uint64_t *chunk_off;
  const uint32_t *chunk_tile;
  const DistItem *items;         // per work item: its header
  int64_t nchunks;               // capacity
  const uint64_t *nchunks_dev;   // actual count (null: the capacity)
  const uint32_t *only;          // per first item: 0 skips the tile (null: all)
  float *d2buf;
  int32_t clen;
  int32_t pad_;
  // pair windows: the sorted pairs (tile_keys, tile_tris), npairs capacity
  const uint32_t *keys;
  int64_t npairs;
  const uint64_t *npairs_dev;
};

// Sign / combine work of one lattice: 8 x 8 x kSignZ bricks.
// Fields of a sign job that only its rare paths read (a line's crossing
// list, the pending lists, error flags); the kernel keeps them in LDS, not
// in scalar registers.
struct FusedCold {
  const uint64_t *line_off;
  const LineFast *fast;
  const uint32_t *meta;
  PendOut pend;
  uint32_t *err;
};

struct FusedJob {
  GridP g;
  TileDims td;
  TileDims bd;  // brick counts
  const uint64_t *chunk_off;
  const uint64_t *tile_end;  // distances by tile (non-null): a tile has work when its run ends > 0
  const float *d2buf;
  const LineInline *inl;
  const BrickPlan *plan;     // per launch position (null: decoded in the kernel)
  const uint32_t *segw;      // segment bytes, 4 per word (SegLayout), then a sentinel
  int64_t segoff1, segoff2;  // first words of the y- and z-line parts
  int64_t segtotal;          // the sentinel word (0x01010101)
  int64_t base1, base2;
  FusedOut out;
  int64_t chunk_cap;
  int64_t nbricks;  // launch size of the brick order
  int32_t do_sign;
  int32_t pad_;
  FusedCold cold;
};
DistJob make_dist_job(const GridP &g, const Tri64 *tri, const Bins &b, float *d2buf,
                      const uint32_t *only);
FusedJob make_fused_job(const GridP &g, const Bins &b, const float *d2buf, bool do_sign,
                        FusedOut out, const PendOut &pend, uint32_t *err);
void build_tile_bins(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, double radius, Bins &b);
// Two-phase forms: *_begin issues the counting kernels and scans (no host
// wait); *_finish reads the total, sizes the lists and fills them.
struct TileBuild {
  double R = 0.0;
  uint32_t *tcnt = nullptr, *big = nullptr, *nbig = nullptr;
  uint64_t *toff = nullptr;
  // per tile: begin and end of its run in the sorted pair list (both 0 for
  // an empty tile; zeroed by the counting pass), ntiles + 1 each, adjacent
  uint64_t *bounds = nullptr;
};
struct LineBuild {
  int64_t nlx = 0, nly = 0, nlines = 0;
  uint32_t *lcnt = nullptr, *lcnt2 = nullptr, *big = nullptr, *nbig = nullptr;
  uint64_t *loff = nullptr;
  uint32_t *hslot = nullptr;  // per triangle: slots of its first hits in lcnt
  int64_t max_blocks = 0;  // grid cap of the per-triangle and per-line kernels (0: none)
  uint8_t *hitn = nullptr;   // per triangle: hits found by the counting pass (255: many)
  uint32_t *hitl = nullptr;  // their line ids, 4 per triangle
  bool same_masks = false;   // the brick rows' byte-1 / byte-7 masks (launch_brick_same)
};
void tile_bins_begin(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, double radius, TileBuild &tb);
// cap < 0: size the lists exactly (reads the count back: host wait).
// d2tile (nullable): the distance buffer by tile ((ntiles + 1) * kTilePts
// floats); its spare slot and the slots of tiles whose runs cross a pair
// window boundary are set to +inf (launch_dist_pairs).  defer_bounds (with
// d2tile): the run bounds (tile_off, tile_end), which the pair windows do
// not read, are left to tile_bins_bounds, which the caller issues on a
// stream of its choice after this one's work.
void tile_bins_finish(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                      const GridP &g, TileBuild &tb, Bins &b, int64_t cap = -1,
                      float *d2tile = nullptr, bool defer_bounds = false);
// The run bounds tile_bins_finish deferred (no-op otherwise).
void tile_bins_bounds(hipStream_t s, const GridP &g, const Bins &b);
// The sign bricks whose lines all carry segment byte 1, or all byte 7, at
// the brick: bits 2 and 3 of their plan flags (after launch_brick_plan and
// line_hits_finish; no-op without the lines' masks).  The sign kernel then
// writes such a brick without reading its lines.
// rest (nullable, one entry per launch position, *nrest zero): the launch
// positions of the other bricks.  False when nothing was marked.
bool launch_brick_same(hipStream_t s, const GridP &g, const Bins &b, BrickPlan *plan,
                       uint32_t *rest = nullptr, uint32_t *nrest = nullptr);
void line_hits_begin(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, LineBuild &lb, Bins &b);
void line_hits_finish(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                      const GridP &g, LineBuild &lb, Bins &b, int64_t cap = -1);
void build_line_hits(hipStream_t s, Arena &ar, const Tri64 *tri, int64_t nel,
                     const GridP &g, Bins &b);
// d2buf (nullable): also allocates the distance buffer (one slot of
// kTilePts floats per work item plus a spare slot of +inf after the last)
// and sets to +inf the slots that take atomicMin (first slots of tiles with
// several work items), the only slots read before they are written.
void build_dist_chunks(hipStream_t s, Arena &ar, const GridP &g, Bins &b, int64_t cap = -1,
                       int chunk_len = kTilePts, float **d2buf = nullptr);

// query side
// Distance work items -> squared distances per point of non-empty tiles
// (d2buf: 64 floats per work item slot, indexed by the tile's first item).
// max_blocks > 0 caps the grid (grid-stride over the items).
void launch_dist_chunks(hipStream_t s, const GridP &g, const Tri64 *tri,
                        const Bins &b, float *d2buf, const uint32_t *only = nullptr,
                        int64_t max_blocks = 0);
// Distance work items with a learned per-point squared radius r2buf (slots
// as d2buf); nodes whose radius held no triangle are appended to `list`.
// stats: [0] fallback nodes, [8..263] exact evaluations per wave (its
// longest candidate list), [264..519] (node, triangle) pairs within the
// learned radius (lower bound), [520..775] the same within h, [776..1031]
// the exact evaluations the plain loop makes (longest list within h).
void launch_dist_chunks_learned(hipStream_t s, const GridP &g, const Tri64 *tri,
                                const Bins &b, const float *r2buf, float *d2buf,
                                uint64_t *list, unsigned long long *stats);
// Squared distances of every non-empty tile's points from the sorted pairs
// in windows of kTilePts consecutive pairs (d2buf by tile, prepared by
// tile_bins_finish).  Grid: per_cu workgroups per CU (at most what the LDS
// allows; <= 0: that many) on ncu CUs.
void launch_dist_pairs(hipStream_t s, const GridP &g, const Tri64 *tri, const Bins &b,
                       float *d2buf, int per_cu, int ncu);
struct Bvh;
// Exact distances of the listed nodes (count in stats[0]) over the LBVH with
// the tile loop's arithmetic; overwrites their d2buf slots.
void launch_learned_fallback(hipStream_t s, const GridP &g, const Tri64 *tri, const Bvh &bvh,
                             const Bins &b, const uint64_t *list,
                             const unsigned long long *stats, float slack, float *d2buf,
                             uint32_t *err);
// select (sdf output with signs and distances by tile only): 1 the bricks
// launch_brick_same marked, 2 the bricks in list[0, *nlist) (the others);
// 0 every brick.
// Per tile: distance from d2buf (or h), six axis rays, combine, outputs.
// Points with a vote inside its rounding margin go to pend.p[0] (restart
// from ray 1 on the LBVH); non-unanimous votes go to pend.p[1] (5:1) or
// pend.p[2] (4:2, 3:3) and continue at ray 7.
// Launch positions of the sign kernel (BrickPlan entries it reads).
int64_t brick_plan_size(const GridP &g);
// The sign kernel's per-brick plan (coordinates, distance slots) from the
// distance items' offsets (d2buf null: coordinates only).
void launch_brick_plan(hipStream_t s, const GridP &g, const Bins &b, const float *d2buf,
                       BrickPlan *plan);
void launch_fused_lattice(hipStream_t s, const GridP &g, const Tri64 *tri,
                          const Bins &b, const float *d2buf, bool do_sign,
                          FusedOut out, const PendOut &pend, uint32_t *err,
                          const BrickPlan *plan, int64_t max_blocks = 0, int select = 0,
                          const uint32_t *list = nullptr, const uint32_t *nlist = nullptr);

struct RayParams {
  double dirs[kNumDirs * 3];
  int32_t min_rays, max_rays;
  double conf_thr;
  int32_t use_winding;
};

struct Bvh;

// The pending lists voted in one phase (at most three).  List k holds
// *n_dev[k] items (capacity cap[k]) whose rays below r0[k] are counted in
// their state; the phase traces rays [r0[k], r1[k]) of its items (an item
// whose state already counts a ray skips it) and the vote replay may use
// rays below avail[k] (kNumDirs: every ray the vote can use, so no item is
// carried).  An item's slot over the lists is its index in list 0, cap[0] +
// its index in list 1, and so on.  Items of list ahead_list with an index
// below ahead_rows have their winding chunk sums at row ahead_base + index
// of the ahead buffer.
struct PendLists {
  Pending *items[kPendLists] = {nullptr, nullptr, nullptr};
  const uint32_t *n_dev[kPendLists] = {nullptr, nullptr, nullptr};
  uint32_t cap[kPendLists] = {0, 0, 0};
  int32_t r0[kPendLists] = {0, 0, 0};
  int32_t r1[kPendLists] = {0, 0, 0};
  int32_t avail[kPendLists] = {0, 0, 0};
  int32_t count = 0;
  int32_t ahead_list = -1;
  uint32_t ahead_rows = 0;
  uint32_t ahead_base = 0;
  void add(Pending *p, const uint32_t *n, uint32_t c, int r) {
    items[count] = p;
    n_dev[count] = n;
    cap[count] = c;
    r0[count] = r;
    r1[count] = r;
    avail[count] = r;
    count++;
  }
  __host__ __device__ uint64_t base(int k) const {
    uint64_t b = 0;
    for (int i = 0; i < k; i++) b += cap[i];
    return b;
  }
  __host__ __device__ uint64_t slots() const { return base(count); }
};

// One phase of the adaptive vote, with no host waits (grids are fixed,
// counts are read on the device): the rays of every list (PendLists) through
// the LBVH (overflowing rays by brute force), then the vote replay.  Items
// whose vote needs a ray at or above their list's avail go to `next` (count
// *nnext, capacity next_cap) with their state and cleared ray bits; items
// whose final confidence is below the threshold go to the winding list
// `wind` (count *nwind, capacity wind_cap).  d_counts[0]: rays redone by
// brute force.  pts == nullptr: lattice points of g.
void resolve_pending(hipStream_t s, Arena &ar, const GridP &g, const double *pts,
                     const Tri64 *tri, int64_t nel, const Bvh &bvh, const RayParams &rp,
                     const PendLists &lists, FusedOut out, Pending *next, uint32_t *nnext,
                     uint32_t next_cap, WindItem *wind, uint32_t *nwind, uint32_t wind_cap,
                     uint32_t *d_counts, uint32_t *err, int stack = 1024);
// The winding fallback of the items in `wind` (count *nwind, at most
// wind_cap used): chunk sums of the items without an ahead row, then the
// totals, signs and outputs (ahead rows from `ahead`, complete once the
// stream reaches this call).
void resolve_winding(hipStream_t s, Arena &ar, const GridP &g, const double *pts,
                     const Tri64 *tri, int64_t nel, const WindItem *wind,
                     const uint32_t *nwind, uint32_t wind_cap, const double *ahead,
                     FusedOut out);
// Chunks of the winding-number sums over nel triangles (chunk: triangles
// per chunk).
int winding_chunks(int64_t nel, int64_t *chunk);
// The winding-number chunk sums of the first min(*count_dev, rows) items of
// one pending list, item i into ahead[(base + i) * chunks ...], ahead of the
// vote that sends some of them to the fallback.
void launch_winding_ahead(hipStream_t s, const GridP &g, const double *pts, const Tri64 *tri,
                          int64_t nel, const Pending *items, const uint32_t *count_dev,
                          uint32_t rows, uint32_t base, double *ahead);

// xp of evalDistancesOnTriMesh: exact FP64 closest point on the mesh for
// every lattice point (pts == nullptr) or caller point, 3 doubles per point.
void launch_closest_points(hipStream_t s, const GridP &g, const double *pts, int64_t npts,
                           const Tri64 *tri, const Bvh &bvh, double *xp, uint32_t *err);
void launch_dist_points_bvh(hipStream_t s, const double *pts, int64_t npts,
                            const Tri64 *tri, const Bvh &bvh, double h,
                            float slack, double *dist, uint32_t *err);

void launch_init_pending_all(hipStream_t s, Pending *pend, int64_t npts);

// connected components (sdf_ccl.hip)
struct CclResult {
  int64_t n = 0;
  int32_t *lab = nullptr;   // canonical root (smallest index) or -1
  uint32_t *size = nullptr; // component size, indexed by root
  unsigned long long *words = nullptr;  // [0] largest packed, [1] ncomp, [2] flipped
};
// Labels and sizes into caller-given buffers (n + 1 entries each, 4 words):
// allocates nothing.
void run_ccl_into(hipStream_t s, const GridP &g, const double *sdf, double thr,
                  CclResult &r);
CclResult run_ccl(hipStream_t s, Arena &ar, const GridP &g, const double *sdf,
                  double thr);
void ccl_flip(hipStream_t s, const CclResult &r, int32_t largest,
              uint32_t min_size, double *sdf);
void ccl_roots(hipStream_t s, const CclResult &r, int64_t *roots, int64_t *sizes,
               uint32_t *nroots, uint32_t cap);
void ccl_labels(hipStream_t s, const CclResult &r, int64_t *out);

// RBF smoothing (sdf_rbf.hip)
void rbf_prepare(hipStream_t s, Arena &ar, const double *d, int64_t n, float *f32);
int rbf_weights(hipStream_t s, Arena &ar, const GridP &g, const float *b,
                double threshold, double reltol, int max_outer, int max_inner,
                float *w_out, double *final_rel);
void rbf_fine(hipStream_t s, Arena &ar, const GridP &g, int smooth,
              double threshold, const float *w, float *fine);

}  // namespace sdf
