// C ABI of the MI355X SDF library (declared in include/sdf_mi355x.h).
//
// Orchestration of one call on the lattice path:
//   gather   1-based IEN + X -> per-triangle FP64 vertices
//   bin      triangles -> 8x4x2 point tiles within h; axis-line hit lists
//   fused    per tile: exact distance (LDS-staged triangles) + six axis rays
//            + combine; undecided points appended to a pending list
//   resolve  adaptive vote (rays 7..30) and winding fallback for pending
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sdf_internal.h"
#include "sdf_lbvh.h"
#include "sdf_mi355x.h"
#include "sdf_mlp.h"

// Host arithmetic that restates reference formulas (grid setup, ray
// directions) must round every operation separately.
#pragma clang fp contract(off)

using sdf::Error;

namespace sdf {
bool import_stl_file(const char *path, std::vector<double> &X, std::vector<int64_t> &IEN,
                     std::string &err);  // sdf_stl.cpp
}

namespace {

thread_local std::string g_last_error;

// device counters of a learned-radius run (layout: launch_dist_chunks_learned)
constexpr int kLearnedStats = 1032;

template <typename F>
sdf_status guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return SDF_OK;
  } catch (const Error &e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = std::string("internal error: ") + e.what();
    return SDF_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "internal error";
    return SDF_ERR_INTERNAL;
  }
}

[[noreturn]] void fail(int code, const std::string &msg) {
  throw Error{code, msg};
}

// generate_ray_directions_prioritized (RaycastingSignDetection.jl:162-227):
// 6 axes, 12 face diagonals, 8 body diagonals, each v * inv(norm(v)); then
// 10 golden-angle directions [r cos(theta), r sin(theta), z], not normalised.
void ray_directions(double *d) {
  static const double raw[26][3] = {
      {1, 0, 0},   {0, 1, 0},   {0, 0, 1},   {-1, 0, 0},  {0, -1, 0},
      {0, 0, -1},  {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0},
      {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},  {-1, 0, -1}, {0, 1, 1},
      {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},   {1, 1, -1},
      {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1},
      {-1, -1, -1}};
  for (int i = 0; i < 26; i++) {
    double s = raw[i][0] * raw[i][0];
    s = s + raw[i][1] * raw[i][1];
    s = s + raw[i][2] * raw[i][2];
    double inv = 1.0 / sqrt(s);
    for (int c = 0; c < 3; c++) d[3 * i + c] = raw[i][c] * inv;
  }
  const double phi = (1.0 + sqrt(5.0)) / 2.0;
  const double two_pi = 6.283185307179586;  // Float64(2pi)
  for (int i = 1; i <= 10; i++) {
    double x = (double)i * phi;
    double theta = two_pi * fmod(x, 1.0);
    double z = 1.0 - 2.0 * ((double)i / 10.0);
    double r = sqrt(1.0 - z * z);
    double *o = d + 3 * (25 + i);
    o[0] = r * cos(theta);
    o[1] = r * sin(theta);
    o[2] = z;
  }
}

// ---------------------------------------------------------------- static scratch
// Device memory of sdf_remove_artifacts_device_noalloc: component labels
// (int32) and sizes (uint32) for up to SDF_NOALLOC_MAX_NODES + 1 nodes, then
// four 64-bit words, each part 256-byte aligned.  One copy per GPU, held
// until the process ends; every call on a GPU shares it.
constexpr size_t kScrNodes = (size_t)SDF_NOALLOC_MAX_NODES + 1;
constexpr size_t kScrPart = (kScrNodes * 4 + 255) / 256 * 256;
constexpr size_t kScrLab = 0, kScrSize = kScrPart, kScrWords = 2 * kScrPart;
constexpr size_t kScratchBytes = 2 * kScrPart + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address, per device (the host-side name is only a
// stand-in): looked up once per device by the first context created on it,
// with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

void *scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) fail(SDF_ERR_ARG, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      fail(SDF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace

struct sdf_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  // helper streams for independent build chains (line hits, LBVH); default
  // priority, joined to the main stream with events
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  // [0] pass start, [1] line hits done, [2] LBVH done, [3] tile chain done
  hipEvent_t join[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_plan = nullptr;  // the sign kernel's brick plan is written
  // split sign stage: the distances are written (main stream), the marked
  // bricks are written (helper stream 1), and that pass's timing events
  hipEvent_t ev_dist = nullptr, ev_same = nullptr, ev_same_t[2] = {nullptr, nullptr};
  bool same_timed = false;
  // the vote's last phase: its item list is known, its winding sums are done
  hipEvent_t ev_wind[2] = {nullptr, nullptr};
  std::mutex mu;
  sdf::Arena ar;
  void *scratch = nullptr;  // this device's static scratch array
  // device mesh
  sdf::Tri64 *tri = nullptr;
  int64_t nel = 0, nnp = 0;
  bool mesh_ready = false;
  bool mesh_word_zero = false;  // the mesh error word was cleared by the last reset
  sdf::Bvh bvh;
  sdf::LineBuild lb;  // the line chain between its counting and emit passes
  bool bvh_ready = false;  // the ray hierarchy of the current mesh is built
  bool bvh_full = false;   // and the per-triangle one too
  // small device words: [0] error flags, [1] pending count (5:1 splits), [2]
  // pending count of ambiguous axis decisions, [3] the points path's item
  // count, [4..15] resolve-phase counters, [16] pending count (4:2, 3:3)
  uint32_t *d_words = nullptr;
  // Capacities of the lattice path's lists, from earlier calls (grow-only;
  // < 0: size exactly, reading the counts back).  With all set, a call
  // issues no host waits until its final status read.
  struct Caps {
    int64_t pairs = -1, hits = -1, chunks = -1;
    uint32_t pend[sdf::kPendLists] = {0, 0, 0};  // PendOut order
    uint32_t wind = 1024;
  } caps;
  uint64_t *d_status = nullptr;  // sizes of the last lattice call (16 words)
  bool own_status = true;        // d_status is this context's own h_status
  // pinned, coherent host block the status kernel writes directly (its
  // device view is d_status while own_status)
  uint64_t *h_status = nullptr;
  // tunables: triangles per distance work item, LDS ray-stack depth,
  // workgroups per CU of the distance and sign kernels (0: one per item).
  // The pair-window distance kernel's LDS allows 11 workgroups per CU; at 7
  // the line hits beside it get room and finish about when the distances do
  // (1M-triangle star, half-row jobs, LBVH built after the line hits: 6..8
  // measured, 1.090 ms per step at 7 against 1.10 at 6 and 8), while the
  // scalar-bound sign kernel needs all its waves.  chunk_len applies to the
  // learned radius's work items.
  int chunk_len = 64;
  int ray_stack = 1024;
  int dist_per_cu = 7, fused_per_cu = 0;
  int lines_early = 0;  // 1: the line chain starts with the tile chain; 2: its count pass
  int line_per_cu = 0;  // workgroups per CU of the line kernels (0: no cap)
  // LBVH build: 3 on helper stream 2 from the start of the pass, beside the
  // latency-bound tile chain (which it stretches by 0.03 ms), so the
  // distance kernel and the line hits later share the GPU with nothing
  // else (1.045-1.056 against 1.065-1.077 ms per step for 0); on helper
  // stream 1 after the tiles' run bounds and the brick plan: 0 issued after
  // the distance kernel and the line hits, 1 issued before them (no change
  // against 0), 2 after the line hits end (beside the sign stage: not done
  // when the vote needs it, 0.14 ms wait)
  int lbvh_early = 3;
  int sign_split = 1;  // split sign stage when the bricks are marked (0: one pass)
  // winding sums of the vote's likely winding items issued on helper stream
  // 0 beside the ray phases (0: only after the vote, for the items it
  // leaves).  On a closed surface the vote leaves next to none (one point of
  // 82.6 M on the star surface), and the sums issued ahead slow the ray
  // phases and the marked-brick pass beside them: 0.959-0.963 ms per step
  // without against 0.969-0.973 ms with
  int wind_ahead = 0;
  // workgroups per CU of its marked-brick pass (0: one per brick; capping it
  // to leave wave slots for the vote beside it was slower: 8, 16, 24 per CU
  // gave 1.04, 1.01, 1.03 against 0.99 ms per step)
  int same_per_cu = 0;
  int ncu = 256;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<std::string> ev_name;
  int nev = 0;
  int64_t counters[7] = {0, 0, 0, 0, 0, 0, 0};
  // the pending lists of the last lattice call (points handed to the vote)
  const sdf::Pending *last_pend[sdf::kPendLists] = {nullptr, nullptr, nullptr};
  uint64_t last_npend[sdf::kPendLists] = {0, 0, 0};
  // learned search radius
  sdf::MlpDev mlp;
  bool mlp_alloc = false;
  bool mlp_ready = false;     // parameters present (fitted or set)
  sdf::MlpNorm mlp_norm{};    // point normalisation of the fit
  double mlp_hc = 0.0;        // radius = (y + delta) * hc
  double mlp_delta = 0.0;
  unsigned long long *lstats = nullptr;  // counters of the last learned run
  bool learned_skipped = false;  // the last learned call ran without the network
  bool learned_force = false;    // tunable: evaluate the network even then
  int64_t lstat_pairs = 0;

  void mark(const char *name) {
    if (!timing) return;
    if (nev >= (int)ev.size()) {
      hipEvent_t e;
      SDF_HIP(hipEventCreate(&e));
      ev.push_back(e);
      ev_name.push_back("");
    }
    SDF_HIP(hipEventRecord(ev[nev], stream));
    ev_name[nev] = name;
    nev++;
  }
};

namespace {

void use_device(sdf_ctx *c) { SDF_HIP(hipSetDevice(c->device)); }

sdf::GridP make_grid(const double aabb_min[3], double h, const int64_t N[3]) {
  if (!aabb_min || !N) fail(SDF_ERR_ARG, "grid: NULL aabb_min or N");
  if (!(h > 0.0) || !isfinite(h)) fail(SDF_ERR_ARG, "grid: cell size must be > 0");
  for (int d = 0; d < 3; d++) {
    if (N[d] < 0 || N[d] + 1 > (int64_t)0x7FFFFFFF)
      fail(SDF_ERR_ARG, "grid: N out of range");
    if (!isfinite(aabb_min[d])) fail(SDF_ERR_ARG, "grid: aabb_min not finite");
  }
  sdf::GridP g;
  g.ox = aabb_min[0];
  g.oy = aabb_min[1];
  g.oz = aabb_min[2];
  g.h = h;
  g.nx = (int32_t)(N[0] + 1);
  g.ny = (int32_t)(N[1] + 1);
  g.nz = (int32_t)(N[2] + 1);
  g.pad_ = 0;
  return g;
}

int64_t grid_points(const sdf::GridP &g) {
  return (int64_t)g.nx * g.ny * g.nz;
}

// d_words[kMeshErrWord] holds the mesh upload's error flags (an IEN index
// outside 1..nnp); only a new mesh clears it, so every call on that mesh
// reports it, and uploading a device mesh needs no host wait.
constexpr int kMeshErrWord = 31;

// Clears the call's device words; with mesh, also the mesh upload's word
// (a mesh upload that follows then issues no fill of its own).
void reset_words(sdf_ctx *c, bool mesh = false) {
  sdf::launch_fill_u32(c->stream, c->d_words, mesh ? kMeshErrWord + 1 : kMeshErrWord, 0u);
  c->mesh_word_zero = mesh;
}

void check_err_flags(sdf_ctx *c) {
  uint32_t w[kMeshErrWord + 1];
  SDF_HIP(hipMemcpyAsync(w, c->d_words, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
  const uint32_t e = w[0] | w[kMeshErrWord];
  if (e & sdf::kErrBadIndex)
    fail(SDF_ERR_ARG, "IEN holds a vertex index outside 1..nnp");
  if (e & sdf::kErrHitOverflow)
    fail(SDF_ERR_LIMIT, "a ray crossed more triangles than the hit buffer holds");
  if (e & sdf::kErrStackOverflow)
    fail(SDF_ERR_LIMIT, "BVH traversal stack overflow");
  if (e & sdf::kErrPendingOverflow)
    fail(SDF_ERR_INTERNAL, "pending list overflow");
}

// Mesh from device X / IEN (layouts of TriangularMesh) into per-triangle
// FP64 vertices held by the context.
// A mesh upload refused after the call's reset cleared the mesh error word:
// the current mesh's error record is gone with it, so that mesh is dropped.
[[noreturn]] void mesh_arg_fail(sdf_ctx *c, const char *msg) {
  if (c->mesh_word_zero) c->mesh_ready = false;
  c->mesh_word_zero = false;
  fail(SDF_ERR_ARG, msg);
}

void set_mesh_device(sdf_ctx *c, const double *dX, int64_t nnp,
                     const int64_t *dIEN, int64_t nel) {
  if (nel < 0 || nnp < 0) mesh_arg_fail(c, "negative mesh size");
  if (nel > 0 && (!dX || !dIEN)) mesh_arg_fail(c, "NULL mesh arrays");
  if (nel > 0x7FFFFFFFll) mesh_arg_fail(c, "too many triangles");
  c->tri = c->ar.get<sdf::Tri64>("tri", nel + 1);
  c->nel = nel;
  c->nnp = nnp;
  c->mark("gather");
  if (!c->mesh_word_zero) sdf::launch_fill_u32(c->stream, c->d_words + kMeshErrWord, 1, 0u);
  c->mesh_word_zero = false;
  sdf::launch_gather_tris(c->stream, dX, nnp, dIEN, nel, c->tri, c->d_words + kMeshErrWord);
  c->mesh_ready = true;
  c->bvh_ready = false;
  c->bvh_full = false;
}

// Both hierarchies (nearest-point queries need the per-triangle one).
const sdf::Bvh &need_bvh(sdf_ctx *c) {
  if (!c->bvh_ready || !c->bvh_full) {
    c->mark("lbvh");
    c->bvh = sdf::build_lbvh(c->stream, c->ar, c->tri, c->nel);
    c->bvh_ready = true;
    c->bvh_full = true;
  }
  return c->bvh;
}

// Build the LBVH on helper stream q after the work recorded in join[3] (the
// end of the current pass's tile chain); joined later through join[2].
// Morton codes are quantised over the lattice's box (it covers the mesh the
// grid was built around; any box gives a valid hierarchy).
void bvh_async(sdf_ctx *c, int q, const sdf::GridP &g, bool full, int after = 3) {
  if (c->bvh_ready && (c->bvh_full || !full)) return;
  if (!c->aux[q]) SDF_HIP(hipStreamCreateWithFlags(&c->aux[q], hipStreamNonBlocking));
  SDF_HIP(hipStreamWaitEvent(c->aux[q], c->join[after], 0));
  c->ar.scope = "bvh/";
  sdf::QBox qb;
  const double o[3] = {g.ox, g.oy, g.oz};
  const int32_t nn[3] = {g.nx, g.ny, g.nz};
  for (int d = 0; d < 3; d++) {
    qb.lo[d] = (float)o[d];
    const double e = (double)(nn[d] > 1 ? nn[d] - 1 : 1) * g.h;
    qb.ext[d] = e > 0.0 ? (float)e : 1.0f;
  }
  c->bvh = sdf::build_lbvh(c->aux[q], c->ar, c->tri, c->nel, &qb, full);
  c->ar.scope.clear();
  c->bvh_ready = true;
  c->bvh_full = full;
  SDF_HIP(hipEventRecord(c->join[2], c->aux[q]));
}

// Adaptive vote for the items of the first lists (PendLists, r0 set).  List
// `spec` (-1: none) holds items that cannot finish by the first phase's end:
// its rays up to max_rays are traced in the first phase, so none of its
// items is handed on.  Phases, from a = the largest r0: the first min_rays
// rays (when a < min_rays), the next batch of 4, then the rest; every list
// takes part in the first phase, tracing its own rays from its r0, and each
// phase hands the items it cannot finish to the next, so later (and longer)
// phases trace rays only for the survivors.  Items below the confidence
// threshold after their last ray take the winding fallback after the last
// phase; the chunk sums of the spec list's items and of the last phase's
// items (the few the earlier phases left open) are computed on helper
// stream 0 beside the rays, rows of at most kWindAheadSlots items each.
// cnt: 4 device words per phase (at most 12), zero at the call's start:
// [0] rays redone by brute force, [1] items handed on; cnt[2] counts the
// winding items.
constexpr uint32_t kWindAheadSlots = 8192;

void resolve_items(sdf_ctx *c, const sdf::GridP &g, const double *d_pts,
                   const sdf::RayParams &rp, const sdf::PendLists &first, int spec,
                   sdf::FusedOut out, uint32_t *cnt) {
  const uint64_t slots = first.slots();
  if (slots == 0) return;
  const uint32_t cap = slots < 0xFFFFFFF0ull ? (uint32_t)slots : 0xFFFFFFF0u;
  const int rmax = rp.max_rays < sdf::kNumDirs ? rp.max_rays : sdf::kNumDirs;
  int a = first.r0[0];
  for (int k = 1; k < first.count; k++) a = first.r0[k] > a ? first.r0[k] : a;
  int ph[3], np = 0;  // phase ends
  if (a < rp.min_rays) {
    ph[np++] = rp.min_rays;
    a = rp.min_rays;
  }
  if (a + 4 < rmax) {
    ph[np++] = a + 4;
    a += 4;
  }
  ph[np++] = rmax;
  const bool wind = rp.use_winding != 0 && c->caps.wind > 0;
  // at most one winding item per item of the first phase
  sdf::WindItem *wl = c->ar.get<sdf::WindItem>("res_wind", cap);
  uint32_t *nwind = cnt + 2;
  int64_t chunk = 0;
  const int chunks = sdf::winding_chunks(c->nel, &chunk);
  const bool early = wind && c->wind_ahead;
  const uint32_t rows_spec =
      early && spec >= 0 ? std::min<uint32_t>(first.cap[spec], kWindAheadSlots) : 0u;
  const uint32_t rows_last = early && np > 1 ? std::min<uint32_t>(cap, kWindAheadSlots) : 0u;
  double *ahead = rows_spec + rows_last > 0
                      ? c->ar.get<double>("res_wahead", (size_t)(rows_spec + rows_last) * chunks)
                      : nullptr;
  bool ahead_issued = false;
  auto issue_ahead = [&](const sdf::Pending *items, const uint32_t *n, uint32_t rows,
                         uint32_t base) {
    SDF_HIP(hipEventRecord(c->ev_wind[0], c->stream));
    SDF_HIP(hipStreamWaitEvent(c->aux[0], c->ev_wind[0], 0));
    sdf::launch_winding_ahead(c->aux[0], g, d_pts, c->tri, c->nel, items, n, rows, base, ahead);
    ahead_issued = true;
  };
  static const char *names[2] = {"res_next0", "res_next1"};
  sdf::PendLists cur = first;
  for (int p = 0; p < np; p++) {
    const bool last = p + 1 == np;
    for (int k = 0; k < cur.count; k++) {
      const bool all = p == 0 && k == spec;
      cur.r1[k] = all ? rmax : ph[p];
      cur.avail[k] = last || all ? sdf::kNumDirs : ph[p];
    }
    if (p == 0 && rows_spec > 0) {
      issue_ahead(first.items[spec], first.n_dev[spec], rows_spec, 0u);
      cur.ahead_list = spec;
      cur.ahead_rows = rows_spec;
      cur.ahead_base = 0u;
    }
    if (last && p > 0 && rows_last > 0) {
      issue_ahead(cur.items[0], cur.n_dev[0], rows_last, rows_spec);
      cur.ahead_list = 0;
      cur.ahead_rows = rows_last;
      cur.ahead_base = rows_spec;
    }
    sdf::Pending *next = last ? nullptr : c->ar.get<sdf::Pending>(names[p], cap);
    sdf::resolve_pending(c->stream, c->ar, g, d_pts, c->tri, c->nel, c->bvh, rp, cur, out, next,
                         cnt + 4 * p + 1, cap, wl, nwind, cap, cnt + 4 * p, c->d_words,
                         c->ray_stack);
    cur = sdf::PendLists();
    cur.add(next, cnt + 4 * p + 1, cap, ph[p]);
  }
  if (wind) {
    if (ahead_issued) {
      SDF_HIP(hipEventRecord(c->ev_wind[1], c->aux[0]));
      SDF_HIP(hipStreamWaitEvent(c->stream, c->ev_wind[1], 0));
    }
    sdf::resolve_winding(c->stream, c->ar, g, d_pts, c->tri, c->nel, wl, nwind, c->caps.wind,
                         ahead, out);
  }
}

void set_mesh_host(sdf_ctx *c, const double *X, int64_t nnp, const int64_t *IEN,
                   int64_t nel) {
  if (nel < 0 || nnp < 0) mesh_arg_fail(c, "negative mesh size");
  if (nel > 0 && (!X || !IEN)) mesh_arg_fail(c, "NULL mesh arrays");
  double *dX = c->ar.get<double>("in_X", 3 * nnp + 3);
  int64_t *dI = c->ar.get<int64_t>("in_IEN", 3 * nel + 3);
  if (nnp > 0)
    SDF_HIP(hipMemcpyAsync(dX, X, (size_t)(3 * nnp) * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  if (nel > 0)
    SDF_HIP(hipMemcpyAsync(dI, IEN, (size_t)(3 * nel) * sizeof(int64_t),
                           hipMemcpyHostToDevice, c->stream));
  set_mesh_device(c, dX, nnp, dI, nel);
}

sdf::RayParams ray_params(int32_t min_rays, int32_t max_rays, double thr,
                          int32_t use_winding) {
  if (min_rays < 1 || min_rays > sdf::kNumDirs)
    fail(SDF_ERR_ARG, "min_rays must be in 1..36");
  if (max_rays < 0) fail(SDF_ERR_ARG, "max_rays must be >= 0");
  sdf::RayParams rp;
  ray_directions(rp.dirs);
  rp.min_rays = min_rays;
  // rays past the 36th are skipped by the reference's vote loop
  // (RaycastingSignDetection.jl:265-271): they change neither the count nor
  // the confidence, so the vote stops at 36 rays with the same result
  rp.max_rays = max_rays < sdf::kNumDirs ? max_rays : sdf::kNumDirs;
  rp.conf_thr = thr;
  rp.use_winding = use_winding ? 1 : 0;
  return rp;
}

// Lattice path: binning, fused kernel, adaptive finisher, as one pass with
// no host waits once the capacities are known (the first call on a context
// sizes its lists exactly).  The pass ends with the sizes gathered into
// d_status; run_lattice reads them once and repeats the pass with larger
// capacities when a list overflowed.
struct LatticeRun {
  sdf::GridP g;
  bool do_dist = false, do_sign = false, learned = false;
  sdf::RayParams rp;
  sdf::FusedOut out;
  sdf::Bins bins;
  float *d2buf = nullptr;
  bool bvh_was_ready = false;
  // the sign kernel's pending lists (PendOut): margin points, 5:1, 4:2 / 3:3
  uint32_t cap[sdf::kPendLists] = {0, 0, 0};
  sdf::Pending *pend[sdf::kPendLists] = {nullptr, nullptr, nullptr};
  sdf::BrickPlan *plan = nullptr;  // the sign kernel's per-brick plan
  // split sign stage: the bricks whose lines all agree on helper stream 1
  // (beside the others and the vote), the others listed in rest
  bool split = false;
  uint32_t *rest = nullptr;
};

enum : int {
  kStErr = 0, kStPend, kStPend0, kStPairs, kStHits, kStChunks, kStWind, kStPend2,
  kStMeshErr, kStNext0, kStCount
};

// device words of the pending lists' counts (PendOut order) and of the
// split sign stage's list of other bricks
constexpr int kPendWord[sdf::kPendLists] = {2, 1, 16};
constexpr int kRestWord = 17;

// Stage 1: the build chains.  Tile bins and distance work items run on the
// main stream; the axis-line hits (needed by the sign stage) on helper
// stream 0 and the LBVH (needed only by the vote stage) on helper stream 1,
// both from the end of the tile chain, beside the distance and sign
// kernels.  Issue order: the tile chain (stage_bins_tiles), the distance
// kernel (stage_dist), the line chain (stage_bins_lines), the LBVH
// (stage_bins_bvh), then the sign kernel: the distance kernel is queued
// before the host spends its time issuing the helper chains, so the main
// stream does not idle after the tile chain.
void stage_bins_lines(sdf_ctx *c, LatticeRun &L, int part);

void stage_bins_tiles(sdf_ctx *c, LatticeRun &L) {
  const sdf::GridP &g = L.g;
  const int64_t npts = grid_points(g);
  sdf::Bins &bins = L.bins;
  bins = sdf::Bins();
  SDF_HIP(hipEventRecord(c->join[0], c->stream));
  L.bvh_was_ready = c->bvh_ready;
  if (c->lines_early == 1) stage_bins_lines(c, L, 3);
  if (c->lines_early == 2) stage_bins_lines(c, L, 1);
  c->mark("bin_tiles");
  L.d2buf = nullptr;
  if (L.do_dist) {
    sdf::TileBuild tb;
    sdf::tile_bins_begin(c->stream, c->ar, c->tri, c->nel, g, g.h, tb);
    // the distance buffer, with a spare slot of +inf (the sign kernel reads
    // it for tiles without distance work, so its loads are unconditional):
    // by tile for the pair windows of the plain path, by work item for the
    // learned radius's chunks
    float *d2tile = nullptr;
    if (!L.learned)
      d2tile = c->ar.get<float>("d2tile", (size_t)(sdf::tile_dims(g).ntiles + 1) * sdf::kTilePts);
    // (plain path: the run bounds go to the brick plan's stream below, off
    // the distance kernel's path)
    sdf::tile_bins_finish(c->stream, c->ar, c->tri, c->nel, g, tb, bins, c->caps.pairs, d2tile,
                          !L.learned);
    if (L.learned)
      sdf::build_dist_chunks(c->stream, c->ar, g, bins, c->caps.chunks, c->chunk_len,
                             &L.d2buf);
    else
      L.d2buf = d2tile;
  }
  // the helper chains start when the tile chain is done: beside the
  // distance kernel (whose grid leaves room on every CU) they cost little,
  // beside the latency-bound tile chain they stretch every small kernel
  // (measured: about 35 us per step)
  SDF_HIP(hipEventRecord(c->join[3], c->stream));
  // the sign kernel's brick plan (coordinates and distance slots, a few
  // microseconds): with signs on helper stream 1 ahead of the LBVH, so the
  // distance kernel starts at once; the sign stage waits for ev_plan
  L.plan = c->ar.get<sdf::BrickPlan>("sign_plan", sdf::brick_plan_size(g) + 1);
  if (L.do_sign) {
    SDF_HIP(hipStreamWaitEvent(c->aux[1], c->join[3], 0));
    sdf::tile_bins_bounds(c->aux[1], g, bins);
    sdf::launch_brick_plan(c->aux[1], g, bins, L.d2buf, L.plan);
    SDF_HIP(hipEventRecord(c->ev_plan, c->aux[1]));
  } else {
    sdf::tile_bins_bounds(c->stream, g, bins);
    sdf::launch_brick_plan(c->stream, g, bins, L.d2buf, L.plan);
  }
  // issued after the tile chain (so that chain's launches are not delayed)
  // but free to start beside it
  if (L.do_sign && c->lbvh_early == 1) bvh_async(c, 1, L.g, L.learned, 0);
  // (issued after the tile chain's launches, which it would delay)
  if (L.do_sign && c->lbvh_early == 3) bvh_async(c, 2, L.g, L.learned, 0);
  const uint32_t dflt[sdf::kPendLists] = {1u << 20, 1u << 22, 1u << 20};
  static const char *pnames[sdf::kPendLists] = {"pend0", "pend", "pend2"};
  for (int k = 0; k < sdf::kPendLists; k++) {
    const uint32_t d = (uint32_t)(npts < (int64_t)dflt[k] ? (npts > 0 ? npts : 1) : dflt[k]);
    L.cap[k] = c->caps.pend[k] ? c->caps.pend[k] : d;
    L.pend[k] = c->ar.get<sdf::Pending>(pnames[k], L.cap[k]);
  }
}

// The axis-line hits on helper stream 0, joined through join[1].
// lines_early 0: all of it from the end of the tile chain (join[3]);
// 1: all of it from its start (join[0]); 2: the counting pass from its
// start, the emit and summary passes from its end.
void stage_bins_lines(sdf_ctx *c, LatticeRun &L, int part) {
  if (!L.do_sign) return;
  c->ar.scope = "line/";
  if (part & 1) {
    SDF_HIP(hipStreamWaitEvent(c->aux[0], c->join[c->lines_early ? 0 : 3], 0));
    c->lb = sdf::LineBuild();
    c->lb.max_blocks = (int64_t)c->line_per_cu * c->ncu;
    sdf::line_hits_begin(c->aux[0], c->ar, c->tri, c->nel, L.g, c->lb, L.bins);
  }
  if (part & 2) {
    if (c->lines_early == 2) SDF_HIP(hipStreamWaitEvent(c->aux[0], c->join[3], 0));
    c->lb.same_masks = c->sign_split != 0;
    sdf::line_hits_finish(c->aux[0], c->ar, c->tri, c->nel, L.g, c->lb, L.bins, c->caps.hits);
    // the bricks whose lines all agree: marked in the brick plan (written on
    // helper stream 1), the others listed, for a split sign stage (sdf
    // output with distances by tile only)
    const bool splittable = c->sign_split && L.d2buf && L.bins.d2_by_tile && L.out.sdf &&
                            !L.out.dist && !L.out.signs && !L.out.conf;
    if (splittable) {
      SDF_HIP(hipStreamWaitEvent(c->aux[0], c->ev_plan, 0));
      L.rest = c->ar.get<uint32_t>("sign_rest", (size_t)sdf::brick_plan_size(L.g) + 1);
      L.split = sdf::launch_brick_same(c->aux[0], L.g, L.bins, L.plan, L.rest,
                                       c->d_words + kRestWord);
    }
    SDF_HIP(hipEventRecord(c->join[1], c->aux[0]));
  }
  c->ar.scope.clear();
}

void stage_bins_bvh(sdf_ctx *c, LatticeRun &L) {
  // the vote stage needs only the rays' one; lbvh_early 2: built after the
  // axis-line chain (beside the sign stage rather than the distance kernel)
  if (L.do_sign) bvh_async(c, 1, L.g, false, c->lbvh_early == 2 ? 1 : 3);
}

void stage_bins(sdf_ctx *c, LatticeRun &L) {
  stage_bins_tiles(c, L);
  if (c->lines_early != 1) stage_bins_lines(c, L, c->lines_early == 2 ? 2 : 3);
  stage_bins_bvh(c, L);
}

// Stage 2: exact distances of the tile points (plain or learned radius).
void stage_dist(sdf_ctx *c, LatticeRun &L) {
  if (!L.do_dist) return;
  const sdf::GridP &g = L.g;
  sdf::Bins &bins = L.bins;
  c->mark("dist");
  if (L.learned) {
    // network radius per tile point, then the radius-limited distance
    // loop; nodes whose radius held no triangle get exact LBVH queries
    float *r2 = c->ar.get<float>("l_r2", (size_t)bins.nchunks * sdf::kTilePts + 1);
    uint64_t *list = c->ar.get<uint64_t>("l_list", (size_t)bins.nchunks * sdf::kTilePts + 1);
    c->lstats = c->ar.get<unsigned long long>("l_stats", kLearnedStats);
    sdf::launch_fill_u64(c->stream, (uint64_t *)c->lstats, kLearnedStats, 0ull);
    c->mark("radius");
    sdf::mlp_tile_radius(c->stream, c->mlp, c->mlp_norm, g, bins, (float)c->mlp_hc,
                         (float)c->mlp_delta, r2);
    c->mark("dist_learned");
    sdf::launch_dist_chunks_learned(c->stream, g, c->tri, bins, r2, L.d2buf, list, c->lstats);
    if (!L.bvh_was_ready || !c->bvh_full) {  // the fallback queries the per-triangle one
      bvh_async(c, 1, g, true);  // (issued already when it started with the tile chain)
      SDF_HIP(hipStreamWaitEvent(c->stream, c->join[2], 0));
    }
    c->mark("fallback");
    const float slack = 1e-6f * (float)(fabs(g.ox) + fabs(g.oy) + fabs(g.oz) +
                                        (double)(g.nx + g.ny + g.nz) * g.h + 1e-30);
    sdf::launch_learned_fallback(c->stream, g, c->tri, c->bvh, bins, list, c->lstats, slack,
                                 L.d2buf, c->d_words);
  } else {
    sdf::launch_dist_pairs(c->stream, g, c->tri, bins, L.d2buf, c->dist_per_cu, c->ncu);
  }
}

// Stage 3: signs and combine (waits for the axis-line hits).
void stage_sign(sdf_ctx *c, LatticeRun &L) {
  c->mark("bin_lines_wait");
  if (L.do_sign) {
    SDF_HIP(hipStreamWaitEvent(c->stream, c->join[1], 0));
    SDF_HIP(hipStreamWaitEvent(c->stream, c->ev_plan, 0));
  }
  sdf::PendOut po{};
  for (int k = 0; k < sdf::kPendLists; k++) {
    po.p[k] = L.pend[k];
    po.n[k] = c->d_words + kPendWord[k];
    po.cap[k] = L.cap[k];
  }
  c->same_timed = false;
  if (L.split) {
    // the marked bricks (store-bound, most of the grid) on helper stream 1
    // once the distances and the marks are done; the others, which alone
    // hand points to the vote, here, so the vote runs beside the former
    SDF_HIP(hipEventRecord(c->ev_dist, c->stream));
    SDF_HIP(hipStreamWaitEvent(c->aux[1], c->ev_dist, 0));
    SDF_HIP(hipStreamWaitEvent(c->aux[1], c->join[1], 0));
    if (c->timing) SDF_HIP(hipEventRecord(c->ev_same_t[0], c->aux[1]));
    sdf::launch_fused_lattice(c->aux[1], L.g, c->tri, L.bins, L.d2buf, L.do_sign, L.out, po,
                              c->d_words, L.plan, (int64_t)c->same_per_cu * c->ncu, 1);
    if (c->timing) SDF_HIP(hipEventRecord(c->ev_same_t[1], c->aux[1]));
    c->same_timed = c->timing;
    SDF_HIP(hipEventRecord(c->ev_same, c->aux[1]));
    c->mark("fused");
    sdf::launch_fused_lattice(c->stream, L.g, c->tri, L.bins, L.d2buf, L.do_sign, L.out, po,
                              c->d_words, L.plan, 0, 2, L.rest, c->d_words + kRestWord);
    return;
  }
  c->mark("fused");
  sdf::launch_fused_lattice(c->stream, L.g, c->tri, L.bins, L.d2buf, L.do_sign, L.out, po,
                            c->d_words, L.plan, (int64_t)c->fused_per_cu * c->ncu);
}

// Stage 4: adaptive vote for the pending points (waits for the LBVH), then
// the sizes of this pass into d_status.
void stage_resolve(sdf_ctx *c, LatticeRun &L) {
  uint32_t *cnt = c->d_words + 4;
  if (L.do_sign) {
    if (!L.bvh_was_ready) {
      c->mark("lbvh_wait");
      SDF_HIP(hipStreamWaitEvent(c->stream, c->join[2], 0));
    }
    c->mark("resolve");
    // points whose axis-ray decisions fell inside a rounding margin vote
    // from ray 1, in the same launches as the points the six axis rays left
    // undecided (which vote from ray min_rays + 1); those split 4:2 or 3:3
    // take all their rays in the first phase (list 2)
    sdf::PendLists lists;
    for (int k = 0; k < sdf::kPendLists; k++)
      lists.add(L.pend[k], c->d_words + kPendWord[k], L.cap[k], k == 0 ? 0 : L.rp.min_rays);
    resolve_items(c, L.g, nullptr, L.rp, lists, 2, L.out, cnt);
  }
  if (L.split) {  // the marked bricks' pass on helper stream 1
    c->mark("same_wait");
    SDF_HIP(hipStreamWaitEvent(c->stream, c->ev_same, 0));
  }
  c->mark("end");
  sdf::StatusSrc src;
  src.n = kStCount;
  src.set(kStErr, c->d_words, false);
  src.set(kStMeshErr, c->d_words + kMeshErrWord, false);
  src.set(kStPend, c->d_words + kPendWord[1], false);
  src.set(kStPend0, c->d_words + kPendWord[0], false);
  src.set(kStPend2, c->d_words + kPendWord[2], false);
  src.set(kStPairs, L.bins.npairs_dev, true);
  src.set(kStHits, L.bins.nhits_dev, true);
  src.set(kStChunks, L.bins.nchunks_dev, true);
  if (L.do_sign) {
    src.set(kStNext0, cnt + 1, false);
    src.set(kStWind, cnt + 2, false);
  }
  sdf::launch_status(c->stream, src, c->d_status);
}

void lattice_pass(sdf_ctx *c, LatticeRun &L) {
  stage_bins_tiles(c, L);
  stage_dist(c, L);  // the learned variant builds the LBVH itself, first
  if (c->lines_early != 1) stage_bins_lines(c, L, c->lines_early == 2 ? 2 : 3);
  stage_bins_bvh(c, L);
  stage_sign(c, L);
  stage_resolve(c, L);
}

// Reads the status of the last pass; grows the capacities that overflowed
// (true: the pass must be repeated).  Reports errors otherwise.
bool lattice_settle(sdf_ctx *c, LatticeRun &L, const uint64_t *st) {
  bool again = false;
  auto grow64 = [&](int64_t &cap, uint64_t need, bool used) {
    if (!used) return;
    const int64_t want = (int64_t)(need + need / 8 + 1024);
    if (cap >= 0 && (int64_t)need > cap) again = true;
    if (cap < want && ((int64_t)need > cap || cap < 0)) cap = want;
  };
  grow64(c->caps.pairs, st[kStPairs], L.do_dist && c->nel > 0);
  grow64(c->caps.chunks, st[kStChunks], L.do_dist && L.learned);
  grow64(c->caps.hits, st[kStHits], L.do_sign);
  const int pst[sdf::kPendLists] = {kStPend0, kStPend, kStPend2};
  for (int k = 0; k < sdf::kPendLists; k++) {
    const uint64_t n = st[pst[k]];
    if (n > L.cap[k]) {
      c->caps.pend[k] = (uint32_t)std::min<uint64_t>(n + n / 8 + 1024, 0xFFFFFFF0u);
      again = true;
    }
  }
  const uint64_t w = st[kStWind];
  if (w > c->caps.wind) {
    c->caps.wind = (uint32_t)std::min<uint64_t>(w + w / 8 + 64, 0xFFFFFFF0u);
    again = true;
  }
  const uint32_t e = ((uint32_t)st[kStErr] | (uint32_t)st[kStMeshErr]) &
                     ~(uint32_t)sdf::kErrPendingOverflow;
  if (e & sdf::kErrBadIndex) fail(SDF_ERR_ARG, "IEN holds a vertex index outside 1..nnp");
  if (again) return true;
  if (e & sdf::kErrHitOverflow)
    fail(SDF_ERR_LIMIT, "a ray crossed more triangles than the hit buffer holds");
  if (e & sdf::kErrStackOverflow) fail(SDF_ERR_LIMIT, "BVH traversal stack overflow");
  if (st[kStErr] & sdf::kErrPendingOverflow) fail(SDF_ERR_INTERNAL, "pending list overflow");
  return false;
}

// The status block of the last pass: the context's own block is pinned host
// memory the status kernel wrote directly (no copy); a batch lane's lives in
// the batch's device block.
void read_status(sdf_ctx *c, uint64_t *st) {
  if (c->own_status) {
    SDF_HIP(hipStreamSynchronize(c->stream));
    const volatile uint64_t *h = c->h_status;
    for (int i = 0; i < 16; i++) st[i] = h[i];
    return;
  }
  SDF_HIP(hipMemcpyAsync(st, c->d_status, 16 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
}

void run_lattice(sdf_ctx *c, const sdf::GridP &g, bool do_dist, bool do_sign,
                 const sdf::RayParams &rp, sdf::FusedOut out, bool learned = false) {
  if (!c->mesh_ready) fail(SDF_ERR_ARG, "no mesh");
  if (learned) {
    // the learned radius is min(max((max(y, 0) + delta) * hc, h / 100), h):
    // when delta * hc >= h it is h at every point whatever the network
    // outputs, so it would prune nothing and the plain distance loop (whose
    // radius is h) runs instead, without the network
    c->learned_skipped = !c->learned_force && c->mlp_delta * c->mlp_hc >= g.h;
    if (c->learned_skipped) {
      learned = false;
      c->lstats = nullptr;  // no counters for this call
    }
  }
  LatticeRun L;
  L.g = g;
  L.do_dist = do_dist;
  L.do_sign = do_sign;
  L.learned = learned;
  L.rp = rp;
  L.out = out;
  for (int attempt = 0;; attempt++) {
    if (attempt > 0) {
      reset_words(c);
      c->nev = 0;
    }
    lattice_pass(c, L);
    uint64_t st[16];
    read_status(c, st);
    if (!lattice_settle(c, L, st)) {
      c->counters[0] = (int64_t)(st[kStPend] + st[kStPend0] + st[kStPend2]);
      const uint64_t pn[sdf::kPendLists] = {st[kStPend0], st[kStPend], st[kStPend2]};
      for (int k = 0; k < sdf::kPendLists; k++) {
        c->last_pend[k] = L.do_sign ? L.pend[k] : nullptr;
        c->last_npend[k] = L.do_sign ? std::min<uint64_t>(pn[k], L.cap[k]) : 0;
      }
      c->counters[1] = (int64_t)st[kStPairs];
      c->counters[2] = (int64_t)st[kStHits];
      c->counters[3] = (int64_t)c->ar.bytes();
      c->counters[4] = L.do_sign ? (int64_t)st[kStNext0] : 0;
      c->counters[5] = L.do_sign ? (int64_t)st[kStWind] : 0;
      c->counters[6] = L.do_sign ? (int64_t)st[kStPend0] : 0;
      c->lstat_pairs = (int64_t)st[kStPairs];
      return;
    }
    if (attempt >= 3) fail(SDF_ERR_INTERNAL, "list capacities did not settle");
  }
}

bool is_lattice(const sdf::GridP &g, const double *pts, int64_t npts) {
  if (!pts) return true;
  if (npts != grid_points(g)) return false;
  int64_t a = 0;
  for (int64_t k = 0; k < g.nz; k++) {
    double z = sdf::lattice_coord(g.oz, g.h, k);
    for (int64_t j = 0; j < g.ny; j++) {
      double y = sdf::lattice_coord(g.oy, g.h, j);
      for (int64_t i = 0; i < g.nx; i++, a++) {
        const double *p = pts + 3 * a;
        if (p[1] != y || p[2] != z || p[0] != sdf::lattice_coord(g.ox, g.h, i))
          return false;
      }
    }
  }
  return true;
}

// mesh: the call uploads a mesh next (its error word is cleared here too)
void begin_call(sdf_ctx *c, bool mesh = false) {
  use_device(c);
  c->nev = 0;
  reset_words(c, mesh);
}


// ---------------------------------------------------------------- learned radius
void need_mlp(sdf_ctx *c) {
  if (!c->mlp_alloc) {
    sdf::mlp_alloc(c->ar, c->mlp);
    c->mlp_alloc = true;
  }
}

// SplitMix64: deterministic stream for initialisation and shuffles.
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// He-uniform weights (fan-in 100 active features for layer 1, 128 after),
// zero hidden biases, output bias 0.5.
void mlp_init_params(uint64_t seed, std::vector<float> &p) {
  p.assign(sdf::kMlpParams, 0.f);
  SplitMix r{seed};
  const int offs[3] = {sdf::kOffW1, sdf::kOffW2, sdf::kOffW3};
  const double fan[3] = {100.0, 128.0, 128.0};
  for (int L = 0; L < 3; L++) {
    const double a = sqrt(6.0 / fan[L]);
    for (int i = 0; i < 16384; i++) p[offs[L] + i] = (float)((2.0 * r.uniform() - 1.0) * a);
  }
  const double a4 = sqrt(3.0 / 128.0);
  for (int i = 0; i < 128; i++) p[sdf::kOffW4 + i] = (float)((2.0 * r.uniform() - 1.0) * a4);
  p[sdf::kOffB4] = 0.5f;
}

void mlp_upload(sdf_ctx *c, const float *params) {
  need_mlp(c);
  SDF_HIP(hipMemcpyAsync(c->mlp.params, params, sizeof(float) * sdf::kMlpParams,
                         hipMemcpyHostToDevice, c->stream));
  sdf::launch_fill_u32(c->stream, (uint32_t *)c->mlp.m, sdf::kMlpParams, 0u);
  sdf::launch_fill_u32(c->stream, (uint32_t *)c->mlp.v, sdf::kMlpParams, 0u);
  c->mlp.step = 0;
  sdf::mlp_pack(c->stream, c->mlp);
  c->mlp_ready = true;
}

// Batch order for `steps` steps of `batch` samples out of n: concatenated
// Fisher-Yates shuffles (seed != 0) or the samples in order (seed == 0).
std::vector<uint32_t> batch_order(int64_t n, int64_t steps, int64_t batch, uint64_t seed) {
  std::vector<uint32_t> perm((size_t)n), out((size_t)(steps * batch));
  for (int64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  SplitMix r{seed ^ 0xA5A5A5A5DEADBEEFull};
  int64_t pos = n;  // exhausted: reshuffle
  for (int64_t k = 0; k < steps * batch; k++) {
    if (pos == n) {
      if (seed != 0)
        for (int64_t i = n - 1; i > 0; i--) {
          const int64_t j = (int64_t)(r.next() % (uint64_t)(i + 1));
          std::swap(perm[i], perm[j]);
        }
      pos = 0;
    }
    out[k] = perm[pos++];
  }
  return out;
}

// Adam over `steps` batches with a cosine learning-rate decay to lr/20.
// Returns the mean loss of the last min(20, steps) steps.
double mlp_train(sdf_ctx *c, const float4 *d_samples, int64_t n, int32_t steps, int32_t batch,
                 float lr, uint64_t seed, float *loss_hist_host) {
  if (steps < 1 || batch < 1) fail(SDF_ERR_ARG, "steps and batch must be >= 1");
  if (lr <= 0.f) fail(SDF_ERR_ARG, "lr must be > 0");
  if ((int64_t)batch > n) batch = (int32_t)n;
  std::vector<uint32_t> order = batch_order(n, steps, batch, seed);
  uint32_t *d_idx = c->ar.get<uint32_t>("mlp_idx", order.size());
  SDF_HIP(hipMemcpyAsync(d_idx, order.data(), order.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice, c->stream));
  float *d_loss = c->ar.get<float>("mlp_loss", steps);
  for (int32_t t = 0; t < steps; t++) {
    const double frac = steps > 1 ? (double)t / (double)(steps - 1) : 0.0;
    const float lrt = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(3.14159265358979323846 * frac))));
    sdf::mlp_step(c->stream, c->ar, c->mlp, d_samples, d_idx + (int64_t)t * batch, batch,
                  nullptr, d_loss + t, lrt);
  }
  std::vector<float> hist((size_t)steps);
  SDF_HIP(hipMemcpyAsync(hist.data(), d_loss, sizeof(float) * steps, hipMemcpyDeviceToHost,
                         c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
  if (loss_hist_host) memcpy(loss_hist_host, hist.data(), sizeof(float) * steps);
  const int k = steps < 20 ? steps : 20;
  double m = 0.0;
  for (int i = steps - k; i < steps; i++) m += hist[i];
  return m / k;
}

float4 *upload_samples(sdf_ctx *c, const float *u, const float *t, int64_t n) {
  std::vector<float> hs((size_t)(4 * n));
  for (int64_t i = 0; i < n; i++) {
    hs[4 * i] = u[3 * i];
    hs[4 * i + 1] = u[3 * i + 1];
    hs[4 * i + 2] = u[3 * i + 2];
    hs[4 * i + 3] = t ? t[i] : 0.f;
  }
  float4 *d = c->ar.get<float4>("mlp_user_samples", n + 1);
  SDF_HIP(hipMemcpyAsync(d, hs.data(), hs.size() * sizeof(float), hipMemcpyHostToDevice,
                         c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
  return d;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}


// Context set-up and release shared by sdf_ctx_create / sdf_ctx_destroy
// and the batch's worker contexts (own stream, two helper streams, events,
// status words).
void ctx_init(sdf_ctx *c, int32_t device) {
  c->device = device;
  SDF_HIP(hipSetDevice(device));
  SDF_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
  c->stream = c->own;
  // (helper stream 2, for the LBVH beside the tile chain, is created on
  // first use: batch lanes never use it, and every stream shares the
  // device's few hardware queues)
  for (int q = 0; q < 2; q++) SDF_HIP(hipStreamCreateWithFlags(&c->aux[q], hipStreamNonBlocking));
  for (int q = 0; q < 4; q++) SDF_HIP(hipEventCreateWithFlags(&c->join[q], hipEventDisableTiming));
  SDF_HIP(hipEventCreateWithFlags(&c->ev_plan, hipEventDisableTiming));
  SDF_HIP(hipEventCreateWithFlags(&c->ev_dist, hipEventDisableTiming));
  SDF_HIP(hipEventCreateWithFlags(&c->ev_same, hipEventDisableTiming));
  for (int q = 0; q < 2; q++) SDF_HIP(hipEventCreate(&c->ev_same_t[q]));
  for (int q = 0; q < 2; q++) SDF_HIP(hipEventCreateWithFlags(&c->ev_wind[q], hipEventDisableTiming));
  SDF_HIP(hipMalloc(&c->d_words, 256));
  SDF_HIP(hipHostMalloc((void **)&c->h_status, 16 * sizeof(uint64_t),
                        hipHostMallocMapped | hipHostMallocCoherent));
  SDF_HIP(hipHostGetDevicePointer((void **)&c->d_status, c->h_status, 0));
  SDF_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
  c->scratch = scratch_address(device);  // with `device` current
}

void ctx_free(sdf_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  // the helper streams' work is done before the buffers it uses go
  for (int q = 0; q < 3; q++)
    if (c->aux[q]) (void)hipStreamSynchronize(c->aux[q]);
  c->ar.release();
  for (auto e : c->ev) (void)hipEventDestroy(e);
  for (int q = 0; q < 3; q++)
    if (c->aux[q]) (void)hipStreamDestroy(c->aux[q]);
  for (int q = 0; q < 4; q++)
    if (c->join[q]) (void)hipEventDestroy(c->join[q]);
  if (c->ev_plan) (void)hipEventDestroy(c->ev_plan);
  if (c->ev_dist) (void)hipEventDestroy(c->ev_dist);
  if (c->ev_same) (void)hipEventDestroy(c->ev_same);
  for (int q = 0; q < 2; q++)
    if (c->ev_same_t[q]) (void)hipEventDestroy(c->ev_same_t[q]);
  for (int q = 0; q < 2; q++)
    if (c->ev_wind[q]) (void)hipEventDestroy(c->ev_wind[q]);
  if (c->d_words) (void)hipFree(c->d_words);
  if (c->h_status) (void)hipHostFree(c->h_status);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

// ---------------------------------------------------------------- batches
// Many meshes and grids in one submission.  The meshes are dealt round-robin
// to `lanes` worker contexts (each set up as sdf_ctx_create sets up a
// context: own stream, helper streams, buffers); every mesh keeps its own
// list capacities, which its lane takes on before the mesh's pass.  A run
// issues every mesh's lattice pass with no host wait (capacities known from
// earlier runs), so the lanes' passes overlap on the GPU, then reads every
// mesh's status block once; a mesh whose lists outgrew their capacities is
// computed again on its own.  Results equal sdf_signed_distance_device's
// bit for bit.
struct BatchTune {
  int lanes = 4;            // worker contexts
  int threads = 0;          // 1: each lane's passes issued by a host thread of its own
  int chunk_len = 64;       // triangles per distance work item
  int ray_stack = 1024;     // LDS ray-stack depth of the vote stage
};

struct BatchMesh {
  const double *dX = nullptr;
  const int64_t *dIEN = nullptr;
  int64_t nnp = 0, nel = 0;
  sdf::GridP g;
  double *out = nullptr;
  sdf_ctx::Caps caps;  // this mesh's list capacities
};

constexpr int kMaxLanes = 8;
constexpr int kMaxBatchMeshes = 4096;

}  // namespace

struct sdf_batch {
  sdf_ctx *parent = nullptr;
  std::vector<sdf_ctx *> lanes;  // kMaxLanes worker contexts
  std::vector<BatchMesh> mesh;
  std::vector<LatticeRun> runs;
  uint64_t *d_status = nullptr;  // 16 words per mesh
  BatchTune tune;
  std::vector<double> tune_log;  // (candidate id, ms) pairs
  int64_t total_points = 0;
  double last_ms = 0.0;
  double last_issue_ms = 0.0;  // host time to issue the last run's passes
};

namespace {

void batch_free(sdf_batch *b) {
  if (!b) return;
  for (sdf_ctx *k : b->lanes) ctx_free(k);
  if (b->d_status) (void)hipFree(b->d_status);
  delete b;
}

// LBVH start of the batch's lanes: after the tile chain, on helper stream 1
// (the small meshes' passes already overlap across lanes, and a third
// helper stream per lane crowds the device's hardware queues: with one
// created in every lane, 7.5-7.6 against 9.2 G points/s).
constexpr int kBatchLbvh = 0;

// Issue mesh i's lattice pass on its lane, status into the batch's block.
void batch_issue(sdf_batch *b, int i) {
  sdf_ctx *k = b->lanes[i % b->tune.lanes];
  BatchMesh &m = b->mesh[(size_t)i];
  k->caps = m.caps;
  k->chunk_len = b->tune.chunk_len;
  k->ray_stack = b->tune.ray_stack;
  k->lbvh_early = kBatchLbvh;
  k->sign_split = 0;  // (the batch's lanes overlap their passes already)
  k->d_status = b->d_status + 16 * (int64_t)i;
  k->nev = 0;
  reset_words(k, true);
  set_mesh_device(k, m.dX, m.nnp, m.dIEN, m.nel);
  LatticeRun &L = b->runs[(size_t)i];
  L = LatticeRun();
  L.g = m.g;
  L.do_dist = true;
  L.do_sign = true;
  L.rp = ray_params(6, 30, 0.6, 1);
  L.out.sdf = m.out;
  lattice_pass(k, L);
  m.caps = k->caps;  // capacities a first pass sized exactly
}

// Mesh i alone, until its capacities hold (after an overflow).
void batch_redo(sdf_batch *b, int i) {
  sdf_ctx *k = b->lanes[i % b->tune.lanes];
  BatchMesh &m = b->mesh[(size_t)i];
  k->caps = m.caps;
  k->chunk_len = b->tune.chunk_len;
  k->ray_stack = b->tune.ray_stack;
  k->lbvh_early = kBatchLbvh;
  k->sign_split = 0;  // (the batch's lanes overlap their passes already)
  k->d_status = b->d_status + 16 * (int64_t)i;
  k->nev = 0;
  reset_words(k, true);
  set_mesh_device(k, m.dX, m.nnp, m.dIEN, m.nel);
  sdf::FusedOut out;
  out.sdf = m.out;
  run_lattice(k, m.g, true, true, ray_params(6, 30, 0.6, 1), out);
  m.caps = k->caps;
}

// Every mesh alone on its lane, waiting for each: sizes every mesh's lists
// and grows each lane's buffers to the largest of its meshes, so that the
// overlapped runs after it allocate nothing while other lanes' passes are
// in flight.  Needed whenever the mesh-to-lane assignment changes.
void batch_plan(sdf_batch *b) {
  for (int i = 0; i < (int)b->mesh.size(); i++) batch_redo(b, i);
}

// One run over every mesh; returns its wall time in ms (the issue, the
// passes and the status read; re-runs of overflowed meshes excluded).
double batch_run(sdf_batch *b) {
  const auto t0 = std::chrono::steady_clock::now();
  const int n = (int)b->mesh.size();
  const int nl = b->tune.lanes < n ? b->tune.lanes : n;
  if (b->tune.threads && nl > 1) {
    // lane l issues meshes l, l + lanes, ...: lanes share no host state
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs((size_t)nl);
    const int dev = b->parent->device;
    for (int l = 0; l < nl; l++)
      th.emplace_back([&, l] {
        try {
          SDF_HIP(hipSetDevice(dev));
          for (int i = l; i < n; i += b->tune.lanes) batch_issue(b, i);
        } catch (...) {
          errs[(size_t)l] = std::current_exception();
        }
      });
    for (auto &t : th) t.join();
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);
  } else {
    for (int i = 0; i < n; i++) batch_issue(b, i);
  }
  b->last_issue_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int l = 0; l < b->tune.lanes && l < n; l++)
    SDF_HIP(hipStreamSynchronize(b->lanes[(size_t)l]->stream));
  std::vector<uint64_t> st((size_t)16 * n);
  SDF_HIP(hipMemcpy(st.data(), b->d_status, st.size() * sizeof(uint64_t),
                    hipMemcpyDeviceToHost));
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < n; i++) {
    sdf_ctx *k = b->lanes[i % b->tune.lanes];
    BatchMesh &m = b->mesh[(size_t)i];
    k->caps = m.caps;
    bool again;
    try {
      again = lattice_settle(k, b->runs[(size_t)i], st.data() + 16 * (int64_t)i);
    } catch (const Error &e) {
      throw Error{e.code, "mesh " + std::to_string(i) + ": " + e.msg};
    }
    m.caps = k->caps;
    if (again) batch_redo(b, i);
  }
  return ms;
}

// Times candidate issue modes (one host thread, or one per lane), lane
// counts, work-item lengths and ray-stack depths, in that order (best of
// two runs each, after a planning pass for the candidate), and keeps the
// fastest of each; ties keep the earlier candidate.
void batch_autotune(sdf_batch *b) {
  auto measure = [&](int id) {
    batch_plan(b);
    const double ms = std::min(batch_run(b), batch_run(b));
    b->tune_log.push_back((double)id);
    b->tune_log.push_back(ms);
    return ms;
  };
  auto pick = [&](int *field, std::initializer_list<int> cand, int base) {
    double best = 1e30;
    int keep = *field;
    for (int v : cand) {
      *field = v;
      const double ms = measure(base + v);
      if (ms < best) {
        best = ms;
        keep = v;
      }
    }
    *field = keep;
  };
  pick(&b->tune.threads, {0, 1}, 10000);
  pick(&b->tune.lanes, {1, 2, 3, 4, 6, 8}, 0);
  pick(&b->tune.chunk_len, {64, 32, 16}, 100);
  pick(&b->tune.ray_stack, {1024, 2048}, 1000);
}

}  // namespace

extern "C" {

int32_t sdf_abi_version(void) { return 100; }

const char *sdf_last_error(void) { return g_last_error.c_str(); }

sdf_status sdf_ctx_create(int32_t device, sdf_ctx **out) {
  return guarded([&] {
    if (!out) fail(SDF_ERR_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    SDF_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
      fail(SDF_ERR_ARG, "device " + std::to_string(device) + " not present (" +
                            std::to_string(n) + " devices)");
    sdf_ctx *c = new sdf_ctx();
    try {
      ctx_init(c, device);
    } catch (...) {
      ctx_free(c);
      throw;
    }
    *out = c;
  });
}

sdf_status sdf_ctx_destroy(sdf_ctx *c) {
  return guarded([&] { ctx_free(c); });
}

sdf_status sdf_ctx_set_stream(sdf_ctx *c, void *s) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    c->stream = s ? (hipStream_t)s : c->own;
  });
}

sdf_status sdf_ctx_synchronize(sdf_ctx *c) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    use_device(c);
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_ctx_tune(sdf_ctx *c, const char *name, int64_t value) {
  return guarded([&] {
    if (!c || !name) fail(SDF_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    const std::string n = name;
    if (n == "chunk_len") {
      if (value < 1 || value > sdf::kTilePts) fail(SDF_ERR_ARG, "chunk_len must be in 1..64");
      c->chunk_len = (int)value;
    } else if (n == "ray_stack") {
      if (value != 1024 && value != 2048) fail(SDF_ERR_ARG, "ray_stack must be 1024 or 2048");
      c->ray_stack = (int)value;
    } else if (n == "lbvh_early") {
      if (value < 0 || value > 3) fail(SDF_ERR_ARG, "lbvh_early must be in 0..3");
      c->lbvh_early = (int)value;
    } else if (n == "lines_early") {
      if (value < 0 || value > 2) fail(SDF_ERR_ARG, "lines_early must be 0, 1 or 2");
      c->lines_early = (int)value;
    } else if (n == "sign_split") {
      if (value != 0 && value != 1) fail(SDF_ERR_ARG, "sign_split must be 0 or 1");
      c->sign_split = (int)value;
    } else if (n == "wind_ahead") {
      if (value != 0 && value != 1) fail(SDF_ERR_ARG, "wind_ahead must be 0 or 1");
      c->wind_ahead = (int)value;
    } else if (n == "learned_force") {
      if (value != 0 && value != 1) fail(SDF_ERR_ARG, "learned_force must be 0 or 1");
      c->learned_force = value != 0;
    } else if (n == "dist_per_cu" || n == "fused_per_cu" || n == "line_per_cu" ||
               n == "same_per_cu") {
      if (value < 0 || value > 64) fail(SDF_ERR_ARG, n + " must be in 0..64");
      (n == "dist_per_cu"    ? c->dist_per_cu
       : n == "fused_per_cu" ? c->fused_per_cu
       : n == "line_per_cu"  ? c->line_per_cu
                             : c->same_per_cu) = (int)value;
    } else {
      fail(SDF_ERR_ARG, "unknown tunable " + n);
    }
  });
}

sdf_status sdf_ctx_enable_timing(sdf_ctx *c, int32_t on) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    c->timing = on != 0;
  });
}

sdf_status sdf_ctx_stage_times(sdf_ctx *c, char *names, int32_t names_cap,
                               double *ms, int32_t cap, int32_t *n) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    use_device(c);
    std::string all;
    int w = 0;
    if (c->nev > 0) SDF_HIP(hipEventSynchronize(c->ev[c->nev - 1]));
    for (int i = 0; i + 1 < c->nev; i++) {
      float t = 0.f;
      SDF_HIP(hipEventElapsedTime(&t, c->ev[i], c->ev[i + 1]));
      if (w < cap && ms) ms[w] = t;
      if (!all.empty()) all += ";";
      all += c->ev_name[i];
      w++;
    }
    if (c->same_timed) {  // the split sign stage's pass on helper stream 1
      float t = 0.f;
      SDF_HIP(hipEventSynchronize(c->ev_same_t[1]));
      SDF_HIP(hipEventElapsedTime(&t, c->ev_same_t[0], c->ev_same_t[1]));
      if (w < cap && ms) ms[w] = t;
      if (!all.empty()) all += ";";
      all += "fused_same";
      w++;
    }
    if (names && names_cap > 0) {
      size_t m = all.size() < (size_t)(names_cap - 1) ? all.size() : (size_t)(names_cap - 1);
      memcpy(names, all.data(), m);
      names[m] = 0;
    }
    if (n) *n = w;
  });
}

sdf_status sdf_ctx_counters(sdf_ctx *c, int64_t *out, int32_t cap, int32_t *n) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    int w = 0;
    for (int i = 0; i < 7 && i < cap; i++, w++) out[i] = c->counters[i];
    if (n) *n = w;
  });
}

sdf_status sdf_ctx_vote_points(sdf_ctx *c, int64_t *idx_out, int64_t cap, int64_t *n) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    if (cap < 0 || (cap > 0 && !idx_out)) fail(SDF_ERR_ARG, "idx_out is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    std::vector<int64_t> all;
    for (int k = 0; k < sdf::kPendLists; k++) {
      if (!c->last_pend[k] || c->last_npend[k] == 0) continue;
      std::vector<sdf::Pending> h((size_t)c->last_npend[k]);
      SDF_HIP(hipMemcpyAsync(h.data(), c->last_pend[k], h.size() * sizeof(sdf::Pending),
                             hipMemcpyDeviceToHost, c->stream));
      SDF_HIP(hipStreamSynchronize(c->stream));
      for (const sdf::Pending &p : h) all.push_back(p.idx);
    }
    std::sort(all.begin(), all.end());
    for (int64_t i = 0; i < cap && i < (int64_t)all.size(); i++) idx_out[i] = all[(size_t)i];
    if (n) *n = (int64_t)all.size();
  });
}

sdf_status sdf_grid_setup(const double xmin[3], const double xmax[3],
                          int64_t n_max, int64_t margin, double aabb_min[3],
                          double aabb_max[3], int64_t N[3], double *cell_size,
                          int64_t *ngp) {
  return guarded([&] {
    if (!xmin || !xmax || !aabb_min || !aabb_max || !N || !cell_size || !ngp)
      fail(SDF_ERR_ARG, "NULL argument");
    if (n_max < 1) fail(SDF_ERR_ARG, "N_max must be >= 1");
    // Grid.jl:15-29, operation for operation
    double ext = xmax[0] - xmin[0];
    for (int d = 1; d < 3; d++) {
      double e = xmax[d] - xmin[d];
      if (e > ext) ext = e;
    }
    double h = ext / (double)n_max;
    double mh = (double)margin * h;
    int64_t prod = 1;
    for (int d = 0; d < 3; d++) {
      double lo = xmin[d] - mh;
      double hi = xmax[d] + mh;
      double cells = ceil((hi - lo) / h);
      N[d] = (int64_t)cells;
      double nh = (double)N[d] * h;
      aabb_min[d] = lo;
      aabb_max[d] = lo + nh;
      prod *= N[d] + 1;
    }
    *cell_size = h;
    *ngp = prod;
  });
}

namespace {

sdf::GridP lattice_only(const int64_t N[3]) {
  const double o[3] = {0.0, 0.0, 0.0};
  return make_grid(o, 1.0, N);
}

// scratch: NULL takes the buffers from the context's arena; otherwise the
// static scratch array (no allocation, no stage events).
int64_t remove_artifacts_dev(sdf_ctx *c, double *d_sdf, const sdf::GridP &g,
                             double thr, double ratio, void *scratch = nullptr) {
  if (grid_points(g) > 0x7FFFFFFFll)
    fail(SDF_ERR_LIMIT, "artifact removal supports < 2^31 grid points");
  sdf::CclResult r;
  if (scratch) {
    unsigned char *base = (unsigned char *)scratch;
    r.lab = (int32_t *)(base + kScrLab);
    r.size = (uint32_t *)(base + kScrSize);
    r.words = (unsigned long long *)(base + kScrWords);
    sdf::run_ccl_into(c->stream, g, d_sdf, thr, r);
  } else {
    c->mark("ccl");
    r = sdf::run_ccl(c->stream, c->ar, g, d_sdf, thr);
  }
  unsigned long long w[2] = {0, 0};
  SDF_HIP(hipMemcpyAsync(w, r.words, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
  if (w[1] == 0) return 0;  // no interior nodes (SdfArtifactRemoval.jl:152-155)
  const uint32_t largest_size = (uint32_t)(w[0] >> 32);
  const int32_t largest = (int32_t)(0xFFFFFFFFu - (uint32_t)(w[0] & 0xFFFFFFFFull));
  // max(1, round(Int, ratio * largest)) with round-half-even (:209)
  double x = nearbyint(ratio * (double)largest_size);
  uint32_t min_size = x < 1.0 ? 1u : (x > 4.0e9 ? 0xFFFFFFFFu : (uint32_t)x);
  if (!scratch) c->mark("flip");
  sdf::ccl_flip(c->stream, r, largest, min_size, d_sdf);
  unsigned long long f = 0;
  SDF_HIP(hipMemcpyAsync(&f, r.words + 2, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  SDF_HIP(hipStreamSynchronize(c->stream));
  if (!scratch) c->mark("end");
  return (int64_t)f;
}

}  // namespace

sdf_status sdf_remove_artifacts(sdf_ctx *c, double *sdf_inout, const int64_t N[3],
                                double thr, double ratio, int64_t *nflipped) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    const int64_t n = grid_points(g);
    if (n > 0 && !sdf_inout) fail(SDF_ERR_ARG, "sdf is NULL");
    begin_call(c);
    double *d = c->ar.get<double>("art_sdf", n + 1);
    if (n > 0)
      SDF_HIP(hipMemcpyAsync(d, sdf_inout, (size_t)n * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    int64_t f = remove_artifacts_dev(c, d, g, thr, ratio);
    if (f > 0)
      SDF_HIP(hipMemcpyAsync(sdf_inout, d, (size_t)n * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
    if (nflipped) *nflipped = f;
  });
}

sdf_status sdf_remove_artifacts_device(sdf_ctx *c, double *d_sdf, const int64_t N[3],
                                       double thr, double ratio, int64_t *nflipped) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    if (grid_points(g) > 0 && !d_sdf) fail(SDF_ERR_ARG, "sdf is NULL");
    begin_call(c);
    int64_t f = remove_artifacts_dev(c, d_sdf, g, thr, ratio);
    if (nflipped) *nflipped = f;
  });
}

sdf_status sdf_remove_artifacts_device_noalloc(sdf_ctx *c, double *d_sdf, const int64_t N[3],
                                               double thr, double ratio,
                                               int64_t *nflipped) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    const int64_t n = grid_points(g);
    if (n > (int64_t)SDF_NOALLOC_MAX_NODES)
      fail(SDF_ERR_LIMIT, "grid has " + std::to_string(n) + " nodes; the allocation-free "
                          "entry point takes at most " +
                              std::to_string((int64_t)SDF_NOALLOC_MAX_NODES));
    if (n > 0 && !d_sdf) fail(SDF_ERR_ARG, "sdf is NULL");
    if (!c->scratch) fail(SDF_ERR_INTERNAL, "no scratch address for this device");
    use_device(c);
    const int64_t f = remove_artifacts_dev(c, d_sdf, g, thr, ratio, c->scratch);
    if (nflipped) *nflipped = f;
  });
}

sdf_status sdf_analyze_components(sdf_ctx *c, const double *sdf, const int64_t N[3],
                                  double thr, int64_t *labels_out, int64_t *roots_out,
                                  int64_t *sizes_out, int64_t cap, int64_t *ncomp) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    const int64_t n = grid_points(g);
    if (n > 0x7FFFFFFFll) fail(SDF_ERR_LIMIT, "component analysis supports < 2^31 grid points");
    if (n > 0 && !sdf) fail(SDF_ERR_ARG, "sdf is NULL");
    if (cap < 0 || (cap > 0 && (!roots_out || !sizes_out)))
      fail(SDF_ERR_ARG, "roots_out / sizes_out is NULL");
    begin_call(c);
    double *d = c->ar.get<double>("art_sdf", n + 1);
    if (n > 0)
      SDF_HIP(hipMemcpyAsync(d, sdf, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    sdf::CclResult r = sdf::run_ccl(c->stream, c->ar, g, d, thr);
    unsigned long long w[2] = {0, 0};
    SDF_HIP(hipMemcpyAsync(w, r.words, sizeof(w), hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
    const uint32_t nc = (uint32_t)w[1];
    int64_t *droots = c->ar.get<int64_t>("cc_roots", (size_t)nc + 1);
    int64_t *dsizes = c->ar.get<int64_t>("cc_sizes", (size_t)nc + 1);
    uint32_t *dn = c->ar.get<uint32_t>("cc_n", 1);
    sdf::launch_fill_u32(c->stream, dn, 1, 0u);
    sdf::ccl_roots(c->stream, r, droots, dsizes, dn, nc);
    std::vector<int64_t> hr(nc), hs(nc);
    if (nc > 0) {
      SDF_HIP(hipMemcpyAsync(hr.data(), droots, nc * sizeof(int64_t), hipMemcpyDeviceToHost,
                             c->stream));
      SDF_HIP(hipMemcpyAsync(hs.data(), dsizes, nc * sizeof(int64_t), hipMemcpyDeviceToHost,
                             c->stream));
    }
    if (labels_out && n > 0) {
      int64_t *dl = c->ar.get<int64_t>("cc_labels", n);
      sdf::ccl_labels(c->stream, r, dl);
      SDF_HIP(hipMemcpyAsync(labels_out, dl, (size_t)n * sizeof(int64_t),
                             hipMemcpyDeviceToHost, c->stream));
    }
    SDF_HIP(hipStreamSynchronize(c->stream));
    std::vector<size_t> ord(nc);
    for (size_t i = 0; i < nc; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return hr[a] < hr[b]; });
    for (int64_t i = 0; i < cap && i < (int64_t)nc; i++) {
      roots_out[i] = hr[ord[i]];
      sizes_out[i] = hs[ord[i]];
    }
    if (ncomp) *ncomp = nc;
  });
}

namespace {

// RBFs_smoothing on device data: dd (ngp doubles) -> fine (prod(N*smooth+1)
// floats), both device buffers, on the context stream.  Returns after the
// last kernel is queued; the caller synchronises.
void rbf_dev(sdf_ctx *c, const double *dd, const sdf::GridP &g, int32_t is_interp,
             int32_t smooth, double threshold, double reltol, float *fine, int32_t *iters_out,
             double *rel_out) {
  const int64_t n = grid_points(g);
  if (reltol <= 0.0) reltol = 3.4526698300124393e-4;  // sqrt(eps(Float32))
  float *b = c->ar.get<float>("rbf_b", n + 1);
  float *w = c->ar.get<float>("rbf_w", n + 1);
  c->mark("rbf_prepare");
  sdf::rbf_prepare(c->stream, c->ar, dd, n, b);
  int iters = 0;
  double rel = 0.0;
  c->mark("rbf_weights");
  if (is_interp) {
    iters = sdf::rbf_weights(c->stream, c->ar, g, b, threshold, reltol, 50, 2000, w, &rel);
  } else {
    SDF_HIP(hipMemcpyAsync(w, b, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice,
                           c->stream));
  }
  c->mark("rbf_fine");
  sdf::rbf_fine(c->stream, c->ar, g, smooth, threshold, w, fine);
  c->mark("end");
  if (iters_out) *iters_out = iters;
  if (rel_out) *rel_out = rel;
}

// Argument checks shared by both RBF entry points; returns the fine count.
int64_t rbf_check(const sdf::GridP &g, int32_t smooth, double threshold) {
  if (smooth < 1 || smooth > 8) fail(SDF_ERR_ARG, "smooth must be in 1..8");
  if (!(threshold > 0.0 && threshold < 1.0)) fail(SDF_ERR_ARG, "threshold must be in (0, 1)");
  return ((int64_t)(g.nx - 1) * smooth + 1) * ((int64_t)(g.ny - 1) * smooth + 1) *
         ((int64_t)(g.nz - 1) * smooth + 1);
}

}  // namespace

sdf_status sdf_rbf_smoothing(sdf_ctx *c, const double *dist, const int64_t N[3],
                             double h, int32_t is_interp, int32_t smooth,
                             double threshold, double reltol, float *fine_out,
                             int32_t *iters_out, double *rel_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    (void)h;  // the kernel is expressed in cells: sigma = h
    const int64_t nf = rbf_check(g, smooth, threshold);
    const int64_t n = grid_points(g);
    if (n > 0 && (!dist || !fine_out)) fail(SDF_ERR_ARG, "NULL array");
    begin_call(c);
    double *dd = c->ar.get<double>("rbf_in", n + 1);
    float *fine = c->ar.get<float>("rbf_fine", nf + 1);
    if (n > 0)
      SDF_HIP(hipMemcpyAsync(dd, dist, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    rbf_dev(c, dd, g, is_interp, smooth, threshold, reltol, fine, iters_out, rel_out);
    SDF_HIP(hipMemcpyAsync(fine_out, fine, (size_t)nf * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_rbf_smoothing_device(sdf_ctx *c, const double *d_dist, const int64_t N[3],
                                    double h, int32_t is_interp, int32_t smooth,
                                    double threshold, double reltol, float *d_fine_out,
                                    int32_t *iters_out, double *rel_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = lattice_only(N);
    (void)h;
    rbf_check(g, smooth, threshold);
    if (grid_points(g) > 0 && (!d_dist || !d_fine_out)) fail(SDF_ERR_ARG, "NULL array");
    begin_call(c);
    rbf_dev(c, d_dist, g, is_interp, smooth, threshold, reltol, d_fine_out, iters_out, rel_out);
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_import_stl(const char *path, double *X, int64_t *nnp, int64_t *IEN,
                          int64_t *nel) {
  return guarded([&] {
    if (!path || !nnp || !nel) fail(SDF_ERR_ARG, "NULL argument");
    if ((X == nullptr) != (IEN == nullptr)) fail(SDF_ERR_ARG, "X and IEN must both be given");
    std::vector<double> hx;
    std::vector<int64_t> hi;
    std::string err;
    if (!sdf::import_stl_file(path, hx, hi, err)) fail(SDF_ERR_ARG, err);
    const int64_t nv = (int64_t)(hx.size() / 3), nt = (int64_t)(hi.size() / 3);
    if (X) {
      if (*nnp < nv || *nel < nt)
        fail(SDF_ERR_ARG, "X / IEN smaller than the mesh (" + std::to_string(nv) +
                              " vertices, " + std::to_string(nt) + " triangles)");
      memcpy(X, hx.data(), hx.size() * sizeof(double));
      memcpy(IEN, hi.data(), hi.size() * sizeof(int64_t));
    }
    *nnp = nv;
    *nel = nt;
  });
}

sdf_status sdf_ray_directions(double *dirs_out) {
  return guarded([&] {
    if (!dirs_out) fail(SDF_ERR_ARG, "NULL");
    ray_directions(dirs_out);
  });
}

namespace {

// evalDistancesOnTriMesh: dist always, xp (3 doubles per point) when asked.
sdf_status eval_distances(sdf_ctx *c, const double *X, int64_t nnp, const int64_t *IEN,
                          int64_t nel, const double aabb_min[3], double h, const int64_t N[3],
                          const double *points, int64_t npts, double *dist_out,
                          double *xp_out, bool want_xp) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    if (!points) npts = grid_points(g);
    if (npts < 0) fail(SDF_ERR_ARG, "npts < 0");
    if (npts > 0 && !dist_out) fail(SDF_ERR_ARG, "dist_out is NULL");
    if (want_xp && npts > 0 && !xp_out) fail(SDF_ERR_ARG, "xp_out is NULL");
    begin_call(c, true);
    set_mesh_host(c, X, nnp, IEN, nel);
    double *d_dist = c->ar.get<double>("out_dist", npts + 1);
    const bool lattice = is_lattice(g, points, npts);
    double *d_pts = nullptr;
    if (lattice) {
      sdf::FusedOut out;
      out.dist = d_dist;
      sdf::RayParams rp = ray_params(6, 30, 0.6, 1);
      run_lattice(c, g, true, false, rp, out);
    } else {
      d_pts = c->ar.get<double>("in_pts", 3 * npts + 3);
      SDF_HIP(hipMemcpyAsync(d_pts, points, (size_t)(3 * npts) * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
      const sdf::Bvh &bvh = need_bvh(c);
      c->mark("dist_points");
      float slack = 1e-6f * (float)(fabs(aabb_min[0]) + fabs(aabb_min[1]) +
                                    fabs(aabb_min[2]) + (double)(N[0] + N[1] + N[2]) * h + 1e-30);
      sdf::launch_dist_points_bvh(c->stream, d_pts, npts, c->tri, bvh, h, slack,
                                  d_dist, c->d_words);
      c->mark("end");
      check_err_flags(c);
    }
    double *d_xp = nullptr;
    if (want_xp && npts > 0) {
      d_xp = c->ar.get<double>("out_xp", 3 * npts + 3);
      if (c->nel == 0) {
        // no triangle: the reference's zero-initialised xp
        SDF_HIP(hipMemsetAsync(d_xp, 0, (size_t)(3 * npts) * sizeof(double), c->stream));
      } else {
        const sdf::Bvh &bvh = need_bvh(c);
        c->mark("closest");
        sdf::launch_closest_points(c->stream, g, lattice ? nullptr : d_pts, npts, c->tri, bvh,
                                   d_xp, c->d_words);
        c->mark("end");
        check_err_flags(c);
      }
    }
    if (npts > 0)
      SDF_HIP(hipMemcpyAsync(dist_out, d_dist, (size_t)npts * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    if (d_xp)
      SDF_HIP(hipMemcpyAsync(xp_out, d_xp, (size_t)(3 * npts) * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // namespace

sdf_status sdf_eval_distances(sdf_ctx *c, const double *X, int64_t nnp,
                              const int64_t *IEN, int64_t nel,
                              const double aabb_min[3], double h,
                              const int64_t N[3], const double *points,
                              int64_t npts, double *dist_out) {
  return eval_distances(c, X, nnp, IEN, nel, aabb_min, h, N, points, npts, dist_out, nullptr,
                        false);
}

sdf_status sdf_eval_distances_xp(sdf_ctx *c, const double *X, int64_t nnp,
                                 const int64_t *IEN, int64_t nel,
                                 const double aabb_min[3], double h,
                                 const int64_t N[3], const double *points,
                                 int64_t npts, double *dist_out, double *xp_out) {
  return eval_distances(c, X, nnp, IEN, nel, aabb_min, h, N, points, npts, dist_out, xp_out,
                        true);
}

sdf_status sdf_raycast_signs(sdf_ctx *c, const double *X, int64_t nnp,
                             const int64_t *IEN, int64_t nel,
                             const double aabb_min[3], double h,
                             const int64_t N[3], const double *points,
                             int64_t npts, int32_t min_rays, int32_t max_rays,
                             double conf_thr, int32_t use_winding,
                             double *signs_out, double *conf_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    if (!points) npts = grid_points(g);
    if (npts < 0) fail(SDF_ERR_ARG, "npts < 0");
    if (npts > 0 && (!signs_out || !conf_out))
      fail(SDF_ERR_ARG, "signs_out / conf_out is NULL");
    sdf::RayParams rp = ray_params(min_rays, max_rays, conf_thr, use_winding);
    begin_call(c, true);
    set_mesh_host(c, X, nnp, IEN, nel);
    double *d_s = c->ar.get<double>("out_signs", npts + 1);
    double *d_c = c->ar.get<double>("out_conf", npts + 1);
    sdf::FusedOut out;
    out.signs = d_s;
    out.conf = d_c;
    bool lattice = is_lattice(g, points, npts);
    if (lattice && min_rays == 6 && conf_thr <= 1.0) {
      run_lattice(c, g, false, true, rp, out);
    } else {
      const double *d_pts = nullptr;
      if (!lattice) {
        double *p = c->ar.get<double>("in_pts", 3 * npts + 3);
        SDF_HIP(hipMemcpyAsync(p, points, (size_t)(3 * npts) * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
        d_pts = p;
      }
      if (npts > 0xFFFFFFFFll) fail(SDF_ERR_ARG, "too many points");
      sdf::Pending *pend = c->ar.get<sdf::Pending>("pend_all", npts + 1);
      need_bvh(c);
      for (int attempt = 0;; attempt++) {
        if (attempt > 0) reset_words(c);
        sdf::launch_init_pending_all(c->stream, pend, npts);
        uint32_t *nd = c->d_words + 3;
        sdf::launch_fill_u32(c->stream, nd, 1, (uint32_t)npts);
        uint32_t *cnt = c->d_words + 4;
        c->mark("resolve");
        sdf::PendLists lists;
        lists.add(pend, nd, (uint32_t)npts, 0);
        resolve_items(c, g, d_pts, rp, lists, -1, out, cnt);
        c->mark("end");
        uint32_t hc[12];
        SDF_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
        SDF_HIP(hipStreamSynchronize(c->stream));
        const uint32_t w = hc[2];
        if (w <= c->caps.wind || attempt >= 2) break;
        c->caps.wind = w + w / 8 + 64;
      }
      check_err_flags(c);
    }
    if (npts > 0) {
      SDF_HIP(hipMemcpyAsync(signs_out, d_s, (size_t)npts * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
      SDF_HIP(hipMemcpyAsync(conf_out, d_c, (size_t)npts * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    }
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_signed_distance(sdf_ctx *c, const double *X, int64_t nnp,
                               const int64_t *IEN, int64_t nel,
                               const double aabb_min[3], double h,
                               const int64_t N[3], double *sdf_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    int64_t npts = grid_points(g);
    if (npts > 0 && !sdf_out) fail(SDF_ERR_ARG, "sdf_out is NULL");
    begin_call(c, true);
    set_mesh_host(c, X, nnp, IEN, nel);
    double *d_out = c->ar.get<double>("out_sdf", npts + 1);
    sdf::FusedOut out;
    out.sdf = d_out;
    run_lattice(c, g, true, true, ray_params(6, 30, 0.6, 1), out);
    if (npts > 0)
      SDF_HIP(hipMemcpyAsync(sdf_out, d_out, (size_t)npts * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_mesh_set_device(sdf_ctx *c, const double *d_X, int64_t nnp,
                               const int64_t *d_IEN, int64_t nel) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    begin_call(c, true);
    set_mesh_device(c, d_X, nnp, d_IEN, nel);
  });
}

sdf_status sdf_signed_distance_device(sdf_ctx *c, const double aabb_min[3],
                                      double h, const int64_t N[3],
                                      double *d_sdf_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    if (grid_points(g) > 0 && !d_sdf_out) fail(SDF_ERR_ARG, "d_sdf_out is NULL");
    begin_call(c);
    sdf::FusedOut out;
    out.sdf = d_sdf_out;
    run_lattice(c, g, true, true, ray_params(6, 30, 0.6, 1), out);
  });
}

sdf_status sdf_signed_fused(sdf_ctx *c, const double aabb_min[3], double h,
                            const int64_t N[3], int32_t remove_artifacts,
                            double min_component_ratio, double *d_sdf_out,
                            int64_t *nflipped) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    if (grid_points(g) > 0 && !d_sdf_out) fail(SDF_ERR_ARG, "d_sdf_out is NULL");
    if (remove_artifacts && !(min_component_ratio > 0.0 && min_component_ratio < 1.0))
      fail(SDF_ERR_ARG, "artifact_ratio must be in range (0, 1)");
    begin_call(c);
    sdf::FusedOut out;
    out.sdf = d_sdf_out;
    run_lattice(c, g, true, true, ray_params(6, 30, 0.6, 1), out);
    int64_t f = 0;
    if (remove_artifacts) f = remove_artifacts_dev(c, d_sdf_out, g, 0.0, min_component_ratio);
    if (nflipped) *nflipped = f;
  });
}


// ---------------------------------------------------------------- learned radius
sdf_status sdf_mlp_set_params(sdf_ctx *c, const float *params) {
  return guarded([&] {
    if (!c || !params) fail(SDF_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    mlp_upload(c, params);
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_mlp_get_params(sdf_ctx *c, float *params_out) {
  return guarded([&] {
    if (!c || !params_out) fail(SDF_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    use_device(c);
    SDF_HIP(hipMemcpyAsync(params_out, c->mlp.params, sizeof(float) * sdf::kMlpParams,
                           hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_mlp_forward(sdf_ctx *c, const float *u, int64_t n, float *y_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    if (n < 0) fail(SDF_ERR_ARG, "n < 0");
    if (n > 0 && (!u || !y_out)) fail(SDF_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    use_device(c);
    if (n == 0) return;
    float *du = c->ar.get<float>("mlp_fwd_u", 3 * n);
    float *dy = c->ar.get<float>("mlp_fwd_y", n);
    SDF_HIP(hipMemcpyAsync(du, u, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
    sdf::mlp_forward(c->stream, c->mlp, du, 3, n, dy);
    SDF_HIP(hipMemcpyAsync(y_out, dy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_mlp_loss_grad(sdf_ctx *c, const float *u, const float *t, int64_t n,
                             float *grad_out, double *loss_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    if (n < 1 || !u || !t) fail(SDF_ERR_ARG, "need n >= 1 samples");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    use_device(c);
    float4 *ds = upload_samples(c, u, t, n);
    float *dg = c->ar.get<float>("mlp_grad", sdf::kMlpParams);
    float *dl = c->ar.get<float>("mlp_loss1", 1);
    sdf::mlp_step(c->stream, c->ar, c->mlp, ds, nullptr, n, dg, dl, 0.f);
    float l = 0.f;
    if (grad_out)
      SDF_HIP(hipMemcpyAsync(grad_out, dg, sizeof(float) * sdf::kMlpParams,
                             hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = l;
  });
}

sdf_status sdf_mlp_train(sdf_ctx *c, const float *u, const float *t, int64_t n, int32_t steps,
                         int32_t batch, float lr, uint64_t seed, float *loss_hist) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    if (n < 1 || !u || !t) fail(SDF_ERR_ARG, "need n >= 1 samples");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    use_device(c);
    float4 *ds = upload_samples(c, u, t, n);
    mlp_train(c, ds, n, steps, batch, lr, seed, loss_hist);
  });
}

sdf_status sdf_learned_fit(sdf_ctx *c, const double aabb_min[3], double h, const int64_t N[3],
                           int32_t n_coarse, int32_t steps, int32_t batch, float lr,
                           uint64_t seed, double quantile, double *stats_out,
                           int32_t stats_cap) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    if (n_coarse < 4) fail(SDF_ERR_ARG, "n_coarse must be >= 4");
    if (quantile <= 0.0) quantile = 0.99;
    if (quantile > 1.0) fail(SDF_ERR_ARG, "quantile must be in (0, 1]");
    std::lock_guard<std::mutex> lk(c->mu);
    sdf::GridP g = make_grid(aabb_min, h, N);
    begin_call(c);
    if (!c->mesh_ready) fail(SDF_ERR_ARG, "no mesh (set one first)");
    // coarse lattice over the same box, n_coarse points on the longest axis
    const int64_t maxN = std::max(N[0], std::max(N[1], N[2]));
    if (maxN < 1) fail(SDF_ERR_ARG, "empty grid");
    const double hc = h * (double)maxN / (double)(n_coarse - 1);
    int64_t Nc[3];
    for (int a = 0; a < 3; a++) {
      Nc[a] = (int64_t)ceil((double)N[a] * h / hc - 1e-9);
      if (Nc[a] < 1) Nc[a] = 1;
    }
    sdf::GridP gc = make_grid(aabb_min, hc, Nc);
    const double t0 = now_ms();
    double *dc = c->ar.get<double>("fit_dc", grid_points(gc) + 1);
    sdf::FusedOut o;
    o.dist = dc;
    run_lattice(c, gc, true, false, ray_params(6, 30, 0.6, 1), o);
    SDF_HIP(hipStreamSynchronize(c->stream));
    const double t1 = now_ms();
    sdf::MlpNorm nm;
    nm.c[0] = g.ox + 0.5 * h * (double)N[0];
    nm.c[1] = g.oy + 0.5 * h * (double)N[1];
    nm.c[2] = g.oz + 0.5 * h * (double)N[2];
    nm.s = 2.0 / (h * (double)maxN);
    float4 *smp = nullptr;
    int64_t nband = 0;
    const int64_t ns = sdf::mlp_fit_samples(c->stream, c->ar, gc, nm, dc, hc, &smp, &nband);
    if (ns < 1) fail(SDF_ERR_ARG, "no training samples");
    std::vector<float> p;
    mlp_init_params(seed, p);
    mlp_upload(c, p.data());
    const double loss = mlp_train(c, smp, ns, steps, batch, lr, seed, nullptr);
    // calibration: delta = `quantile` of (target - output) over the band,
    // the output clamped at 0 as the radius kernel clamps it
    float *dy = c->ar.get<float>("fit_y", ns);
    sdf::mlp_forward(c->stream, c->mlp, (const float *)smp, 4, ns, dy);
    std::vector<float> y((size_t)ns), hs((size_t)(4 * ns));
    SDF_HIP(hipMemcpyAsync(y.data(), dy, sizeof(float) * ns, hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipMemcpyAsync(hs.data(), smp, sizeof(float) * 4 * ns, hipMemcpyDeviceToHost,
                           c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
    const double t2 = now_ms();
    std::vector<double> res;
    double mae = 0.0;
    for (int64_t i = 0; i < ns; i++)
      if (hs[4 * i + 3] < 1.f) {
        const double e = (double)hs[4 * i + 3] - (double)(y[i] > 0.f ? y[i] : 0.f);
        res.push_back(e);
        mae += fabs(e);
      }
    double delta = 0.0;
    if (!res.empty()) {
      mae /= (double)res.size();
      size_t k = (size_t)(quantile * (double)(res.size() - 1));
      std::nth_element(res.begin(), res.begin() + k, res.end());
      delta = res[k];
    }
    c->mlp_norm = nm;
    c->mlp_hc = hc;
    c->mlp_delta = delta;
    const double st[9] = {loss, mae, delta, (double)ns, (double)nband, t2 - t1, t1 - t0, hc,
                          (double)steps};
    for (int i = 0; i < 9 && i < stats_cap && stats_out; i++) stats_out[i] = st[i];
  });
}

sdf_status sdf_learned_signed_distance_device(sdf_ctx *c, const double aabb_min[3], double h,
                                              const int64_t N[3], double *d_sdf_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    sdf::GridP g = make_grid(aabb_min, h, N);
    if (grid_points(g) > 0 && !d_sdf_out) fail(SDF_ERR_ARG, "d_sdf_out is NULL");
    begin_call(c);
    sdf::FusedOut out;
    out.sdf = d_sdf_out;
    run_lattice(c, g, true, true, ray_params(6, 30, 0.6, 1), out, true);
  });
}

sdf_status sdf_learned_signed_distance(sdf_ctx *c, const double *X, int64_t nnp,
                                       const int64_t *IEN, int64_t nel,
                                       const double aabb_min[3], double h, const int64_t N[3],
                                       double *sdf_out) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->mlp_ready) fail(SDF_ERR_ARG, "no network parameters (fit or set them first)");
    sdf::GridP g = make_grid(aabb_min, h, N);
    int64_t npts = grid_points(g);
    if (npts > 0 && !sdf_out) fail(SDF_ERR_ARG, "sdf_out is NULL");
    begin_call(c, true);
    set_mesh_host(c, X, nnp, IEN, nel);
    double *d_out = c->ar.get<double>("out_sdf", npts + 1);
    sdf::FusedOut out;
    out.sdf = d_out;
    run_lattice(c, g, true, true, ray_params(6, 30, 0.6, 1), out, true);
    if (npts > 0)
      SDF_HIP(hipMemcpyAsync(sdf_out, d_out, (size_t)npts * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    SDF_HIP(hipStreamSynchronize(c->stream));
  });
}

sdf_status sdf_learned_stats(sdf_ctx *c, double *out, int32_t cap, int32_t *n) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    double st[9] = {0, 0, (double)c->lstat_pairs, 0, 0, c->mlp_delta, c->mlp_hc, 0,
                    c->learned_skipped ? 1.0 : 0.0};
    if (c->lstats) {
      std::vector<unsigned long long> h(kLearnedStats);
      SDF_HIP(hipMemcpyAsync(h.data(), c->lstats, sizeof(unsigned long long) * kLearnedStats,
                             hipMemcpyDeviceToHost, c->stream));
      SDF_HIP(hipStreamSynchronize(c->stream));
      unsigned long long v = 0, lr = 0, lh = 0, vp = 0;
      for (int i = 0; i < 256; i++) {
        v += h[8 + i];
        lr += h[264 + i];
        lh += h[520 + i];
        vp += h[776 + i];
      }
      st[0] = (double)h[0];
      st[1] = (double)v;
      st[3] = (double)lr;
      st[4] = (double)lh;
      st[7] = (double)vp;
    }
    int w = 0;
    for (; w < 9 && w < cap && out; w++) out[w] = st[w];
    if (n) *n = w;
  });
}

sdf_status sdf_mesh_set(sdf_ctx *c, const double *X, int64_t nnp, const int64_t *IEN,
                        int64_t nel) {
  return guarded([&] {
    if (!c) fail(SDF_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    begin_call(c, true);
    set_mesh_host(c, X, nnp, IEN, nel);
    SDF_HIP(hipStreamSynchronize(c->stream));
    check_err_flags(c);
  });
}

// ---------------------------------------------------------------- batches
sdf_status sdf_batch_create(sdf_ctx *c, int32_t nmesh, const double *const *d_X,
                            const int64_t *nnp, const int64_t *const *d_IEN, const int64_t *nel,
                            const double *aabb_min, const double *h, const int64_t *N,
                            double *const *d_sdf_out, int32_t autotune, sdf_batch **out) {
  return guarded([&] {
    if (!c || !out) fail(SDF_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (nmesh < 1 || nmesh > kMaxBatchMeshes)
      fail(SDF_ERR_ARG, "nmesh must be in 1.." + std::to_string(kMaxBatchMeshes));
    if (!d_X || !nnp || !d_IEN || !nel || !aabb_min || !h || !N || !d_sdf_out)
      fail(SDF_ERR_ARG, "NULL array");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    sdf_batch *b = new sdf_batch();
    b->parent = c;
    try {
      for (int i = 0; i < nmesh; i++) {
        if (nel[i] < 0 || nnp[i] < 0) fail(SDF_ERR_ARG, "mesh " + std::to_string(i) + ": negative mesh size");
        if (nel[i] > 0 && (!d_X[i] || !d_IEN[i]))
          fail(SDF_ERR_ARG, "mesh " + std::to_string(i) + ": NULL mesh arrays");
        BatchMesh m;
        m.dX = d_X[i];
        m.dIEN = d_IEN[i];
        m.nnp = nnp[i];
        m.nel = nel[i];
        m.g = make_grid(aabb_min + 3 * i, h[i], N + 3 * i);
        m.out = d_sdf_out[i];
        if (grid_points(m.g) > 0 && !m.out)
          fail(SDF_ERR_ARG, "mesh " + std::to_string(i) + ": NULL output");
        b->mesh.push_back(m);
        b->total_points += grid_points(m.g);
      }
      b->runs.resize((size_t)nmesh);
      SDF_HIP(hipMalloc(&b->d_status, (size_t)16 * nmesh * sizeof(uint64_t)));
      for (int l = 0; l < kMaxLanes; l++) {
        sdf_ctx *k = new sdf_ctx();
        b->lanes.push_back(k);
        ctx_init(k, c->device);
        k->d_status = nullptr;  // statuses go to the batch's block
        k->own_status = false;
      }
      if (autotune) batch_autotune(b);
      batch_plan(b);
      b->last_ms = batch_run(b);
    } catch (...) {
      batch_free(b);
      throw;
    }
    *out = b;
  });
}

sdf_status sdf_batch_run(sdf_batch *b) {
  return guarded([&] {
    if (!b) fail(SDF_ERR_ARG, "batch is NULL");
    std::lock_guard<std::mutex> lk(b->parent->mu);
    use_device(b->parent);
    b->last_ms = batch_run(b);
  });
}

sdf_status sdf_batch_info(sdf_batch *b, double *out, int32_t cap, int32_t *n) {
  return guarded([&] {
    if (!b) fail(SDF_ERR_ARG, "batch is NULL");
    std::vector<double> v = {(double)b->mesh.size(),  (double)b->total_points,
                             b->last_ms,              (double)b->tune.chunk_len,
                             (double)b->tune.ray_stack, (double)b->tune.lanes,
                             b->last_issue_ms, (double)b->tune.threads,
                             (double)(b->tune_log.size() / 2)};
    v.insert(v.end(), b->tune_log.begin(), b->tune_log.end());
    int w = 0;
    for (; w < (int)v.size() && w < cap && out; w++) out[w] = v[w];
    if (n) *n = (int32_t)v.size();
  });
}

sdf_status sdf_batch_destroy(sdf_batch *b) {
  return guarded([&] {
    if (!b) return;
    std::lock_guard<std::mutex> lk(b->parent->mu);
    batch_free(b);
  });
}

}  // extern "C"
