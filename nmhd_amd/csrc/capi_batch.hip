// capi_batch.hip -- the nmhd_batch_* entry points of include/nmhd.h: many
// independent boxes (periodic or shearing) in shared pools, per-box step
// records, per-box ghost fills, a work
// table of tiles for the stage kernels (grouped by tile shape: one launch per
// shape with a workgroup per tile, or one persistent launch over the whole
// table), HIP-graph replay of the step, and the autotuner of tile shapes and
// launch parameters.
//
// Pools: state[k] (k = 0..2, the rotating W^n / W* / W^{n+1} copies) holds
// the 8 fields of every box back to back (box b at BatchBox::off8, field f at
// off8 + f*ncell_b); flux holds the 21 face-flux arrays of every box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "capi_internal.h"

using namespace nmhd;

namespace {

// per thread: the message of the calling thread's last failed create
thread_local std::string g_batch_create_error;

constexpr int kZChunks[5] = {2, 4, 8, 16, 32};
constexpr int kWgPerCu[5] = {0, 1, 2, 3, 4};  // 0: one workgroup per item
constexpr int kTuneItems = 48;  // work items timed per shape and tile shape

}  // namespace

struct nmhd_batch {
  int nbox = 0;
  int device = 0;
  int cus = 256;
  std::string arch;
  hipStream_t stream = nullptr;
  std::vector<nmhd_desc> desc;
  std::vector<nmhd_geometry> geo;
  std::vector<BatchBox> boxes;
  BatchBox *d_boxes = nullptr;
  long total8 = 0, total21 = 0;
  float *state[3] = {};
  float *flux = nullptr;
  StepState *st = nullptr;
  StepState *st_tune = nullptr;
  float *d_goal = nullptr;
  int cur = 0;
  int integrator = NMHD_MIDPOINT2;
  std::vector<int> ftile, utile;
  // workgroups per CU of the persistent launches; 0: one workgroup per work
  // item, one launch per tile shape
  int fzc = 16, uzc = 16, fwg = 0, uwg = 0;
  WorkItem *d_fitems = nullptr, *d_uitems = nullptr, *d_titems = nullptr;
  int nf = 0, nu = 0;
  int nfv[TV_COUNT] = {}, nuv[TV_COUNT] = {};  // items per tile shape (the lists are grouped)
  // side streams and events that let the launches of different tile shapes
  // run concurrently (forked from and joined back into stream)
  hipStream_t side[TV_COUNT - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[TV_COUNT - 1] = {};
  size_t cap_f = 0, cap_u = 0, cap_t = 0;
  bool items_dirty = true;
  bool cfl_valid = false;
  // blocks per box of the ghost fills and their CFL pass (0: no box has
  // that ghost layer)
  int grid_gy = 0, grid_gz = 0, grid_gcfl = 0;
  hipGraphExec_t exec[3] = {};
  std::string err;
};

namespace {

int bfail(nmhd_batch *b, int code, const std::string &msg) {
  if (b)
    b->err = msg;
  else
    g_batch_create_error = msg;
  return code;
}

#define BHIP(b, call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bfail((b), NMHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

// Work items of the given boxes: tiles of tile shape tile[box], z chunks of
// zc planes, grouped by tile shape, then box by box, x fastest, then y, then
// z; counts[v] (optional) receives the number of items of shape v.
std::vector<WorkItem> make_items(const nmhd_batch *b, const std::vector<int> &tile, int zc,
                                 const std::vector<int> &which, int *counts = nullptr) {
  std::vector<WorkItem> out;
  for (int v = 0; v < TV_COUNT; ++v) {
    const size_t start = out.size();
    const int TX = kTileTX[v], TY = kTileTY[v];
    for (int box : which) {
      if (tile[box] != v) continue;
      const Box &bx = b->boxes[box].b;
      for (int z0 = 0; z0 < bx.n[2]; z0 += zc)
        for (int y0 = 0; y0 < bx.n[1]; y0 += TY)
          for (int x0 = 0; x0 < bx.n[0]; x0 += TX) {
            WorkItem w;
            w.box = box;
            w.xy = x0 | (y0 << 16);
            w.z = z0 | (std::min(z0 + zc, bx.n[2]) << 16);
            w.variant = v;
            out.push_back(w);
          }
    }
    if (counts) counts[v] = (int)(out.size() - start);
  }
  return out;
}

std::vector<int> all_boxes(const nmhd_batch *b) {
  std::vector<int> v(b->nbox);
  for (int i = 0; i < b->nbox; ++i) v[i] = i;
  return v;
}

int upload_items(nmhd_batch *b, const std::vector<WorkItem> &items, WorkItem **dst,
                 size_t *cap) {
  if (items.size() > *cap) {
    if (*dst) BHIP(b, hipFree(*dst));
    *dst = nullptr;
    *cap = 0;
    BHIP(b, hipMalloc(dst, sizeof(WorkItem) * items.size()));
    *cap = items.size();
  }
  if (!items.empty())
    BHIP(b, hipMemcpyAsync(*dst, items.data(), sizeof(WorkItem) * items.size(),
                           hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return NMHD_OK;
}

void drop_graphs(nmhd_batch *b) {
  for (auto &e : b->exec)
    if (e) (void)hipGraphExecDestroy(e), e = nullptr;
}

int refresh_items(nmhd_batch *b) {
  if (!b->items_dirty) return NMHD_OK;
  const std::vector<int> all = all_boxes(b);
  const std::vector<WorkItem> f = make_items(b, b->ftile, b->fzc, all, b->nfv);
  const std::vector<WorkItem> u = make_items(b, b->utile, b->uzc, all, b->nuv);
  int r = upload_items(b, f, &b->d_fitems, &b->cap_f);
  if (r) return r;
  r = upload_items(b, u, &b->d_uitems, &b->cap_u);
  if (r) return r;
  b->nf = (int)f.size();
  b->nu = (int)u.size();
  drop_graphs(b);
  b->items_dirty = false;
  return NMHD_OK;
}

int grid_of(int nitems, int wg_per_cu, int cus) {
  return std::max(1, std::min(nitems, wg_per_cu * cus));
}

// One launch per tile shape that has items, launch(stream, shape, offset):
// the first on s, the others on side streams forked from s and joined back
// into it.  The shapes' items belong to different boxes, so the launches are
// independent, and running them side by side fills the GPU where a small
// launch alone would not (and hides each launch's tail).  Works eagerly and
// under graph capture.
template <class F>
int over_shapes(nmhd_batch *b, hipStream_t s, const int *counts, F &&launch) {
  int used[TV_COUNT], off[TV_COUNT], nused = 0;
  for (int v = 0, o = 0; v < TV_COUNT; o += counts[v], ++v)
    if (counts[v] > 0) used[nused] = v, off[nused++] = o;
  if (nused == 0) return NMHD_OK;
  if (nused > 1) BHIP(b, hipEventRecord(b->ev_fork, s));
  launch(s, used[0], off[0]);
  for (int k = 1; k < nused; ++k) {
    hipStream_t t = b->side[k - 1];
    BHIP(b, hipStreamWaitEvent(t, b->ev_fork, 0));
    launch(t, used[k], off[k]);
    BHIP(b, hipEventRecord(b->ev_join[k - 1], t));
    BHIP(b, hipStreamWaitEvent(s, b->ev_join[k - 1], 0));
  }
  return NMHD_OK;
}

// The face / update launches of one stage over a grouped item list: the
// persistent kernel with wg workgroups per CU, or (wg == 0) one launch per
// tile shape with one workgroup per item.
int faces_over(nmhd_batch *b, hipStream_t s, int wg, const WorkItem *items, int n,
               const int *counts, float *state, const StepState *st, int hsel) {
  if (wg > 0) {
    launch_batch_faces(s, grid_of(n, wg, b->cus), b->d_boxes, items, n, state, b->flux, st, hsel);
    return NMHD_OK;
  }
  return over_shapes(b, s, counts, [&](hipStream_t t, int v, int off) {
    launch_batch_faces_items(t, v, b->d_boxes, items + off, counts[v], state, b->flux, st, hsel);
  });
}

int update_over(nmhd_batch *b, hipStream_t s, int wg, const WorkItem *items, int n,
                const int *counts, float *in, float *base, float *out, StepState *st, int hsel,
                bool cfl) {
  if (wg > 0) {
    launch_batch_update(s, grid_of(n, wg, b->cus), b->d_boxes, items, n, in, b->flux, base, out,
                        st, hsel, cfl);
    return NMHD_OK;
  }
  return over_shapes(b, s, counts, [&](hipStream_t t, int v, int off) {
    launch_batch_update_items(t, v, b->d_boxes, items + off, counts[v], in, b->flux, base, out,
                              st, hsel, cfl);
  });
}

// One iteration of the proceed() loop for every box.
int enqueue_step(nmhd_batch *b, bool full_cfl, float goal) {
  hipStream_t s = b->stream;
  const int A = b->cur, n = b->nbox;
  launch_batch_begin(s, b->st, n, full_cfl);
  // ghost fills of the boxes that have ghost layers, then the CFL minimum:
  // a full pass, or the ghost cells added to the stage-2 epilogue's minimum
  if (b->grid_gy) launch_batch_ghost_y(s, b->d_boxes, n, b->grid_gy, b->state[A], b->st);
  if (b->grid_gz) launch_batch_ghost_z(s, b->d_boxes, n, b->grid_gz, b->state[A], b->st);
  if (full_cfl)
    launch_batch_cfl(s, b->d_boxes, n, b->state[A], goal, b->st);
  else if (b->grid_gcfl)
    launch_batch_cfl_ghosts(s, b->d_boxes, n, b->grid_gcfl, b->state[A], b->st);
  launch_batch_dt(s, b->st, n, b->d_goal);
  auto stage = [&](int in, int base, int out, int hsel, bool last) {
    const int r = faces_over(b, s, b->fwg, b->d_fitems, b->nf, b->nfv, b->state[in], b->st, hsel);
    if (r) return r;
    return update_over(b, s, b->uwg, b->d_uitems, b->nu, b->nuv, b->state[in], b->state[base],
                       b->state[out], b->st, hsel, last);
  };
  int next, r;
  if (b->integrator == NMHD_EULER1) {
    next = (A + 1) % 3;
    r = stage(A, A, next, H_FULL, true);
  } else {
    const int B = (A + 1) % 3;
    next = (A + 2) % 3;
    r = stage(A, A, B, H_HALF, false);
    if (!r) r = stage(B, A, next, H_FULL, true);
  }
  if (r) return r;
  launch_batch_copy(s, b->d_boxes, n, b->state[A], b->state[next], b->st);
  launch_batch_commit(s, b->st, n);
  b->cur = next;
  return NMHD_OK;
}

int ensure_graph(nmhd_batch *b) {
  const int phase = b->cur;
  if (b->exec[phase]) return NMHD_OK;
  hipGraph_t g = nullptr;
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int re = enqueue_step(b, false, 0.0f);
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  b->cur = phase;  // capture records; it does not advance the state
  if (re) {
    if (g) (void)hipGraphDestroy(g);
    return re;
  }
  if (e != hipSuccess)
    return bfail(b, NMHD_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&b->exec[phase], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    b->exec[phase] = nullptr;
    return bfail(b, NMHD_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  return NMHD_OK;
}

int read_records(nmhd_batch *b, std::vector<StepState> &h) {
  h.resize(b->nbox);
  BHIP(b, hipMemcpyAsync(h.data(), b->st, sizeof(StepState) * b->nbox, hipMemcpyDeviceToHost,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return NMHD_OK;
}

int write_records(nmhd_batch *b, const std::vector<StepState> &h) {
  BHIP(b, hipMemcpyAsync(b->st, h.data(), sizeof(StepState) * b->nbox, hipMemcpyHostToDevice,
                         b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return NMHD_OK;
}

void free_batch(nmhd_batch *b) {
  drop_graphs(b);
  for (auto &p : b->state)
    if (p) (void)hipFree(p), p = nullptr;
  void *ptrs[] = {b->flux, b->st, b->st_tune, b->d_goal, b->d_boxes,
                  b->d_fitems, b->d_uitems, b->d_titems};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  b->flux = nullptr;
  b->st = b->st_tune = nullptr;
  b->d_goal = nullptr;
  b->d_boxes = nullptr;
  b->d_fitems = b->d_uitems = b->d_titems = nullptr;
  for (auto &t : b->side)
    if (t) (void)hipStreamDestroy(t), t = nullptr;
  for (auto &e : b->ev_join)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork), b->ev_fork = nullptr;
  if (b->stream) (void)hipStreamDestroy(b->stream), b->stream = nullptr;
}

int default_tile(int nx) { return nx >= 64 ? TV_64x4 : (nx >= 32 ? TV_32x8 : TV_16x16); }

// ------------------------------------------------------------ autotuner

std::string shape_key(const nmhd_batch *b, int box) {
  const Box &x = b->boxes[box].b;
  std::ostringstream o;
  o << x.n[0] << " " << x.n[1] << " " << x.n[2] << " " << x.recon << " " << x.source;
  return o.str();
}

std::string cache_file(const char *path) {
  if (path && *path) return path;
  const char *env = std::getenv("NMHD_TUNE_CACHE");
  if (env && *env) return env;
  const char *home = std::getenv("HOME");
  return std::string(home && *home ? home : ".") + "/.cache/nmhd_amd/batch_tuning.txt";
}

void make_parent_dirs(const std::string &file) {
  for (size_t i = 1; i < file.size(); ++i)
    if (file[i] == '/') (void)mkdir(file.substr(0, i).c_str(), 0755);
}

// Cache lines: "shape <arch> <nx> <ny> <nz> <recon> <source> : <faces> <update>"
// and "launch <arch> <hash> : <fzc> <uzc> <fwg> <uwg>".
struct TuneCache {
  std::map<std::string, std::vector<int>> entries;
  void load(const std::string &file) {
    std::ifstream in(file);
    std::string line;
    while (std::getline(in, line)) {
      const size_t colon = line.find(" : ");
      if (line.empty() || line[0] == '#' || colon == std::string::npos) continue;
      std::istringstream vals(line.substr(colon + 3));
      std::vector<int> v;
      int x;
      while (vals >> x) v.push_back(x);
      entries[line.substr(0, colon)] = v;
    }
  }
  bool save(const std::string &file) const {
    make_parent_dirs(file);
    std::ofstream out(file, std::ios::trunc);
    if (!out) return false;
    out << "# nmhd batch autotuner cache (tile shapes per box shape, launch parameters)\n";
    for (auto &kv : entries) {
      out << kv.first << " :";
      for (int v : kv.second) out << " " << v;
      out << "\n";
    }
    return (bool)out;
  }
};

std::string launch_key(const nmhd_batch *b) {
  std::vector<std::string> shapes;
  for (int i = 0; i < b->nbox; ++i) shapes.push_back(shape_key(b, i));
  std::sort(shapes.begin(), shapes.end());
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over the sorted shapes
  for (auto &s : shapes)
    for (char c : s + ";") h = (h ^ (unsigned char)c) * 1099511628211ull;
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%016llx", h);
  return "launch " + b->arch + " " + buf;
}

// Milliseconds of one launch of fn (after one untimed warm-up launch).
template <class F>
int time_launch(nmhd_batch *b, F &&fn, float *ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  fn();
  BHIP(b, hipEventRecord(e0, b->stream));
  fn();
  BHIP(b, hipEventRecord(e1, b->stream));
  BHIP(b, hipEventSynchronize(e1));
  BHIP(b, hipGetLastError());
  BHIP(b, hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return NMHD_OK;
}

int prepare_tune_records(nmhd_batch *b) {
  // every box active with a zero step: the update kernels then reproduce
  // the base state into the scratch rotation slot
  std::vector<StepState> h(b->nbox);
  std::memset(h.data(), 0, sizeof(StepState) * h.size());
  for (auto &s : h) {
    s.active = 1;
    s.cfl_bits = s.cfl_next_bits = 0x7f800000u;
  }
  BHIP(b, hipMemcpyAsync(b->st_tune, h.data(), sizeof(StepState) * h.size(),
                         hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return NMHD_OK;
}

int tune(nmhd_batch *b, const char *path, int *timed) {
  const std::string file = cache_file(path);
  TuneCache cache;
  cache.load(file);
  int r = prepare_tune_records(b);
  if (r) return r;
  float *in = b->state[b->cur];
  float *scratch = b->state[(b->cur + 1) % 3];
  int ntimed = 0;

  // per shape: the workgroup-time of each tile shape
  std::map<std::string, std::vector<int>> groups;
  for (int i = 0; i < b->nbox; ++i) groups[shape_key(b, i)].push_back(i);
  for (auto &g : groups) {
    const std::string key = "shape " + b->arch + " " + g.first;
    auto hit = cache.entries.find(key);
    if (hit == cache.entries.end() || hit->second.size() != 2) {
      float best_f = 1e30f, best_u = 1e30f;
      int vf = TV_16x16, vu = TV_16x16;
      for (int v = 0; v < TV_COUNT; ++v) {
        std::vector<int> tile(b->nbox, v);
        std::vector<WorkItem> items = make_items(b, tile, 16, g.second);
        const double total = (double)items.size();
        if ((int)items.size() > kTuneItems) items.resize(kTuneItems);
        r = upload_items(b, items, &b->d_titems, &b->cap_t);
        if (r) return r;
        const int n = (int)items.size();
        float msf = 0.0f, msu = 0.0f;
        r = time_launch(
            b,
            [&] {
              launch_batch_faces(b->stream, 1, b->d_boxes, b->d_titems, n, in, b->flux,
                                 b->st_tune, H_HALF);
            },
            &msf);
        if (r) return r;
        r = time_launch(
            b,
            [&] {
              launch_batch_update(b->stream, 1, b->d_boxes, b->d_titems, n, in, b->flux, in,
                                  scratch, b->st_tune, H_HALF, false);
            },
            &msu);
        if (r) return r;
        const float cf = (float)(msf * total / n), cu = (float)(msu * total / n);
        if (cf < best_f) best_f = cf, vf = v;
        if (cu < best_u) best_u = cu, vu = v;
      }
      cache.entries[key] = {vf, vu};
      ++ntimed;
    }
    for (int box : g.second) {
      b->ftile[box] = cache.entries[key][0];
      b->utile[box] = cache.entries[key][1];
    }
  }

  // whole batch: z chunk and workgroups per CU of each persistent launch
  const std::string lkey = launch_key(b);
  auto hit = cache.entries.find(lkey);
  if (hit == cache.entries.end() || hit->second.size() != 4) {
    const std::vector<int> all = all_boxes(b);
    float best_f = 1e30f, best_u = 1e30f;
    int fzc = 16, uzc = 16, fwg = 0, uwg = 0;
    for (int zc : kZChunks) {
      int cf[TV_COUNT], cu[TV_COUNT];
      const std::vector<WorkItem> fi = make_items(b, b->ftile, zc, all, cf);
      r = upload_items(b, fi, &b->d_titems, &b->cap_t);
      if (r) return r;
      for (int wg : kWgPerCu) {
        float ms = 0.0f;
        const int n = (int)fi.size();
        int rl = NMHD_OK;
        r = time_launch(
            b,
            [&] {
              if (!rl) rl = faces_over(b, b->stream, wg, b->d_titems, n, cf, in, b->st_tune, H_HALF);
            },
            &ms);
        if (!r) r = rl;
        if (r) return r;
        if (ms < best_f) best_f = ms, fzc = zc, fwg = wg;
      }
      const std::vector<WorkItem> ui = make_items(b, b->utile, zc, all, cu);
      r = upload_items(b, ui, &b->d_titems, &b->cap_t);
      if (r) return r;
      for (int wg : kWgPerCu) {
        float ms = 0.0f;
        const int n = (int)ui.size();
        int rl = NMHD_OK;
        r = time_launch(
            b,
            [&] {
              if (!rl)
                rl = update_over(b, b->stream, wg, b->d_titems, n, cu, in, in, scratch,
                                 b->st_tune, H_HALF, false);
            },
            &ms);
        if (!r) r = rl;
        if (r) return r;
        if (ms < best_u) best_u = ms, uzc = zc, uwg = wg;
      }
    }
    cache.entries[lkey] = {fzc, uzc, fwg, uwg};
  }
  const std::vector<int> &lp = cache.entries[lkey];
  b->fzc = lp[0];
  b->uzc = lp[1];
  b->fwg = lp[2];
  b->uwg = lp[3];
  b->items_dirty = true;
  (void)cache.save(file);  // a read-only cache location only costs a re-tune
  if (timed) *timed = ntimed;
  return NMHD_OK;
}

bool box_ok(const nmhd_batch *b, int box) { return b && box >= 0 && box < b->nbox; }

}  // namespace

extern "C" {

const char *nmhd_batch_last_error(const nmhd_batch *b) {
  return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int nmhd_batch_create(const nmhd_desc *descs, int nbox, nmhd_batch **out) {
  if (!descs || !out || nbox <= 0) return bfail(nullptr, NMHD_EINVAL, "bad argument");
  *out = nullptr;
  for (int i = 0; i < nbox; ++i) {
    std::string why;
    const nmhd_desc &d = descs[i];
    if (!desc_valid(&d, why)) return bfail(nullptr, NMHD_EINVAL, "box " + std::to_string(i) + ": " + why);
    if (d.integrator != descs[0].integrator)
      return bfail(nullptr, NMHD_EINVAL, "all boxes of a batch share one integrator");
    for (int a = 0; a < 3; ++a)
      if (d.global_size[a] + 2 * d.margin[a] > 32767)
        return bfail(nullptr, NMHD_EINVAL,
                     "batched boxes are at most 32767 cells per axis (ghosts included)");
  }
  nmhd_batch *b = new nmhd_batch();
  auto bail = [&](int code, const std::string &msg) {
    free_batch(b);
    delete b;
    return bfail(nullptr, code, msg);
  };
  b->nbox = nbox;
  b->integrator = descs[0].integrator;
  b->desc.assign(descs, descs + nbox);
  b->geo.resize(nbox);
  b->boxes.resize(nbox);
  b->ftile.resize(nbox);
  b->utile.resize(nbox);
  long off8 = 0, off21 = 0;
  for (int i = 0; i < nbox; ++i) {
    derive_box(&descs[i], b->boxes[i].b, b->geo[i]);
    b->boxes[i].off8 = off8;
    b->boxes[i].off21 = off21;
    off8 += 8 * b->boxes[i].b.ncell;
    off21 += 21 * b->boxes[i].b.ncell;
    b->ftile[i] = default_tile(b->boxes[i].b.n[0]);
    b->utile[i] = b->boxes[i].b.n[0] >= 32 ? TV_32x8 : TV_16x16;
  }
  b->total8 = off8;
  b->total21 = off21;
  // ghost grids: blocks of 256 for the largest slab, at most 512 per box
  {
    long gy = 0, gz = 0, gc = 0;
    for (const BatchBox &bb : b->boxes) {
      const Box &x = bb.b;
      gy = std::max(gy, (long)x.n[0] * x.m[1] * x.n[2]);
      gz = std::max(gz, (long)x.n[0] * x.n[1] * x.m[2]);
      long c = 0;
      if (x.m[1]) c += (long)x.n[0] * 2 * x.m[1] * x.n[2];
      if (x.m[2]) c += (long)x.n[0] * x.n[1] * (2 * x.m[2] + 1);
      gc = std::max(gc, c);
    }
    auto blocks = [](long cells, long cap) {
      return cells == 0 ? 0 : (int)std::min(cap, (cells + 255) / 256);
    };
    b->grid_gy = blocks(gy, 512);
    b->grid_gz = blocks(gz, 512);
    // each block of the CFL pass ends in one atomic on its box's record
    b->grid_gcfl = blocks(gc / 4, 128);
  }
  if (hipGetDevice(&b->device) != hipSuccess) return bail(NMHD_EHIP, "no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) != hipSuccess)
    return bail(NMHD_EHIP, "hipGetDeviceProperties failed");
  b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  b->arch = prop.gcnArchName;
  const size_t colon = b->arch.find(':');
  if (colon != std::string::npos) b->arch = b->arch.substr(0, colon);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(NMHD_EHIP, "hipStreamCreate failed");
  for (auto &t : b->side)
    if (hipStreamCreateWithFlags(&t, hipStreamNonBlocking) != hipSuccess)
      return bail(NMHD_EHIP, "hipStreamCreate failed");
  if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess)
    return bail(NMHD_EHIP, "hipEventCreate failed");
  for (auto &e : b->ev_join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bail(NMHD_EHIP, "hipEventCreate failed");
  for (auto &p : b->state)
    if (hipMalloc(&p, sizeof(float) * (size_t)off8) != hipSuccess)
      return bail(NMHD_ENOMEM, "state pool allocation failed");
  if (hipMalloc(&b->flux, sizeof(float) * (size_t)off21) != hipSuccess)
    return bail(NMHD_ENOMEM, "flux pool allocation failed");
  if (hipMalloc(&b->st, sizeof(StepState) * nbox) != hipSuccess ||
      hipMalloc(&b->st_tune, sizeof(StepState) * nbox) != hipSuccess ||
      hipMalloc(&b->d_goal, sizeof(float)) != hipSuccess ||
      hipMalloc(&b->d_boxes, sizeof(BatchBox) * nbox) != hipSuccess)
    return bail(NMHD_ENOMEM, "batch record allocation failed");
  for (auto &p : b->state)
    if (hipMemsetAsync(p, 0, sizeof(float) * (size_t)off8, b->stream) != hipSuccess)
      return bail(NMHD_EHIP, "state clear failed");
  std::vector<StepState> h(nbox);
  std::memset(h.data(), 0, sizeof(StepState) * nbox);
  for (auto &s : h) s.cfl_bits = s.cfl_next_bits = 0x7f800000u;
  if (hipMemcpyAsync(b->st, h.data(), sizeof(StepState) * nbox, hipMemcpyHostToDevice,
                     b->stream) != hipSuccess ||
      hipMemcpyAsync(b->d_boxes, b->boxes.data(), sizeof(BatchBox) * nbox,
                     hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return bail(NMHD_EHIP, "batch record init failed");
  *out = b;
  return NMHD_OK;
}

int nmhd_batch_destroy(nmhd_batch *b) {
  if (!b) return NMHD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  delete b;
  return NMHD_OK;
}

int nmhd_batch_get_geometry(const nmhd_batch *b, int box, nmhd_geometry *geo) {
  if (!box_ok(b, box) || !geo) return NMHD_EINVAL;
  *geo = b->geo[box];
  return NMHD_OK;
}

int nmhd_batch_set_state(nmhd_batch *b, int box, const float *const fields[8]) {
  if (!box_ok(b, box) || !fields) return bfail(b, NMHD_EINVAL, "bad argument");
  for (int f = 0; f < 8; ++f)
    if (!fields[f]) return bfail(b, NMHD_EINVAL, "null field pointer");
  BHIP(b, hipStreamSynchronize(b->stream));
  const FieldPtrs W = box_fields(b->state[b->cur], b->boxes[box]);
  const size_t bytes = sizeof(float) * (size_t)b->boxes[box].b.ncell;
  for (int f = 0; f < 8; ++f)
    BHIP(b, hipMemcpyAsync(W.f[f], fields[f], bytes, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->cfl_valid = false;
  return NMHD_OK;
}

int nmhd_batch_get_state(nmhd_batch *b, int box, float *const fields[8]) {
  if (!box_ok(b, box) || !fields) return bfail(b, NMHD_EINVAL, "bad argument");
  for (int f = 0; f < 8; ++f)
    if (!fields[f]) return bfail(b, NMHD_EINVAL, "null field pointer");
  BHIP(b, hipStreamSynchronize(b->stream));
  const FieldPtrs W = box_fields(b->state[b->cur], b->boxes[box]);
  const size_t bytes = sizeof(float) * (size_t)b->boxes[box].b.ncell;
  for (int f = 0; f < 8; ++f)
    BHIP(b, hipMemcpyAsync(fields[f], W.f[f], bytes, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return NMHD_OK;
}

int nmhd_batch_set_time(nmhd_batch *b, int box, double t, int generation) {
  if (!box_ok(b, box)) return bfail(b, NMHD_EINVAL, "bad box");
  std::vector<StepState> h;
  int r = read_records(b, h);
  if (r) return r;
  StepState &s = h[box];
  std::memset(&s, 0, sizeof(s));
  s.t = t;
  s.t_next = t;
  s.generation = generation;
  s.cfl_bits = s.cfl_next_bits = 0x7f800000u;
  b->cfl_valid = false;
  return write_records(b, h);
}

int nmhd_batch_get_time(nmhd_batch *b, int box, double *t, int *generation) {
  if (!box_ok(b, box)) return bfail(b, NMHD_EINVAL, "bad box");
  std::vector<StepState> h;
  int r = read_records(b, h);
  if (r) return r;
  if (t) *t = h[box].t;
  if (generation) *generation = h[box].generation;
  return NMHD_OK;
}

int nmhd_batch_get_status(nmhd_batch *b, int box, int *steps_taken, int *goal_reached,
                          float *last_dt) {
  if (!box_ok(b, box)) return bfail(b, NMHD_EINVAL, "bad box");
  std::vector<StepState> h;
  int r = read_records(b, h);
  if (r) return r;
  if (steps_taken) *steps_taken = h[box].steps_done;
  if (goal_reached) *goal_reached = h[box].goal_reached;
  if (last_dt) *last_dt = h[box].dt;
  return NMHD_OK;
}

int nmhd_batch_set_tiles(nmhd_batch *b, int box, int faces_tile, int update_tile) {
  if (!b || box < -1 || box >= b->nbox || faces_tile < 0 || faces_tile >= TV_COUNT ||
      update_tile < 0 || update_tile >= TV_COUNT)
    return bfail(b, NMHD_EINVAL, "bad tile selection");
  for (int i = 0; i < b->nbox; ++i)
    if (box == -1 || box == i) b->ftile[i] = faces_tile, b->utile[i] = update_tile;
  b->items_dirty = true;
  return NMHD_OK;
}

int nmhd_batch_set_launch(nmhd_batch *b, int fzc, int uzc, int fwg, int uwg) {
  if (!b || fzc < 1 || uzc < 1 || fzc > 32767 || uzc > 32767 || fwg < 0 || uwg < 0 || fwg > 64 ||
      uwg > 64)
    return bfail(b, NMHD_EINVAL, "bad launch parameters");
  b->fzc = fzc;
  b->uzc = uzc;
  b->fwg = fwg;
  b->uwg = uwg;
  b->items_dirty = true;
  return NMHD_OK;
}

int nmhd_batch_get_launch(const nmhd_batch *b, int *fzc, int *uzc, int *fwg, int *uwg) {
  if (!b) return NMHD_EINVAL;
  if (fzc) *fzc = b->fzc;
  if (uzc) *uzc = b->uzc;
  if (fwg) *fwg = b->fwg;
  if (uwg) *uwg = b->uwg;
  return NMHD_OK;
}

int nmhd_batch_autotune(nmhd_batch *b, const char *cache_path, int *timed_shapes) {
  if (!b) return NMHD_EINVAL;
  BHIP(b, hipStreamSynchronize(b->stream));
  return tune(b, cache_path, timed_shapes);
}

int nmhd_batch_run(nmhd_batch *b, float goal_time, int nsteps, int use_graph, int *min_steps,
                   int *all_reached) {
  if (!b || nsteps < 0) return bfail(b, NMHD_EINVAL, "bad argument");
  std::vector<StepState> h;
  int r = read_records(b, h);
  if (r) return r;
  for (auto &s : h) s.finished = s.active = s.goal_reached = s.steps_done = 0;
  r = write_records(b, h);
  if (r) return r;
  r = refresh_items(b);
  if (r) return r;
  BHIP(b, hipMemcpyAsync(b->d_goal, &goal_time, sizeof(float), hipMemcpyHostToDevice,
                         b->stream));
  for (int k = 0; k < nsteps; ++k) {
    if (!b->cfl_valid) {
      r = enqueue_step(b, true, goal_time);
      if (r) return r;
      b->cfl_valid = true;
    } else if (use_graph) {
      r = ensure_graph(b);
      if (r) return r;
      BHIP(b, hipGraphLaunch(b->exec[b->cur], b->stream));
      b->cur = b->integrator == NMHD_EULER1 ? (b->cur + 1) % 3 : (b->cur + 2) % 3;
    } else {
      r = enqueue_step(b, false, goal_time);
      if (r) return r;
    }
  }
  BHIP(b, hipGetLastError());
  r = read_records(b, h);
  if (r) return r;
  int mn = nsteps, all = 1;
  for (auto &s : h) {
    mn = std::min(mn, s.steps_done);
    all = all && s.goal_reached;
  }
  if (min_steps) *min_steps = mn;
  if (all_reached) *all_reached = all;
  return NMHD_OK;
}

}  // extern "C"
