// pool.hip - many independent shots (grids of any size, own media, sources,
// detectors and step counts) advanced together: the batched and served forms
// of the stepper. No reference counterpart: the reference runs one grid per
// process (F/FDTD_cuda.cu:384-441).
//
// Every pool launch advances each unfinished shot by min(T, steps left).
// The shots' tile entries are concatenated into two packed lists (interior,
// CPML class), one work-group of 4 entries per shot at a time, and two
// persistent kernels (one per class, on two streams) drain them through a
// ticket queue in device memory. Per-shot parameters live in a descriptor
// table; a shot's step index and buffer set follow from the pool's launch
// counter g, which a one-wave tick kernel advances on the device. Nothing a
// launch reads changes its address, so a chunk of launches can be captured
// once as a hipGraph and replayed while shots come and go.
//
// Shots whose steps left are fewer than T run in separate tail launches of
// their exact step count at the same g. Device memory of shots comes from one
// arena when the pool has a byte budget (admission control), else from one
// hipMalloc per shot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"
#include "errors.h"
#include "plan_impl.h"
#include "plan_internal.h"
#include "pool_policy.h"
#include "validate.h"

using fdtd2d::set_error;
using namespace fdtd2d::detail;

namespace {

constexpr size_t kArenaAlign = 256;

// First-fit allocator over one device allocation.
struct Arena {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  std::map<size_t, size_t> used_;  // offset -> length
  size_t in_use = 0;

  bool init(size_t bytes) {
    if (hipMalloc(reinterpret_cast<void **>(&base), bytes) != hipSuccess) return false;
    size = bytes;
    free_[0] = bytes;
    return true;
  }
  char *take(size_t bytes) {
    bytes = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      used_[off] = bytes;
      in_use += bytes;
      return base + off;
    }
    return nullptr;
  }
  void give(char *p) {
    auto it = used_.find(static_cast<size_t>(p - base));
    if (it == used_.end()) return;
    size_t off = it->first, len = it->second;
    used_.erase(it);
    in_use -= len;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + len == nx->first) {
      len += nx->second;
      free_.erase(nx);
    }
    auto pv = free_.lower_bound(off);
    if (pv != free_.begin()) {
      --pv;
      if (pv->first + pv->second == off) {
        off = pv->first;
        len += pv->second;
        free_.erase(pv);
      }
    }
    free_[off] = len;
  }
  size_t largest_free() const {
    size_t m = 0;
    for (auto &f : free_) m = std::max(m, f.second);
    return m;
  }
};

struct Shot {
  fdtd2d_plan *pl = nullptr;
  bool live = false;      // slot in use
  bool finished = false;  // all steps issued
  int last_g = -1;        // pool launch in which its last steps were issued
  size_t bytes = 0;
  int medium_kind = 0;
};

template <class T>
int grow(T *&buf, size_t &cap, size_t need, bool &moved) {
  if (need <= cap) return FDTD2D_OK;
  size_t nc = std::max<size_t>(need, cap * 2);
  nc = std::max<size_t>(nc, 1024);
  T *nb = nullptr;
  if (hipMalloc(reinterpret_cast<void **>(&nb), nc * sizeof(T)) != hipSuccess)
    return set_error(FDTD2D_ENOMEM, "hipMalloc pool lists");
  if (buf) hipFree(buf);
  buf = nb;
  cap = nc;
  moved = true;
  return FDTD2D_OK;
}

}  // namespace

struct fdtd2d_pool {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr, aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // launch shape
  // (defaults: the shape the autotuner picks for mixed shot sizes)
  int T = 8, band = 128, band_cpml = 64, D = 1, C = 1;
  int wg_per_cu = 3, wg_per_cu_cpml = 2;
  int fused = 0;  // 0: one kernel per tile class on two streams; 1: one kernel, CPML-class work first
  int chunk = 8;  // launches per captured graph; 0 = no graphs
  int medium_kind = -1;
  // memory
  size_t budget = 0;
  Arena arena;
  // shots
  std::vector<Shot> shots;
  std::vector<int> free_slots;
  // device state
  fdtd2d::StepArgs *d_descs = nullptr;
  std::vector<fdtd2d::StepArgs> h_descs;  // host copy of the descriptor table
  size_t cap_slots = 0;
  // main lists in a ring of `nbufs` buffers, each remembering the shot set
  // it holds: a launch whose set one of them holds switches to it (two sets
  // alternate under time slicing); otherwise the lists are rebuilt into the
  // least recently read buffer once the last launch that read it has
  // finished, so the launches of the other list versions keep running
  // meanwhile (lists: 0 interior, 1 CPML class or its fusable kinds, 2 the
  // leftover CPML-class entries of a fused pool)
  static constexpr int kListBufs = 4;
  struct ListBuf {
    int4 *d[fdtd2d::kPoolLists] = {nullptr, nullptr, nullptr};
    size_t cap[fdtd2d::kPoolLists] = {0, 0, 0};
    int last_g = -1;  // last pool launch that read it
    bool valid = false;
    std::vector<int> set;  // slots whose tiles it holds
    int count[fdtd2d::kPoolLists] = {0, 0, 0};
  } bufs[kListBufs];
  int nbufs = kListBufs;
  int cur = 0;
  int main_count[fdtd2d::kPoolLists] = {0, 0, 0};  // work-groups per list of the current main lists
  // tail lists: one copy per tail group, taken in turn; a launch has at most
  // T - 1 <= 7 groups, so a copy is reused only after launches whose
  // completion events exist, and is waited for through them
  static constexpr int kTailBufs = 8;
  struct TailBuf {
    int4 *d[fdtd2d::kPoolLists] = {nullptr, nullptr, nullptr};
    size_t cap[fdtd2d::kPoolLists] = {0, 0, 0};
    int last_g = -1;
  } tailbufs[kTailBufs];
  int tail_next = 0;
  fdtd2d::PoolQueue *d_q = nullptr, *d_qt = nullptr;
  // host bookkeeping
  int g = 0;                 // pool launches issued
  std::vector<int> main_set;  // slots in the current main lists
  bool lists_valid = false;
  // graph of `chunk` launches
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  // completion events: one recorded after every launch
  std::vector<std::pair<int, hipEvent_t>> done_events;  // (g after launch, event)
  std::vector<hipEvent_t> spare_events;
  int g_done_known = 0;  // launches known complete on the device
  // stats
  double cell_updates = 0;
  int graph_replays = 0, eager_launches = 0, tail_launches = 0;
  int rebuilds = 0, captures = 0, list_switches = 0;
  // time slicing: shots of at least slice_cells cells run only in launches g
  // with g % slice_every == 0 (0 cells: off)
  double slice_cells = 0;
  int slice_every = 1;
  // band and CPML sub-band rows of the shots that are not sliced while
  // slicing is on (0: band, band_cpml)
  int band_small = 0, band_cpml_small = 0;
  // host seconds: rebuilding and uploading main lists, capturing graphs,
  // blocked on launches still running (outside fdtd2d_pool_wait/_sync),
  // and adding shots
  double rebuild_s = 0, capture_s = 0, blocked_s = 0, add_s = 0;
};

namespace {

void retire(fdtd2d_pool *pp, int wait_g);

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Appends shot s's tile entries to the pool lists, tagged with its slot, each
// list padded to whole work-groups (one shot per work-group). A fused pool
// sends the CPML-class entries of the kinds its fused kernel carries to list
// 1 and the rest to list 2; otherwise every CPML-class entry goes to list 1.
void append_shot(const fdtd2d_pool *pp, int s, std::vector<int4> *lists) {
  const fdtd2d_plan *pl = pp->shots[s].pl;
  const Tiling &t = pl->tiling;
  const bool captures = pl->a.nprobe > 0 || pl->line_rows > 0;
  for (int c = 0; c < 2; ++c) {
    const size_t start[fdtd2d::kPoolLists] = {lists[0].size(), lists[1].size(), lists[2].size()};
    for (const int4 &e : t.lists[c]) {
      if (e.y >= e.z) continue;  // padding
      int to = c;
      if (c == 1 && pp->fused && !fdtd2d::fusable_kind(fdtd2d::entry_kind(e.w), pl->a.medium_kind, captures))
        to = 2;
      lists[to].push_back(make_int4(e.x, e.y, e.z, (e.w & 0xff) | (s << 8)));
    }
    for (int l = 0; l < fdtd2d::kPoolLists; ++l)
      while ((lists[l].size() - start[l]) % fdtd2d::kWavesPerWG) lists[l].push_back(make_int4(0, 0, 0, s << 8));
  }
}

// Copies the lists to device buffers d[], growing them as needed.
int upload_lists(const std::vector<int4> *lists, int4 **d, size_t *cap, int count[fdtd2d::kPoolLists]) {
  bool moved = false;
  for (int c = 0; c < fdtd2d::kPoolLists; ++c) {
    int rc = grow(d[c], cap[c], std::max<size_t>(lists[c].size(), 4), moved);
    if (rc) return rc;
    if (!lists[c].empty())
      HIP_TRY(hipMemcpy(d[c], lists[c].data(), lists[c].size() * sizeof(int4), hipMemcpyHostToDevice),
              "hipMemcpy pool lists");
    count[c] = static_cast<int>(lists[c].size() / fdtd2d::kWavesPerWG);
  }
  return FDTD2D_OK;
}

// Makes ring buffer `nb` (holding the lists of `slots`) the current lists.
int use_lists(fdtd2d_pool *pp, int nb, const std::vector<int> &slots) {
  const fdtd2d_pool::ListBuf &b = pp->bufs[nb];
  hipError_t e = fdtd2d::launch_pool_set(pp->d_q, b.d, b.count, -1, pp->stream);
  if (e != hipSuccess) return hip_fail(e, "pool set launch");
  // a captured graph holds the leftover kernel's grid (none when list 2 is empty)
  const int gc = pp->ncu * pp->wg_per_cu_cpml;
  if (std::min(gc, b.count[2]) != std::min(gc, pp->main_count[2])) pp->graph_valid = false;
  for (int c = 0; c < fdtd2d::kPoolLists; ++c) pp->main_count[c] = b.count[c];
  pp->cur = nb;
  pp->main_set = slots;
  pp->lists_valid = true;
  return FDTD2D_OK;
}

int rebuild_main(fdtd2d_pool *pp, const std::vector<int> &slots) {
  for (int i = 0; i < pp->nbufs; ++i)
    if (pp->bufs[i].valid && pp->bufs[i].set == slots) {
      pp->list_switches++;
      return use_lists(pp, i, slots);
    }
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int4> lists[fdtd2d::kPoolLists];
  for (int s : slots) append_shot(pp, s, lists);
  bool valid[fdtd2d_pool::kListBufs];
  int last_g[fdtd2d_pool::kListBufs];
  for (int i = 0; i < pp->nbufs; ++i) {
    valid[i] = pp->bufs[i].valid;
    last_g[i] = pp->bufs[i].last_g;
  }
  const int nb = fdtd2d::policy::pick_copy(valid, last_g, pp->nbufs, pp->cur);
  fdtd2d_pool::ListBuf &b = pp->bufs[nb];
  const auto tw = std::chrono::steady_clock::now();
  retire(pp, b.last_g + 1);  // launches that read this buffer have finished
  pp->blocked_s += seconds_since(tw);
  b.valid = false;
  int rc = upload_lists(lists, b.d, b.cap, b.count);
  if (rc) return rc;
  b.set = slots;
  b.valid = true;
  if ((rc = use_lists(pp, nb, slots))) return rc;
  pp->rebuilds++;
  pp->rebuild_s += seconds_since(t0);
  return FDTD2D_OK;
}

// Forgets every list copy (tilings or routing changed).
void invalidate_lists(fdtd2d_pool *pp) {
  for (auto &b : pp->bufs) b.valid = false;
  pp->lists_valid = false;
}

bool slicing(const fdtd2d_pool *pp) { return pp->slice_every > 1 && pp->slice_cells > 0; }

bool sliced_cells(const fdtd2d_pool *pp, double cells) { return slicing(pp) && cells >= pp->slice_cells; }

bool sliced(const fdtd2d_pool *pp, const fdtd2d_plan *pl) {
  return sliced_cells(pp, static_cast<double>(pl->L.nx) * pl->L.ny);
}

// Band (and CPML sub-band) rows of a shot of `cells` cells: shorter bands
// for the shots that run in every launch while large ones are sliced (a
// launch carrying only them lasts about one band's row loop).
int shot_band(const fdtd2d_pool *pp, double cells) {
  return slicing(pp) && pp->band_small > 0 && !sliced_cells(pp, cells) ? pp->band_small : pp->band;
}

int shot_band_cpml(const fdtd2d_pool *pp, double cells) {
  return slicing(pp) && pp->band_cpml_small > 0 && !sliced_cells(pp, cells) ? pp->band_cpml_small
                                                                              : pp->band_cpml;
}

// Fills the host copy of shot `slot`'s descriptor: its state at the next
// pool launch.
void fill_desc(fdtd2d_pool *pp, int slot) {
  const fdtd2d_plan *pl = pp->shots[slot].pl;
  fdtd2d::StepArgs &a = pp->h_descs[slot];
  a = pl->a;
  // a sliced shot runs in every slice_every-th launch from the next one on
  const int k = sliced(pp, pl) ? pp->slice_every : 1;
  a.g_stride = k;
  a.g_base = fdtd2d::policy::first_turn(pp->g, k);
  a.n_base = pl->n;
  a.src_base = pl->src;
}

// Copies descriptors [lo, hi) of the host table to the device table.
int copy_descs(fdtd2d_pool *pp, size_t lo, size_t hi) {
  if (hi > lo)
    HIP_TRY(hipMemcpy(pp->d_descs + lo, pp->h_descs.data() + lo, (hi - lo) * sizeof(fdtd2d::StepArgs),
                      hipMemcpyHostToDevice),
            "hipMemcpy descriptors");
  return FDTD2D_OK;
}

int upload_desc(fdtd2d_pool *pp, int slot) {
  fill_desc(pp, slot);
  return copy_descs(pp, slot, slot + 1);
}

hipEvent_t take_event(fdtd2d_pool *pp) {
  if (!pp->spare_events.empty()) {
    hipEvent_t e = pp->spare_events.back();
    pp->spare_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

// Retires completion events in launch order: waits for those of launches up
// to wait_g (launch counts), takes the others only if they have passed.
void retire(fdtd2d_pool *pp, int wait_g) {
  size_t k = 0;
  for (; k < pp->done_events.size(); ++k) {
    auto &de = pp->done_events[k];
    if (de.first <= wait_g) hipEventSynchronize(de.second);
    else if (hipEventQuery(de.second) != hipSuccess) break;
    pp->g_done_known = std::max(pp->g_done_known, de.first);
    pp->spare_events.push_back(de.second);
  }
  pp->done_events.erase(pp->done_events.begin(), pp->done_events.begin() + k);
}

int record_done(fdtd2d_pool *pp) {
  hipEvent_t e = take_event(pp);
  if (!e) return set_error(FDTD2D_EDEVICE, "hipEventCreate");
  HIP_TRY(hipEventRecord(e, pp->stream), "hipEventRecord");
  pp->done_events.push_back({pp->g, e});
  if (pp->done_events.size() > 256) retire(pp, -1);
  return FDTD2D_OK;
}

// One pool launch of the main lists at chunk position k (every class).
// Unfused: the CPML-class kernel on the aux stream beside the interior one.
// Fused: one kernel takes the fusable CPML-class work-groups first, then the
// interior ones; the leftover CPML-class list, when it has work, runs in its
// own kernel on the aux stream beside it.
int launch_main(fdtd2d_pool *pp, int k) {
  const int gi = pp->ncu * pp->wg_per_cu, gc = pp->ncu * pp->wg_per_cu_cpml;
  const bool side = !pp->fused || pp->main_count[2] > 0;
  hipError_t e = hipSuccess;
  if (side) {
    HIP_TRY(hipEventRecord(pp->ev_fork, pp->stream), "hipEventRecord fork");
    HIP_TRY(hipStreamWaitEvent(pp->aux, pp->ev_fork, 0), "hipStreamWaitEvent fork");
    e = fdtd2d::launch_pool(pp->d_descs, pp->d_q, k, pp->T, pp->fused ? 3 : 1, pp->T, 1, 1, pp->medium_kind,
                            pp->fused ? std::min(gc, pp->main_count[2]) : gc, pp->aux);
    if (e != hipSuccess) return hip_fail(e, "pool CPML kernel launch");
    HIP_TRY(hipEventRecord(pp->ev_join, pp->aux), "hipEventRecord join");
  }
  e = fdtd2d::launch_pool(pp->d_descs, pp->d_q, k, pp->T, pp->fused ? 2 : 0, pp->T, pp->C, pp->D, pp->medium_kind,
                          gi,
                          pp->stream);
  if (e != hipSuccess) return hip_fail(e, "pool kernel launch");
  if (side) HIP_TRY(hipStreamWaitEvent(pp->stream, pp->ev_join, 0), "hipStreamWaitEvent join");
  return FDTD2D_OK;
}

int capture_graph(fdtd2d_pool *pp) {
  const auto t0 = std::chrono::steady_clock::now();
  pp->captures++;
  if (pp->gexec) {
    hipGraphExecDestroy(pp->gexec);
    pp->gexec = nullptr;
  }
  HIP_TRY(hipStreamBeginCapture(pp->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  int rc = FDTD2D_OK;
  for (int k = 0; k < pp->chunk && !rc; ++k) rc = launch_main(pp, k);
  if (!rc) {
    hipError_t e = fdtd2d::launch_pool_tick(pp->d_q, pp->chunk, pp->chunk, pp->stream);
    if (e != hipSuccess) rc = hip_fail(e, "pool tick launch");
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(pp->stream, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&pp->gexec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  pp->graph_valid = true;
  pp->capture_s += seconds_since(t0);
  return FDTD2D_OK;
}

// Tail launches at the current g: shots with 0 < steps left < T, grouped by
// steps left.
int run_tails(fdtd2d_pool *pp, const std::vector<int> &tails) {
  std::map<int, std::vector<int>> by_t;
  for (int s : tails) by_t[pp->shots[s].pl->nsteps - pp->shots[s].pl->n].push_back(s);
  for (auto &kv : by_t) {
    const int t = kv.first;
    std::vector<int4> lists[fdtd2d::kPoolLists];
    for (int s : kv.second) append_shot(pp, s, lists);
    fdtd2d_pool::TailBuf &tb = pp->tailbufs[pp->tail_next];
    pp->tail_next = (pp->tail_next + 1) % fdtd2d_pool::kTailBufs;
    const auto tw = std::chrono::steady_clock::now();
    retire(pp, tb.last_g + 1);  // the launch that last read this copy has finished
    pp->blocked_s += seconds_since(tw);
    tb.last_g = pp->g;
    int count[fdtd2d::kPoolLists];
    int rc = upload_lists(lists, tb.d, tb.cap, count);
    if (rc) return rc;
    const int n0 = count[0], n1 = count[1], n2 = count[2];
    const int gi = pp->ncu * pp->wg_per_cu, gc = pp->ncu * pp->wg_per_cu_cpml;
    hipError_t e = fdtd2d::launch_pool_set(pp->d_qt, tb.d, count, pp->g, pp->stream);
    if (pp->fused) {
      // in stream order: the leftover kernel's tiles are disjoint from the
      // fused kernel's, both read set src and write set 1 - src
      if (e == hipSuccess && n2 > 0)
        e = fdtd2d::launch_pool(pp->d_descs, pp->d_qt, 0, pp->T, 3, t, 1, 1, pp->medium_kind, std::min(gc, n2),
                                pp->stream);
      if (e == hipSuccess && n0 + n1 > 0)
        e = fdtd2d::launch_pool(pp->d_descs, pp->d_qt, 0, pp->T, 2, t, pp->C, pp->D, pp->medium_kind,
                                std::min(gi, n0 + n1), pp->stream);
    } else {
      if (e == hipSuccess && n1 > 0)
        e = fdtd2d::launch_pool(pp->d_descs, pp->d_qt, 0, pp->T, 1, t, 1, 1, pp->medium_kind, std::min(gc, n1),
                                pp->stream);
      if (e == hipSuccess && n0 > 0)
        e = fdtd2d::launch_pool(pp->d_descs, pp->d_qt, 0, pp->T, 0, t, pp->C, pp->D, pp->medium_kind,
                                std::min(gi, n0), pp->stream);
    }
    if (e == hipSuccess) e = fdtd2d::launch_pool_tick(pp->d_qt, 1, 0, pp->stream);
    if (e != hipSuccess) return hip_fail(e, "pool tail launch");
    pp->tail_launches++;
    for (int s : kv.second) {
      fdtd2d_plan *pl = pp->shots[s].pl;
      pp->cell_updates += static_cast<double>(pl->L.nx) * pl->L.ny * t;
      pl->n += t;
      pl->src ^= 1;
      pp->shots[s].finished = true;
      pp->shots[s].last_g = pp->g;
    }
  }
  return FDTD2D_OK;
}

// Retiles every live shot at the pool's launch shape (no launch in flight);
// the descriptors go to the device in one copy.
int reshape(fdtd2d_pool *pp) {
  for (size_t s = 0; s < pp->shots.size(); ++s) {
    Shot &sh = pp->shots[s];
    if (!sh.live) continue;
    sh.pl->T = pp->T;
    sh.pl->C = pp->C;
    const double cells = static_cast<double>(sh.pl->L.nx) * sh.pl->L.ny;
    sh.pl->band = shot_band(pp, cells);
    sh.pl->band_cpml = shot_band_cpml(pp, cells);
    int rc = set_tiling(sh.pl);
    if (rc) return rc;
    fill_desc(pp, static_cast<int>(s));
  }
  int rc = copy_descs(pp, 0, pp->shots.size());
  if (rc) return rc;
  invalidate_lists(pp);
  pp->graph_valid = false;
  return FDTD2D_OK;
}

int step_impl(fdtd2d_pool *pp, int nlaunch) {
  HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
  while (nlaunch > 0) {
    std::vector<int> main_set, tails;
    int min_left = 1 << 30;
    // time slicing: sliced shots sit out the launches between their turns
    const bool slice_turn = pp->g % std::max(pp->slice_every, 1) == 0;
    bool any_sliced = false;
    for (size_t s = 0; s < pp->shots.size(); ++s) {
      const Shot &sh = pp->shots[s];
      if (!sh.live || sh.finished) continue;
      if (sliced(pp, sh.pl)) {
        any_sliced = true;
        if (!slice_turn) continue;
      }
      const int left = sh.pl->nsteps - sh.pl->n;
      if (left >= pp->T) {
        main_set.push_back(static_cast<int>(s));
        min_left = std::min(min_left, left);
      } else {
        tails.push_back(static_cast<int>(s));
      }
    }
    if (main_set.empty() && tails.empty()) {
      if (!any_sliced) break;
      // only sliced shots run and this is not their turn: move the launch
      // counter (host and device) on to their next turn, launching nothing
      const int skip = fdtd2d::policy::to_next_turn(pp->g, pp->slice_every);
      hipError_t e = fdtd2d::launch_pool_tick(pp->d_q, 0, skip, pp->stream);
      if (e != hipSuccess) return hip_fail(e, "pool tick launch");
      pp->g += skip;
      continue;
    }
    if (!pp->lists_valid || main_set != pp->main_set) {
      int rc = rebuild_main(pp, main_set);
      if (rc) return rc;
    }
    const bool use_graph = pp->chunk > 0 && nlaunch >= pp->chunk && tails.empty() && !main_set.empty() &&
                           min_left >= pp->chunk * pp->T && !any_sliced;
    if (use_graph) {
      if (!pp->graph_valid) {
        int rc = capture_graph(pp);
        if (rc) return rc;
      }
      HIP_TRY(hipGraphLaunch(pp->gexec, pp->stream), "hipGraphLaunch");
      pp->graph_replays++;
      pp->bufs[pp->cur].last_g = pp->g + pp->chunk - 1;
      for (int s : main_set) {
        fdtd2d_plan *pl = pp->shots[s].pl;
        pp->cell_updates += static_cast<double>(pl->L.nx) * pl->L.ny * pp->T * pp->chunk;
        pl->n += pp->T * pp->chunk;
        pl->src ^= pp->chunk & 1;
        if (pl->n >= pl->nsteps) {
          pp->shots[s].finished = true;
          pp->shots[s].last_g = pp->g + pp->chunk - 1;
        }
      }
      pp->g += pp->chunk;
      nlaunch -= pp->chunk;
    } else {
      if (!main_set.empty()) {
        int rc = launch_main(pp, 0);
        if (rc) return rc;
        pp->eager_launches++;
        pp->bufs[pp->cur].last_g = pp->g;
      }
      if (!tails.empty()) {
        int rc = run_tails(pp, tails);
        if (rc) return rc;
      }
      hipError_t e = fdtd2d::launch_pool_tick(pp->d_q, 1, 1, pp->stream);
      if (e != hipSuccess) return hip_fail(e, "pool tick launch");
      for (int s : main_set) {
        fdtd2d_plan *pl = pp->shots[s].pl;
        pp->cell_updates += static_cast<double>(pl->L.nx) * pl->L.ny * pp->T;
        pl->n += pp->T;
        pl->src ^= 1;
        if (pl->n >= pl->nsteps) {
          pp->shots[s].finished = true;
          pp->shots[s].last_g = pp->g;
        }
      }
      pp->g += 1;
      nlaunch -= 1;
    }
    int rc = record_done(pp);
    if (rc) return rc;
  }
  return FDTD2D_OK;
}

Shot *get_shot(fdtd2d_pool *pp, int id) {
  if (!pp || id < 0 || id >= static_cast<int>(pp->shots.size()) || !pp->shots[id].live) return nullptr;
  return &pp->shots[id];
}

// Waits until the launch that issued the shot's last steps (or, for a shot
// still running, everything issued) has finished on the device.
int wait_shot(fdtd2d_pool *pp, const Shot &sh) {
  const auto t0 = std::chrono::steady_clock::now();
  if (sh.finished) {
    retire(pp, sh.last_g + 1);
    if (pp->g_done_known > sh.last_g) {
      pp->blocked_s += seconds_since(t0);
      return FDTD2D_OK;
    }
  }
  HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
  retire(pp, -1);
  pp->blocked_s += seconds_since(t0);
  return FDTD2D_OK;
}

}  // namespace

extern "C" {

int fdtd2d_pool_create(int device, fdtd2d_pool **out) {
  if (!out) return set_error(FDTD2D_EINVAL, "null out");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev) return set_error(FDTD2D_EINVAL, "no such device");
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  auto *pp = new (std::nothrow) fdtd2d_pool();
  if (!pp) return set_error(FDTD2D_ENOMEM, "out of host memory");
  pp->device = device;
  hipDeviceGetAttribute(&pp->ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (pp->ncu <= 0) pp->ncu = 256;
  auto fail = [&](int rc) {
    fdtd2d_pool_destroy(pp);
    return rc;
  };
  if (hipStreamCreateWithFlags(&pp->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&pp->aux, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(FDTD2D_EDEVICE, "hipStreamCreate"));
  if (hipEventCreateWithFlags(&pp->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&pp->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(FDTD2D_EDEVICE, "hipEventCreate"));
  if (hipMalloc(reinterpret_cast<void **>(&pp->d_q), sizeof(fdtd2d::PoolQueue)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&pp->d_qt), sizeof(fdtd2d::PoolQueue)) != hipSuccess)
    return fail(set_error(FDTD2D_ENOMEM, "hipMalloc pool queue"));
  if (hipMemset(pp->d_q, 0, sizeof(fdtd2d::PoolQueue)) != hipSuccess ||
      hipMemset(pp->d_qt, 0, sizeof(fdtd2d::PoolQueue)) != hipSuccess)
    return fail(set_error(FDTD2D_EDEVICE, "hipMemset"));
  *out = pp;
  return FDTD2D_OK;
}

int fdtd2d_pool_set_option(fdtd2d_pool *pp, const char *name, double value) {
  if (!pp || !name) return set_error(FDTD2D_EINVAL, "null argument");
  const std::string n = name;
  const long long v = static_cast<long long>(value);
  bool shape = false;
  if (n == "budget_bytes") {
    if (!pp->shots.empty() || pp->arena.base) return set_error(FDTD2D_ESTATE, "budget must be set before any shot");
    if (value < 0) return set_error(FDTD2D_EINVAL, "negative budget");
    pp->budget = static_cast<size_t>(value);
    if (pp->budget > 0) {
      HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
      if (!pp->arena.init(pp->budget)) return set_error(FDTD2D_ENOMEM, "hipMalloc of the pool arena");
    }
    return FDTD2D_OK;
  }
  if (n == "tblock") {
    if (v < 1 || v > fdtd2d::kMaxT) return set_error(FDTD2D_EINVAL, "tblock must be 1..8");
    pp->T = static_cast<int>(v);
    shape = true;
  } else if (n == "band") {
    if (v < 0) return set_error(FDTD2D_EINVAL, "band must be >= 0");
    pp->band = static_cast<int>(v);
    shape = true;
  } else if (n == "band_small" || n == "band_cpml_small") {
    if (v < 0) return set_error(FDTD2D_EINVAL, n + " must be >= 0");
    (n == "band_small" ? pp->band_small : pp->band_cpml_small) = static_cast<int>(v);
    shape = true;
  } else if (n == "band_cpml") {
    if (v < 0) return set_error(FDTD2D_EINVAL, "band_cpml must be >= 0");
    pp->band_cpml = static_cast<int>(v);
    shape = true;
  } else if (n == "cols_per_lane") {
    if (v != 1 && v != 2) return set_error(FDTD2D_EINVAL, "cols_per_lane must be 1 or 2");
    pp->C = static_cast<int>(v);
    shape = true;
  } else if (n == "prefetch") {
    if (v != 1 && v != 2) return set_error(FDTD2D_EINVAL, "prefetch must be 1 or 2");
    pp->D = static_cast<int>(v);
    pp->graph_valid = false;
  } else if (n == "fused") {
    if (v != 0 && v != 1) return set_error(FDTD2D_EINVAL, "fused must be 0 or 1");
    pp->fused = static_cast<int>(v);
    // the CPML-class entries are routed to lists 1 and 2 by the arrangement
    invalidate_lists(pp);
    pp->graph_valid = false;
  } else if (n == "wg_per_cu") {
    if (v < 1 || v > 8) return set_error(FDTD2D_EINVAL, "wg_per_cu must be 1..8");
    pp->wg_per_cu = static_cast<int>(v);
    pp->graph_valid = false;
  } else if (n == "wg_per_cu_cpml") {
    if (v < 1 || v > 8) return set_error(FDTD2D_EINVAL, "wg_per_cu_cpml must be 1..8");
    pp->wg_per_cu_cpml = static_cast<int>(v);
    pp->graph_valid = false;
  } else if (n == "graph_chunk") {
    if (v < 0 || v > fdtd2d::kMaxChunk) return set_error(FDTD2D_EINVAL, "graph_chunk must be 0..16");
    pp->chunk = static_cast<int>(v);
    pp->graph_valid = false;
  } else if (n == "list_buffers") {
    if (v < 2 || v > fdtd2d_pool::kListBufs) return set_error(FDTD2D_EINVAL, "list_buffers must be 2..4");
    pp->nbufs = static_cast<int>(v);  // from the next rebuild on; each buffer keeps its last reader
  } else if (n == "slice_cells" || n == "slice_every") {
    if (n == "slice_cells" && !(v >= 0)) return set_error(FDTD2D_EINVAL, "slice_cells must be >= 0");
    if (n == "slice_every" && (v < 1 || v > 64)) return set_error(FDTD2D_EINVAL, "slice_every must be 1..64");
    // the shots' descriptors count their turns from the next launch on
    HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
    retire(pp, pp->g);
    if (n == "slice_cells") pp->slice_cells = v;
    else pp->slice_every = static_cast<int>(v);
    if (pp->band_small > 0 || pp->band_cpml_small > 0) return reshape(pp);  // a shot's bands follow its class
    for (size_t s = 0; s < pp->shots.size(); ++s)
      if (pp->shots[s].live) fill_desc(pp, static_cast<int>(s));
    return copy_descs(pp, 0, pp->shots.size());
  } else {
    return set_error(FDTD2D_EINVAL, "unknown pool option " + n);
  }
  if (shape) {
    HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
    return reshape(pp);
  }
  return FDTD2D_OK;
}

int fdtd2d_pool_get_info(const fdtd2d_pool *pp, const char *name, double *value) {
  if (!pp || !name || !value) return set_error(FDTD2D_EINVAL, "null argument");
  const std::string n = name;
  int live = 0, running = 0;
  for (const Shot &s : pp->shots) {
    live += s.live;
    running += s.live && !s.finished;
  }
  if (n == "tblock") *value = pp->T;
  else if (n == "band") *value = pp->band;
  else if (n == "band_cpml") *value = pp->band_cpml;
  else if (n == "band_small") *value = pp->band_small;
  else if (n == "band_cpml_small") *value = pp->band_cpml_small;
  else if (n == "prefetch") *value = pp->D;
  else if (n == "cols_per_lane") *value = pp->C;
  else if (n == "wg_per_cu") *value = pp->wg_per_cu;
  else if (n == "fused") *value = pp->fused;
  else if (n == "wg_per_cu_cpml") *value = pp->wg_per_cu_cpml;
  else if (n == "graph_chunk") *value = pp->chunk;
  else if (n == "list_buffers") *value = pp->nbufs;
  else if (n == "slice_cells") *value = pp->slice_cells;
  else if (n == "slice_every") *value = pp->slice_every;
  else if (n == "list_switches") *value = pp->list_switches;
  else if (n == "budget_bytes") *value = static_cast<double>(pp->budget);
  else if (n == "bytes_in_use") *value = static_cast<double>(pp->arena.in_use);
  else if (n == "largest_free") *value = static_cast<double>(pp->budget ? pp->arena.largest_free() : 0);
  else if (n == "shots") *value = live;
  else if (n == "running") *value = running;
  else if (n == "launches") *value = pp->g;
  else if (n == "cell_updates") *value = pp->cell_updates;
  else if (n == "graph_replays") *value = pp->graph_replays;
  else if (n == "eager_launches") *value = pp->eager_launches;
  else if (n == "tail_launches") *value = pp->tail_launches;
  else if (n == "rebuilds") *value = pp->rebuilds;
  else if (n == "captures") *value = pp->captures;
  else if (n == "rebuild_s") *value = pp->rebuild_s;
  else if (n == "capture_s") *value = pp->capture_s;
  else if (n == "blocked_s") *value = pp->blocked_s;
  else if (n == "add_s") *value = pp->add_s;
  else if (n == "cus") *value = pp->ncu;
  else return set_error(FDTD2D_EINVAL, "unknown pool info " + n);
  return FDTD2D_OK;
}

int fdtd2d_pool_add(fdtd2d_pool *pp, const fdtd2d_problem *p, int *shot) {
  if (!pp || !p || !shot) return set_error(FDTD2D_EINVAL, "null argument");
  *shot = -1;
  int rc = fdtd2d::validate_problem(p);
  if (rc) return rc;
  if (p->nrc > 0) return set_error(FDTD2D_EUNSUPPORTED, "RC cutoff lines run on single plans only");
  if (pp->medium_kind >= 0 && p->medium_kind != pp->medium_kind)
    return set_error(FDTD2D_EUNSUPPORTED, "a pool holds shots of one medium kind");
  HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t bytes = plan_bytes(p);
  char *base = nullptr;
  if (pp->budget > 0) {
    base = pp->arena.take(bytes);
    if (!base) return set_error(FDTD2D_ENOMEM, "over the pool's memory budget");
  }
  fdtd2d_plan *pl = nullptr;
  rc = create_plan(p, pp->device, true, pp->stream, pp->T, pp->C,
                   shot_band(pp, static_cast<double>(p->nx) * p->ny),
                   shot_band_cpml(pp, static_cast<double>(p->nx) * p->ny), base, &pl);
  if (rc) {
    if (base) pp->arena.give(base);
    return rc;
  }
  int slot;
  if (!pp->free_slots.empty()) {
    slot = pp->free_slots.back();
    pp->free_slots.pop_back();
  } else {
    slot = static_cast<int>(pp->shots.size());
    if (slot >= (1 << 23)) {
      fdtd2d_plan_destroy(pl);
      if (base) pp->arena.give(base);
      return set_error(FDTD2D_ENOMEM, "too many shots");
    }
    pp->shots.emplace_back();
  }
  if (static_cast<size_t>(slot) >= pp->cap_slots) {
    // grow the descriptor table (keeps the existing descriptors)
    HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
    const size_t nc = std::max<size_t>(64, pp->cap_slots * 2);
    fdtd2d::StepArgs *nd = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&nd), nc * sizeof(fdtd2d::StepArgs)) != hipSuccess)
      return set_error(FDTD2D_ENOMEM, "hipMalloc descriptors");
    if (pp->d_descs) {
      HIP_TRY(hipMemcpy(nd, pp->d_descs, pp->cap_slots * sizeof(fdtd2d::StepArgs), hipMemcpyDeviceToDevice),
              "hipMemcpy descriptors");
      hipFree(pp->d_descs);
    }
    pp->d_descs = nd;
    pp->cap_slots = nc;
    pp->h_descs.resize(nc);
    pp->graph_valid = false;
  }
  Shot &sh = pp->shots[slot];
  sh = Shot();
  sh.pl = pl;
  sh.live = true;
  sh.bytes = bytes;
  sh.medium_kind = p->medium_kind;
  sh.finished = p->nsteps == 0;
  sh.last_g = pp->g - 1;
  pp->medium_kind = p->medium_kind;
  if ((rc = upload_desc(pp, slot))) return rc;
  *shot = slot;
  pp->add_s += seconds_since(t0);
  return FDTD2D_OK;
}

int fdtd2d_pool_step(fdtd2d_pool *pp, int nlaunch) {
  if (!pp) return set_error(FDTD2D_EINVAL, "null pool");
  if (nlaunch < 0) return set_error(FDTD2D_EINVAL, "negative launch count");
  return step_impl(pp, nlaunch);
}

int fdtd2d_pool_sync(fdtd2d_pool *pp) {
  if (!pp) return set_error(FDTD2D_EINVAL, "null pool");
  HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
  retire(pp, pp->g);
  return FDTD2D_OK;
}

int fdtd2d_pool_wait(fdtd2d_pool *pp, int max_pending) {
  if (!pp) return set_error(FDTD2D_EINVAL, "null pool");
  if (max_pending < 0) return set_error(FDTD2D_EINVAL, "negative launch count");
  HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
  retire(pp, pp->g - max_pending);
  return FDTD2D_OK;
}

int fdtd2d_pool_shot_state(fdtd2d_pool *pp, int id, int *steps_done, int *steps_total, int *complete) {
  Shot *sh = get_shot(pp, id);
  if (!sh) return set_error(FDTD2D_EINVAL, "no such shot");
  if (steps_done) *steps_done = sh->pl->n;
  if (steps_total) *steps_total = sh->pl->nsteps;
  if (complete) {
    retire(pp, -1);
    *complete = sh->finished && pp->g_done_known > sh->last_g;
  }
  return FDTD2D_OK;
}

int fdtd2d_pool_read_detectors(fdtd2d_pool *pp, int id, float *det, size_t det_count, float *comp,
                               size_t comp_count) {
  Shot *sh = get_shot(pp, id);
  if (!sh) return set_error(FDTD2D_EINVAL, "no such shot");
  int rc = wait_shot(pp, *sh);
  if (rc) return rc;
  return fdtd2d_read_detectors(sh->pl, det, det_count, comp, comp_count);
}

int fdtd2d_pool_read_field(fdtd2d_pool *pp, int id, int which, float *out, size_t count) {
  Shot *sh = get_shot(pp, id);
  if (!sh) return set_error(FDTD2D_EINVAL, "no such shot");
  int rc = wait_shot(pp, *sh);
  if (rc) return rc;
  return fdtd2d_read_field(sh->pl, which, out, count);
}

int fdtd2d_pool_read_line(fdtd2d_pool *pp, int id, float *line, size_t count) {
  Shot *sh = get_shot(pp, id);
  if (!sh) return set_error(FDTD2D_EINVAL, "no such shot");
  int rc = wait_shot(pp, *sh);
  if (rc) return rc;
  return fdtd2d_read_line(sh->pl, line, count);
}

int fdtd2d_pool_remove(fdtd2d_pool *pp, int id) {
  Shot *sh = get_shot(pp, id);
  if (!sh) return set_error(FDTD2D_EINVAL, "no such shot");
  HIP_TRY(hipSetDevice(pp->device), "hipSetDevice");
  // A running shot may be named by launches in flight: wait for them.
  int rc = wait_shot(pp, *sh);
  if (rc) return rc;
  if (!sh->finished) {
    HIP_TRY(hipStreamSynchronize(pp->stream), "hipStreamSynchronize");
  }
  char *base = sh->pl->owns_base ? nullptr : sh->pl->base;
  fdtd2d_plan *pl = sh->pl;
  sh->pl = nullptr;
  sh->live = false;
  // the plan's own teardown syncs the pool stream only when it owns memory
  pl->stream = nullptr;
  fdtd2d_plan_destroy(pl);
  if (base) pp->arena.give(base);
  pp->free_slots.push_back(id);
  auto it = std::find(pp->main_set.begin(), pp->main_set.end(), id);
  if (it != pp->main_set.end()) pp->lists_valid = false;
  // the slot may come back with another shot: no list copy may name it
  for (auto &b : pp->bufs)
    if (b.valid && std::find(b.set.begin(), b.set.end(), id) != b.set.end()) b.valid = false;
  return FDTD2D_OK;
}

int fdtd2d_pool_destroy(fdtd2d_pool *pp) {
  if (!pp) return FDTD2D_OK;
  hipSetDevice(pp->device);
  if (pp->stream) hipStreamSynchronize(pp->stream);
  if (pp->aux) hipStreamSynchronize(pp->aux);
  for (Shot &s : pp->shots)
    if (s.live && s.pl) {
      s.pl->stream = nullptr;
      fdtd2d_plan_destroy(s.pl);
    }
  for (auto &de : pp->done_events) hipEventDestroy(de.second);
  for (auto e : pp->spare_events) hipEventDestroy(e);
  if (pp->gexec) hipGraphExecDestroy(pp->gexec);
  for (int c = 0; c < fdtd2d::kPoolLists; ++c) {
    for (int b = 0; b < fdtd2d_pool::kListBufs; ++b)
      if (pp->bufs[b].d[c]) hipFree(pp->bufs[b].d[c]);
    for (auto &tb : pp->tailbufs)
      if (tb.d[c]) hipFree(tb.d[c]);
  }
  if (pp->d_descs) hipFree(pp->d_descs);
  if (pp->d_q) hipFree(pp->d_q);
  if (pp->d_qt) hipFree(pp->d_qt);
  if (pp->arena.base) hipFree(pp->arena.base);
  if (pp->ev_fork) hipEventDestroy(pp->ev_fork);
  if (pp->ev_join) hipEventDestroy(pp->ev_join);
  if (pp->aux) hipStreamDestroy(pp->aux);
  if (pp->stream) hipStreamDestroy(pp->stream);
  delete pp;
  return FDTD2D_OK;
}

}  // extern "C"
