// Batched time steps of many independent boxes (the serving path), with
// continuous batching at the granularity of Krylov iterations.
//
// A pf_batch holds the descriptors of up to `capacity` resident boxes.  Each
// box is idle or inside a time step.  pf_batch_advance
//   1. starts a step for the boxes the caller names (one batched launch per
//      kernel over just those boxes: RK4 stages, Poisson source and Krylov
//      start, the first KA);
//   2. replays the Krylov chunk graph - kChunk BiCGStab iterations over all
//      resident boxes, ending with every box's done flag copied to pinned
//      host memory - until at least one running box's solve is done;
//   3. finishes the step of those boxes (batched correction, pressure update,
//      monitor values) and returns which boxes finished.
// A box whose solve ends early therefore starts its next step while slower
// solves continue; no box waits for another.  Boxes that are idle or done
// return at the top of every chunk kernel (their control block says so).
//
// Every launch covers its boxes at once and each box keeps the block
// decomposition of its single-box launch, so its results are bit-identical
// alone, in any mix and through the single-box entry points.  The chunk
// graph is captured once per composition (pf_batch_set) and replayed; it
// does not depend on iteration counts.  Steps start at chunk boundaries,
// which keeps every box's iteration parity aligned with the graph's.
//
// A box whose persistent solve fills the GPU (pf_persist.hip: 128^3 FP32
// takes 512 blocks) solves its Poisson problem as that one cooperative
// launch, after the batched Krylov start of its step, in place of its share
// of the chunk kernels: its control block then says done and the next chunk
// boundary finishes its step.  At most persist_cap such solves run per
// pf_batch_advance call (default 1); a big box whose step starts beyond the
// cap waits, its control block held at done so that the chunk kernels pass
// over it, and solves at the start of a later call, first come first served
// - so the other boxes get a chunk between persistent solves.  The
// persistent solve's iterates are the chunk kernels' bits, so results stay
// bit-identical.  A box whose cooperative launch cannot start runs on the
// chunk kernels.
//
// Inside a chunk the boxes run in cache-residency groups: boxes are packed,
// by tile shape and then size, into groups whose Krylov vectors together fit
// a byte budget sized for the Infinity Cache (MALL), and each group runs all
// kChunk iterations before the next group starts.  A box alone on the GPU
// keeps its Krylov set in that cache from one kernel to the next; sweeping
// every resident box in each launch instead makes the sets of a mix stream
// from HBM.  Grouping changes only the order of independent boxes' launches,
// not any box's arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "pf_kernels.h"

namespace {

// BiCGStab iterations per graph replay (even: buffer parity).  With one
// persistent solve per advance call, a smaller box's step spans fewer calls,
// each holding a big solve, when chunks are longer; a big box then waits for
// a longer chunk.  On the 24-request serving stream (FP32) 8 gave p50 / p99
// step latency 35.9 / 146 ms, 16 35.3 / 85.7, 32 25.6 / 69.4, 48 20.8 / 82.4
// and 64 21.3 / 94.5 ms, at 344-361 Mcell-steps/s (64 lowest).
constexpr int kChunk = 32;
// Default cache-residency budget of a group's Krylov vectors.  The nine
// vectors per box overstate what one kernel streams (KA, KB and KC touch six,
// two and seven of them), so the best budget exceeds the 256 MB Infinity
// Cache: on the 24-request serving stream (FP32) 0 (one group) gave 251,
// 192 MiB 260, 256 MiB 268, 320 MiB 272, 384 MiB 258 and 512 MiB 256
// Mcell-steps/s.
constexpr size_t kGroupBytes = (size_t)320 << 20;
// Boxes whose persistent solve takes at least this many blocks (every CU
// busy) solve persistently in a batch; smaller ones gain more from sharing
// the chunk kernels with the other boxes (64^3 FP32 alone: 128 blocks).
constexpr int kPersistMinBlocks = 256;

struct List {
  std::vector<int> box, begin;
  int total = 0;
  int tz = 0, ty = 0;
  int* dev = nullptr;        // box ids, then begins (n + 1), on the device
  pf::BatchList view() const {
    pf::BatchList L;
    L.box = dev;
    L.begin = dev + box.size();
    L.n = (int)box.size();
    return L;
  }
};

// Kernel-family lists of one set of boxes.
struct Lists {
  List flat, red, pad;
  std::vector<List> tiles;
  std::vector<int> ids;      // the boxes, in order
  int* ids_dev = nullptr;
  std::vector<List*> all() {
    std::vector<List*> v = {&flat, &red, &pad};
    for (auto& t : tiles) v.push_back(&t);
    return v;
  }
};

}  // namespace

struct pf_batch {
  int dtype;
  int capacity;
  int n;
  size_t desc_size;
  void* desc;                 // device BoxDesc<T>[capacity]
  std::vector<pf_box*> boxes;
  std::vector<char> running;  // per box: inside a time step
  std::vector<char> pers;     // per box: solves with the persistent kernel
  std::vector<char> chunked;  // per box: its current step solves on the chunk kernels
  std::vector<int64_t> pend;  // per box: 0, or when its persistent solve began to wait (held done)
  int64_t pend_seq;           // stamps of pend, in the order the waits began
  int persist_cap;            // persistent solves per advance call (0: none)
  int* done_vals;             // pinned {0, 1}: control-block done values to copy
  double* out_dev;            // 8 doubles per box
  double* out_host;           // pinned
  int* flags_dev;             // done flag per box
  int* flags_host;            // pinned
  hipStream_t st;
  hipEvent_t ev;
  hipEvent_t wev;             // completion marker the chunk loop's waits spin on
  Lists full, pre, post, ka;
  int* list_mem;              // device lists: full, pre, post, groups, first-KA regions
  int* list_stage;            // pinned staging of the three regions
  size_t list_cap;            // ints per region
  hipGraphExec_t g_chunk;     // chunk over the groups of boxes that do not solve persistently
  hipGraphExec_t g_chunk_all; // chunk over every group (null if no box solves persistently)
  size_t plain_groups;        // groups before the persistent boxes' groups
  int maxit_max;
  size_t group_bytes;         // cache-residency budget of one group (0: one group)
  std::vector<Lists> groups;  // tile lists of each group (list region 3)
};

using namespace pf;

namespace {

template <typename T>
BoxDesc<T> make_desc(pf_box* h, void* const* f, void* src, const double* prm, double* out) {
  // f: u, v, w, ut, vt, wt, DP, p ; prm: dt, mu, rho, tol, maxit
  BoxDesc<T> d;
  d.g = h->g;
  d.tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
  d.nblk = h->nblk;
  for (int q = 0; q < 3; ++q) {
    d.u[q] = (T*)f[q];
    d.ut[q] = (T*)f[3 + q];
    d.yp[q] = (T*)h->ypong[q];
    d.acc[q] = (T*)h->acc[q];
  }
  d.DP = (T*)f[6];
  d.p = (T*)f[7];
  d.src = (T*)src;
  d.x = (T*)h->x;
  d.rk[0] = (T*)h->r;
  d.rk[1] = (T*)h->s;   // second residual buffer
  d.rh = (T*)h->rh;
  d.pk[0] = (T*)h->p[0];
  d.pk[1] = (T*)h->p[1];
  d.vk[0] = (T*)h->v[0];
  d.vk[1] = (T*)h->v[1];
  d.ctl = h->ctl;
  d.pA = part_buf(h, 0);
  d.pB = part_buf(h, 1);
  d.pC = part_buf(h, 2);
  d.pS = part_buf(h, 3);
  d.out = out;
  const double dt = prm[0], mu = prm[1], rho = prm[2];
  d.k = make_step_coef<T>(h, mu, rho, dt);
  d.L = make_lap<T>(h);
  const double vol = h->dx * h->dy * h->dz;
  d.sc1 = (T)(rho / dt);
  d.sc2 = (T)(-1.0 * vol);
  d.dxi = (T)(1.0 / h->dx);
  d.dyi = (T)(1.0 / h->dy);
  d.dzi = (T)(1.0 / h->dz);
  d.neg_c = (T)(-(dt / rho));
  d.tol = prm[3];
  d.maxit = (int)prm[4];
  return d;
}

void add(List& L, int box, int nblocks) {
  L.box.push_back(box);
  L.begin.push_back(L.total);
  L.total += nblocks;
}

// Lists of the boxes `ids` (indices into b->boxes).
void build_lists(const pf_batch* b, const std::vector<int>& ids, Lists& out) {
  out = Lists();
  out.ids = ids;
  for (int i : ids) {
    const pf_box* h = b->boxes[i];
    add(out.flat, i, launch_blocks(h->g));
    add(out.red, i, h->nblk);
    const int64_t np = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
    const int64_t npb = (np + kBlock - 1) / kBlock;
    add(out.pad, i, (int)(npb > 65536 ? 65536 : npb));
    const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
    List* tl = nullptr;
    for (auto& t : out.tiles)
      if (t.tz == tg.tz && t.ty == tg.ty) tl = &t;
    if (!tl) {
      out.tiles.push_back(List());
      tl = &out.tiles.back();
      tl->tz = tg.tz;
      tl->ty = tg.ty;
    }
    add(*tl, i, tg.nblk);
  }
}

// Ints the lists of n boxes take: each family lists every box once with its
// first block (2n + 1 per list; flat, red, pad and at most 8 tile shapes),
// then the id list.
size_t list_ints(int n) { return (size_t)11 * (2 * (size_t)n + 1) + (size_t)n; }

// Serialise `L` into host[0..) and point its lists at dev[0..); returns ints used.
size_t pack_lists(Lists& L, int* host, int* dev, size_t cap) {
  size_t off = 0;
  for (List* l : L.all()) {
    const size_t need = 2 * l->box.size() + 1;
    if (off + need > cap) return cap + 1;
    l->dev = dev + off;
    for (int v : l->box) host[off++] = v;
    for (int v : l->begin) host[off++] = v;
    host[off++] = l->total;
  }
  if (off + L.ids.size() > cap) return cap + 1;
  L.ids_dev = dev + off;
  for (int v : L.ids) host[off++] = v;
  return off;
}

// Upload the lists of a set of boxes through the pinned staging area of
// `region` (0: all boxes, 1: boxes starting a step, 2: boxes finishing one).
int upload_lists(pf_batch* b, Lists& L, int region) {
  int* host = b->list_stage + (size_t)region * b->list_cap;
  int* dev = b->list_mem + (size_t)region * b->list_cap;
  const size_t used = pack_lists(L, host, dev, b->list_cap);
  if (used > b->list_cap) return fail(PF_EINVAL, "pf_batch: more than 8 tile shapes in one batch");
  if (used) PF_HIP(hipMemcpyAsync(dev, host, used * sizeof(int), hipMemcpyHostToDevice, b->st));
  return 0;
}

// Bytes one box's BiCGStab iterations stream: the nine interior Krylov vectors.
size_t krylov_set_bytes(const pf_box* h) { return (size_t)9 * (size_t)h->g.ncell * h->esize; }

// Pack the resident boxes into cache-residency groups (greedy, boxes ordered
// by tile shape, then Krylov set size) and upload the groups' tile lists.
int build_groups(pf_batch* b) {
  b->groups.clear();
  const int n = b->n;
  if (n == 0) return 0;
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  auto shape = [&](int i) {
    const pf_box* h = b->boxes[i];
    const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
    return tg.tz * 1000 + tg.ty;
  };
  // boxes that solve persistently go last, in groups of their own: the chunk
  // kernels only pass over them (unless their cooperative launch failed)
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    if (b->pers[a] != b->pers[c]) return b->pers[a] < b->pers[c];
    const int sa = shape(a), sc = shape(c);
    if (sa != sc) return sa < sc;
    return krylov_set_bytes(b->boxes[a]) < krylov_set_bytes(b->boxes[c]);
  });
  std::vector<std::vector<int>> ids(1);
  size_t used = 0;
  for (size_t k = 0; k < order.size(); ++k) {
    const int i = order[k];
    const size_t by = krylov_set_bytes(b->boxes[i]);
    const bool split = k > 0 && b->pers[i] != b->pers[order[k - 1]];
    if (!ids.back().empty() && (split || (b->group_bytes && used + by > b->group_bytes))) {
      ids.emplace_back();
      used = 0;
    }
    ids.back().push_back(i);
    used += by;
  }
  int* host = b->list_stage + (size_t)3 * b->list_cap;
  int* dev = b->list_mem + (size_t)3 * b->list_cap;
  size_t off = 0;
  b->plain_groups = 0;
  for (auto& g : ids) {
    if (!b->pers[g[0]]) ++b->plain_groups;
    std::sort(g.begin(), g.end());
    b->groups.emplace_back();
    Lists& L = b->groups.back();
    build_lists(b, g, L);
    const size_t u = pack_lists(L, host + off, dev + off, b->list_cap - off);
    if (off + u > b->list_cap) return fail(PF_EINVAL, "pf_batch: group lists exceed their region");
    off += u;
  }
  if (off) PF_HIP(hipMemcpyAsync(dev, host, off * sizeof(int), hipMemcpyHostToDevice, b->st));
  return 0;
}

int tiles_launch(pf_batch* b, Lists& L, int what, int parity) {
  for (const List& t : L.tiles)
    if (int rc = batch_krylov_launch(b->dtype, what, parity, b->desc, t.view(), t.total, t.tz,
                                     t.ty, b->st, b->group_bytes == 0))
      return rc;
  return 0;
}

int launch_pre(pf_batch* b, Lists& L) {
  const int m = (int)L.ids.size();
  if (int rc = batch_ctl_launch(b->dtype, kBatchCtlInit, b->desc, L.ids_dev, m, nullptr, b->st))
    return rc;
  for (int s = 1; s <= 4; ++s)
    if (int rc = batch_stage_launch(b->dtype, s, b->desc, L.flat.view(), L.flat.total, b->st))
      return rc;
  return batch_krylov_launch(b->dtype, kBatchInit, 0, b->desc, L.red.view(), L.red.total, 0, 0,
                             b->st);
}

// The whole Poisson solve of box i as one persistent launch on the batch's
// stream (after its batched Krylov start); *used = false if it did not start.
int persist_start(pf_batch* b, int i, bool* used) {
  pf_box* h = b->boxes[i];
  const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
  if (b->dtype == PF_F64)
    return persist_solve<double>(h, b->st, tg, make_lap<double>(h), (const double*)h->r,
                                 (const double*)h->rh, (double*)h->x, h->nblk, used, nullptr);
  return persist_solve<float>(h, b->st, tg, make_lap<float>(h), (const float*)h->r,
                              (const float*)h->rh, (float*)h->x, h->nblk, used, nullptr);
}

int launch_post(pf_batch* b, Lists& L) {
  const int m = (int)L.ids.size();
  if (int rc = batch_krylov_launch(b->dtype, kBatchScatter, 0, b->desc, L.pad.view(),
                                   L.pad.total, 0, 0, b->st))
    return rc;
  if (int rc = batch_krylov_launch(b->dtype, kBatchProject, 0, b->desc, L.flat.view(),
                                   L.flat.total, 0, 0, b->st))
    return rc;
  if (int rc = batch_krylov_launch(b->dtype, kBatchStats, 0, b->desc, L.red.view(), L.red.total,
                                   0, 0, b->st))
    return rc;
  return batch_ctl_launch(b->dtype, kBatchFinalize, b->desc, L.ids_dev, m, nullptr, b->st);
}

int record_chunk(pf_batch* b, size_t ngroups) {
  for (size_t k = 0; k < ngroups; ++k) {
    Lists& L = b->groups[k];
    for (int q = 0; q < kChunk; ++q) {
      const int par = q & 1;
      if (int rc = tiles_launch(b, L, kBatchKB, par)) return rc;
      if (int rc = tiles_launch(b, L, kBatchKC, par)) return rc;
      if (int rc = tiles_launch(b, L, kBatchKA, 1 - par)) return rc;
    }
  }
  if (int rc = batch_ctl_launch(b->dtype, kBatchCtlFlags, b->desc, nullptr, b->n, b->flags_dev,
                                b->st))
    return rc;
  PF_HIP(hipMemcpyAsync(b->flags_host, b->flags_dev, sizeof(int) * b->n, hipMemcpyDeviceToHost,
                        b->st));
  return 0;
}

int capture_groups(pf_batch* b, size_t ngroups, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  PF_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = record_chunk(b, ngroups);
  hipError_t e = hipStreamEndCapture(b->st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  return 0;
}

// The chunk graph leaves out the groups of the boxes that solve
// persistently: their blocks would only read a done control block, and a
// launch over a 128^3 box's 1024 tiles costs microseconds even so.  The graph
// over every group serves the steps of such boxes whose cooperative launch
// could not start.
int capture_chunk(pf_batch* b) {
  if (b->g_chunk) hipGraphExecDestroy(b->g_chunk);
  if (b->g_chunk_all) hipGraphExecDestroy(b->g_chunk_all);
  b->g_chunk = b->g_chunk_all = nullptr;
  if (int rc = capture_groups(b, b->plain_groups, &b->g_chunk)) return rc;
  if (b->plain_groups < b->groups.size())
    return capture_groups(b, b->groups.size(), &b->g_chunk_all);
  return 0;
}

}  // namespace

extern "C" {

int pf_batch_create(pf_batch** out, int capacity, int dtype) {
  if (!out) return fail(PF_EINVAL, "pf_batch_create: NULL out");
  *out = nullptr;
  if (capacity < 1) return fail(PF_EINVAL, "pf_batch_create: capacity must be >= 1");
  if (dtype != PF_F64 && dtype != PF_F32) return fail(PF_EINVAL, "pf_batch_create: bad dtype");
  pf_batch* b = new pf_batch();
  b->dtype = dtype;
  b->capacity = capacity;
  b->n = 0;
  b->desc_size = dtype == PF_F64 ? sizeof(BoxDesc<double>) : sizeof(BoxDesc<float>);
  b->desc = nullptr;
  b->out_dev = nullptr;
  b->out_host = nullptr;
  b->flags_dev = nullptr;
  b->flags_host = nullptr;
  b->st = nullptr;
  b->ev = nullptr;
  b->wev = nullptr;
  b->list_mem = nullptr;
  b->list_stage = nullptr;
  b->list_cap = list_ints(capacity);
  b->g_chunk = nullptr;
  b->g_chunk_all = nullptr;
  b->plain_groups = 0;
  b->maxit_max = 0;
  b->group_bytes = kGroupBytes;
  b->persist_cap = 1;
  b->pend_seq = 0;
  b->done_vals = nullptr;
  hipError_t e = hipMalloc(&b->desc, b->desc_size * capacity);
  if (e == hipSuccess) e = hipMalloc((void**)&b->out_dev, sizeof(double) * 8 * capacity);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->out_host, sizeof(double) * 8 * capacity);
  if (e == hipSuccess) e = hipMalloc((void**)&b->flags_dev, sizeof(int) * capacity);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->flags_host, sizeof(int) * capacity);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->done_vals, sizeof(int) * 2);
  if (e == hipSuccess) {
    b->done_vals[0] = 0;
    b->done_vals[1] = 1;
  }
  // list regions: all boxes, boxes starting a step, boxes finishing one,
  // groups, boxes starting a step on the chunk kernels
  if (e == hipSuccess) e = hipMalloc((void**)&b->list_mem, sizeof(int) * 5 * b->list_cap);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->list_stage, sizeof(int) * 5 * b->list_cap);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->wev, hipEventDisableTiming);
  if (e != hipSuccess) {
    pf_batch_destroy(b);
    return check_hip(e, "pf_batch_create");
  }
  *out = b;
  return 0;
}

int pf_batch_destroy(pf_batch* b) {
  if (!b) return 0;
  if (b->st) hipStreamSynchronize(b->st);
  if (b->g_chunk) hipGraphExecDestroy(b->g_chunk);
  if (b->g_chunk_all) hipGraphExecDestroy(b->g_chunk_all);
  if (b->desc) hipFree(b->desc);
  if (b->out_dev) hipFree(b->out_dev);
  if (b->out_host) hipHostFree(b->out_host);
  if (b->flags_dev) hipFree(b->flags_dev);
  if (b->flags_host) hipHostFree(b->flags_host);
  if (b->done_vals) hipHostFree(b->done_vals);
  if (b->list_mem) hipFree(b->list_mem);
  if (b->list_stage) hipHostFree(b->list_stage);
  if (b->ev) hipEventDestroy(b->ev);
  if (b->wev) hipEventDestroy(b->wev);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
  return 0;
}

int pf_batch_set(pf_batch* b, int n, pf_box* const* boxes, void* const* fields,
                 void* const* sources, const double* params) {
  if (!b) return fail(PF_EINVAL, "pf_batch_set: NULL batch");
  if (n < 0 || n > b->capacity) return fail(PF_EINVAL, "pf_batch_set: n exceeds the capacity");
  if (n > 0 && (!boxes || !fields || !sources || !params))
    return fail(PF_EINVAL, "pf_batch_set: NULL array");
  PF_HIP(hipStreamSynchronize(b->st));
  // boxes keep their phase across compositions (a running solve continues)
  std::unordered_map<const pf_box*, char> was;
  std::unordered_map<const pf_box*, int64_t> wasp;
  std::unordered_map<const pf_box*, char> wasc;
  for (int i = 0; i < b->n; ++i) {
    was[b->boxes[i]] = b->running[i];
    wasp[b->boxes[i]] = b->pend[i];
    wasc[b->boxes[i]] = b->chunked[i];
  }
  std::vector<char> host(b->desc_size * (n > 0 ? n : 1));
  std::vector<pf_box*> bx(n);
  std::vector<char> run(n, 0);
  std::vector<int64_t> pnd(n, 0);
  std::vector<char> chk(n, 0);
  int maxit = 0;
  for (int i = 0; i < n; ++i) {
    pf_box* h = boxes[i];
    if (int rc = check_box(h)) return rc;
    if (h->dtype != b->dtype) return fail(PF_EINVAL, "pf_batch_set: box dtype differs from the batch");
    for (int j = 0; j < i; ++j)
      if (bx[j] == h) return fail(PF_EINVAL, "pf_batch_set: a box appears twice");
    void* const* f = fields + 8 * (size_t)i;
    for (int q = 0; q < 8; ++q)
      if (!f[q]) return fail(PF_EINVAL, "pf_batch_set: NULL field pointer");
    if (!sources[i]) return fail(PF_EINVAL, "pf_batch_set: NULL source pointer");
    const double* prm = params + 5 * (size_t)i;
    if (!(prm[2] != 0.0) || !(prm[0] != 0.0))
      return fail(PF_EINVAL, "pf_batch_set: dt and rho must be nonzero");
    double* out = b->out_dev + 8 * (size_t)i;
    if (b->dtype == PF_F64) {
      BoxDesc<double> d = make_desc<double>(h, f, sources[i], prm, out);
      memcpy(host.data() + i * b->desc_size, &d, sizeof(d));
    } else {
      BoxDesc<float> d = make_desc<float>(h, f, sources[i], prm, out);
      memcpy(host.data() + i * b->desc_size, &d, sizeof(d));
    }
    if ((int)prm[4] > maxit) maxit = (int)prm[4];
    bx[i] = h;
    auto it = was.find(h);
    if (it != was.end()) {
      run[i] = it->second;
      pnd[i] = wasp[h];
      chk[i] = wasc[h];
    } else {
      // a box joins idle: mark its control block done so that the chunk
      // kernels skip it until its first step starts
      const int one = 1;
      PF_HIP(hipMemcpy(&h->ctl->done, &one, sizeof(int), hipMemcpyHostToDevice));
    }
  }
  b->boxes = bx;
  b->running = run;
  b->pend = pnd;
  b->chunked = chk;
  b->pers.assign(n, 0);
  for (int i = 0; i < n; ++i)
    b->pers[i] = b->persist_cap > 0 && persist_blocks(bx[i]) >= kPersistMinBlocks;
  b->n = n;
  b->maxit_max = maxit;
  if (n > 0) PF_HIP(hipMemcpy(b->desc, host.data(), b->desc_size * n, hipMemcpyHostToDevice));
  std::vector<int> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  build_lists(b, all, b->full);
  if (int rc = upload_lists(b, b->full, 0)) return rc;
  if (int rc = build_groups(b)) return rc;
  PF_HIP(hipStreamSynchronize(b->st));
  if (n > 0) return capture_chunk(b);
  if (b->g_chunk) hipGraphExecDestroy(b->g_chunk);
  if (b->g_chunk_all) hipGraphExecDestroy(b->g_chunk_all);
  b->g_chunk = b->g_chunk_all = nullptr;
  return 0;
}

int pf_batch_advance(pf_batch* b, const int* start, int* finished, double* out, int* out_chunks,
                     void* stream) {
  if (!b) return fail(PF_EINVAL, "pf_batch_advance: NULL batch");
  if (out_chunks) *out_chunks = 0;
  for (int i = 0; i < b->n; ++i) {
    if (finished) finished[i] = 0;
    if (start && start[i] && b->running[i])
      return fail(PF_EINVAL, "pf_batch_advance: a box was started while inside a step");
  }
  if (b->n == 0) return 0;
  PF_HIP(hipEventRecord(b->ev, (hipStream_t)stream));
  PF_HIP(hipStreamWaitEvent(b->st, b->ev, 0));
  std::vector<int> ids;
  for (int i = 0; i < b->n; ++i)
    if (start && start[i]) ids.push_back(i);
  // hold (1) or release (0) a box's control block; from pinned memory, so
  // the copy does not wait for the work queued before it
  auto set_done = [&](int i, int v) -> int {
    PF_HIP(hipMemcpyAsync(&b->boxes[i]->ctl->done, b->done_vals + v, sizeof(int),
                          hipMemcpyHostToDevice, b->st));
    return 0;
  };
  // the cap lets the other boxes run between persistent solves; with no
  // other box in a step, every waiting and starting big box solves now
  bool others = false;
  for (int i = 0; i < b->n; ++i)
    others = others || (!b->pers[i] && (b->running[i] || (start && start[i])));
  const size_t cap = others ? (size_t)b->persist_cap : (size_t)b->n;
  std::vector<int> psolved;   // boxes whose solve ran persistently in this call
  std::vector<int> ka;        // boxes whose first KA the chunk kernels' start needs
  // waiting boxes that pf_batch_set_persistent has since moved to the chunk
  // kernels (with the cap at 0 the loop below would never release them)
  for (int i = 0; i < b->n; ++i) {
    if (!b->pend[i] || b->pers[i]) continue;
    b->pend[i] = 0;
    if (int rc = set_done(i, 0)) return rc;
    ka.push_back(i);
  }
  // the steps, started in earlier calls, whose persistent solves have waited
  // longest (first come, first served: a box that restarts behind the others
  // cannot overtake them)
  while (psolved.size() < cap) {
    int i = -1;
    for (int k = 0; k < b->n; ++k)
      if (b->pend[k] && (i < 0 || b->pend[k] < b->pend[i])) i = k;
    if (i < 0) break;
    b->pend[i] = 0;
    bool used = false;
    if (int rc = persist_start(b, i, &used)) return rc;
    if (used) {
      psolved.push_back(i);
    } else {   // onto the chunk kernels
      if (int rc = set_done(i, 0)) return rc;
      ka.push_back(i);
    }
  }
  if (!ids.empty()) {
    build_lists(b, ids, b->pre);
    if (int rc = upload_lists(b, b->pre, 1)) return rc;
    if (int rc = launch_pre(b, b->pre)) return rc;
    for (int i : ids) {
      if (b->pers[i]) {
        if (psolved.size() >= cap) {   // over the cap: this one waits
          if (int rc = set_done(i, 1)) return rc;
          b->pend[i] = ++b->pend_seq;
          continue;
        }
        bool used = false;
        if (int rc = persist_start(b, i, &used)) return rc;
        if (used) {
          psolved.push_back(i);
          continue;
        }
      }
      ka.push_back(i);
    }
    for (int i : ids) b->running[i] = 1;
  }
  if (!ka.empty()) {   // the first KA of the boxes on the chunk kernels
    for (int i : ka) b->chunked[i] = 1;
    build_lists(b, ka, b->ka);
    if (int rc = upload_lists(b, b->ka, 4)) return rc;
    if (int rc = tiles_launch(b, b->ka, kBatchKA, 0)) return rc;
  }
  bool any = false;
  for (int i = 0; i < b->n; ++i) any = any || b->running[i];
  if (!any) {
    PF_HIP(hipStreamSynchronize(b->st));
    return 0;
  }
  const int max_chunks = b->maxit_max / kChunk + 2;
  int chunks = 0;
  std::vector<int> fin;
  // with every running box solved persistently in this call or waiting, the
  // chunk kernels have nothing to do: read the done flags alone
  bool chunk_work = false;
  for (int i = 0; i < b->n; ++i)
    chunk_work = chunk_work || (b->running[i] && !b->pend[i] &&
                                std::find(psolved.begin(), psolved.end(), i) == psolved.end());
  // a big box whose cooperative launch did not start needs its group's launches
  bool all_groups = false;
  for (int i = 0; i < b->n; ++i) all_groups = all_groups || (b->pers[i] && b->chunked[i]);
  hipGraphExec_t graph = all_groups && b->g_chunk_all ? b->g_chunk_all : b->g_chunk;
  for (;;) {
    if (chunk_work) {
      PF_HIP(hipGraphLaunch(graph, b->st));
      ++chunks;
    } else {
      if (int rc = batch_ctl_launch(b->dtype, kBatchCtlFlags, b->desc, nullptr, b->n,
                                    b->flags_dev, b->st))
        return rc;
      PF_HIP(hipMemcpyAsync(b->flags_host, b->flags_dev, sizeof(int) * b->n,
                            hipMemcpyDeviceToHost, b->st));
      chunk_work = true;   // (a solve that did not finish goes on in the chunks)
    }
    if (int rc = wait_stream(b->st, b->wev)) return rc;
    for (int i = 0; i < b->n; ++i)
      if (b->running[i] && b->flags_host[i] && !b->pend[i]) fin.push_back(i);
    if (!fin.empty()) break;
    if (chunks > max_chunks)
      return fail(PF_EHIP, "pf_batch_advance: no solve finished within the iteration cap");
  }
  for (int i : psolved) {
    bool aborted = false;
    if (int rc = persist_aborted(b->boxes[i], b->st, &aborted)) return rc;
    if (aborted)
      return fail(PF_EHIP, "pf_batch_advance: a block of a persistent solve stopped waiting for "
                           "the others (grid sync timed out)");
  }
  build_lists(b, fin, b->post);
  if (int rc = upload_lists(b, b->post, 2)) return rc;
  if (int rc = launch_post(b, b->post)) return rc;
  PF_HIP(hipMemcpyAsync(b->out_host, b->out_dev, sizeof(double) * 8 * b->n,
                        hipMemcpyDeviceToHost, b->st));
  PF_HIP(hipStreamSynchronize(b->st));
  for (int i : fin) {
    b->running[i] = 0;
    b->chunked[i] = 0;
    if (finished) finished[i] = 1;
    if (out) memcpy(out + 8 * (size_t)i, b->out_host + 8 * (size_t)i, sizeof(double) * 8);
  }
  if (out_chunks) *out_chunks = chunks;
  return 0;
}

int pf_batch_chunk_iterations(void) { return kChunk; }

int pf_batch_set_persistent(pf_batch* b, int on) {
  if (!b) return fail(PF_EINVAL, "pf_batch_set_persistent: NULL batch");
  PF_HIP(hipStreamSynchronize(b->st));
  if (on < 0) return fail(PF_EINVAL, "pf_batch_set_persistent: on must be >= 0");
  b->persist_cap = on;
  int any = 0;
  for (int i = 0; i < b->n; ++i) {
    b->pers[i] = b->persist_cap > 0 && persist_blocks(b->boxes[i]) >= kPersistMinBlocks;
    any |= b->pers[i];
  }
  if (b->n > 0) {
    if (int rc = build_groups(b)) return rc;
    PF_HIP(hipStreamSynchronize(b->st));
    if (int rc = capture_chunk(b)) return rc;
  }
  return any;
}

int pf_batch_set_group_bytes(pf_batch* b, int64_t bytes) {
  if (!b) return fail(PF_EINVAL, "pf_batch_set_group_bytes: NULL batch");
  if (bytes < 0) return fail(PF_EINVAL, "pf_batch_set_group_bytes: bytes must be >= 0");
  PF_HIP(hipStreamSynchronize(b->st));
  b->group_bytes = (size_t)bytes;
  if (b->n == 0) return 0;
  if (int rc = build_groups(b)) return rc;
  PF_HIP(hipStreamSynchronize(b->st));
  return capture_chunk(b);
}

}  // extern "C"
