// Batched time steps of many independent boxes (the serving path).
//
// A pf_batch holds the descriptors of up to `capacity` resident boxes and
// advances all of them by one time step per pf_batch_step call: the RK4
// stages, the Poisson source and Krylov start, BiCGStab iterations until
// every box's solve is done, then the correction, the pressure update and
// the monitor values.  Every kernel launch covers all boxes at once; each
// box keeps the block decomposition of its single-box launch, so its
// results are bit-identical alone and in any mix, and the same as the
// single-box entry points'.
//
// The launches of one composition are captured once into three hipGraphs
// (before the solve; one chunk of kChunk BiCGStab iterations ending with a
// count of finished solves copied to pinned host memory; after the solve)
// and replayed every step.  Boxes whose solve is done skip the rest of the
// chunk on the device (their control block says so), so the graphs do not
// depend on iteration counts.  Changing the composition (pf_batch_set)
// re-captures them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "pf_kernels.h"

namespace {

constexpr int kChunk = 16;    // BiCGStab iterations per graph replay (even: buffer parity)

struct List {
  std::vector<int> box, begin;
  int total = 0;
  int tz = 0, ty = 0;
  int* dev = nullptr;        // box ids then begins, device copy
  pf::BatchList view() const {
    pf::BatchList L;
    L.box = dev;
    L.begin = dev + box.size();
    L.n = (int)box.size();
    return L;
  }
};

}  // namespace

struct pf_batch {
  int dtype;
  int capacity;
  int n;
  size_t desc_size;
  void* desc;                 // device BoxDesc<T>[capacity]
  double* out_dev;            // 8 doubles per box
  double* out_host;           // pinned
  int* done_dev;
  int* done_host;             // pinned
  hipStream_t st;
  hipEvent_t ev;
  List flat, red, pad;
  std::vector<List> tiles;
  int* list_mem;
  size_t list_cap;
  hipGraphExec_t g_pre, g_chunk, g_post;
  bool graphs;
  int maxit_max;
  int last_chunks;            // chunk replays the previous step needed
};

using namespace pf;

namespace {

template <typename T>
BoxDesc<T> make_desc(pf_box* h, void* const* f, void* src, const double* prm, double* out) {
  // f: u, v, w, ut, vt, wt, DP, p ; prm: dt, mu, rho, tol, maxit
  BoxDesc<T> d;
  d.g = h->g;
  d.tg = make_tile_grid(h->g, h->tile_rows);
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
  d.r = (T*)h->r;
  d.rh = (T*)h->rh;
  d.pk[0] = (T*)h->p[0];
  d.pk[1] = (T*)h->p[1];
  d.vk[0] = (T*)h->v[0];
  d.vk[1] = (T*)h->v[1];
  d.s = (T*)h->s;
  d.t = (T*)h->t;
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

void destroy_graphs(pf_batch* b) {
  if (b->g_pre) hipGraphExecDestroy(b->g_pre);
  if (b->g_chunk) hipGraphExecDestroy(b->g_chunk);
  if (b->g_post) hipGraphExecDestroy(b->g_post);
  b->g_pre = b->g_chunk = b->g_post = nullptr;
  b->graphs = false;
}

std::vector<List*> all_lists(pf_batch* b) {
  std::vector<List*> v = {&b->flat, &b->red, &b->pad};
  for (auto& t : b->tiles) v.push_back(&t);
  return v;
}

int upload_lists(pf_batch* b) {
  size_t need = 0;
  for (List* L : all_lists(b)) need += L->box.size() * 2 + 1;
  if (need > b->list_cap) {
    if (b->list_mem) PF_HIP(hipFree(b->list_mem));
    b->list_mem = nullptr;
    b->list_cap = 0;
    PF_HIP(hipMalloc((void**)&b->list_mem, need * sizeof(int)));
    b->list_cap = need;
  }
  std::vector<int> host;
  host.reserve(need);
  size_t off = 0;
  for (List* L : all_lists(b)) {
    L->dev = b->list_mem + off;
    host.insert(host.end(), L->box.begin(), L->box.end());
    host.insert(host.end(), L->begin.begin(), L->begin.end());
    host.push_back(L->total);
    off += L->box.size() * 2 + 1;
  }
  if (!host.empty())
    PF_HIP(hipMemcpy(b->list_mem, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int tiles_launch(pf_batch* b, int what, int parity) {
  for (const List& t : b->tiles)
    if (int rc = batch_krylov_launch(b->dtype, what, parity, b->desc, t.view(), t.total, t.tz,
                                     t.ty, b->st))
      return rc;
  return 0;
}

int record_pre(pf_batch* b) {
  if (int rc = batch_ctl_launch(b->dtype, kBatchCtlInit, b->desc, b->n, nullptr, b->st)) return rc;
  for (int s = 1; s <= 4; ++s)
    if (int rc = batch_stage_launch(b->dtype, s, b->desc, b->flat.view(), b->flat.total, b->st))
      return rc;
  if (int rc = batch_krylov_launch(b->dtype, kBatchInit, 0, b->desc, b->red.view(), b->red.total,
                                   0, 0, b->st))
    return rc;
  return tiles_launch(b, kBatchKA, 0);
}

int record_chunk(pf_batch* b) {
  for (int q = 0; q < kChunk; ++q) {
    const int par = q & 1;
    if (int rc = tiles_launch(b, kBatchKB, par)) return rc;
    if (int rc = batch_krylov_launch(b->dtype, kBatchKC, par, b->desc, b->red.view(),
                                     b->red.total, 0, 0, b->st))
      return rc;
    if (int rc = tiles_launch(b, kBatchKA, 1 - par)) return rc;
  }
  if (int rc = batch_ctl_launch(b->dtype, kBatchCtlCount, b->desc, b->n, b->done_dev, b->st))
    return rc;
  PF_HIP(hipMemcpyAsync(b->done_host, b->done_dev, sizeof(int), hipMemcpyDeviceToHost, b->st));
  return 0;
}

int record_post(pf_batch* b) {
  if (int rc = batch_krylov_launch(b->dtype, kBatchScatter, 0, b->desc, b->pad.view(),
                                   b->pad.total, 0, 0, b->st))
    return rc;
  if (int rc = batch_krylov_launch(b->dtype, kBatchProject, 0, b->desc, b->flat.view(),
                                   b->flat.total, 0, 0, b->st))
    return rc;
  if (int rc = batch_krylov_launch(b->dtype, kBatchStats, 0, b->desc, b->red.view(), b->red.total,
                                   0, 0, b->st))
    return rc;
  if (int rc = batch_ctl_launch(b->dtype, kBatchFinalize, b->desc, b->n, nullptr, b->st))
    return rc;
  PF_HIP(hipMemcpyAsync(b->out_host, b->out_dev, sizeof(double) * 8 * b->n,
                        hipMemcpyDeviceToHost, b->st));
  return 0;
}

int capture(pf_batch* b, int (*rec)(pf_batch*), hipGraphExec_t* exec) {
  hipGraph_t g = nullptr;
  PF_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int rc = rec(b);
  hipError_t e = hipStreamEndCapture(b->st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return check_hip(e, "hipGraphInstantiate");
  return 0;
}

int build_graphs(pf_batch* b) {
  destroy_graphs(b);
  if (int rc = capture(b, record_pre, &b->g_pre)) return rc;
  if (int rc = capture(b, record_chunk, &b->g_chunk)) return rc;
  if (int rc = capture(b, record_post, &b->g_post)) return rc;
  b->graphs = true;
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
  b->done_dev = nullptr;
  b->done_host = nullptr;
  b->st = nullptr;
  b->ev = nullptr;
  b->list_mem = nullptr;
  b->list_cap = 0;
  b->g_pre = b->g_chunk = b->g_post = nullptr;
  b->graphs = false;
  b->maxit_max = 0;
  b->last_chunks = 1;
  hipError_t e = hipMalloc(&b->desc, b->desc_size * capacity);
  if (e == hipSuccess) e = hipMalloc((void**)&b->out_dev, sizeof(double) * 8 * capacity);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->out_host, sizeof(double) * 8 * capacity);
  if (e == hipSuccess) e = hipMalloc((void**)&b->done_dev, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->done_host, sizeof(int));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
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
  destroy_graphs(b);
  if (b->desc) hipFree(b->desc);
  if (b->out_dev) hipFree(b->out_dev);
  if (b->out_host) hipHostFree(b->out_host);
  if (b->done_dev) hipFree(b->done_dev);
  if (b->done_host) hipHostFree(b->done_host);
  if (b->list_mem) hipFree(b->list_mem);
  if (b->ev) hipEventDestroy(b->ev);
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
  destroy_graphs(b);
  b->flat = List();
  b->red = List();
  b->pad = List();
  b->tiles.clear();
  b->maxit_max = 0;
  std::vector<char> host(b->desc_size * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    pf_box* h = boxes[i];
    if (int rc = check_box(h)) return rc;
    if (h->dtype != b->dtype) return fail(PF_EINVAL, "pf_batch_set: box dtype differs from the batch");
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
    if ((int)prm[4] > b->maxit_max) b->maxit_max = (int)prm[4];
    add(b->flat, i, launch_blocks(h->g));
    add(b->red, i, h->nblk);
    const int64_t np = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
    int64_t npb = (np + kBlock - 1) / kBlock;
    add(b->pad, i, (int)(npb > 65536 ? 65536 : npb));
    const TileGrid tg = make_tile_grid(h->g, h->tile_rows);
    List* tl = nullptr;
    for (auto& t : b->tiles)
      if (t.tz == tg.tz && t.ty == tg.ty) tl = &t;
    if (!tl) {
      b->tiles.push_back(List());
      tl = &b->tiles.back();
      tl->tz = tg.tz;
      tl->ty = tg.ty;
    }
    add(*tl, i, tg.nblk);
  }
  b->n = n;
  if (n > 0)
    PF_HIP(hipMemcpy(b->desc, host.data(), b->desc_size * n, hipMemcpyHostToDevice));
  return upload_lists(b);
}

int pf_batch_step(pf_batch* b, double* out, int* out_chunks, void* stream) {
  if (!b) return fail(PF_EINVAL, "pf_batch_step: NULL batch");
  if (b->n == 0) return 0;
  if (!b->graphs)
    if (int rc = build_graphs(b)) return rc;
  // order after the caller's work on its stream
  PF_HIP(hipEventRecord(b->ev, (hipStream_t)stream));
  PF_HIP(hipStreamWaitEvent(b->st, b->ev, 0));
  PF_HIP(hipGraphLaunch(b->g_pre, b->st));
  const int max_chunks = b->maxit_max / kChunk + 2;
  int chunks = 0;
  // queue the chunks the last step needed, less one, before the first poll
  int ahead = b->last_chunks - 1;
  if (ahead < 1) ahead = 1;
  for (;;) {
    for (int q = 0; q < ahead; ++q) PF_HIP(hipGraphLaunch(b->g_chunk, b->st));
    chunks += ahead;
    ahead = 1;
    PF_HIP(hipStreamSynchronize(b->st));
    if (*b->done_host >= b->n) break;
    if (chunks > max_chunks)
      return fail(PF_EHIP, "pf_batch_step: solves did not finish within the iteration cap");
  }
  b->last_chunks = chunks;
  PF_HIP(hipGraphLaunch(b->g_post, b->st));
  PF_HIP(hipStreamSynchronize(b->st));
  if (out) memcpy(out, b->out_host, sizeof(double) * 8 * b->n);
  if (out_chunks) *out_chunks = chunks;
  return 0;
}

int pf_batch_chunk_iterations(void) { return kChunk; }

}  // extern "C"
