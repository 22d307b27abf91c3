// Box handle lifetime, error reporting and overlap fill.
#include <hip/hip_runtime.h>

#include <chrono>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "pf_kernels.h"

namespace pf {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int wait_stream(hipStream_t s, hipEvent_t ev) {
  if (!ev) {
    const hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : check_hip(e, "hipStreamSynchronize");
  }
  hipError_t e = hipEventRecord(ev, s);
  if (e != hipSuccess) return check_hip(e, "hipEventRecord");
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    e = hipEventQuery(ev);
    if (e != hipErrorNotReady) break;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      e = hipEventSynchronize(ev);
      break;
    }
    __builtin_ia32_pause();
  }
  return e == hipSuccess ? 0 : check_hip(e, "wait_stream");
}

int check_hip(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  set_error(m);
  return (e == hipErrorOutOfMemory) ? PF_ENOMEM : PF_EHIP;
}

int check_box(const pf_box* h) {
  if (!h) return fail(PF_EINVAL, "null box handle");
  return 0;
}

// ---------------------------------------------------------------------------
// Overlap fill: each interior cell writes its periodic images.  Only cells
// within two planes of a face have images, but a flat launch over the whole
// interior keeps the kernel trivial; the images are the only stores.
template <typename T>
__global__ void __launch_bounds__(kBlock) fill_halo_kernel(Geom g, T* __restrict__ a) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    bool bi = (i < kNover) || (i >= g.nx - kNover);
    bool bj = (j < kNover) || (j >= g.ny - kNover);
    bool bk = (k < kNover) || (k >= g.nz - kNover);
    if (!(bi || bj || bk)) continue;
    T v = a[pidx(g, i, j, k)];
    store_padded(g, a, i, j, k, v);
  }
}

int launch_blocks(const Geom& g) {
  int64_t b = (g.ncell + kBlock - 1) / kBlock;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_abi_version(void) { return PF_ABI_VERSION; }

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_box_create(pf_box** out, int nx, int ny, int nz, double dx, double dy, double dz,
                  int dtype) {
  if (!out) return fail(PF_EINVAL, "pf_box_create: out is NULL");
  *out = nullptr;
  if (nx < 4 || ny < 4 || nz < 4)
    return fail(PF_EINVAL, "pf_box_create: every dimension must be >= 4 (two overlap cells)");
  if ((int64_t)nx * ny * nz >= (1ll << 31))
    return fail(PF_EINVAL, "pf_box_create: nx*ny*nz must be < 2^31");
  if (dtype != PF_F64 && dtype != PF_F32) return fail(PF_EINVAL, "pf_box_create: bad dtype");
  if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(PF_EINVAL, "pf_box_create: spacings must be > 0");

  pf_box* h = new pf_box();
  memset(h, 0, sizeof(*h));
  h->srct_budget = -1;
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->dx = dx; h->dy = dy; h->dz = dz;
  h->dtype = dtype;
  h->esize = dtype == PF_F64 ? 8 : 4;
  h->g = make_geom(nx, ny, nz);
  hipGetDevice(&h->device);
  int64_t nb = (h->g.ncell + kBlock - 1) / kBlock;
  h->nblk = (int)(nb < 1024 ? nb : 1024);
  h->chunk = 16;
  h->mp_inner_tol = 1e-1;
  h->mp_inner_maxit = 500;
  h->mp_outer_maxit = 100;
  // 256-thread blocks; taller tiles come from two cells per thread (512-
  // thread blocks measured slower on the 128^3 solve)
  h->tile_rows = 1;
  {
    // the tiled Krylov kernels may use more blocks than the flat ones
    const int64_t tiles = make_tile_grid(h->g, (int)h->esize, h->tile_rows).nblk;
    const int64_t mx = tiles > 1024 ? tiles : 1024;   // >= every reducing grid
    h->part_stride = mx * kMaxPartSlots;
  }

  const size_t ni = (size_t)h->g.ncell * h->esize;
  const size_t np = (size_t)(nx + 4) * (ny + 4) * (nz + 4) * h->esize;
  auto alloc = [&](void** p, size_t n) -> int {
    hipError_t e = hipMalloc(p, n);
    if (e != hipSuccess) return check_hip(e, "hipMalloc");
    e = hipMemset(*p, 0, n);
    if (e != hipSuccess) return check_hip(e, "hipMemset");
    h->bytes += (int64_t)n;
    return 0;
  };
  int rc = 0;
  void** ivecs[] = {&h->x, &h->r, &h->rh, &h->p[0], &h->p[1], &h->v[0], &h->v[1], &h->s, &h->t,
                    &h->acc[0], &h->acc[1], &h->acc[2]};
  for (void** p : ivecs)
    if (!rc) rc = alloc(p, ni);
  for (int c = 0; c < 3 && !rc; ++c) rc = alloc(&h->ypong[c], np);
  if (!rc) rc = alloc((void**)&h->part, sizeof(double) * kNumPartBufs * h->part_stride);
  if (!rc) rc = alloc((void**)&h->ctl, sizeof(KrylovCtl));
  if (!rc) rc = alloc((void**)&h->stats_dev, sizeof(double) * 8);
  if (!rc) {
    const size_t pb = persist_bytes(h);
    if (pb) rc = alloc(&h->pers, pb);
    h->persist_on = 1;
  }
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&h->ctl_host, sizeof(KrylovCtl), hipHostMallocDefault);
    if (e != hipSuccess) rc = check_hip(e, "hipHostMalloc");
  }
  if (!rc) {
    hipError_t e = hipHostMalloc((void**)&h->stats_host, sizeof(double) * 8, hipHostMallocDefault);
    if (e != hipSuccess) rc = check_hip(e, "hipHostMalloc");
  }
  if (!rc) {
    hipError_t e = hipEventCreateWithFlags(&h->wait_ev, hipEventDisableTiming);
    if (e != hipSuccess) rc = check_hip(e, "hipEventCreateWithFlags");
  }
  if (rc) {
    pf_box_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int pf_box_destroy(pf_box* h) {
  if (!h) return 0;
  void* dev[] = {h->x, h->r, h->rh, h->p[0], h->p[1], h->v[0], h->v[1], h->s, h->t,
                 h->acc[0], h->acc[1], h->acc[2], h->ypong[0], h->ypong[1], h->ypong[2],
                 h->part, h->ctl, h->stats_dev, h->hist, h->mp_x, h->mp_b, h->mp_hist,
                 h->src_ws, h->srct_ws, h->pers};
  for (void* p : dev)
    if (p) hipFree(p);
  if (h->prof_ev) {
    for (int q = 0; q < 3 * h->prof_cap * 2; ++q) hipEventDestroy(h->prof_ev[q]);
    delete[] h->prof_ev;
    hipEventDestroy(h->prof_pev[0]);
    hipEventDestroy(h->prof_pev[1]);
  }
  if (h->ctl_host) hipHostFree(h->ctl_host);
  if (h->stats_host) hipHostFree(h->stats_host);
  if (h->wait_ev) hipEventDestroy(h->wait_ev);
  delete h;
  return 0;
}

int64_t pf_box_device_bytes(const pf_box* h) { return h ? h->bytes : 0; }

int pf_box_set_persistent(pf_box* h, int on) {
  if (int rc = check_box(h)) return rc;
  h->persist_on = on ? 1 : 0;
  return h->pers ? 1 : 0;
}

int pf_solve_last_persistent(const pf_box* h) {
  if (int rc = check_box(h)) return rc;
  return h->last_persist;
}

int pf_box_set_spacing(pf_box* h, double dx, double dy, double dz) {
  if (int rc = check_box(h)) return rc;
  if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(PF_EINVAL, "pf_box_set_spacing: spacings must be > 0");
  h->dx = dx; h->dy = dy; h->dz = dz;
  return 0;
}

int pf_fill_halo(pf_box* h, void* field, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!field) return fail(PF_EINVAL, "pf_fill_halo: field is NULL");
  hipStream_t s = (hipStream_t)stream;
  int nb = launch_blocks(h->g);
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(fill_halo_kernel<double>, dim3(nb), dim3(kBlock), 0, s, h->g, (double*)field);
  else
    hipLaunchKernelGGL(fill_halo_kernel<float>, dim3(nb), dim3(kBlock), 0, s, h->g, (float*)field);
  PF_LAUNCH_CHECK("fill_halo_kernel");
  return 0;
}

}  // extern "C"
