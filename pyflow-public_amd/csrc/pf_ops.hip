// Field-level operations of state_P (data/state.py:142-211) on their own:
// ZAXPY, update/copy and vel_corr, for callers that drive the advancer or the
// correction step piecewise instead of through the fused entry points.
#include <hip/hip_runtime.h>

#include "pf_box.h"

namespace pf {

struct Operand {
  const void* p;
  int layout;   // PF_LAYOUT_PADDED / PF_LAYOUT_INTERIOR
};

template <typename T>
__device__ __forceinline__ T ld(const Geom& g, const T* base, int layout, uint32_t idx, int i,
                                int j, int k) {
  return layout == PF_LAYOUT_PADDED ? base[pidx(g, i, j, k)] : base[idx];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) elementwise_kernel(Geom g, int op, T a, T* out,
                                                             const T* X, int xl, const T* Y,
                                                             int yl, int halo) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const T x = ld(g, X, xl, idx, i, j, k);
    T r;
    if (op == PF_EW_ZAXPY) {
      r = (a * x) + ld(g, Y, yl, idx, i, j, k);        // state.py:160-162: A*X + Y
    } else if (op == PF_EW_COPY) {
      r = x;                                            // state.py:168-172
    } else {
      r = fma(x, -a, out[pidx(g, i, j, k)]);            // state.py:205-208: sub_(s, grad)
    }
    if (halo)
      store_padded(g, out, i, j, k, r);
    else
      out[pidx(g, i, j, k)] = r;
  }
}

template <typename T>
int elementwise_impl(pf_box* h, hipStream_t s, int op, double a, void* out, const void* X, int xl,
                     const void* Y, int yl, int halo) {
  hipLaunchKernelGGL(elementwise_kernel<T>, dim3(launch_blocks(h->g)), dim3(kBlock), 0, s, h->g,
                     op, (T)a, (T*)out, (const T*)X, xl, (const T*)Y, yl, halo);
  PF_LAUNCH_CHECK("elementwise_kernel");
  return 0;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_state_op(pf_box* h, int op, double a, void* out, const void* X, int x_layout,
                           const void* Y, int y_layout, int fill_halo, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (op != PF_EW_ZAXPY && op != PF_EW_COPY && op != PF_EW_SUBSCALED)
    return fail(PF_EINVAL, "pf_state_op: unknown op");
  if (!out || !X || (op == PF_EW_ZAXPY && !Y)) return fail(PF_EINVAL, "pf_state_op: NULL operand");
  auto okl = [](int l) { return l == PF_LAYOUT_PADDED || l == PF_LAYOUT_INTERIOR; };
  if (!okl(x_layout) || (op == PF_EW_ZAXPY && !okl(y_layout)))
    return fail(PF_EINVAL, "pf_state_op: bad layout");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return elementwise_impl<double>(h, s, op, a, out, X, x_layout, Y, y_layout, fill_halo);
  return elementwise_impl<float>(h, s, op, a, out, X, x_layout, Y, y_layout, fill_halo);
}

// ---------------------------------------------------------------------------
// metric_uniform building blocks (metric/metric_staggered.py:112-526) with the
// reference's slice semantics over the whole padded array, overlap included.
namespace pf {

struct PBox {                 // padded sub-box [lo, hi) per axis
  int lo[3], hi[3];
  int n[3];                   // padded sizes
};

__device__ __forceinline__ bool pbox_cell(const PBox& b, int64_t q, int& a0, int& a1, int& a2) {
  const int e0 = b.hi[0] - b.lo[0], e1 = b.hi[1] - b.lo[1], e2 = b.hi[2] - b.lo[2];
  if (q >= (int64_t)e0 * e1 * e2) return false;
  a2 = (int)(q % e2) + b.lo[2];
  const int64_t t = q / e2;
  a1 = (int)(t % e1) + b.lo[1];
  a0 = (int)(t / e1) + b.lo[0];
  return true;
}

// dst[p] = coef*(src[p+o1] + src[p+o2])   (kind 0, interpolation)
// dst[p] = (src[p+o1] - src[p+o2])*coef   (kind 1, difference)
template <typename T>
__global__ void __launch_bounds__(kBlock) pair_kernel(PBox b, int kind, int64_t o1, int64_t o2,
                                                      T coef, T* __restrict__ dst,
                                                      const T* __restrict__ src) {
  const int64_t sx = (int64_t)b.n[1] * b.n[2], sy = b.n[2];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;; q += (int64_t)gridDim.x * blockDim.x) {
    int a0, a1, a2;
    if (!pbox_cell(b, q, a0, a1, a2)) break;
    const int64_t p = a0 * sx + a1 * sy + a2;
    dst[p] = kind == 0 ? coef * (src[p + o1] + src[p + o2]) : (src[p + o1] - src[p + o2]) * coef;
  }
}

// dst[p] = 0.5*0.5*(((S[p] + S[p-ea]) + S[p-eb]) + S[p-ea-eb])
template <typename T>
__global__ void __launch_bounds__(kBlock) edge4_kernel(PBox b, int64_t ea, int64_t eb,
                                                       T* __restrict__ dst, const T* __restrict__ S) {
  const int64_t sx = (int64_t)b.n[1] * b.n[2], sy = b.n[2];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;; q += (int64_t)gridDim.x * blockDim.x) {
    int a0, a1, a2;
    if (!pbox_cell(b, q, a0, a1, a2)) break;
    const int64_t p = a0 * sx + a1 * sy + a2;
    dst[p] = (T)0.25 * (((S[p] + S[p - ea]) + S[p - eb]) + S[p - ea - eb]);
  }
}

// div_visc (metric_staggered.py:406-418): six fused multiply-adds per cell.
template <typename T>
__global__ void __launch_bounds__(kBlock) div_visc_kernel(Geom g, const T* __restrict__ FX,
                                                          const T* __restrict__ FY,
                                                          const T* __restrict__ FZ,
                                                          T* __restrict__ rhs, T dxi, T dyi, T dzi) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    T r = rhs[idx];
    r = fma(FX[c], dxi, r);
    r = fma(FX[c - g.psx], -dxi, r);
    r = fma(FY[c], dyi, r);
    r = fma(FY[c - g.psy], -dyi, r);
    r = fma(FZ[c], dzi, r);
    r = fma(FZ[c - 1], -dzi, r);
    rhs[idx] = r;
  }
}

// div_vel_over (metric_staggered.py:424-433) over the whole padded array.
template <typename T>
__global__ void __launch_bounds__(kBlock) div_over_kernel(PBox b, const T* __restrict__ U,
                                                          const T* __restrict__ V,
                                                          const T* __restrict__ W,
                                                          T* __restrict__ D, T dxi, T dyi, T dzi) {
  const int64_t sx = (int64_t)b.n[1] * b.n[2], sy = b.n[2];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;; q += (int64_t)gridDim.x * blockDim.x) {
    int a0, a1, a2;
    if (!pbox_cell(b, q, a0, a1, a2)) break;
    const int64_t p = a0 * sx + a1 * sy + a2;
    T d = (T)0;
    if (a0 < b.n[0] - 1) d = (U[p + sx] - U[p]) * dxi;
    if (a1 < b.n[1] - 1) d = d + ((V[p + sy] - V[p]) * dyi);
    if (a2 < b.n[2] - 1) d = d + ((W[p + 1] - W[p]) * dzi);
    D[p] = d;
  }
}

// vel_conv_* (metric_staggered.py:458-526): rhs -= ((A1*B1 - A0*B0)*h)*sign
template <typename T>
__global__ void __launch_bounds__(kBlock) vel_conv_kernel(Geom g, int64_t o1, int64_t o0,
                                                          const T* __restrict__ A,
                                                          const T* __restrict__ B,
                                                          T* __restrict__ rhs, T h, T sign) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    const T t = ((A[c + o1] * B[c + o1] - A[c + o0] * B[c + o0]) * h) * sign;
    rhs[idx] = rhs[idx] - t;
  }
}

// adj_conv_* (metric_staggered.py:529-800): for s in {0, s2}
//   rhs -= ((a[m - d + s] - a[m + s]) * vi[m + s + vsh]) * coef
template <typename T>
__global__ void __launch_bounds__(kBlock) adj_conv_kernel(Geom g, int64_t d, int64_t s2,
                                                          int64_t vsh, const T* __restrict__ a,
                                                          const T* __restrict__ vi,
                                                          T* __restrict__ rhs, T coef) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    T r = rhs[idx];
    r = r - (((a[c - d] - a[c]) * vi[c + vsh]) * coef);
    r = r - (((a[c - d + s2] - a[c + s2]) * vi[c + s2 + vsh]) * coef);
    rhs[idx] = r;
  }
}

static PBox full_box(const pf_box* h) {
  PBox b;
  b.n[0] = h->nx + 4; b.n[1] = h->ny + 4; b.n[2] = h->nz + 4;
  for (int a = 0; a < 3; ++a) { b.lo[a] = 0; b.hi[a] = b.n[a]; }
  return b;
}

static int grid_for(int64_t n) {
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  return (int)(nb < 1 ? 1 : nb);
}

static int64_t axis_stride(const PBox& b, int axis) {
  return axis == 0 ? (int64_t)b.n[1] * b.n[2] : (axis == 1 ? b.n[2] : 1);
}

template <typename T>
int pair_impl(pf_box* h, hipStream_t s, int kind, int axis, int upper, void* dst, const void* src,
              double coef) {
  PBox b = full_box(h);
  const int64_t e = axis_stride(b, axis);
  int64_t o1, o2;
  if (upper) {            // box [0, n-1): reads p+e and p
    b.hi[axis] = b.n[axis] - 1;
    o1 = e; o2 = 0;
  } else {                // box [1, n): reads p and p-e
    b.lo[axis] = 1;
    o1 = 0; o2 = -e;
  }
  const T* in = (const T*)src;
  if (dst == src) {       // in-place (interp_uvwi_*): read from a copy, like torch's temporary
    const size_t bytes = (size_t)b.n[0] * b.n[1] * b.n[2] * sizeof(T);
    PF_HIP(hipMemcpyAsync(h->ypong[0], src, bytes, hipMemcpyDeviceToDevice, s));
    in = (const T*)h->ypong[0];
  }
  const int64_t n = (int64_t)(b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
  hipLaunchKernelGGL(pair_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, b, kind, o1, o2,
                     (T)coef, (T*)dst, in);
  PF_LAUNCH_CHECK("pair_kernel");
  return 0;
}

template <typename T>
int edge4_impl(pf_box* h, hipStream_t s, int plane, void* dst, const void* src) {
  PBox b = full_box(h);
  const int ax[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  const int a = ax[plane][0], c = ax[plane][1];
  b.lo[a] = 1;
  b.lo[c] = 1;
  const int64_t n = (int64_t)(b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
  const T* in = (const T*)src;
  if (dst == src) {
    const size_t bytes = (size_t)b.n[0] * b.n[1] * b.n[2] * sizeof(T);
    PF_HIP(hipMemcpyAsync(h->ypong[0], src, bytes, hipMemcpyDeviceToDevice, s));
    in = (const T*)h->ypong[0];
  }
  hipLaunchKernelGGL(edge4_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, b,
                     axis_stride(b, a), axis_stride(b, c), (T*)dst, in);
  PF_LAUNCH_CHECK("edge4_kernel");
  return 0;
}

}  // namespace pf

extern "C" {

int pf_metric_pair(pf_box* h, int kind, int axis, int upper, void* dst, const void* src,
                   double coef, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!dst || !src) return fail(PF_EINVAL, "pf_metric_pair: NULL field");
  if (kind != PF_PAIR_INTERP && kind != PF_PAIR_DIFF) return fail(PF_EINVAL, "pf_metric_pair: bad kind");
  if (axis < 0 || axis > 2) return fail(PF_EINVAL, "pf_metric_pair: bad axis");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? pair_impl<double>(h, s, kind, axis, upper, dst, src, coef)
                            : pair_impl<float>(h, s, kind, axis, upper, dst, src, coef);
}

int pf_metric_edge4(pf_box* h, int plane, void* dst, const void* src, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!dst || !src) return fail(PF_EINVAL, "pf_metric_edge4: NULL field");
  if (plane < 0 || plane > 2) return fail(PF_EINVAL, "pf_metric_edge4: plane is 0 (xy), 1 (xz) or 2 (yz)");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? edge4_impl<double>(h, s, plane, dst, src)
                            : edge4_impl<float>(h, s, plane, dst, src);
}

int pf_div_visc(pf_box* h, const void* FX, const void* FY, const void* FZ, void* rhs, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!FX || !FY || !FZ || !rhs) return fail(PF_EINVAL, "pf_div_visc: NULL field");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(div_visc_kernel<double>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g,
                       (const double*)FX, (const double*)FY, (const double*)FZ, (double*)rhs,
                       1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz);
  else
    hipLaunchKernelGGL(div_visc_kernel<float>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g,
                       (const float*)FX, (const float*)FY, (const float*)FZ, (float*)rhs,
                       (float)(1.0 / h->dx), (float)(1.0 / h->dy), (float)(1.0 / h->dz));
  PF_LAUNCH_CHECK("div_visc_kernel");
  return 0;
}

int pf_div_vel_over(pf_box* h, const void* u, const void* v, const void* w, void* divg_o,
                    void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !divg_o) return fail(PF_EINVAL, "pf_div_vel_over: NULL field");
  hipStream_t s = (hipStream_t)stream;
  PBox b = full_box(h);
  const int64_t n = (int64_t)b.n[0] * b.n[1] * b.n[2];
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(div_over_kernel<double>, dim3(grid_for(n)), dim3(kBlock), 0, s, b,
                       (const double*)u, (const double*)v, (const double*)w, (double*)divg_o,
                       1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz);
  else
    hipLaunchKernelGGL(div_over_kernel<float>, dim3(grid_for(n)), dim3(kBlock), 0, s, b,
                       (const float*)u, (const float*)v, (const float*)w, (float*)divg_o,
                       (float)(1.0 / h->dx), (float)(1.0 / h->dy), (float)(1.0 / h->dz));
  PF_LAUNCH_CHECK("div_over_kernel");
  return 0;
}

int pf_vel_conv(pf_box* h, int axis, int diagonal, const void* A, const void* B, void* rhs,
                double sign, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!A || !B || !rhs) return fail(PF_EINVAL, "pf_vel_conv: NULL field");
  if (axis < 0 || axis > 2) return fail(PF_EINVAL, "pf_vel_conv: bad axis");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  const int64_t e = axis == 0 ? g.psx : (axis == 1 ? g.psy : 1);
  // xx/yy/zz read cells i and i-1, the off-diagonal fluxes i+1 and i
  const int64_t o1 = diagonal ? 0 : e, o0 = diagonal ? -e : 0;
  const double hinv = axis == 0 ? 1.0 / h->dx : (axis == 1 ? 1.0 / h->dy : 1.0 / h->dz);
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(vel_conv_kernel<double>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g, o1,
                       o0, (const double*)A, (const double*)B, (double*)rhs,
                       hinv, sign);
  else
    hipLaunchKernelGGL(vel_conv_kernel<float>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g, o1,
                       o0, (const float*)A, (const float*)B, (float*)rhs,
                       (float)hinv, (float)sign);
  PF_LAUNCH_CHECK("vel_conv_kernel");
  return 0;
}

int pf_adj_conv(pf_box* h, int d_axis, int s2_axis, int centre, const void* a, const void* vi,
                void* rhs, double coef, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!a || !vi || !rhs) return fail(PF_EINVAL, "pf_adj_conv: NULL field");
  if (d_axis < 0 || d_axis > 2 || s2_axis < 0 || s2_axis > 2)
    return fail(PF_EINVAL, "pf_adj_conv: bad axis");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  const int64_t st[3] = {g.psx, g.psy, 1};
  const int64_t d = st[d_axis], s2 = st[s2_axis], vsh = centre ? -d : 0;
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(adj_conv_kernel<double>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g, d,
                       s2, vsh, (const double*)a, (const double*)vi, (double*)rhs, coef);
  else
    hipLaunchKernelGGL(adj_conv_kernel<float>, dim3(launch_blocks(g)), dim3(kBlock), 0, s, g, d,
                       s2, vsh, (const float*)a, (const float*)vi, (float*)rhs, (float)coef);
  PF_LAUNCH_CHECK("adj_conv_kernel");
  return 0;
}

}  // extern "C"
