// constraints.hip — the object hard constraints of the reconstruction loop and their gradient.
//
// ObjectConstraints.apply_hard_constraints for obj_type "complex" / "pure_phase"
// (quantem/diffractive_imaging/object_models.py:289-301):
//   O' = a e^{i phi},  a = clamp(|O|, 0, 1) (or 1),  phi = theta - mean(theta),  theta = angle(O).
// Gradient (torch convention g = dL/dRe + i dL/dIm, g' that of O'):
//   dL/da = Re(conj(g') e^{i phi}),  dL/dphi = a (Im g' cos phi - Re g' sin phi),
//   dL/dtheta_j = dL/dphi_j - mean_k dL/dphi_k,  dL/dr = dL/da [r <= 1] (0 for pure phase),
//   g = (O / r) (dL/dr + i dL/dtheta / r),  0 where O = 0 (torch's abs / angle backward there).
// Each pass is two launches: fixed-size block partial sums (double), then the elementwise kernel,
// whose blocks add the partials in index order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.h"
#include "ptycho_common.h"
#include "qem_constraints.h"

namespace qem {
namespace {

constexpr int kT = 256;
constexpr int kMaxParts = 1024;
constexpr int kPerThread = 8;
// workspace: part[kMaxParts] (double), then the mean phase, then the sum of dL/dphi
constexpr size_t kWsBytes = (kMaxParts + 2) * sizeof(double);

int parts_for(int64_t n) {
  const int64_t p = (n + (int64_t)kT * kPerThread - 1) / ((int64_t)kT * kPerThread);
  return p < 1 ? 1 : (p > kMaxParts ? kMaxParts : (int)p);
}

__device__ __forceinline__ void chunk_of(int64_t n, int parts, int b, int64_t& lo, int64_t& hi) {
  const int64_t c = (n + parts - 1) / parts;
  lo = (int64_t)b * c;
  hi = lo + c < n ? lo + c : n;
}

struct Polar {
  float r, a, c, s;  // |O|, amplitude after the constraint, cos / sin of the constrained phase
};

__device__ __forceinline__ Polar polar(cf o, float mean, int pure) {
  Polar q;
  q.r = hypotf(o.x, o.y);
  q.a = pure ? 1.0f : fminf(q.r, 1.0f);
  sincosf(atan2f(o.y, o.x) - mean, &q.s, &q.c);
  return q;
}

// Sum of angle(O) over one chunk per block.
__global__ __launch_bounds__(kT) void k_angle_part(int64_t n, const cf* __restrict__ O, double* __restrict__ ws) {
  __shared__ double red[kT / 64];
  int64_t lo, hi;
  chunk_of(n, gridDim.x, blockIdx.x, lo, hi);
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kT) {
    const cf o = O[i];
    acc += (double)atan2f(o.y, o.x);
  }
  const double t = block_sum<kT>(acc, red);
  if (threadIdx.x == 0) ws[blockIdx.x] = t;
}

__global__ __launch_bounds__(kT) void k_cons_forward(int64_t n, int pure, int parts, const cf* __restrict__ O,
                                                       cf* __restrict__ out, double* __restrict__ ws) {
  __shared__ float s_mean;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < parts; ++i) t += ws[i];
    const double mean = t / (double)n;
    s_mean = (float)mean;
    if (blockIdx.x == 0) ws[kMaxParts] = mean;  // read by the backward pass
  }
  __syncthreads();
  const float mean = s_mean;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const Polar q = polar(O[i], mean, pure);
    out[i] = mk(q.a * q.c, q.a * q.s);
  }
}

// Sum of dL/dphi over one chunk per block.
__global__ __launch_bounds__(kT) void k_dphi_part(int64_t n, int pure, const cf* __restrict__ O,
                                                    const cf* __restrict__ g2, double* __restrict__ ws) {
  __shared__ double red[kT / 64];
  const float mean = (float)ws[kMaxParts];
  int64_t lo, hi;
  chunk_of(n, gridDim.x, blockIdx.x, lo, hi);
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kT) {
    const Polar q = polar(O[i], mean, pure);
    const cf g = g2[i];
    acc += (double)(q.a * (g.y * q.c - g.x * q.s));
  }
  const double t = block_sum<kT>(acc, red);  // partial slots (< kMaxParts) never alias the mean's
  if (threadIdx.x == 0) ws[blockIdx.x] = t;
}

__global__ __launch_bounds__(kT) void k_cons_backward(int64_t n, int pure, int parts, const cf* __restrict__ O,
                                                        const cf* __restrict__ g2, cf* __restrict__ gout,
                                                        const double* __restrict__ ws) {
  __shared__ float s_dmean;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < parts; ++i) t += ws[i];
    s_dmean = (float)(t / (double)n);
  }
  __syncthreads();
  const float dmean = s_dmean;
  const float mean = (float)ws[kMaxParts];
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const cf o = O[i];
    const Polar q = polar(o, mean, pure);
    const cf g = g2[i];
    const float dphi = q.a * (g.y * q.c - g.x * q.s);
    const float dth = dphi - dmean;
    const float dr = (pure || q.r > 1.0f) ? 0.0f : g.x * q.c + g.y * q.s;
    cf res = mk(0.0f, 0.0f);
    if (q.r > 0.0f) {
      const float inv_r = 1.0f / q.r;
      const float u = dr, w = dth * inv_r;  // (O / r) (u + i w)
      const float ex = o.x * inv_r, ey = o.y * inv_r;
      res = mk(ex * u - ey * w, ex * w + ey * u);
    }
    gout[i] = res;
  }
}

int check_ptrs(const void* a, const char* an, const void* b, const char* bn, const void* c, const char* cn,
               const void* ws, size_t ws_bytes) {
  const void* p[3] = {a, b, c};
  const char* nm[3] = {an, bn, cn};
  for (int i = 0; i < 3; ++i) {
    if (p[i] == nullptr) return set_error(QEM_ERR_NULL, "%s is NULL", nm[i]);
    if (((uintptr_t)p[i]) & 7u) return set_error(QEM_ERR_ALIGN, "%s is misaligned", nm[i]);
  }
  if (ws == nullptr) return set_error(QEM_ERR_NULL, "workspace is NULL");
  if (((uintptr_t)ws) & 7u) return set_error(QEM_ERR_ALIGN, "workspace is misaligned");
  if (ws_bytes < kWsBytes) return set_error(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", ws_bytes, kWsBytes);
  return QEM_OK;
}

int grid_for(int64_t n) {
  const int64_t b = (n + kT - 1) / kT;
  return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace
}  // namespace qem

extern "C" {

int qem_object_constraints_workspace_bytes(int64_t n, size_t* bytes) {
  qem::clear_error();
  if (n < 1) return qem::set_error(QEM_ERR_ARG, "n must be >= 1 (got %lld)", (long long)n);
  if (bytes == nullptr) return qem::set_error(QEM_ERR_NULL, "bytes is NULL");
  *bytes = qem::kWsBytes;
  return QEM_OK;
}

int qem_object_constraints_forward(int64_t n, int32_t pure_phase, const void* obj, void* out, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (n < 1) return qem::set_error(QEM_ERR_ARG, "n must be >= 1 (got %lld)", (long long)n);
  if (int rc = qem::check_ptrs(obj, "obj", out, "out", out, "out", workspace, workspace_bytes)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int parts = qem::parts_for(n);
  double* ws = (double*)workspace;
  hipLaunchKernelGGL(qem::k_angle_part, dim3(parts), dim3(qem::kT), 0, st, n, (const qem::cf*)obj, ws);
  if (int rc = qem::check_last_launch("angle_part")) return rc;
  hipLaunchKernelGGL(qem::k_cons_forward, dim3(qem::grid_for(n)), dim3(qem::kT), 0, st, n, (int)(pure_phase != 0),
                     parts, (const qem::cf*)obj, (qem::cf*)out, ws);
  return qem::check_last_launch("constraints_forward");
}

int qem_object_constraints_backward(int64_t n, int32_t pure_phase, const void* obj, const void* grad_out,
                                    void* grad_obj, void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (n < 1) return qem::set_error(QEM_ERR_ARG, "n must be >= 1 (got %lld)", (long long)n);
  if (int rc = qem::check_ptrs(obj, "obj", grad_out, "grad_out", grad_obj, "grad_obj", workspace, workspace_bytes))
    return rc;
  hipStream_t st = (hipStream_t)stream;
  const int parts = qem::parts_for(n);
  double* ws = (double*)workspace;
  hipLaunchKernelGGL(qem::k_dphi_part, dim3(parts), dim3(qem::kT), 0, st, n, (int)(pure_phase != 0),
                     (const qem::cf*)obj, (const qem::cf*)grad_out, ws);
  if (int rc = qem::check_last_launch("dphi_part")) return rc;
  hipLaunchKernelGGL(qem::k_cons_backward, dim3(qem::grid_for(n)), dim3(qem::kT), 0, st, n,
                     (int)(pure_phase != 0), parts, (const qem::cf*)obj, (const qem::cf*)grad_out,
                     (qem::cf*)grad_obj, (const double*)ws);
  return qem::check_last_launch("constraints_backward");
}

}  // extern "C"
