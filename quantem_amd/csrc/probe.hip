// probe.hip — orthogonalisation of the probe modes (the probe's hard constraint) and its gradient.
//
// ProbeConstraints._probe_orthogonalization_constraint (quantem/diffractive_imaging/probe_models.py:
// 484-525): modified Gram-Schmidt over the M modes, each result rescaled to its input mode's norm, then
// the modes sorted by intensity (strongest first).  For full-rank input that is the thin QR
// factorisation A = Q R (A: pixels x M, R upper triangular with positive real diagonal) followed by
// out = Q diag(n), n_i = ||a_i||.  Here R comes from the Cholesky factor of the Gram matrix
// G = A^H A (accumulated in double), so every pass over the N^2 x M data is a parallel one:
//
//   forward   k_cross   partial sums of A^H A per block (fixed pixel order, double)
//             k_small   G = sum of partials (block order), R = chol(G), T = R^-1, sort, coefficients
//             k_apply   out_k = sum_j a_j C_jk                          (C = T diag(n), sorted columns)
//   backward  k_cross   partial sums of A^H Obar (Obar: gradient of the sorted output, unsorted)
//             k_small   W = T^H A^H Obar, nbar_i = Re W_ii, B = W diag(n), H = B's Hermitian part
//                       built from its upper triangle with real diagonal, then
//                       E = diag(n) T^H,  F = -T H T^H + diag(nbar / n)
//             k_apply   Abar = Obar E + A F
//
// The backward is the QR gradient: with C = Q^H dA R^-1 = X + U (X anti-Hermitian, U upper triangular
// with real diagonal), dQ = (I - Q Q^H) dA R^-1 + Q X, so Abar = (Qbar - Q herm_upper(Q^H Qbar)) R^-H,
// plus the norm term a_i nbar_i / n_i.  Degenerate modes (Gram-Schmidt residual below 1e-12, where the
// reference clamps the norm) keep R_ii = 1e-12 in the forward; their gradient is not defined there.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "common.h"
#include "qem_probe.h"
#include "qem_ptycho.h"

namespace qem {
namespace {

typedef float2 c32;

constexpr int kMaxProbeModes = QEM_PROBE_MAX_MODES;
constexpr int kT = 256;          // threads of the cross-product and apply kernels
constexpr int kTile = 64;        // pixels staged in LDS per step of the cross-product kernel
constexpr int kChunk = 2048;     // pixels per cross-product block

struct cd {
  double x, y;
};

__device__ __forceinline__ cd cdmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cdconj(cd a) { return {a.x, -a.y}; }
__device__ __forceinline__ cd cdadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd cdsub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cdscale(cd a, double s) { return {a.x * s, a.y * s}; }

// Workspace layout (bytes): partials [blocks][M][M] cd | state: T [M][M] cd, n [M] double, perm [M] int
// | coefficients C1 [M][M] c32, C2 [M][M] c32.
struct WsLayout {
  size_t parts, T, n, perm, c1, c2, total;
  int blocks;
};

WsLayout ws_layout(int M, int64_t pixels) {
  WsLayout L;
  const int64_t blocks = (pixels + kChunk - 1) / kChunk;
  L.blocks = (int)(blocks < 1 ? 1 : blocks);
  size_t o = 0;
  L.parts = o;
  o += (size_t)L.blocks * M * M * sizeof(cd);
  L.T = o;
  o += (size_t)M * M * sizeof(cd);
  L.n = o;
  o += (size_t)M * sizeof(double);
  L.perm = o;
  o += (size_t)M * sizeof(int) + 8;
  o = (o + 15) & ~(size_t)15;
  L.c1 = o;
  o += (size_t)M * M * sizeof(c32);
  L.c2 = o;
  o += (size_t)M * M * sizeof(c32);
  L.total = o;
  return L;
}

// parts[blk][j][k] = sum over the block's pixels of conj(X_j) Y_k.  Thread t handles the pair
// t % (M*M) for the pixels of the tile congruent to t / (M*M) modulo the number of pair groups; the
// groups are added in index order at the end.
__global__ __launch_bounds__(kT) void k_cross(int M, int64_t pixels, const c32* __restrict__ X,
                                              const c32* __restrict__ Y, cd* __restrict__ parts) {
  __shared__ c32 sx[kMaxProbeModes][kTile];
  __shared__ c32 sy[kMaxProbeModes][kTile];
  __shared__ cd red[kT];
  const int P2 = M * M;
  const int groups = kT / P2;
  const int t = threadIdx.x;
  const int pair = t % P2, grp = t / P2;
  const int j = pair / M, k = pair % M;
  const bool active = grp < groups;
  cd acc = {0.0, 0.0};
  const int64_t lo = (int64_t)blockIdx.x * kChunk;
  const int64_t hi = lo + kChunk < pixels ? lo + kChunk : pixels;
  for (int64_t base = lo; base < hi; base += kTile) {
    const int cnt = (int)(hi - base < kTile ? hi - base : kTile);
    __syncthreads();
    for (int i = t; i < M * kTile; i += kT) {
      const int m = i / kTile, x = i % kTile;
      const c32 zero = {0.0f, 0.0f};
      sx[m][x] = x < cnt ? X[(size_t)m * pixels + base + x] : zero;
      sy[m][x] = x < cnt ? Y[(size_t)m * pixels + base + x] : zero;
    }
    __syncthreads();
    if (active) {
      for (int x = grp; x < cnt; x += groups) {
        const c32 a = sx[j][x], b = sy[k][x];
        // conj(a) b
        acc.x += (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        acc.y += (double)a.x * (double)b.y - (double)a.y * (double)b.x;
      }
    }
  }
  red[t] = active ? acc : cd{0.0, 0.0};
  __syncthreads();
  if (t < P2) {
    cd s = red[t];
    for (int g = 1; g < groups; ++g) s = cdadd(s, red[g * P2 + t]);
    parts[(size_t)blockIdx.x * P2 + t] = s;
  }
}

// Sums the partials of pair (j, k) in block order.
__device__ cd sum_parts(const cd* parts, int blocks, int P2, int pair) {
  cd s = {0.0, 0.0};
  for (int b = 0; b < blocks; ++b) s = cdadd(s, parts[(size_t)b * P2 + pair]);
  return s;
}

// Forward's small step (one block): G, Cholesky, T = R^-1, sort, sorted coefficient columns.
__global__ __launch_bounds__(kT) void k_small_forward(int M, int blocks, const cd* __restrict__ parts,
                                                      cd* __restrict__ Tout, double* __restrict__ nout,
                                                      int* __restrict__ perm, c32* __restrict__ C) {
  // the M x M algebra runs in one thread, on matrices in LDS
  __shared__ cd G[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd R[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd T[kMaxProbeModes][kMaxProbeModes];
  __shared__ double n[kMaxProbeModes], inten[kMaxProbeModes];
  __shared__ int order[kMaxProbeModes];
  const int P2 = M * M;
  for (int pr = threadIdx.x; pr < P2; pr += kT) G[pr / M][pr % M] = sum_parts(parts, blocks, P2, pr);
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < M; ++i) {
    n[i] = sqrt(fmax(G[i][i].x, 0.0));
    for (int k = 0; k < M; ++k) R[i][k] = T[i][k] = cd{0.0, 0.0};
  }
  // G = R^H R: R_ii = sqrt(G_ii - sum_{l<i} |R_li|^2), R_ik = (G_ik - sum_{l<i} conj(R_li) R_lk) / R_ii
  for (int i = 0; i < M; ++i) {
    double d = G[i][i].x;
    for (int l = 0; l < i; ++l) d -= R[l][i].x * R[l][i].x + R[l][i].y * R[l][i].y;
    const double rii = fmax(sqrt(fmax(d, 0.0)), 1e-12);  // the reference's clamp of the residual norm
    R[i][i] = cd{rii, 0.0};
    for (int k = i + 1; k < M; ++k) {
      cd s = G[i][k];
      for (int l = 0; l < i; ++l) s = cdsub(s, cdmul(cdconj(R[l][i]), R[l][k]));
      R[i][k] = cdscale(s, 1.0 / rii);
    }
  }
  // T = R^-1 (upper triangular), column by column: T_ii = 1 / R_ii, T_jk = -(sum_{j<l<=k} R_jl T_lk) / R_jj
  for (int k = 0; k < M; ++k) {
    T[k][k] = cd{1.0 / R[k][k].x, 0.0};
    for (int j = k - 1; j >= 0; --j) {
      cd s = {0.0, 0.0};
      for (int l = j + 1; l <= k; ++l) s = cdadd(s, cdmul(R[j][l], T[l][k]));
      T[j][k] = cdscale(s, -1.0 / R[j][j].x);
    }
  }
  // intensity of output mode i: n_i^2 |q_i|^2, |q_i| = r_i / max(r_i, 1e-12) (1 unless clamped)
  for (int i = 0; i < M; ++i) {
    double r2 = G[i][i].x;
    for (int l = 0; l < i; ++l) r2 -= R[l][i].x * R[l][i].x + R[l][i].y * R[l][i].y;
    const double r = sqrt(fmax(r2, 0.0));
    const double q = r / R[i][i].x;
    inten[i] = n[i] * n[i] * q * q;
  }
  for (int i = 0; i < M; ++i) order[i] = i;
  for (int a = 1; a < M; ++a) {  // stable insertion sort, descending
    const int v = order[a];
    int b = a - 1;
    while (b >= 0 && inten[order[b]] < inten[v]) {
      order[b + 1] = order[b];
      --b;
    }
    order[b + 1] = v;
  }
  for (int i = 0; i < M; ++i) {
    nout[i] = n[i];
    perm[i] = order[i];
    for (int k = 0; k < M; ++k) Tout[i * M + k] = T[i][k];
  }
  // out_s = sum_j a_j C[j][s],  C[j][s] = T[j][order[s]] n[order[s]]
  for (int jj = 0; jj < M; ++jj)
    for (int s = 0; s < M; ++s) {
      const cd c = cdscale(T[jj][order[s]], n[order[s]]);
      C[jj * M + s] = c32{(float)c.x, (float)c.y};
    }
}

// Backward's small step (one block): from K = A^H Obar (Obar unsorted, columns in original order).
__global__ __launch_bounds__(kT) void k_small_backward(int M, int blocks, const cd* __restrict__ parts,
                                                       const cd* __restrict__ Tin, const double* __restrict__ nin,
                                                       const int* __restrict__ perm, c32* __restrict__ E,
                                                       c32* __restrict__ F) {
  // the M x M algebra runs in one thread, on matrices in LDS
  __shared__ cd K[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd T[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd Ko[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd W[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd H[kMaxProbeModes][kMaxProbeModes];
  __shared__ cd TH[kMaxProbeModes][kMaxProbeModes];
  __shared__ double n[kMaxProbeModes], nbar[kMaxProbeModes];
  __shared__ int inv[kMaxProbeModes];
  const int P2 = M * M;
  for (int pr = threadIdx.x; pr < P2; pr += kT) K[pr / M][pr % M] = sum_parts(parts, blocks, P2, pr);
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < M; ++i) {
    n[i] = nin[i];
    inv[perm[i]] = i;
    for (int k = 0; k < M; ++k) T[i][k] = Tin[i * M + k];
  }
  // K here holds A^H Obar_sorted: column s belongs to original mode perm[s]; reorder to original order.
  for (int j = 0; j < M; ++j)
    for (int i = 0; i < M; ++i) Ko[j][i] = K[j][inv[i]];
  // W = T^H K
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      cd s = {0.0, 0.0};
      for (int l = 0; l < M; ++l) s = cdadd(s, cdmul(cdconj(T[l][a]), Ko[l][b]));
      W[a][b] = s;
    }
  for (int i = 0; i < M; ++i) nbar[i] = W[i][i].x;  // Re(q_i^H obar_i)
  // B = W diag(n) = Q^H Qbar;  H: upper triangle of B, its conjugate below, real diagonal
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      if (a < b)
        H[a][b] = cdscale(W[a][b], n[b]);
      else if (a > b)
        H[a][b] = cdconj(cdscale(W[b][a], n[a]));
      else
        H[a][b] = cd{W[a][a].x * n[a], 0.0};
    }
  // F = -T H T^H + diag(nbar / n);  E = diag(n) T^H
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      cd s = {0.0, 0.0};
      for (int l = 0; l < M; ++l) s = cdadd(s, cdmul(T[a][l], H[l][b]));
      TH[a][b] = s;
    }
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      cd s = {0.0, 0.0};
      for (int l = 0; l < M; ++l) s = cdadd(s, cdmul(TH[a][l], cdconj(T[b][l])));
      if (a == b) s = cdadd(cdscale(s, -1.0), cd{n[a] > 0.0 ? nbar[a] / n[a] : 0.0, 0.0});
      else s = cdscale(s, -1.0);
      F[a * M + b] = c32{(float)s.x, (float)s.y};
    }
  // Abar_b = sum_i Obar_i E[i][b]; Obar_i is sorted column inv[i]: rows of E indexed by sorted s
  for (int s = 0; s < M; ++s) {
    const int i = perm[s];
    for (int b = 0; b < M; ++b) {
      const cd e = cdscale(cdconj(T[b][i]), n[i]);  // (diag(n) T^H)[i][b] = n_i conj(T[b][i])
      E[s * M + b] = c32{(float)e.x, (float)e.y};
    }
  }
}

// out_k(x) = sum_j X_j(x) C1[j][k] (+ sum_j Y_j(x) C2[j][k] when Y != NULL), one pixel per thread.
__global__ __launch_bounds__(kT) void k_apply(int M, int64_t pixels, const c32* __restrict__ X,
                                              const c32* __restrict__ C1, const c32* __restrict__ Y,
                                              const c32* __restrict__ C2, c32* __restrict__ out) {
  __shared__ c32 c1[kMaxProbeModes * kMaxProbeModes], c2[kMaxProbeModes * kMaxProbeModes];
  for (int i = threadIdx.x; i < M * M; i += kT) {
    c1[i] = C1[i];
    if (Y != nullptr) c2[i] = C2[i];
  }
  __syncthreads();
  for (int64_t x = (int64_t)blockIdx.x * kT + threadIdx.x; x < pixels; x += (int64_t)gridDim.x * kT) {
    c32 xv[kMaxProbeModes], yv[kMaxProbeModes];
    for (int j = 0; j < M; ++j) {
      xv[j] = X[(size_t)j * pixels + x];
      if (Y != nullptr) yv[j] = Y[(size_t)j * pixels + x];
    }
    for (int k = 0; k < M; ++k) {
      float re = 0.0f, im = 0.0f;
      for (int j = 0; j < M; ++j) {
        const c32 a = xv[j], c = c1[j * M + k];
        re += a.x * c.x - a.y * c.y;
        im += a.x * c.y + a.y * c.x;
      }
      if (Y != nullptr) {
        for (int j = 0; j < M; ++j) {
          const c32 a = yv[j], c = c2[j * M + k];
          re += a.x * c.x - a.y * c.y;
          im += a.x * c.y + a.y * c.x;
        }
      }
      out[(size_t)k * pixels + x] = c32{re, im};
    }
  }
}

int check_args(int32_t M, int64_t pixels) {
  if (M < 1 || M > kMaxProbeModes)
    return set_error(QEM_ERR_ARG, "modes must be in [1, %d] (got %d)", kMaxProbeModes, M);
  if (pixels < 1) return set_error(QEM_ERR_ARG, "pixels must be >= 1 (got %lld)", (long long)pixels);
  return QEM_OK;
}

int check_ptr8(const void* p, const char* name) {
  if (p == nullptr) return set_error(QEM_ERR_NULL, "%s is NULL", name);
  if (((uintptr_t)p) & 7u) return set_error(QEM_ERR_ALIGN, "%s is not 8-byte aligned", name);
  return QEM_OK;
}

int apply_grid(int64_t pixels) {
  const int64_t b = (pixels + kT - 1) / kT;
  return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

}  // namespace
}  // namespace qem

extern "C" int qem_probe_orthogonalize_workspace_bytes(int32_t modes, int64_t pixels, size_t* bytes) {
  qem::clear_error();
  if (int rc = qem::check_args(modes, pixels)) return rc;
  if (bytes == nullptr) return qem::set_error(QEM_ERR_NULL, "bytes is NULL");
  *bytes = qem::ws_layout(modes, pixels).total;
  return QEM_OK;
}

extern "C" int qem_probe_orthogonalize_forward(int32_t modes, int64_t pixels, const void* probe, void* out,
                                               void* workspace, size_t workspace_bytes, void* stream) {
  qem::clear_error();
  if (int rc = qem::check_args(modes, pixels)) return rc;
  if (int rc = qem::check_ptr8(probe, "probe")) return rc;
  if (int rc = qem::check_ptr8(out, "out")) return rc;
  if (int rc = qem::check_ptr8(workspace, "workspace")) return rc;
  const qem::WsLayout L = qem::ws_layout(modes, pixels);
  if (workspace_bytes < L.total)
    return qem::set_error(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  const qem::c32* A = (const qem::c32*)probe;
  hipLaunchKernelGGL(qem::k_cross, dim3(L.blocks), dim3(qem::kT), 0, st, modes, pixels, A, A, (qem::cd*)(w + L.parts));
  if (int rc = qem::check_last_launch("probe_ortho_cross")) return rc;
  hipLaunchKernelGGL(qem::k_small_forward, dim3(1), dim3(qem::kT), 0, st, modes, L.blocks,
                     (const qem::cd*)(w + L.parts), (qem::cd*)(w + L.T), (double*)(w + L.n), (int*)(w + L.perm),
                     (qem::c32*)(w + L.c1));
  if (int rc = qem::check_last_launch("probe_ortho_small")) return rc;
  hipLaunchKernelGGL(qem::k_apply, dim3(qem::apply_grid(pixels)), dim3(qem::kT), 0, st, modes, pixels, A,
                     (const qem::c32*)(w + L.c1), (const qem::c32*)nullptr, (const qem::c32*)nullptr,
                     (qem::c32*)out);
  return qem::check_last_launch("probe_ortho_apply");
}

extern "C" int qem_probe_orthogonalize_backward(int32_t modes, int64_t pixels, const void* probe, const void* grad_out,
                                                void* grad_probe, void* workspace, size_t workspace_bytes,
                                                void* stream) {
  qem::clear_error();
  if (int rc = qem::check_args(modes, pixels)) return rc;
  if (int rc = qem::check_ptr8(probe, "probe")) return rc;
  if (int rc = qem::check_ptr8(grad_out, "grad_out")) return rc;
  if (int rc = qem::check_ptr8(grad_probe, "grad_probe")) return rc;
  if (int rc = qem::check_ptr8(workspace, "workspace")) return rc;
  const qem::WsLayout L = qem::ws_layout(modes, pixels);
  if (workspace_bytes < L.total)
    return qem::set_error(QEM_ERR_WORKSPACE, "workspace has %zu bytes, needs %zu", workspace_bytes, L.total);
  char* w = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  const qem::c32* A = (const qem::c32*)probe;
  const qem::c32* G = (const qem::c32*)grad_out;
  hipLaunchKernelGGL(qem::k_cross, dim3(L.blocks), dim3(qem::kT), 0, st, modes, pixels, A, G, (qem::cd*)(w + L.parts));
  if (int rc = qem::check_last_launch("probe_ortho_cross")) return rc;
  hipLaunchKernelGGL(qem::k_small_backward, dim3(1), dim3(qem::kT), 0, st, modes, L.blocks,
                     (const qem::cd*)(w + L.parts), (const qem::cd*)(w + L.T), (const double*)(w + L.n),
                     (const int*)(w + L.perm), (qem::c32*)(w + L.c1), (qem::c32*)(w + L.c2));
  if (int rc = qem::check_last_launch("probe_ortho_small")) return rc;
  hipLaunchKernelGGL(qem::k_apply, dim3(qem::apply_grid(pixels)), dim3(qem::kT), 0, st, modes, pixels, G,
                     (const qem::c32*)(w + L.c1), A, (const qem::c32*)(w + L.c2), (qem::c32*)grad_probe);
  return qem::check_last_launch("probe_ortho_apply");
}
