// ptycho_common.h — parameters and device helpers shared by the N <= 128 (register/LDS), the
// N = 256 (two-pass) and the generic-size (Nr x Nc, ptycho_gen.hip) multislice kernels.  Buffer shapes
// below write (N, N) for one pattern; the generic chain stores (Nr, Nc) in natural order there.
#pragma once
#include <hip/hip_runtime.h>

#include "fft2.h"
#include "qem_ptycho.h"

namespace qem {

struct Params {
#ifdef QEM_PARAMS_LAYOUT_TEST
  int layout_test_extra;  // only in the CPU test build that checks the layout comparison below
#endif
  int S, M, B, H, W, P, loss_type;
  int NR, NC;  // pattern rows and columns (NR == NC == N for the power-of-two kernels)
  float mean_I;
  unsigned flags;
  const cf* obj;
  const cf* probe;
  const float* pos;
  const cf* prop;
  const float* targets;
  const float* mask;
  const float* descan;
  float* loss_out;
  float* intens_out;
  const float* dloss;
  cf* grad_obj;
  cf* grad_probe;
  float* grad_pos;
  float* grad_descan;
  // workspace
  cf* psi;          // (B*M*S, N, N)  incident waves; reused by the adjoint for per-mode object-gradient patches
  cf* scratch;      // (B*M, N, N)    row-transformed wave of each item (N = 256 chain)
  cf* X;            // (B*M, N, N)    far-field waves; reused by the adjoint for probe-gradient spectra
  float* gI;        // (B, N, N)      dL/dI, unshifted
  double* loss_part;  // per loss block
  cf* pk;           // (M, N, N)      FFT2 of the probe modes
  cf* gsum;         // (M, N, N)      sum over the batch of probe-gradient spectra
  cf* ramps;        // (B, Nr + Nc)   sub-pixel shift ramps exp(-2 pi i f dr_b) (f = fftfreq(Nr)), then
                    //                exp(-2 pi i f dc_b) (f = fftfreq(Nc)); (B, 2, N) when square
  float* gpos_part;   // (B*M, 2)
  float* gdesc_part;  // (B*M, 2)
  // slice range and wave storage of one launch.  psi holds `wslots` slots per item; slice s of the
  // range lives in slot s - s0.  With checkpoints (ck >= 2) the full forward keeps only the wave
  // entering every ck-th slice in ckpt (nck slots per item) and the adjoint runs segment by
  // segment [s0, s1), recomputing the segment's waves from its checkpoint first.
  int s0, s1;
  int wslots;
  int ck, nck;
  int fwd_mode;  // FWD_* below
  cf* ckpt;      // (B*M*nck, N, N)
  // propagator gradient (QEM_FLAG_GRAD_PROPAGATORS): the forward keeps fft2 of each slice's transmitted
  // wave (P3 layout, slots s = 0..S-2 per item); the adjoint overwrites each with that item's
  // contribution conj(X) fft2(g_psi) / N^2 to dL/dP_s, which a fixed-order reduce sums over the items
  cf* pspec;     // (B*M*(S-1), N, N) or nullptr
  // separable propagators (N = 256 chain): per slice its row factor a[kr] = prop[kr][0] and column
  // factor b[kc] = prop[0][kc] / prop[0][0] ((S-1, 2, N)), and whether a[kr] b[kc] reproduces each
  // 32-row part of the slice ((S-1, 8) ints, 1 = yes); nullptr = apply every propagator entry as given
  cf* pfac;
  int* pfac_ok;
  // N <= 128: the propagators times 1 / N^2 ((S-1, N, N), filled by the probe-spectrum launch; a power of
  // two, so every product with them equals the product with the propagator, scaled)
  cf* psc;
};

// Layout signature of Params as compiled in the current translation unit: its size and every field's
// offset folded into one value.  Params crosses from ptycho.hip (the C ABI) into ptycho256.hip and
// ptycho_gen.hip (the N = 256 and generic-size launchers) by reference, so all three must agree on it;
// a build that compiled them with different versions of this header would hand the kernels a struct
// whose pointers sit at other offsets.  The entry points compare params_signature_local() with
// params_signature_256() and params_signature_gen() and refuse to run on a mismatch.
#define QEM_PARAMS_FIELDS(QF)                                                                          \
  QF(S) QF(M) QF(B) QF(H) QF(W) QF(P) QF(loss_type) QF(NR) QF(NC) QF(mean_I) QF(flags) QF(obj) QF(probe) QF(pos) QF(prop)        \
  QF(targets) QF(mask) QF(descan) QF(loss_out) QF(intens_out) QF(dloss) QF(grad_obj) QF(grad_probe)           \
  QF(grad_pos) QF(grad_descan) QF(psi) QF(scratch) QF(X) QF(gI) QF(loss_part) QF(pk) QF(gsum) QF(ramps)          \
  QF(gpos_part) QF(gdesc_part) QF(s0) QF(s1) QF(wslots) QF(ck) QF(nck) QF(fwd_mode) QF(ckpt) QF(pspec)  \
  QF(pfac) QF(pfac_ok) QF(psc)
// internal linkage: each translation unit evaluates its own copy (an inline function would be merged)
static constexpr unsigned long long params_signature_local() {
  unsigned long long h = sizeof(Params);
#define QEM_PARAMS_FOLD(f) h = h * 1000003ull ^ (unsigned long long)__builtin_offsetof(Params, f);
  QEM_PARAMS_FIELDS(QEM_PARAMS_FOLD)
#undef QEM_PARAMS_FOLD
  return h;
}
// params_signature_local() as evaluated in ptycho256.hip and in ptycho_gen.hip.
unsigned long long params_signature_256();
unsigned long long params_signature_gen();

// forward launch modes
enum : int {
  FWD_NO_STORE = 0,   // full chain, far field, no waves kept (no-grad forward)
  FWD_STORE_ALL = 1,  // full chain, far field, every incident wave kept in psi
  FWD_CHECKPOINT = 2, // full chain, far field, only the checkpoint waves kept in ckpt
  FWD_RECOMPUTE = 3   // slices [s0, s1) from checkpoint s0 / ck into psi; no far field
};

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// numpy/torch fftfreq(N, d=1)[k] in float32 (exact for power-of-two N).
template <int N>
__device__ __forceinline__ float fftfreq1(int k) {
  return (float)(k < N / 2 ? k : k - N) * (1.0f / (float)N);
}

// exp(-2 pi i x) for float32 x.
__device__ __forceinline__ cf expm2pi(float x) {
  float s, c;
  sincospif(-2.0f * x, &s, &c);
  return mk(c, s);
}

// Maps a linear block index to a work item so that consecutive items land on the same XCD
// (blocks are dispatched round-robin over the 8 XCDs; each XCD has its own L2).
__device__ __forceinline__ int xcd_item(int blk, int nblk) {
  if ((nblk & 7) != 0) return blk;
  return (blk & 7) * (nblk >> 3) + (blk >> 3);
}

template <int N>
__device__ __forceinline__ void origin_and_shift(const float* pos, int b, int& r0, int& c0, float& dr, float& dc) {
  const float pr = pos[2 * b], pc = pos[2 * b + 1];
  const float rr = rintf(pr), rc = rintf(pc);  // round half to even, as torch.round
  r0 = (int)rr;
  c0 = (int)rc;
  dr = pr - rr;
  dc = pc - rc;
}

// Streaming (nontemporal) access for the per-slice wave and gradient-patch buffers: each element is
// written once and read once, long after, so it should not displace the object slices, propagators
// and probe spectra from L2 / Infinity Cache.
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_stream(cf* p, cf v) {
  f2v x;
  x.x = v.x;
  x.y = v.y;
  __builtin_nontemporal_store(x, (f2v*)p);
}
__device__ __forceinline__ cf ld_stream(const cf* p) {
  const f2v x = __builtin_nontemporal_load((const f2v*)p);
  return mk(x.x, x.y);
}

// Two consecutive complex64 values (16 bytes, 16-byte aligned) streamed out / in.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_stream2(cf* p, cf a, cf b) {
  f4v x;
  x.x = a.x;
  x.y = a.y;
  x.z = b.x;
  x.w = b.y;
  __builtin_nontemporal_store(x, (f4v*)p);
}
__device__ __forceinline__ void ld_stream2(const cf* p, cf& a, cf& b) {
  const f4v x = __builtin_nontemporal_load((const f4v*)p);
  a = mk(x.x, x.y);
  b = mk(x.z, x.w);
}

// Object element at a byte offset from the slice base.
__device__ __forceinline__ cf obj_at(const cf* O, unsigned byte_off) {
  return *(const cf*)((const char*)O + byte_off);
}

// Block-wide sum of one double per thread; result valid in thread 0.
template <int T>
__device__ __forceinline__ double block_sum(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = x;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < T / 64; ++i) t += red[i];
  }
  __syncthreads();
  return t;
}

// Two-pass (N = 256) chain launchers (ptycho256.hip).
int launch_chain256_forward(const Params& p, hipStream_t st);   // probe spectra + forward kernel
int launch_forward256_kernel(const Params& p, hipStream_t st);  // forward kernel only (p.fwd_mode)
int launch_chain256_adjoint(const Params& p, hipStream_t st);
int launch_probe_grad256(const Params& p, hipStream_t st);
// Generic-size chain (ptycho_gen.hip): Nr x Nc patterns, 2..512 per axis, prime factors 2, 3, 5, 7.
bool gen_supported(int nr, int nc);
int launch_gen_forward(const Params& p, hipStream_t st);         // probe spectra + forward kernel
int launch_gen_forward_kernel(const Params& p, hipStream_t st);  // forward kernel only (p.fwd_mode)
int launch_gen_loss(const Params& p, int ntiles, hipStream_t st);  // per-block partials into p.loss_part
int gen_loss_parts(int ntiles);                                  // number of those partials
int launch_gen_adjoint(const Params& p, hipStream_t st);
int launch_gen_obj_reduce(const Params& p, hipStream_t st);
int launch_gen_probe_grad(const Params& p, hipStream_t st);      // batch reduce + inverse transform
int launch_gen_prop_reduce(const Params& p, cf* grad_prop, hipStream_t st);
// Amplitude centring of measured 256 x 256 patterns (preprocess.hip), `blocks` workgroups with one
// 256 x 256 complex64 wave each in ws.
int launch_center256(int P, int blocks, const float* I, const float* com, float* out, float* sums, cf* ws,
                     hipStream_t st);

}  // namespace qem
