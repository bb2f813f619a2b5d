// ptycho256.hip — multislice forward / adjoint for 256 x 256 patterns (row / column passes).
//
// A 256 x 256 complex64 wave (512 KiB) does not fit one workgroup's LDS, so each workgroup owns one
// (pattern, mode) item and keeps its wave in a per-item buffer W, in the "row-transformed" domain
// (r, kc): every row already Fourier transformed.  A slice then costs two passes over W:
//
//   row pass     for 32 rows at a time: inverse row DFT -> natural-order wave (r, c) -> the
//                multislice elementwise work (store the incident wave, transmission, descan, the
//                adjoint's gradient patch) -> forward row DFT -> back to W
//   column pass  for 32 columns at a time: column DFT -> full 2-D spectrum (kr, kc) -> propagator
//                (or far field / probe-gradient epilogue) -> inverse column DFT -> back to W
//
// A length-256 DFT is 16 x 16: a radix-16 DFT in registers (16 lanes per row or column, 16 values
// per lane), a twiddle, an LDS exchange, and a second radix-16 DFT.  All passes update W in place
// (every chunk is read completely before any of it is written) and end with a barrier.  The chain
// follows the same reference lines as ptycho.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "common.h"
#include "ptycho_common.h"

namespace qem {
namespace n256 {

constexpr int N = 256;
constexpr int N2 = N * N;
// Threads per workgroup: 512 (two waves per SIMD, up to 256 VGPRs each, 32 rows or columns per chunk)
// measured 20 % faster in the forward and 6 % in the adjoint than 256 threads with the next chunk's
// values carried in registers; 1024 threads (128 VGPRs) spill and ran slower.
constexpr int T = 512;
constexpr int CH = T / 16;   // rows (or columns) per chunk: 16 lanes per row
constexpr int VB = 272;      // LDS block per 16-lane vector: 16 x 17 (rows) or 16 x 16 + 16 (columns)
static_assert(T % 256 == 0 && T <= 1024, "whole 16-column groups in the column passes");

struct Smem {
  cf ex[CH * VB];  // exchange of the 16 x 16 decomposition
  cf tw[N];        // W^k, W = exp(-2 pi i / 256)
  cf dcol[N];      // column factor exp(-2 pi i fftfreq(c) d_c) of the descan ramp (adjoint)
  double red[T / 64];
};

// The multislice kernels run one workgroup per CU: with 256 items resident the row-transformed waves
// (256 x 512 KiB) stay in the Infinity Cache across the two passes of every slice, which measured
// faster than two or three items per CU streaming twice that through HBM.  Their LDS holds, besides
// the exchange buffer, one chunk of W (64 KiB) that the LDS DMA copies in while the previous chunk is
// transformed; the footprint above half a CU's 160 KiB also keeps one workgroup per CU.
struct SmemOne : Smem {
  cf dma[CH * N];  // the next chunk of W: rows per wave, columns [e][j][column] (see dma_rows / dma_cols)
  cf pfac[2 * N];  // row factor a[kr], then column factor b[kc] / N^2, of a separable propagator slice
};
static_assert(T == 2 * N, "one propagator factor per thread");
static_assert(sizeof(SmemOne) > 80 * 1024 && sizeof(SmemOne) <= 160 * 1024, "one workgroup per CU");

// Orders this wave's LDS accesses: the lanes of one wave issue LDS instructions together and the LDS
// serves one wave's instructions in order, so a read after this sees every write of the wave before
// it; only the compiler has to be kept from moving accesses across (as rocPRIM's wave_barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS DMA (global_load_lds_dwordx4): every lane copies 16 bytes from its own global address to
// consecutive 16-byte slots of a 1 KiB LDS block.  A chunk of W (CH vectors x 256 values, 64 KiB) is
// 64 such blocks, 8 per wave.  The copy bypasses the registers, so the next chunk is fetched while the
// current one is transformed.  The waits are explicit: a pass waits with vmcnt(16) for its own copy,
// allowing only the previous chunk's 16 row / column stores (the last vector-memory instructions of
// the wave) to stay in flight; loads and stores share the one in-order vmcnt counter.
typedef __attribute__((address_space(3))) void* lds_void_ptr;

// One copy instruction: 16 bytes per lane from base + lane_off to the LDS block at lds (its 64 lanes'
// data consecutive).  Written as inline assembly so that the compiler's wait-count pass does not see
// it: it treats an LDS DMA as pending until vmcnt(0) and waited for every outstanding store before
// any read of the block; the passes wait for their copies explicitly (dma_wait).  The instruction
// writes no register.
__device__ __forceinline__ void lds_dma16(const void* base, unsigned lane_off, const cf* lds) {
  const unsigned m0 = __builtin_amdgcn_readfirstlane((unsigned)(size_t)(lds_void_ptr)lds);  // LDS block
  const char* addr = (const char*)base + lane_off;
  asm volatile("s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(addr), "{m0}"(m0) : "memory");
}

// Row chunk [r0, r0 + CH) of src into dma, each wave's four rows (vectors 4 w .. 4 w + 3) in its own
// 8 KiB region laid out [e][row][j] (value j + 16 e of row r0 + 4 w + row): a wave copies and reads only
// its own rows, so the row passes need no workgroup barrier, and a wave's read of its v[e] is 512
// contiguous bytes.  The wave index is made scalar, so that the block offsets are scalar and each copy
// instruction is a scalar base plus one 32-bit lane offset.
__device__ __forceinline__ void dma_rows(const cf* src, int r0, SmemOne& sm) {
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int l = threadIdx.x & 63;
  // lane: e parity l / 32, row (l % 32) / 8, values 2 (l % 8), +1
  unsigned lane = (unsigned)(((l & 31) >> 3) * N + 2 * (l & 7) + 16 * (l >> 5)) * 8u;
  asm volatile("" : "+v"(lane));  // formed here, not kept live through the transforms
  const unsigned wb = (unsigned)((r0 + 4 * w) * N) * 8u;
#pragma unroll
  for (int i = 0; i < 8; ++i) lds_dma16(src, wb + 256u * i + lane, &sm.dma[w * 1024 + i * 128]);  // e = 2 i, 2 i + 1
}
__device__ __forceinline__ cf dma_row_value(const SmemOne& sm, int e) {
  return sm.dma[(int)(threadIdx.x >> 6) * 1024 + e * 64 + (int)(threadIdx.x & 63)];
}
// Column chunk [c0, c0 + CH) of src into dma laid out [e][j][column ^ 16 (j & 1)] (value j + 16 e of
// column c0 + column): the swizzle spreads a wave's four j rows of 128 bytes over both halves of the
// banks.
__device__ __forceinline__ void dma_cols(const cf* src, int c0, SmemOne& sm) {
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int l = threadIdx.x & 63;
  // lane: j offset l / 16 within the block's four j, columns 2 (l % 16) (+1), swizzled for odd j
  unsigned lane = (unsigned)((l >> 4) * N + ((2 * (l & 15)) ^ (16 * ((l >> 4) & 1)))) * 8u;
  asm volatile("" : "+v"(lane));
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int blk = w * 8 + i;  // e = blk / 4, j = (blk % 4) * 4 .. + 3 (j parity = lane's l / 16 parity)
    const unsigned ub = (unsigned)(((blk & 3) * 4 + 16 * (blk >> 2)) * N + c0) * 8u;
    lds_dma16(src, ub + lane, &sm.dma[blk * 128]);
  }
}
__device__ __forceinline__ cf dma_col_value(const SmemOne& sm, int e) {
  const int j = (threadIdx.x >> 4) & 15;
  const int c = (threadIdx.x & 15) + 16 * (threadIdx.x >> 8);
  return sm.dma[((e * 16 + j) * CH) + (c ^ (16 * (j & 1)))];
}
// Waits for this wave's copies (`stores` = the number of vector-memory instructions issued after them,
// the previous chunk's stores, which may stay in flight: vmcnt(16) or vmcnt(0)), then, for a chunk
// that other waves copied (columns), for every wave's (barrier).
template <bool SHARED>
__device__ __forceinline__ void dma_wait(bool stores) {
  if (stores)
    __builtin_amdgcn_s_waitcnt(0x4F70);  // vmcnt(16)
  else
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  if constexpr (SHARED) {
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  } else {
    wave_lds_sync();
  }
}

// Lane roles.  Row passes: vector (row) t / 16, j = t % 16.  Column passes: column t % 16 + 16 (t / 256)
// (one 16-column group per 256 threads), j = (t / 16) % 16.  In both, the lane holds elements
// n = j + 16 i of its vector in v[i].
__device__ __forceinline__ int row_vec() { return threadIdx.x >> 4; }
__device__ __forceinline__ int row_j() { return threadIdx.x & 15; }
__device__ __forceinline__ int col_vec() { return (threadIdx.x & 15) + 16 * (threadIdx.x >> 8); }
__device__ __forceinline__ int col_j() { return (threadIdx.x >> 4) & 15; }

// Exchange slot of value (j, k1) of this lane's vector.  The row layout pads 16-value rows to 17, the
// column layout pads each j block by 16, so that both the write (lanes vary j) and the read (lanes vary
// k1) of a wave touch every bank at most twice (the minimum for 64 x 8 bytes).
template <bool COLS>
__device__ __forceinline__ int ex_idx(int j, int k1) {
  if constexpr (COLS) {
    const int cc = threadIdx.x & 15, grp = threadIdx.x >> 8;
    return (grp * 16 + j) * VB + k1 * 16 + cc;
  } else {
    return row_vec() * VB + k1 * 17 + j;
  }
}

// In-place length-256 DFT (unnormalised; conjugate twiddles if INV) of this lane's vector:
// on entry v[i] = x[j + 16 i], on return v[k2] = X[j + 16 k2].  A row's 16 lanes are 16 consecutive
// lanes of one wave and its exchange block is its own, so the row exchange needs no workgroup barrier;
// a column's 16 lanes lie in four waves.
template <bool INV, bool COLS>
__device__ __forceinline__ void fft256(cf (&v)[16], Smem& sm) {
  const int j = COLS ? col_j() : row_j();
  Dft<16, INV>::template run<1>(v);  // over i -> k1
#pragma unroll
  for (int k1 = 1; k1 < 16; ++k1) {
    const cf w = sm.tw[(j * k1) & 255];
    v[k1] = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
  }
  if constexpr (COLS)
    lds_barrier();  // earlier readers of sm.ex are done
  else
    wave_lds_sync();
#pragma unroll
  for (int k1 = 0; k1 < 16; ++k1) sm.ex[ex_idx<COLS>(j, k1)] = v[k1];
  if constexpr (COLS)
    lds_barrier();
  else
    wave_lds_sync();
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) v[jj] = sm.ex[ex_idx<COLS>(jj, j)];  // this lane now owns k1 = j
  Dft<16, INV>::template run<1>(v);                                     // over jj -> k2
}

struct NoPre {
  __device__ __forceinline__ void operator()(int) const {}
};

// Row pass over the 256 rows of src (CH rows per chunk): v <- row (columns j + 16 e); pre(r) (the op's
// own loads, issued before the inverse row DFT so that it hides their latency); inverse row DFT if
// PRE_INV; op(r, v) on the natural-order row; forward row DFT if POST_FWD; row -> dst (if not null).
// This form serves the single-pass kernels (probe spectra, probe gradient, centring); the multislice
// kernels use the overloads below, which stage every chunk of W through LDS.
template <bool PRE_INV, bool POST_FWD, typename Op, typename Pre = NoPre>
__device__ __forceinline__ void row_pass(const cf* src, cf* dst, Smem& sm, Op&& op, Pre&& pre = Pre()) {
  const int j = row_j();
  for (int r0 = 0; r0 < N; r0 += CH) {
    const int r = r0 + row_vec();
    cf v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = src[r * N + j + 16 * e];
    pre(r);
    if constexpr (PRE_INV) fft256<true, false>(v, sm);
    op(r, v);
    if constexpr (POST_FWD) fft256<false, false>(v, sm);
    if (dst != nullptr) {
#pragma unroll
      for (int e = 0; e < 16; ++e) dst[r * N + j + 16 * e] = v[e];
    }
  }
  __syncthreads();  // dst complete and visible to the whole workgroup
}

// Column pass over the 256 columns of src (CH columns per chunk): v <- column c (rows j + 16 e);
// op(c, v) (which runs the column DFTs it needs); column -> dst.
template <typename Op>
__device__ __forceinline__ void col_pass(const cf* src, cf* dst, Smem& sm, Op&& op) {
  const int j = col_j();
  for (int c0 = 0; c0 < N; c0 += CH) {
    const int c = c0 + col_vec();
    cf v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = src[(j + 16 * e) * N + c];
    op(c, v);
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[(j + 16 * e) * N + c] = v[e];
  }
  __syncthreads();
}

// row_pass / col_pass of the multislice kernels (one workgroup per CU, SmemOne): the chunk's W values
// come from the LDS copy of it started during the previous chunk.  Per chunk: wait for the copy, read
// it, pre (the op's own global loads), then the transforms and op; the next chunk's copy is started
// once op has used its loads (a load still in flight when the copy is issued would be waited for
// together with the copy) and runs during the rest of the chunk's transforms and its stores.
template <bool PRE_INV, bool POST_FWD, typename Op, typename Pre = NoPre>
__device__ __forceinline__ void row_pass(const cf* src, cf* dst, SmemOne& sm, Op&& op, Pre&& pre = Pre()) {
  const int j = row_j();
  dma_rows(src, 0, sm);
  for (int r0 = 0; r0 < N; r0 += CH) {
    const int r = r0 + row_vec();
    dma_wait<false>(r0 > 0 && dst != nullptr);
    cf v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = dma_row_value(sm, e);
    wave_lds_sync();  // the wave's reads stay before its next copy into the same region
    pre(r);
    if constexpr (PRE_INV) fft256<true, false>(v, sm);
    op(r, v);
    __builtin_amdgcn_sched_barrier(0);  // op's use of its loads stays before the copy
    if (r0 + CH < N) dma_rows(src, r0 + CH, sm);
    __builtin_amdgcn_sched_barrier(0);  // and the copy is issued before the transform
    if constexpr (POST_FWD) fft256<false, false>(v, sm);
    if (dst != nullptr) {
#pragma unroll
      for (int e = 0; e < 16; ++e) dst[r * N + j + 16 * e] = v[e];
    }
  }
  __syncthreads();  // dst complete and visible to the whole workgroup
}

// op(c, v, next) calls next() where the next chunk's copy may start (after its last use of a global
// load); an op without that parameter gets the copy started after it returns.
template <typename Op, typename Pre = NoPre>
__device__ __forceinline__ void col_pass(const cf* src, cf* dst, SmemOne& sm, Op&& op, Pre&& pre = Pre()) {
  const int j = col_j();
  dma_cols(src, 0, sm);
  for (int c0 = 0; c0 < N; c0 += CH) {
    const int c = c0 + col_vec();
    dma_wait<true>(c0 > 0);
    cf v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = dma_col_value(sm, e);
    lds_barrier();  // every wave has read the chunk before the next copy overwrites it
    pre(c);
    bool issued = false;  // an op that never calls next() still gets the copy, after it returns
    auto next = [&] {
      if (issued) return;
      issued = true;
      __builtin_amdgcn_sched_barrier(0);  // op's use of its loads stays before the copy
      if (c0 + CH < N) dma_cols(src, c0 + CH, sm);
      __builtin_amdgcn_sched_barrier(0);  // and the copy is issued before what follows
    };
    if constexpr (std::is_invocable_v<Op, int, cf(&)[16], decltype(next)&>) {
      op(c, v, next);
    } else {
      op(c, v);
    }
    next();
#pragma unroll
    for (int e = 0; e < 16; ++e) dst[(j + 16 * e) * N + c] = v[e];
  }
  __syncthreads();
}

__device__ __forceinline__ void init_tw(Smem& sm) {
  if (threadIdx.x < N) sm.tw[threadIdx.x] = kTw256[threadIdx.x];
  __syncthreads();
}

__device__ __forceinline__ cf ramp2(float fr, float fc, float dr, float dc) {
  return cmul(expm2pi(fr * dr), expm2pi(fc * dc));
}

// Object row / column byte offsets (periodic, fftfreq-ordered patch).
__device__ __forceinline__ unsigned obj_row_off(int r0, int r, int H, int W) {
  return (unsigned)(pmod(r0 + (r < N / 2 ? r : r - N), H) * W) * 8u;
}
// Object columns of a row-pass lane: patch column j + 16 e is object column (q + 16 ((e + 8) % 16)) mod W
// with q = (c0 - N/2 + j) mod W; since W >= N the wrap is one conditional subtraction.
__device__ __forceinline__ int obj_col_base(int c0, int W) { return pmod(c0 - N / 2 + row_j(), W); }
__device__ __forceinline__ unsigned obj_col_off(int q, int e, int W) {
  int x = q + 16 * ((e + 8) & 15);
  if (x >= W) x -= W;
  return (unsigned)x * 8u;
}

struct NoOp {
  __device__ __forceinline__ void operator()(int, cf (&)[16]) const {}
};

// pk[m] = fft2(probe[m]).
__global__ __launch_bounds__(T) void k_spectrum256(Params p) {
  if ((int)blockIdx.x >= p.M) {
    // the blocks past the modes fill the shift-ramp table: ramps[b][0][k] = exp(-2 pi i f_k dr_b),
    // ramps[b][1][k] = exp(-2 pi i f_k dc_b), read by the forward's first and the adjoint's last pass
    const int idx = ((int)blockIdx.x - p.M) * T + (int)threadIdx.x;
    if (idx >= p.B * 2 * N) return;
    const int b = idx / (2 * N), r = idx - b * 2 * N, k = r & (N - 1);
    int r0, c0;
    float dr, dc;
    origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
    p.ramps[idx] = expm2pi(fftfreq1<N>(k) * (r < N ? dr : dc));
    return;
  }
  __shared__ Smem sm;
  init_tw(sm);
  const int m = blockIdx.x;
  cf* Wb = p.scratch + (size_t)m * N2;
  row_pass<false, true>(p.probe + (size_t)m * N2, Wb, sm, NoOp());
  col_pass(Wb, p.pk + (size_t)m * N2, sm, [&](int, cf(&v)[16]) { fft256<false, true>(v, sm); });
}

// Separable propagator slices.  The reference's Fresnel propagator exp(-i pi lambda dz (kr^2 + kc^2)),
// with or without the tilt's per-axis phase terms (probe_models.py:376-427), is a[kr] b[kc]; applied as
// that product, a column pass reads a 2 KiB row factor into LDS once and one column factor per lane
// and chunk instead of its 64 KiB of propagator entries.  Block (s, part) checks rows
// [32 part, 32 part + 32) of slice s: a[kr] = P[kr][0] and b[kc] = P[0][kc] / P[0][0] (quotient in
// double, rounded once), a[kr] b[kc] formed in float as the kernels do, and the part marked separable
// when none of its entries differs from it by more than 1e-6 of the largest magnitude among them (a
// non-finite entry or factor, or P[0][0] = 0, marks it not separable, so NaNs reach the results as
// they would entry by entry).  1024 threads: thread t reads column
// t % 256 of every fourth row of the part from t / 256.
constexpr int kFacT = 1024, kFacParts = 8, kFacRows = N / kFacParts;
__global__ __launch_bounds__(kFacT) void k_prop_factors256(Params p) {
  __shared__ cf a_s[kFacRows];
  __shared__ double red[2][kFacT / 64];
  const int s = blockIdx.x, part = blockIdx.y, t = threadIdx.x, c = t & (N - 1), g = t / N;
  const int r0 = part * kFacRows;
  const cf* P = p.prop + (size_t)s * N2;
  cf* fa = p.pfac + (size_t)s * 2 * N;
  const cf p00 = P[0], p0c = P[c];
  const double d = (double)p00.x * p00.x + (double)p00.y * p00.y;
  const cf b_c = d > 0.0 ? mk((float)(((double)p0c.x * p00.x + (double)p0c.y * p00.y) / d),
                              (float)(((double)p0c.y * p00.x - (double)p0c.x * p00.y) / d))
                         : mk(0.0f, 0.0f);
  if (t < kFacRows) {
    const cf a_r = P[(size_t)(r0 + t) * N];
    a_s[t] = a_r;
    fa[r0 + t] = a_r;
  }
  if (part == 0 && g == 0) fa[N + c] = b_c;
  __syncthreads();
  constexpr int G = kFacT / N, RPT = kFacRows / G;  // row groups, rows per thread
  cf pv[RPT];
#pragma unroll
  for (int i = 0; i < RPT; ++i) pv[i] = P[(size_t)(r0 + g + G * i) * N + c];
  double dev = 0.0, mag = 0.0;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const cf q = cmul(a_s[g + G * i], b_c);
    const double ex = (double)q.x - (double)pv[i].x, ey = (double)q.y - (double)pv[i].y;
    const double e = ex * ex + ey * ey, m = (double)pv[i].x * pv[i].x + (double)pv[i].y * pv[i].y;
    // a NaN counts as an infinite deviation (or magnitude), and an infinity stays the maximum
    if (!(e <= dev)) dev = e == e ? e : INFINITY;
    if (!(m <= mag)) mag = m == m ? m : INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double od = __shfl_down(dev, o, 64), om = __shfl_down(mag, o, 64);
    if (!(od <= dev)) dev = od;
    if (!(om <= mag)) mag = om;
  }
  if ((t & 63) == 0) {
    red[0][t >> 6] = dev;
    red[1][t >> 6] = mag;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kFacT / 64; ++w) {
      if (!(red[0][w] <= dev)) dev = red[0][w];
      if (!(red[1][w] <= mag)) mag = red[1][w];
    }
    // squared magnitudes: deviation <= 1e-6 * largest magnitude of the part, both finite
    p.pfac_ok[s * kFacParts + part] = (d > 0.0 && mag < INFINITY && dev <= 1e-12 * mag) ? 1 : 0;
  }
}

// Whether every propagator slice of the launch is separable (k_prop_factors256's verdicts).  The
// multislice kernels come in two instantiations, FAC (every slice through its factors) and not (every
// entry as given, as when p.pfac is null); with factors checked, both are launched and the one that
// does not match the verdicts returns at once, so the host never waits for them.
__device__ __forceinline__ bool all_separable(const Params& p) {
  if (p.pfac == nullptr) return false;
  const int n = (p.S - 1) * kFacParts;
  bool all = true;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + (int)(threadIdx.x & 63);
    all = all && __ballot(i < n && p.pfac_ok[i] == 0) == 0ull;
  }
  return all;
}

// Forward multislice of one (pattern, mode); the modes (p.fwd_mode) are those of k_forward<N>.
template <bool DESCAN, bool PG = false, bool FAC = false>
__global__ __launch_bounds__(T) void k_forward256(Params p) {
  if (all_separable(p) != FAC) return;  // the other instantiation runs this launch
  __shared__ SmemOne sm;
  init_tw(sm);
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
  cf* Wb = p.scratch + (size_t)item * N2;
  const int mode = p.fwd_mode;
  const bool far = mode != FWD_RECOMPUTE;
  const float inv_n2 = 1.0f / (float)N2;
  const size_t HW = (size_t)p.H * p.W;
  const int j = row_j();
  const int q = obj_col_base(c0, p.W);
  // where the incident wave of slice s is kept (nullptr: not kept)
  auto slot = [&](int s) -> cf* {
    if (mode == FWD_STORE_ALL || mode == FWD_RECOMPUTE) return p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N2;
    if (mode == FWD_CHECKPOINT && s % p.ck == 0) return p.ckpt + ((size_t)item * p.nck + s / p.ck) * N2;
    return nullptr;
  };

  if (far) {
    // psi_0 = ifft2(fft2(P_m) * ramp_b): inverse column DFT here, inverse row DFT in the first row pass
    const cf* rr = p.ramps + (size_t)b * 2 * N;  // row factors, then column factors
    cf rw[16], rcol;
    col_pass(
        p.pk + (size_t)m * N2, Wb, sm,
        [&](int, cf(&v)[16], auto&& next) {
#pragma unroll
          for (int e = 0; e < 16; ++e) v[e] = cscale(cmul(v[e], cmul(rw[e], rcol)), inv_n2);
          next();
          fft256<true, true>(v, sm);
        },
        [&](int kc) {
          rcol = rr[N + kc];
#pragma unroll
          for (int e = 0; e < 16; ++e) rw[e] = rr[col_j() + 16 * e];
        });
  }

  for (int s = p.s0; s < p.s1; ++s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* keep = slot(s);
    const bool transmit = far || s + 1 < p.s1;  // a recomputed segment needs only its incident waves
    const bool exit_ramp = DESCAN && far && s == p.S - 1;
    // the transmit pass (op_t) gets its object values from pre_obj, issued before the chunk's inverse
    // row DFT; the other row passes (a recomputed segment's first and last slice) use op, which loads
    // them after it
    cf ov[16];
    auto pre_obj = [&](int r) {
      const unsigned orow = obj_row_off(r0, r, p.H, p.W);
#pragma unroll
      for (int e = 0; e < 16; ++e) ov[e] = obj_at(O, orow + obj_col_off(q, e, p.W));
    };
    auto op_t = [&](int r, cf(&v)[16]) {
      if (keep != nullptr) {
#pragma unroll
        for (int e = 0; e < 16; ++e) st_stream(keep + r * N + j + 16 * e, v[e]);
      }
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = cmul(ov[e], v[e]);
      if (exit_ramp) {
        const float fr = fftfreq1<N>(r);
#pragma unroll
        for (int e = 0; e < 16; ++e)
          v[e] = cmul(v[e], ramp2(fr, fftfreq1<N>(j + 16 * e), p.descan[2 * b], p.descan[2 * b + 1]));
      }
    };
    auto op = [&](int r, cf(&v)[16]) {
      // object values first: a load issued after the wave stores would wait for them (one vmcnt)
      cf o[16];
      if (transmit) {
        const unsigned orow = obj_row_off(r0, r, p.H, p.W);
#pragma unroll
        for (int e = 0; e < 16; ++e) o[e] = obj_at(O, orow + obj_col_off(q, e, p.W));
      }
      if (keep != nullptr) {
#pragma unroll
        for (int e = 0; e < 16; ++e) st_stream(keep + r * N + j + 16 * e, v[e]);
      }
      if (transmit) {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = cmul(o[e], v[e]);
        if (exit_ramp) {
          const float fr = fftfreq1<N>(r);
#pragma unroll
          for (int e = 0; e < 16; ++e)
            v[e] = cmul(v[e], ramp2(fr, fftfreq1<N>(j + 16 * e), p.descan[2 * b], p.descan[2 * b + 1]));
        }
      }
    };
    if (mode == FWD_RECOMPUTE && s == p.s0) {
      // a segment of an earlier forward starts from its checkpoint wave (natural order)
      const cf* ck = p.ckpt + ((size_t)item * p.nck + p.s0 / p.ck) * N2;
      if (transmit)
        row_pass<false, true>(ck, Wb, sm, op);
      else
        row_pass<false, false>(ck, nullptr, sm, op);
    } else if (transmit) {
      row_pass<true, true>(Wb, Wb, sm, op_t, pre_obj);
    } else {
      row_pass<true, false>(Wb, nullptr, sm, op);
    }
    if (s + 1 < p.s1) {
      const cf* pr = p.prop + (size_t)s * N2;
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + s) * N2 : nullptr;
      cf pq[FAC ? 1 : 16];
      if constexpr (FAC) {
        // separable: a[kr] times b[kc] / N^2, both from LDS (one factor per thread, written before the
        // pass's first barrier): no global load in the pass but its copies
        const cf f = p.pfac[(size_t)s * 2 * N + threadIdx.x];
        sm.pfac[threadIdx.x] = threadIdx.x < N ? f : cscale(f, inv_n2);
      }
      col_pass(
          Wb, Wb, sm,
          [&](int kc, cf(&v)[16], auto&& next) {
            fft256<false, true>(v, sm);
            const int jj = col_j();
            if constexpr (PG) {  // fft2 of the transmitted wave, for the propagator gradient
#pragma unroll
              for (int e = 0; e < 16; ++e) st_stream(xs + (jj + 16 * e) * N + kc, v[e]);
            }
            if constexpr (FAC) {
#pragma unroll
              for (int e = 0; e < 16; ++e) v[e] = cmul(v[e], cmul(sm.pfac[jj + 16 * e], sm.pfac[N + kc]));
            } else {
#pragma unroll
              for (int e = 0; e < 16; ++e) v[e] = cscale(cmul(v[e], pq[e]), inv_n2);
            }
            next();
            fft256<true, true>(v, sm);
          },
          [&](int kc) {
            if constexpr (!FAC) {
#pragma unroll
              for (int e = 0; e < 16; ++e) pq[e] = pr[(col_j() + 16 * e) * N + kc];
            }
          });
    } else if (far) {
      const float inv_n = 1.0f / (float)N;
      col_pass(Wb, p.X + (size_t)item * N2, sm, [&](int, cf(&v)[16], auto&& next) {
        next();
        fft256<false, true>(v, sm);
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = cscale(v[e], inv_n);
      });
    }
  }
}

// Adjoint of one (pattern, mode) over the slice segment [s0, s1) (see k_adjoint<N>).  The top segment
// starts from the far field; a lower one from the row-transformed g that the segment above left in X.
// The bottom segment ends with G = fft2(g_psi_0): probe-gradient spectrum conj(ramp) G over X and the
// position-gradient partial sums; any other segment leaves its row-transformed g in X.
template <bool DESCAN, bool PG = false, bool FAC = false>
__global__ __launch_bounds__(T) void k_adjoint256(Params p) {
  if (all_separable(p) != FAC) return;  // the other instantiation runs this launch
  __shared__ SmemOne sm;
  init_tw(sm);
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int b = item / p.M, m = item - b * p.M;
  int r0, c0;
  float dr, dc;
  origin_and_shift<N>(p.pos, b, r0, c0, dr, dc);
  cf* Wb = p.scratch + (size_t)item * N2;
  cf* Xs = p.X + (size_t)item * N2;
  const float inv_n2 = 1.0f / (float)N2;
  const size_t HW = (size_t)p.H * p.W;
  const int j = row_j();
  const int q = obj_col_base(c0, p.W);
  const bool top = p.s1 == p.S;

  const bool analytic = (p.flags & QEM_FLAG_ANALYTIC) != 0;
  if (top && analytic) {
    // the projection kernel left the Fourier-space residual X' - X over X: g = ifft2_ortho(residual)
    const float inv_n = 1.0f / (float)N;
    col_pass(Xs, Wb, sm, [&](int, cf(&v)[16], auto&& next) {
      next();
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = cscale(v[e], inv_n);
      fft256<true, true>(v, sm);
    });
  } else if (top) {
    // g_x = (1/N) F^H (2 dL/dI X): inverse column DFT here, inverse row DFT in the first row pass
    const float* gI = p.gI + (size_t)b * N2;
    const float sc = 2.0f * p.dloss[0] / (float)N;
    float gv[16];
    col_pass(
        Xs, Wb, sm,
        [&](int, cf(&v)[16], auto&& next) {
#pragma unroll
          for (int e = 0; e < 16; ++e) v[e] = cscale(v[e], sc * gv[e]);
          next();
          fft256<true, true>(v, sm);
        },
        [&](int kc) {
#pragma unroll
          for (int e = 0; e < 16; ++e) gv[e] = gI[(col_j() + 16 * e) * N + kc];
        });
  }

  double dar = 0.0, dac = 0.0;
  if (DESCAN && top) {
    // descan (applied after the last slice in the forward), in a pass of its own: its gradient
    // 2 pi sum f Im(conj(g) x_out) with x_out = ramp O_{S-1} psi_{S-1}, and g <- conj(ramp) g
    const cf* O = p.obj + (size_t)(p.S - 1) * HW;
    const cf* A = p.psi + ((size_t)item * p.wslots + (p.S - 1 - p.s0)) * N2;
    const float ddr = p.descan[2 * b];
    if (threadIdx.x < N) sm.dcol[threadIdx.x] = expm2pi(fftfreq1<N>(threadIdx.x) * p.descan[2 * b + 1]);
    __syncthreads();
    row_pass<true, true>(Wb, Wb, sm, [&](int r, cf(&v)[16]) {
      const unsigned orow = obj_row_off(r0, r, p.H, p.W);
      const float fr = fftfreq1<N>(r);
      const cf drow = expm2pi(fr * ddr);
      float sim = 0.0f, sfim = 0.0f;  // this row's sum of Im(...) and of fc Im(...)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const float fc = fftfreq1<N>(j + 16 * e);
        const cf d = cmul(drow, sm.dcol[j + 16 * e]);
        const cf xo = cmul(cmul(obj_at(O, orow + obj_col_off(q, e, p.W)), A[r * N + j + 16 * e]), d);
        const float im = cconjmul(v[e], xo).y;
        sim += im;
        sfim += fc * im;
        // 1/N: the forward row DFT below is not followed by a column pass, so it must leave R g / N
        // (R = unnormalised row DFT), the form every row pass expects (R^H R = N)
        v[e] = cscale(cmulc(v[e], d), 1.0f / (float)N);
      }
      dar += (double)(fr * sim);
      dac += (double)sfim;
    });
  }
  for (int s = p.s1 - 1; s >= p.s0; --s) {
    const cf* O = p.obj + (size_t)s * HW;
    cf* A = p.psi + ((size_t)item * p.wslots + (s - p.s0)) * N2;
    // g_x of slice s in natural order: object-gradient patch conj(psi_s) g over psi_s, g_psi = conj(O_s) g;
    // a row chunk's stored-wave and object values are loaded by pre, before its inverse row DFT
    // (loading them a whole chunk ahead needs 64 more registers and measured slower)
    cf pv[16], ov[16];
    row_pass<true, true>(
        (!top && s == p.s1 - 1) ? (const cf*)Xs : (const cf*)Wb, Wb, sm,
        [&](int r, cf(&v)[16]) {
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            st_stream(A + r * N + j + 16 * e, cconjmul(pv[e], v[e]));
            v[e] = cconjmul(ov[e], v[e]);
          }
        },
        [&](int r) {
#pragma unroll
          for (int e = 0; e < 16; ++e) pv[e] = ld_stream(A + r * N + j + 16 * e);
          const unsigned orow = obj_row_off(r0, r, p.H, p.W);
#pragma unroll
          for (int e = 0; e < 16; ++e) ov[e] = obj_at(O, orow + obj_col_off(q, e, p.W));
        });
    if (s > 0) {
      // g_x <- ifft2(fft2(g_psi) conj(prop_{s-1})); at a segment bottom it goes to X for the segment below
      const cf* pr = p.prop + (size_t)(s - 1) * N2;
      cf* xs = PG ? p.pspec + ((size_t)item * (p.S - 1) + (s - 1)) * N2 : nullptr;
      cf q[FAC ? 1 : 16];
      cf xq[PG ? 16 : 1];
      if constexpr (FAC) {
        // separable: conj(a[kr] b[kc] / N^2), both factors from LDS as in the forward
        const cf f = p.pfac[(size_t)(s - 1) * 2 * N + threadIdx.x];
        sm.pfac[threadIdx.x] = threadIdx.x < N ? f : cscale(f, inv_n2);
      }
      col_pass(
          Wb, s == p.s0 ? Xs : Wb, sm,
          [&](int kc, cf(&v)[16], auto&& next) {
        const int jj = col_j();
        fft256<false, true>(v, sm);
        if constexpr (PG) {  // this item's dL/dP_{s-1}: conj(fft2(x_{s-1})) fft2(g_psi_s) / N^2
#pragma unroll
          for (int e = 0; e < 16; ++e) st_stream(xs + (jj + 16 * e) * N + kc, cscale(cconjmul(xq[e], v[e]), inv_n2));
        }
        if constexpr (FAC) {
#pragma unroll
          for (int e = 0; e < 16; ++e) v[e] = cmulc(v[e], cmul(sm.pfac[jj + 16 * e], sm.pfac[N + kc]));
        } else {
#pragma unroll
          for (int e = 0; e < 16; ++e) v[e] = cscale(cmulc(v[e], q[e]), inv_n2);
        }
        next();
        fft256<true, true>(v, sm);
          },
          [&](int kc) {
            if constexpr (!FAC) {
#pragma unroll
              for (int e = 0; e < 16; ++e) q[e] = pr[(col_j() + 16 * e) * N + kc];
            }
            if constexpr (PG) {
#pragma unroll
              for (int e = 0; e < 16; ++e) xq[e] = ld_stream(xs + (col_j() + 16 * e) * N + kc);
            }
          });
    }
  }
  if constexpr (DESCAN) {
    if (top && (p.flags & QEM_FLAG_GRAD_DESCAN)) {
      const double tr = block_sum<T>(dar, sm.red);
      const double tc = block_sum<T>(dac, sm.red);
      if (threadIdx.x == 0) {
        p.gdesc_part[2 * item] = (float)(2.0 * M_PI * tr);
        p.gdesc_part[2 * item + 1] = (float)(2.0 * M_PI * tc);
      }
    }
  }
  if (p.s0 > 0) return;
  if (analytic) {
    // the reference sums g at the probe plane without the shift's adjoint: G = fft2(g_psi_0) over X
    col_pass(Wb, Xs, sm, [&](int, cf(&v)[16], auto&& next) {
      next();
      fft256<false, true>(v, sm);
    });
    return;
  }
  // G = fft2(g_psi_0); probe-gradient spectrum conj(ramp_b) G over X, position partials
  // (2 pi / N^2) sum f Im(conj(G) P_m ramp_b)
  const cf* pk = p.pk + (size_t)m * N2;
  const cf* rr = p.ramps + (size_t)b * 2 * N;  // row factors, then column factors
  double ar = 0.0, ac = 0.0;
  cf pkv[16], rrv[16], rcv;
  col_pass(
      Wb, Xs, sm,
      [&](int kc, cf(&v)[16]) {
        fft256<false, true>(v, sm);
        const int jj = col_j();
        const float fc = fftfreq1<N>(kc);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int kr = jj + 16 * e;
          const float fr = fftfreq1<N>(kr);
          const cf R = cmul(rrv[e], rcv);
          const cf G = v[e];
          const float im = cconjmul(G, cmul(pkv[e], R)).y;
          ar += (double)(fr * im);
          ac += (double)(fc * im);
          v[e] = cmulc(G, R);
        }
      },
      [&](int kc) {
        rcv = rr[N + kc];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int kr = col_j() + 16 * e;
          rrv[e] = rr[kr];
          pkv[e] = pk[kr * N + kc];
        }
      });
  if (p.flags & QEM_FLAG_GRAD_POSITIONS) {
    const double tr = block_sum<T>(ar, sm.red);
    const double tc = block_sum<T>(ac, sm.red);
    if (threadIdx.x == 0) {
      p.gpos_part[2 * item] = (float)(2.0 * M_PI * tr * (double)inv_n2);
      p.gpos_part[2 * item + 1] = (float)(2.0 * M_PI * tc * (double)inv_n2);
    }
  }
}

// grad_probe[m] = ifft2(gsum[m]) (1/N^2 included).
__global__ __launch_bounds__(T) void k_probe_grad256(Params p) {
  __shared__ Smem sm;
  init_tw(sm);
  const int m = blockIdx.x;
  cf* Wb = p.scratch + (size_t)m * N2;
  const float inv_n2 = 1.0f / (float)N2;
  col_pass(p.gsum + (size_t)m * N2, Wb, sm, [&](int, cf(&v)[16]) { fft256<true, true>(v, sm); });
  row_pass<true, false>(Wb, p.grad_probe + (size_t)m * N2, sm, [&](int, cf(&v)[16]) {
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = cscale(v[e], inv_n2);
  });
}

// Centring of measured amplitudes (preprocess.hip, n = 256): for each pattern, a = sqrt(max(I, 0))
// into W, forward row DFT, column DFT x shift ramp x inverse column DFT, inverse row DFT and the real
// part / N^2, clamped at 0, written fftshifted.  Workgroups loop over the patterns with their own W.
__global__ __launch_bounds__(T) void k_center256(int P, const float* __restrict__ I, const float* __restrict__ com,
                                                 float* __restrict__ out, float* __restrict__ sums, cf* ws) {
  __shared__ Smem sm;
  init_tw(sm);
  cf* Wb = ws + (size_t)blockIdx.x * N2;
  const float inv_n2 = 1.0f / (float)N2;
  for (int pat = blockIdx.x; pat < P; pat += gridDim.x) {
    const float* src = I + (size_t)pat * N2;
    double si = 0.0, sa = 0.0;
    for (int i = threadIdx.x; i < N2; i += T) {
      const float x = fmaxf(src[i], 0.0f);
      const float a = sqrtf(x);
      si += (double)x;
      sa += (double)a;
      Wb[i] = mk(a, 0.0f);
    }
    __syncthreads();  // W complete before the row pass reads it
    row_pass<false, true>(Wb, Wb, sm, NoOp());
    const float dr = -com[2 * pat], dc = -com[2 * pat + 1];
    col_pass(Wb, Wb, sm, [&](int kc, cf(&v)[16]) {
      fft256<false, true>(v, sm);
      const cf pc = expm2pi(fftfreq1<N>(kc) * dc);
      const int jj = col_j();
#pragma unroll
      for (int e = 0; e < 16; ++e) v[e] = cmul(v[e], cmul(expm2pi(fftfreq1<N>(jj + 16 * e) * dr), pc));
      fft256<true, true>(v, sm);
    });
    float* dst = out + (size_t)pat * N2;
    row_pass<true, false>(Wb, nullptr, sm, [&](int r, cf(&v)[16]) {
      const int rs = (r + N / 2) & (N - 1);
      const int j = row_j();
#pragma unroll
      for (int e = 0; e < 16; ++e) dst[rs * N + ((j + 16 * e + N / 2) & (N - 1))] = fmaxf(v[e].x * inv_n2, 0.0f);
    });
    if (sums != nullptr) {
      const double ti = block_sum<T>(si, sm.red);
      const double ta = block_sum<T>(sa, sm.red);
      if (threadIdx.x == 0) {
        sums[2 * pat] = (float)ti;
        sums[2 * pat + 1] = (float)ta;
      }
    }
  }
}

}  // namespace n256

unsigned long long params_signature_256() { return params_signature_local(); }

int launch_chain256_forward(const Params& p, hipStream_t st) {
  if (p.pfac != nullptr && p.S > 1) {
    // the factors and verdicts stay in the workspace for the adjoint and its recomputed segments
    hipLaunchKernelGGL(n256::k_prop_factors256, dim3(p.S - 1, n256::kFacParts), dim3(n256::kFacT), 0, st, p);
    if (int rc = check_last_launch("prop_factors256")) return rc;
  }
  // M probe-spectrum blocks, then the blocks of the (B, 2, N) shift-ramp table
  const int ramp_blocks = (p.B * 2 * n256::N + n256::T - 1) / n256::T;
  hipLaunchKernelGGL(n256::k_spectrum256, dim3(p.M + ramp_blocks), dim3(n256::T), 0, st, p);
  if (int rc = check_last_launch("spectrum256")) return rc;
  return launch_forward256_kernel(p, st);
}

// With checked factors (p.pfac) both instantiations are launched: the FAC one runs when every slice is
// separable, the other otherwise, and the one that does not run returns at its first instruction.
template <bool DESCAN, bool PG>
static void launch_fwd256(const Params& p, hipStream_t st) {
  if (p.pfac != nullptr)
    hipLaunchKernelGGL((n256::k_forward256<DESCAN, PG, true>), dim3(p.B * p.M), dim3(n256::T), 0, st, p);
  hipLaunchKernelGGL((n256::k_forward256<DESCAN, PG, false>), dim3(p.B * p.M), dim3(n256::T), 0, st, p);
}

template <bool DESCAN, bool PG>
static void launch_adj256(const Params& p, hipStream_t st) {
  if (p.pfac != nullptr)
    hipLaunchKernelGGL((n256::k_adjoint256<DESCAN, PG, true>), dim3(p.B * p.M), dim3(n256::T), 0, st, p);
  hipLaunchKernelGGL((n256::k_adjoint256<DESCAN, PG, false>), dim3(p.B * p.M), dim3(n256::T), 0, st, p);
}

int launch_forward256_kernel(const Params& p, hipStream_t st) {
  if (p.pspec != nullptr && p.descan != nullptr)
    launch_fwd256<true, true>(p, st);
  else if (p.pspec != nullptr)
    launch_fwd256<false, true>(p, st);
  else if (p.descan != nullptr && p.fwd_mode != FWD_RECOMPUTE)
    launch_fwd256<true, false>(p, st);
  else
    launch_fwd256<false, false>(p, st);
  return check_last_launch("forward256");
}

int launch_chain256_adjoint(const Params& p, hipStream_t st) {
  if (p.pspec != nullptr && p.descan != nullptr)
    launch_adj256<true, true>(p, st);
  else if (p.pspec != nullptr)
    launch_adj256<false, true>(p, st);
  else if (p.descan != nullptr)
    launch_adj256<true, false>(p, st);
  else
    launch_adj256<false, false>(p, st);
  return check_last_launch("adjoint256");
}

int launch_probe_grad256(const Params& p, hipStream_t st) {
  hipLaunchKernelGGL(n256::k_probe_grad256, dim3(p.M), dim3(n256::T), 0, st, p);
  return check_last_launch("probe_grad256");
}

int launch_center256(int P, int blocks, const float* I, const float* com, float* out, float* sums, cf* ws,
                     hipStream_t st) {
  hipLaunchKernelGGL(n256::k_center256, dim3(blocks), dim3(n256::T), 0, st, P, I, com, out, sums, ws);
  return check_last_launch("center256");
}

}  // namespace qem
