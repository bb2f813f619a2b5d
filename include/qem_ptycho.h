/*
 * qem_ptycho.h — C ABI of the MI355X (gfx950) multislice mixed-probe ptychography kernels.
 *
 * The library replaces, for one mini-batch of scan positions, the reference's per-batch
 * forward/loss call and its torch-autograd gradient:
 *
 *   quantem/diffractive_imaging/ptychography.py:220-229     (dset.forward -> probe/obj forward ->
 *                                                            forward_operator -> detector -> error_estimate)
 *   quantem/diffractive_imaging/ptychography.py:327-329     (loss.backward() through all of the above)
 *
 * Conventions (all tensors row-major, C-contiguous, device memory, little endian):
 *   complex64 = interleaved {float re, float im}.
 *   obj          (S, Ho, Wo)  complex64  transmission of each slice (after the caller's hard constraints)
 *   probe        (M, N, N)    complex64  probe modes, corner-centred real space
 *   positions    (B, 2)       float32    scan positions in object pixels (row, col); the integer origin is
 *                                        round-half-even(pos) and the sub-pixel shift is pos - origin
 *   propagators  (S-1, N, N)  complex64  Fresnel propagators in unshifted fftfreq order
 *   targets      (B, N, N)    float32    measured amplitudes (or intensities), fftshifted (DC at N/2, N/2)
 *   mask         (N, N)       float32    detector mask, fftshifted
 *   descan       (B, 2)       float32    optional descan shifts in pixels (NULL = none)
 * (N, N) stands for the pattern shape (Nr, Nc) = (roi, roi_cols) of the descriptor; fftshift moves index k
 * of an axis of length n to (k + n / 2) mod n, and patch pixel k lies at offset fftfreq(n, 1/n)[k] from the
 * origin (dataset_models.py:514-539, per axis).
 *
 * Every call is asynchronous on `stream` (a hipStream_t passed as void*), allocates nothing,
 * retains no pointer past its return, and gives results that do not depend on launch order
 * (all reductions run in a fixed order).  Return value 0 = success, otherwise a QEM_ERR_* code;
 * qem_last_error() returns a thread-local message describing the last failure.
 */
#ifndef QEM_PTYCHO_H
#define QEM_PTYCHO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QEM_OK = 0,
  QEM_ERR_ARG = 1,        /* invalid descriptor field or size */
  QEM_ERR_NULL = 2,       /* required pointer is NULL */
  QEM_ERR_ALIGN = 3,      /* pointer not aligned to 8 bytes */
  QEM_ERR_WORKSPACE = 4,  /* workspace smaller than qem_ptycho_workspace_bytes() */
  QEM_ERR_LAUNCH = 5,     /* HIP launch failure */
  QEM_ERR_UNSUPPORTED = 6 /* valid request this build does not implement (e.g. roi size) */
};

/* Loss types of PtychographyBase.error_estimate (ptychography_base.py:908-932). */
enum {
  QEM_LOSS_L2_AMPLITUDE = 0,
  QEM_LOSS_L1_AMPLITUDE = 1,
  QEM_LOSS_L2_INTENSITY = 2,
  QEM_LOSS_L1_INTENSITY = 3,
  QEM_LOSS_POISSON = 4
};

/* Descriptor flags. */
enum {
  QEM_FLAG_STORE_WAVES = 1u << 0, /* keep the per-slice incident waves for qem_ptycho_adjoint */
  QEM_FLAG_GRAD_POSITIONS = 1u << 1, /* adjoint also writes d loss / d positions */
  QEM_FLAG_GRAD_DESCAN = 1u << 2,    /* adjoint also writes d loss / d descan */
  QEM_FLAG_ANALYTIC = 1u << 3,       /* set by qem_ptycho_projection_update (internal) */
  QEM_FLAG_GRAD_PROPAGATORS = 1u << 4, /* with QEM_FLAG_STORE_WAVES and segment_slices 0: keep what
                                         qem_ptycho_propagator_gradient needs (one more (S-1)-slice
                                         spectrum per pattern and mode in the workspace) */
  QEM_FLAG_FULL_PROPAGATORS = 1u << 5  /* 256 x 256 patterns (every pass) and 128 x 128 patterns (the
                                         adjoint without descan or propagator gradient): apply every
                                         propagator entry as given.
                                         By default a propagator slice that is the product of a row
                                         and a column factor (the reference's Fresnel propagators,
                                         tilted or not: exp(-i pi lambda dz (kr^2 + kc^2)) times the
                                         tilt's per-axis terms) is applied as that product, the
                                         factors taken from its first column and first row, when for
                                         every slice the product reproduces each entry within 1e-6 of
                                         the largest magnitude among its 32 rows (16 at 128 x 128;
                                         checked on the GPU by
                                         qem_ptycho_forward_loss, kept in the workspace for the
                                         adjoint); otherwise every slice is applied entry by entry. */
};

typedef struct qem_ptycho_desc {
  int32_t num_slices;      /* S  >= 1 */
  int32_t num_modes;       /* M  >= 1 */
  int32_t batch;           /* B  >= 1, patterns in this call; B * M * Nr * Nc < 2^31 */
  int32_t roi;             /* Nr, pattern rows (roi_shape[0], dataset_models.py:519) */
  int32_t obj_rows;        /* Ho >= Nr */
  int32_t obj_cols;        /* Wo >= Nc */
  int32_t num_gpts;        /* P  total scan positions of the dataset (loss normalisation) */
  int32_t loss_type;       /* QEM_LOSS_* */
  float mean_intensity;    /* mean diffraction intensity of the dataset (> 0) */
  uint32_t flags;          /* QEM_FLAG_* */
  int32_t segment_slices;  /* with QEM_FLAG_STORE_WAVES: 0, 1 or >= S keep the incident wave of every
                              slice for the adjoint; k in [2, S) keeps only the wave entering every k-th
                              slice and the adjoint recomputes the waves of one k-slice segment at a
                              time (workspace B*M*(ceil(S/k) + k)*N*N*8 bytes instead of B*M*S*N*N*8,
                              one extra forward propagation per slice).  No reference counterpart:
                              torch autograd keeps every intermediate (ptychography_base.py:934-947). */
  int32_t roi_cols;        /* Nc, pattern columns (roi_shape[1]); 0 = Nc equal to Nr.  Square 32, 64, 128
                              and 256 run the fixed-size kernels; any other Nr x Nc with 2 <= Nr, Nc <= 512
                              and prime factors 2, 3, 5, 7 only runs the generic-size kernels (mixed-radix
                              LDS FFTs; no qem_ptycho_projection_update there); other shapes return
                              QEM_ERR_UNSUPPORTED. */
} qem_ptycho_desc;

/* Bytes of device workspace a forward_loss/adjoint pair needs for this descriptor.
 * No reference counterpart: the reference allocates its temporaries through torch. */
int qem_ptycho_workspace_bytes(const qem_ptycho_desc* desc, size_t* bytes);

/* Forward model + loss for one batch.
 * Replaces ProbePixelated.forward (probe_models.py:718-720), ObjectPixelated.forward
 * (object_models.py:678-680, gather at 252-265), PtychographyBase.forward_operator
 * (ptychography_base.py:892-906, overlap_projection 934-947, _propagate_array 970-989),
 * DetectorPixelated.forward (detector_models.py:22-30) and PtychographyBase.error_estimate
 * (ptychography_base.py:908-932).
 *   loss        device float[1]  the batch loss
 *   intensities device float (B, N, N) fftshifted predicted intensities, or NULL
 * With QEM_FLAG_STORE_WAVES the workspace keeps what qem_ptycho_adjoint needs. */
int qem_ptycho_forward_loss(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                            const float* positions, const void* propagators, const float* targets,
                            const float* mask, const float* descan, float* loss, float* intensities,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Bytes of the static workspace of qem_ptycho_forward_loss_static (per GPU). */
#define QEM_STATIC_WORKSPACE_BYTES ((size_t)64 << 20)

/* qem_ptycho_forward_loss without a workspace argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The no-gradient forward only (desc->flags must be 0 or QEM_FLAG_FULL_PROPAGATORS: nothing is kept
 * for an adjoint); its workspace is a file-scope device array of QEM_STATIC_WORKSPACE_BYTES bytes that
 * each GPU holds from the first call on it until the process ends.  Requests whose workspace
 * (qem_ptycho_workspace_bytes with flags 0) exceeds it return QEM_ERR_WORKSPACE: at most 109 patterns
 * of 128 x 128 with 4 modes and 8 slices per call, 27 of 256 x 256 with 2 modes.  The call allocates nothing and
 * calls no allocating routine; the workspace's address is looked up (hipGetSymbolAddress) on the first
 * call on each device, for the device current at the call, which must be the device of `stream`.
 * Every call on one GPU shares that array and nothing in it needs to start at zero (every call
 * writes all it reads); calls that overlap on one GPU (other streams, other threads) race on it.
 * Arguments, results and error codes otherwise as qem_ptycho_forward_loss. */
int qem_ptycho_forward_loss_static(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                                   const float* positions, const void* propagators, const float* targets,
                                   const float* mask, const float* descan, float* loss, float* intensities,
                                   void* stream);

/* Gradient of the batch loss (torch convention: grad = dL/dRe + i dL/dIm), scaled by *dloss
 * (device float, the upstream gradient).  Replaces torch autograd of the chain above
 * (invoked at ptychography.py:328).  Must follow qem_ptycho_forward_loss with the same
 * descriptor (QEM_FLAG_STORE_WAVES set), inputs and workspace.  The call consumes the workspace
 * (gradient patches overwrite the stored waves, probe-gradient spectra the far field): a second
 * adjoint needs a new qem_ptycho_forward_loss first.
 *   grad_obj      (S, Ho, Wo) complex64, fully written (zero where no patch lands)
 *   grad_probe     (M, N, N)   complex64
 *   grad_positions (B, 2) float32 or NULL (needs QEM_FLAG_GRAD_POSITIONS)
 *   grad_descan    (B, 2) float32 or NULL (needs QEM_FLAG_GRAD_DESCAN and descan != NULL) */
int qem_ptycho_adjoint(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                       const float* positions, const void* propagators, const float* descan,
                       const float* dloss, void* grad_obj, void* grad_probe, float* grad_positions,
                       float* grad_descan, void* workspace, size_t workspace_bytes, void* stream);

/* Gradient of the batch loss with respect to the propagators (torch convention), scaled by *dloss as
 * in qem_ptycho_adjoint, after a forward_loss / adjoint pair run with QEM_FLAG_STORE_WAVES |
 * QEM_FLAG_GRAD_PROPAGATORS (segment_slices 0) on the same descriptor, inputs and workspace:
 *   grad_propagators (S-1, N, N) complex64 = sum over patterns and modes of
 *                    conj(fft2(x_s)) fft2(dL/dpsi_{s+1}) / N^2,  x_s = O_s psi_s the transmitted wave
 * summed in a fixed order.  Replaces torch autograd into the propagator tensor, which the reference
 * recomputes from the probe tilt in every batch when learn_probe_tilt is set
 * (ptychography_base.py:898-899, probe_models.py:376-427); the caller chains it into the tilt. */
int qem_ptycho_propagator_gradient(const qem_ptycho_desc* desc, void* grad_propagators, void* workspace,
                                   size_t workspace_bytes, void* stream);

/* The reference's analytical update (autograd=False, ptychography.py:339-372) for one batch, after a
 * qem_ptycho_forward_loss call with QEM_FLAG_STORE_WAVES on the same descriptor, inputs and workspace:
 *   Fourier projection of every mode's exit wave onto the measured amplitudes `targets`
 *   (fourier_projection, ptychography.py:356-372), residual g = projected - exit wave (gradient_step,
 *   350-354), reverse slice replay with conj(O_s) and conj propagators (ObjectPixelated.backward,
 *   object_models.py:686-725).  Outputs (unnormalised; the caller forms the reference's updates):
 *     obj_update     (S, Ho, Wo) complex64  sum over modes and patterns of conj(psi_s) g_s scattered
 *     probe_update   (M, N, N)   complex64  sum over patterns of g at the probe plane
 *     probe_overlap  (S, M, Ho, Wo) float32  sum over patterns of |psi_s,m|^2 scattered (or NULL)
 *     patch_energy   (S, B)      float32  sum over each object patch of |O|^2 (or NULL)
 * The stored waves are consumed: run qem_ptycho_forward_loss again before another call. */
int qem_ptycho_projection_update(const qem_ptycho_desc* desc, const void* obj, const void* probe,
                                 const float* positions, const void* propagators, const float* targets,
                                 void* obj_update, void* probe_update, float* probe_overlap, float* patch_energy,
                                 void* workspace, size_t workspace_bytes, void* stream);

/* Batched 2-D complex64 FFT of `count` N x N arrays (N in {32, 64, 128}) with torch.fft.fft2 / ifft2
 * default normalisation (forward unscaled, inverse 1/N^2).  Replaces af.fft2 / af.ifft2
 * (core/utils/array_funcs.py:263-285) for those sizes. */
int qem_fft2_c64(int32_t n, int32_t count, int32_t inverse, const void* in, void* out, void* stream);

/* Measurement hook (no reference counterpart; the reference has no timing code): while active,
 * every forward_loss / adjoint call records a HIP event pair on its stream immediately around its
 * dominant kernel (the fused multislice forward, resp. adjoint), for at most max_launches calls
 * each.  qem_profile_end waits for the recorded events and returns the summed kernel times in
 * milliseconds and the number of timed launches.  Not thread-safe; one session at a time. */
int qem_profile_begin(int32_t max_launches);
int qem_profile_end(float* fwd_ms, int32_t* n_fwd, float* adj_ms, int32_t* n_adj);

/* Thread-local description of the last error (empty string if none). */
const char* qem_last_error(void);

/* ABI version (major * 100 + minor): 104. */
int32_t qem_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QEM_PTYCHO_H */
