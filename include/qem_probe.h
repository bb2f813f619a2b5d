/*
 * qem_probe.h — C ABI of the probe's hard constraint in the reconstruction loop: orthogonalisation of
 * the probe modes, and its gradient.
 *
 * Replaces ProbeConstraints._probe_orthogonalization_constraint (quantem/diffractive_imaging/
 * probe_models.py:484-525), which the reference applies to the probe parameter on the autograd graph
 * every time the loop reads probe_model.probe (probe_models.py:452-459, 660-663): modified
 * Gram-Schmidt over the M modes, each result rescaled to its input mode's L2 norm, the modes sorted by
 * intensity (strongest first); torch autograd differentiates through it.
 *
 * Here the same map is computed as a thin QR factorisation from the Cholesky factor of the M x M Gram
 * matrix (accumulated in double, fixed summation order), so that each pass over the modes is parallel:
 * the results equal Gram-Schmidt's to float32 rounding for modes of full rank (a mode whose Gram-Schmidt
 * residual falls below 1e-12 is clamped as the reference clamps it, and has no defined gradient).
 *
 * Conventions as in qem_ptycho.h: device pointers, complex64 interleaved, asynchronous on `stream`, no
 * allocation, deterministic.  probe, out, grad_out, grad_probe: (modes, pixels) complex64, pixels = N*N.
 * Return 0 or a QEM_ERR_* code of qem_ptycho.h (QEM_ERR_ARG: modes outside [1, QEM_PROBE_MAX_MODES] or
 * pixels < 1; QEM_ERR_NULL / QEM_ERR_ALIGN: pointers (8-byte alignment); QEM_ERR_WORKSPACE).
 */
#ifndef QEM_PROBE_H
#define QEM_PROBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QEM_PROBE_MAX_MODES 16

/* Bytes of workspace of one forward / backward pair (a few KiB). */
int qem_probe_orthogonalize_workspace_bytes(int32_t modes, int64_t pixels, size_t* bytes);

/* out = the orthogonalised, rescaled and sorted modes.  The workspace keeps what the backward needs. */
int qem_probe_orthogonalize_forward(int32_t modes, int64_t pixels, const void* probe, void* out, void* workspace,
                                    size_t workspace_bytes, void* stream);

/* grad_probe = the gradient (torch convention, dL/dRe + i dL/dIm) with respect to `probe` of a loss
 * whose gradient with respect to `out` is grad_out.  Needs the workspace of the forward call on the same
 * probe (it is not consumed: the backward may run again). */
int qem_probe_orthogonalize_backward(int32_t modes, int64_t pixels, const void* probe, const void* grad_out,
                                     void* grad_probe, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEM_PROBE_H */
