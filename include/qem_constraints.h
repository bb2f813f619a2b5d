/*
 * qem_constraints.h — C ABI of the object hard constraints the reconstruction loop applies to the
 * object before every batch, on the autograd graph.
 *
 * Replaces ObjectConstraints.apply_hard_constraints for obj_type "complex" and "pure_phase"
 * (quantem/diffractive_imaging/object_models.py:289-301, no FOV mask):
 *     amp   = clamp(|O|, 0, 1)        ("complex")   or 1   ("pure_phase")
 *     phase = angle(O) - mean(angle(O))   (mean over every element of every slice)
 *     O'    = amp * exp(i * phase)
 * and torch autograd of it.  Conventions as in qem_ptycho.h: complex64 interleaved device arrays,
 * asynchronous on `stream`, no allocation; the caller passes a workspace of the queried size, and
 * the backward call needs the workspace its forward call filled (it holds the mean phase).  All
 * reductions run in a fixed order.
 */
#ifndef QEM_CONSTRAINTS_H
#define QEM_CONSTRAINTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Workspace bytes for n complex elements. */
int qem_object_constraints_workspace_bytes(int64_t n, size_t* bytes);

/* out = constrained obj (n complex64 elements).  pure_phase: 0 = "complex", 1 = "pure_phase". */
int qem_object_constraints_forward(int64_t n, int32_t pure_phase, const void* obj, void* out, void* workspace,
                                   size_t workspace_bytes, void* stream);

/* grad_obj = d loss / d obj (torch convention dL/dRe + i dL/dIm) from grad_out = d loss / d out. */
int qem_object_constraints_backward(int64_t n, int32_t pure_phase, const void* obj, const void* grad_out,
                                    void* grad_obj, void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QEM_CONSTRAINTS_H */
