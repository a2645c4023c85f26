/*
 * formura_adjoint.h -- forward and adjoint steps of one generated program,
 * over caller-owned device memory (for torch.autograd and checkpointing).
 *
 * No counterpart in the reference: Formura has no derivatives.  The forward
 * step here evaluates the same per-point instruction list as the generated
 * Formura_Step (Templates.hs:426-487) -- with the named literals listed in
 * Formura_AD_Describe() passed at run time -- and keeps the reference's
 * drifting frame (output i reads input i + d + off - 2s).  The adjoint step is
 * its exact transpose: the reverse-mode derivative of the instruction list,
 * gathered per input point.
 *
 * State and adjoint buffers hold NF fields of NTOT elements each (the byte
 * layout of the program's formura_data, in the library's real type).  Params
 * is a device array of doubles in the order of Formura_AD_Describe().  All
 * work is enqueued on `stream` (a hipStream_t, e.g. torch's current stream).
 * Errors: message on stderr, exit(1).
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

const char *Formura_AD_Describe(void);
long long Formura_AD_State_Bytes(void);
/* Number of per-workgroup slots of the parameter-gradient accumulator
 * (pgrad has Formura_AD_Param_Slots() * nparams doubles). */
int Formura_AD_Param_Slots(void);
/* nsteps forward steps from `in` into `out` (tmp: scratch of the same size). */
void Formura_AD_Forward(const void *in, void *out, void *tmp, int nsteps, const double *params_dev, void *stream);
/* lam_prev = (d step / d state)^T lam_next at state_prev; pgrad[slot][p] += dL/dparam_p. */
void Formura_AD_Adjoint(const void *state_prev, const void *lam_next, void *lam_prev, const double *params_dev,
                        double *pgrad, void *stream);

#ifdef __cplusplus
}
#endif
