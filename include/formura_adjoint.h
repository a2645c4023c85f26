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

/* Most steps one Formura_AD_Adjoint_Steps call reverses (1 for libraries
 * with one-point-per-work-item kernels). */
int Formura_AD_Levels(void);
/* nsteps (1..Formura_AD_Levels()) adjoint steps in one call from the state at
 * their start; scratch is work space of nsteps-1 state buffers for the inner
 * states the call recomputes (a library whose adjoint reads the state at the
 * point only writes the state two steps in and recomputes the others inside
 * its adjoint launch); lam_prev = dL/d(state0); pgrad += dL/dparams of the
 * nsteps steps. */
void Formura_AD_Adjoint_Steps(const void *state0, void *scratch, const void *lam_next, void *lam_prev, int nsteps,
                              const double *params_dev, double *pgrad, void *stream);
/* The two halves of Formura_AD_Adjoint_Steps, for a caller that runs the
 * forward work of one unit on one stream while another stream reverses the
 * previous unit: the inner states (nlevels = nsteps-1) ... */
void Formura_AD_Inner_States(const void *state0, void *scratch, int nlevels, const double *params_dev, void *stream);
/* ... and the adjoint steps over them. */
void Formura_AD_Adjoint_Levels(const void *state0, const void *scratch, const void *lam_next, void *lam_prev,
                               int nsteps, const double *params_dev, double *pgrad, void *stream);

/* Executes a reversal schedule: nops operations of 4 ints (kind, a, b, k):
 *   kind 0: bufs[b] = k forward steps from bufs[a]                (stream_f)
 *   kind 1: the adjoint of the k (<= Formura_AD_Levels()) steps that start
 *           at bufs[a]: inner states (or the one state the fused adjoint
 *           needs) on stream_f into one half of scratch (2*(levels-1) state
 *           buffers, halves used in turn), adjoint steps on stream_a from
 *           lam to lam2 (then the two swap)                      (stream_a)
 * Events order the two streams on every shared buffer, so the forward work of
 * later operations overlaps the adjoint ones (stream_f == stream_a: one
 * stream, no events; a library with the fused adjoint unit then runs up to two
 * forward launches of later operations -- a unit's level-2 state, a 2- or
 * 4-step recomputation -- inside each adjoint launch, when every earlier
 * operation touching their buffers has been issued before it).  lam holds dL/d(output) on entry; returns 0 if the
 * result dL/d(bufs of the first adjoint's state...) ends in lam, 1 if in lam2.
 * pgrad accumulates as in Formura_AD_Adjoint.  Returns after stream_a has
 * finished when the two streams differ. */
int Formura_AD_Reverse(const int *ops, int nops, void *const *bufs, int nbufs, void *tmp, void *scratch, void *lam,
                       void *lam2, const double *params_dev, double *pgrad, void *stream_a, void *stream_f);

#ifdef __cplusplus
}
#endif
