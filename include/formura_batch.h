/*
 * formura_batch.h -- batched ensembles of one generated 2-D Formura program.
 *
 * No counterpart in the reference: the reference runs one instance of one
 * program per link unit (all state in the static formura_data,
 * Templates.hs:53-61).  A batch library advances many independent instances
 * -- uneven grids, per-instance values of the program's named literals -- in
 * one persistent launch per K steps, with the step loop captured in a
 * hipGraph.  Each instance follows the reference's semantics: periodic grid,
 * drifting frame (offset -= sleeve per step, Templates.hs:346).
 *
 * Errors: message on stderr, then exit(1), as in the reference.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef struct FormuraBatch FormuraBatch;

/* JSON: program, precision ("f64", "f32k" or "f32c"), fields, parameter names, tile,
 * steps per launch. */
const char *Formura_Batch_Describe(void);
/* ninst instances; instance i has grid n1[i] x n2[i] (multiples of the tile)
 * and parameters params[i*np .. i*np+np-1] (np = number of parameter names). */
FormuraBatch *Formura_Batch_Create(int ninst, const int *n1, const int *n2, const double *params);
/* Instance state as nfields row-major arrays of n1*n2 doubles.  Converted to
 * the library's precision; in the Kahan-compensated FP32 forms a value x is
 * stored as the FP32 pair (q, c) = (RN(x), RN(q - x)) and read back as q - c
 * (formed in FP64), the value the compensated accumulation carries. */
void Formura_Batch_Set_State(FormuraBatch *b, int inst, const double *fields);
void Formura_Batch_Get_State(FormuraBatch *b, int inst, double *fields);
/* Frame offsets (offset_x, offset_y) of an instance after the steps so far. */
void Formura_Batch_Offsets(FormuraBatch *b, int inst, int *offsets);
/* nsteps steps of every instance; returns device milliseconds (HIP events). */
double Formura_Batch_Step(FormuraBatch *b, int nsteps);
long long Formura_Batch_Last_Launches(FormuraBatch *b);
long long Formura_Batch_Cells(FormuraBatch *b);
void Formura_Batch_Destroy(FormuraBatch *b);

#ifdef __cplusplus
}
#endif
