/*
 * formura_hip.h -- device-resident extension of the generated Formura API.
 *
 * Every library that formura_amd generates for a program exports, besides the
 * reference API of the generated header (see the per-program headers in this
 * directory), the entry points below.  They are declared after the program's
 * Formura_Navi typedef (the generated header includes this file last).
 *
 * The reference API they sit beside (emitted by src/Formura/Generator/Templates.hs):
 *   double to_pos_<axis>(int, Formura_Navi);        Templates.hs:221-225
 *   void   Formura_Init(int *, char ***, Formura_Navi *);   Templates.hs:227-250
 *   void   Formura_Forward(Formura_Navi *);         Templates.hs:295-334
 *   void   Formura_Finalize();                      Templates.hs:383-386
 *   extern Formura_Grid_Struct formura_data;        Templates.hs:53-61
 *
 * The reference is CPU-only: it has no device and no counterpart of these
 * functions.  Its state is always host-visible, which this library reproduces
 * by default (Formura_Forward uploads formura_data before the step and
 * downloads it after).  Callers that keep the state in HBM switch that off.
 * Errors follow the reference convention: a message on stderr, then exit(1).
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

/* 1: the state stays in HBM between calls; Formura_Forward no longer copies
 * formura_data.  0 (default): every Formura_Forward syncs both ways.  Going
 * from 1 to 0 downloads the device state into formura_data first (host
 * writes made while resident and not uploaded are overwritten). */
void Formura_HIP_Set_Resident(int on);
/* Copy host formura_data to the device state / the device state to host. */
void Formura_HIP_Upload(void);
void Formura_HIP_Download(void);
/* Exactly k calls of Formura_Forward, with no host synchronisation in
 * between (one upload before and one download after when not resident).
 * The device time of the k calls is timed with HIP events on the library's
 * stream and returned by Formura_HIP_Last_Elapsed_Ms. */
void Formura_Forward_Steps(Formura_Navi *n, int k);
double Formura_HIP_Last_Elapsed_Ms(void);
/* JSON description of the program: grid, sleeve, fields, Navi layout. */
const char *Formura_HIP_Describe(void);
/* Device address of the current state image (the byte layout of
 * formura_data), and the HIP stream the library launches on. */
void *Formura_HIP_State_Device(void);
void *Formura_HIP_Stream(void);
/* Kernel launches issued by the last Formura_Forward_Steps call. */
long long Formura_HIP_Last_Launches(void);
/* FP64 totals of every state field over the grid (out[f], fields in the order
 * of formura_data), summed on the device in a fixed order (deterministic). */
void Formura_HIP_Totals(double *out);
/* Allocation-free variant of Formura_HIP_Totals.  No counterpart in the
 * reference (it has no device); for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Sums nfields (1..64) fields of `cells` (0..2^40) values
 * each, of elem_bytes 8 (FP64) or 4 (FP32), stored back to back at the device
 * address `state` (for instance Formura_HIP_State_Device()), into out[0..nfields)
 * in device memory, in Formura_HIP_Totals's order (the same bits), enqueued on
 * `stream` (a hipStream_t; 0 for the null stream); no host synchronisation.
 * Partial sums use a file-scope device array of 8 KiB per GPU, held until the
 * process ends; calls that overlap on one GPU race.  Needs no Formura_Init.
 * Sizes outside the limits: message on stderr, exit(1). */
#define FMR_TOTALS_STATIC_MAX_FIELDS 64
#define FMR_TOTALS_STATIC_MAX_CELLS (1LL << 40)
void Formura_HIP_Totals_Static(const void *state, int nfields, long long cells, int elem_bytes, double *out,
                               void *stream);
/* Mixed-precision libraries: FP64 square-root arguments (wave speeds from
 * pressure and density) found not > 0 since the library was loaded; 0 for
 * libraries without such checks. */
long long Formura_HIP_Positivity_Violations(void);

#ifdef __cplusplus
}
#endif
