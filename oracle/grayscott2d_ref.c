/*
 * ORACLE -- test infrastructure only.  Not part of the product; only tests/,
 * the smoke check and bench.py's cpu_baseline leg may build, load or run it.
 *
 * Hand transcription of the C that formura v2.3.2 emits for
 * formura_amd/programs/grayscott2d.fmr (Gray-Scott, fields u and v, the form
 * of the reference's examples/pearson2/pearson2.fmr), one rank, no MPI,
 * no temporal blocking:
 *
 *   struct layout (fields in Data.Map key order u, v) .. Templates.hs:53-73
 *   Formura_Navi ....................................... Templates.hs:81-115
 *   Formura_Setup (init graph, sleeve 0) ............... Templates.hs:272-280
 *   Formura_Step: one loop nest per manifest node, u then v, each
 *     re-reading both fields (no cross-field fusion) .. Templates.hs:426-487
 *   Formura_Forward (noBlocking) ....................... Templates.hs:300-310, 336-346
 *   halo (no-MPI sendrecv: pack, send->recv, unpack) ... Functions.hs:265-290
 *
 * Per point (every literal a per-point double, ** -> pow, Translate.hs:175):
 *   d_xx(q) = ((q[i+1,j] + q[i-1,j]) - 2*q)/(dx*dx)      (same for y)
 *   u' = u + dt*(((ru*(d_xx u + d_yy u)) - (u*v)*v) + f*(1-u))
 *   v' = v + dt*(((rv*(d_xx v + d_yy v)) + (u*v)*v) - (f+k)*v)
 *   ru = 2*10**(-5), rv = 1*10**(-5), f = 0.04, k = 0.06, dt = 1,
 *   dx = length/N with length_per_node = N/100 (dx = 1/100 exactly).
 * Compile-time parameter: NX (= NY) grid_per_node.
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include <math.h>

#ifndef NX
#define NX 64
#endif
#ifndef NY
#define NY NX
#endif

#define Ns 1
#define L1 NX
#define L2 NY

typedef struct {
  double u[L1][L2];
  double v[L1][L2];
} Formura_Grid_Struct;
typedef struct {
  int time_step;
  int lower_x;
  int lower_y;
  int upper_x;
  int upper_y;
  double space_interval_x;
  double space_interval_y;
  int my_rank;
  int offset_x;
  int offset_y;
  double length_x;
  double length_y;
  int total_grid_x;
  int total_grid_y;
} Formura_Navi;

typedef struct {
  double u[L1 + 2 * Ns][L2 + 2 * Ns];
  double v[L1 + 2 * Ns][L2 + 2 * Ns];
} Formura_Buff;
typedef struct {
  double u;
  double v;
} Formura_Comm_Buff;
static Formura_Buff buff;
static Formura_Comm_Buff send_buf1_p1_0[2 * Ns][L2], recv_buf1_m1_0[2 * Ns][L2];
static Formura_Comm_Buff send_buf1_0_p1[L1][2 * Ns], recv_buf1_0_m1[L1][2 * Ns];
static Formura_Comm_Buff send_buf1_p1_p1[2 * Ns][2 * Ns], recv_buf1_m1_m1[2 * Ns][2 * Ns];
Formura_Grid_Struct formura_data;

double to_pos_x(int i, Formura_Navi n) { return n.space_interval_x * ((i + n.offset_x) % n.total_grid_x); }
double to_pos_y(int i, Formura_Navi n) { return n.space_interval_y * ((i + n.offset_y) % n.total_grid_y); }

/* init: u = if isCenter(i,j) then 1/2 else 1.0, v = if isCenter then 1/4 else 0.0,
 * isCenter = (fabs(total_grid_x/2 - i) < w) && (fabs(total_grid_y/2 - j) < w),
 * w = total_grid_x/20 + 1/2; one nest per field (u, then v). */
static void Formura_Setup(Formura_Navi n, int block_offset_1, int block_offset_2) {
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) {
      double w = (double)L1 / 20.0 + 1.0 / 2.0;
      double ci = (double)L1 / 2.0 - (double)(i1 + n.offset_x + block_offset_1);
      double cj = (double)L2 / 2.0 - (double)(i2 + n.offset_y + block_offset_2);
      bool c = (fabs(ci) < w) && (fabs(cj) < w);
      formura_data.u[i1][i2] = c ? 1.0 / 2.0 : 1.0;
    }
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) {
      double w = (double)L1 / 20.0 + 1.0 / 2.0;
      double ci = (double)L1 / 2.0 - (double)(i1 + n.offset_x + block_offset_1);
      double cj = (double)L2 / 2.0 - (double)(i2 + n.offset_y + block_offset_2);
      bool c = (fabs(ci) < w) && (fabs(cj) < w);
      formura_data.v[i1][i2] = c ? 1.0 / 4.0 : 0.0;
    }
}

static void Formura_Step(Formura_Buff *buff, Formura_Grid_Struct *rslt, Formura_Navi n, int block_offset_1,
                         int block_offset_2) {
  (void)n; (void)block_offset_1; (void)block_offset_2;
  /* manifest node u_next */
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1) {
    for (int i2 = 0; i2 < L2; i2 += 1) {
      double a0 = buff->u[i1 + 1][i2 + 1];
      double a1 = 1.0;                       /* dt */
      double a2 = 2.0;
      double a3 = 10.0;
      double a4 = 5.0;
      double a5 = -a4;
      double a6 = pow(a3, a5);
      double a7 = a2 * a6;                   /* ru */
      double a8 = buff->u[i1 + 2][i2 + 1];
      double a9 = buff->u[i1 + 0][i2 + 1];
      double a10 = a8 + a9;
      double a11 = 2.0;
      double a12 = a11 * a0;
      double a13 = a10 - a12;
      double a14 = 1.0 / 100.0;              /* dx */
      double a15 = a14 * a14;
      double a16 = a13 / a15;
      double a17 = buff->u[i1 + 1][i2 + 2];
      double a18 = buff->u[i1 + 1][i2 + 0];
      double a19 = a17 + a18;
      double a20 = 2.0;
      double a21 = a20 * a0;
      double a22 = a19 - a21;
      double a23 = 1.0 / 100.0;              /* dy */
      double a24 = a23 * a23;
      double a25 = a22 / a24;
      double a26 = a16 + a25;
      double a27 = a7 * a26;
      double a28 = buff->v[i1 + 1][i2 + 1];
      double a29 = a0 * a28;
      double a30 = a29 * a28;
      double a31 = a27 - a30;
      double a32 = 0.04;                     /* f */
      double a33 = 1.0;
      double a34 = a33 - a0;
      double a35 = a32 * a34;
      double a36 = a31 + a35;
      double a37 = a1 * a36;
      double a38 = a0 + a37;
      rslt->u[i1][i2] = a38;
    }
  }
  /* manifest node v_next */
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1) {
    for (int i2 = 0; i2 < L2; i2 += 1) {
      double a0 = buff->v[i1 + 1][i2 + 1];
      double a1 = 1.0;
      double a2 = 1.0;
      double a3 = 10.0;
      double a4 = 5.0;
      double a5 = -a4;
      double a6 = pow(a3, a5);
      double a7 = a2 * a6;                   /* rv */
      double a8 = buff->v[i1 + 2][i2 + 1];
      double a9 = buff->v[i1 + 0][i2 + 1];
      double a10 = a8 + a9;
      double a11 = 2.0;
      double a12 = a11 * a0;
      double a13 = a10 - a12;
      double a14 = 1.0 / 100.0;
      double a15 = a14 * a14;
      double a16 = a13 / a15;
      double a17 = buff->v[i1 + 1][i2 + 2];
      double a18 = buff->v[i1 + 1][i2 + 0];
      double a19 = a17 + a18;
      double a20 = 2.0;
      double a21 = a20 * a0;
      double a22 = a19 - a21;
      double a23 = 1.0 / 100.0;
      double a24 = a23 * a23;
      double a25 = a22 / a24;
      double a26 = a16 + a25;
      double a27 = a7 * a26;
      double a28 = buff->u[i1 + 1][i2 + 1];
      double a29 = a28 * a0;
      double a30 = a29 * a0;
      double a31 = a27 + a30;
      double a32 = 0.04;
      double a33 = 0.06;
      double a34 = a32 + a33;                /* f+k */
      double a35 = a34 * a0;
      double a36 = a31 - a35;
      double a37 = a1 * a36;
      double a38 = a0 + a37;
      rslt->v[i1][i2] = a38;
    }
  }
}

void Formura_Init(int *argc, char ***argv, Formura_Navi *n) {
  (void)argc; (void)argv;
  n->time_step = 0;
  n->lower_x = 0;
  n->lower_y = 0;
  n->upper_x = L1;
  n->upper_y = L2;
  n->space_interval_x = ((double)L1 / 100.0) / L1; /* double(length)/N, NumericalConfig.hs:127 */
  n->space_interval_y = ((double)L2 / 100.0) / L2;
  n->my_rank = 0;
  n->offset_x = 0;
  n->offset_y = 0;
  n->length_x = (double)L1 / 100.0;
  n->length_y = (double)L2 / 100.0;
  n->total_grid_x = L1;
  n->total_grid_y = L2;
  Formura_Setup(*n, 0, 0);
}

void Formura_Forward(Formura_Navi *n) {
  const Formura_Grid_Struct *src = &formura_data;
  /* isendrecv: pack the high edges, send -> recv (Functions.hs:270-282) */
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) {
      send_buf1_p1_0[i1][i2].u = src->u[i1 + L1 - 2 * Ns][i2];
      send_buf1_p1_0[i1][i2].v = src->v[i1 + L1 - 2 * Ns][i2];
    }
  for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) recv_buf1_m1_0[i1][i2] = send_buf1_p1_0[i1][i2];
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < 2 * Ns; i2 += 1) {
      send_buf1_0_p1[i1][i2].u = src->u[i1][i2 + L2 - 2 * Ns];
      send_buf1_0_p1[i1][i2].v = src->v[i1][i2 + L2 - 2 * Ns];
    }
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < 2 * Ns; i2 += 1) recv_buf1_0_m1[i1][i2] = send_buf1_0_p1[i1][i2];
  for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
    for (int i2 = 0; i2 < 2 * Ns; i2 += 1) {
      send_buf1_p1_p1[i1][i2].u = src->u[i1 + L1 - 2 * Ns][i2 + L2 - 2 * Ns];
      send_buf1_p1_p1[i1][i2].v = src->v[i1 + L1 - 2 * Ns][i2 + L2 - 2 * Ns];
      recv_buf1_m1_m1[i1][i2] = send_buf1_p1_p1[i1][i2];
    }
  /* waitAndCopy: recv -> buff at (0 if b_k else 2s) (Functions.hs:284-290) */
  for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) {
      buff.u[i1][i2 + 2 * Ns] = recv_buf1_m1_0[i1][i2].u;
      buff.v[i1][i2 + 2 * Ns] = recv_buf1_m1_0[i1][i2].v;
    }
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < 2 * Ns; i2 += 1) {
      buff.u[i1 + 2 * Ns][i2] = recv_buf1_0_m1[i1][i2].u;
      buff.v[i1 + 2 * Ns][i2] = recv_buf1_0_m1[i1][i2].v;
    }
  for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
    for (int i2 = 0; i2 < 2 * Ns; i2 += 1) {
      buff.u[i1][i2] = recv_buf1_m1_m1[i1][i2].u;
      buff.v[i1][i2] = recv_buf1_m1_m1[i1][i2].v;
    }
  /* interior copy: buff[i+2s] = data[i] (Templates.hs:343) */
#pragma omp parallel for collapse(2)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1) {
      buff.u[i1 + 2 * Ns][i2 + 2 * Ns] = formura_data.u[i1][i2];
      buff.v[i1 + 2 * Ns][i2 + 2 * Ns] = formura_data.v[i1][i2];
    }
  Formura_Step(&buff, &formura_data, *n, 0, 0);
  n->offset_x = (n->offset_x - Ns + n->total_grid_x) % n->total_grid_x;
  n->offset_y = (n->offset_y - Ns + n->total_grid_y) % n->total_grid_y;
  n->time_step += 1;
}

void Formura_Finalize() {}
