/*
 * ORACLE -- test infrastructure only.  Not part of the product; only tests/,
 * the smoke check and bench.py's cpu_baseline leg may build, load or run it.
 *
 * Hand transcription of the C that formura v2.3.2 emits for
 * formura_amd/programs/heat3d.fmr (7-point FTCS heat equation, one field q),
 * one rank, no MPI (the no-MPI path copies send -> recv in place,
 * src/Formura/Generator/Functions.hs:279-281).  Written by reading the
 * emitter, not by running it (the Haskell generator cannot be built here):
 *
 *   header / struct layout ........ Templates.hs:23-116, Encode.hs:11-27
 *   Formura_Init / Setup .......... Templates.hs:227-250, 272-280
 *   Formura_Step (mkKernel) ....... Templates.hs:282-293, 426-487
 *   Formura_Forward, NoBlocking ... Templates.hs:300-310, 336-346
 *   Formura_Forward, TB ........... Templates.hs:311-334, updateWithTB 348-381
 *   halo sendrecv/isendrecv ....... Functions.hs:265-290, copy :250-256
 *
 * Compile-time parameters: NX NY NZ (grid_per_node), and for the temporally
 * blocked variant TB=1 with NT (temporal_blocking_interval) and GB1 GB2 GB3
 * (grid_per_block).  The per-point instruction list is the manifestation of
 * q_next = q + dt*(d_xx(q) + d_yy(q) + d_zz(q)) with
 * d_xx(q) = (q[i+1,j,k] + q[i-1,j,k] - 2*q[i,j,k])/dx**2, dt = dx*dx/8 and
 * length_per_node = [1,1,1]: every literal is a per-point double, ** is pow.
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
#ifndef NZ
#define NZ NX
#endif

#define Ns 1
#define L1 NX
#define L2 NY
#define L3 NZ

typedef struct {
  double q[L1][L2][L3];
} Formura_Grid_Struct;
typedef struct {
  int time_step;
  int lower_x;
  int lower_y;
  int lower_z;
  int upper_x;
  int upper_y;
  int upper_z;
  double space_interval_x;
  double space_interval_y;
  double space_interval_z;
  int my_rank;
  int offset_x;
  int offset_y;
  int offset_z;
  double length_x;
  double length_y;
  double length_z;
  int total_grid_x;
  int total_grid_y;
  int total_grid_z;
} Formura_Navi;

typedef struct {
  double q;
} Formura_Comm_Buff;

#if TB
/* TemporalBlocking: buff [gpb+2s], rslt [gpb], floor [N+2*s*nt], walls
 * [bpn(others)][nt][2s or gpb+2s] (Templates.hs:141-172) */
#define S_TB (Ns * NT)
#define BP1 ((L1 + 2 * S_TB) / GB1)
#define BP2 ((L2 + 2 * S_TB) / GB2)
#define BP3 ((L3 + 2 * S_TB) / GB3)
#if (L1 + 2 * S_TB) % GB1 || (L2 + 2 * S_TB) % GB2 || (L3 + 2 * S_TB) % GB3
#error "Inconsistent config"
#endif
typedef struct { double q[GB1 + 2 * Ns][GB2 + 2 * Ns][GB3 + 2 * Ns]; } Formura_Buff;
typedef struct { double q[GB1][GB2][GB3]; } Formura_Rslt;
typedef struct { double q[L1 + 2 * S_TB][L2 + 2 * S_TB][L3 + 2 * S_TB]; } Formura_Tmp_Floor;
typedef struct { double q[BP2][BP3][NT][2 * Ns][GB2 + 2 * Ns][GB3 + 2 * Ns]; } Formura_Tmp_Wall_1;
typedef struct { double q[BP1][BP3][NT][GB1 + 2 * Ns][2 * Ns][GB3 + 2 * Ns]; } Formura_Tmp_Wall_2;
typedef struct { double q[BP1][BP2][NT][GB1 + 2 * Ns][GB2 + 2 * Ns][2 * Ns]; } Formura_Tmp_Wall_3;
static Formura_Buff buff;
static Formura_Rslt rslt;
static Formura_Tmp_Floor tmp_floor;
static Formura_Tmp_Wall_1 tmp_wall_1;
static Formura_Tmp_Wall_2 tmp_wall_2;
static Formura_Tmp_Wall_3 tmp_wall_3;
#define SB S_TB
#define OUT1 GB1
#define OUT2 GB2
#define OUT3 GB3
typedef Formura_Rslt Formura_Out;
#else
typedef struct { double q[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns]; } Formura_Buff;
static Formura_Buff buff;
#define SB Ns
#define OUT1 L1
#define OUT2 L2
#define OUT3 L3
typedef Formura_Grid_Struct Formura_Out;
#endif

/* communication buffers for the sleeve SB, one per positive basis b
 * (Templates.hs:174-186): shape [2*SB if b_k else N_k] */
static Formura_Comm_Buff send_buf_p1_0_0[2 * SB][L2][L3], recv_buf_m1_0_0[2 * SB][L2][L3];
static Formura_Comm_Buff send_buf_0_p1_0[L1][2 * SB][L3], recv_buf_0_m1_0[L1][2 * SB][L3];
static Formura_Comm_Buff send_buf_0_0_p1[L1][L2][2 * SB], recv_buf_0_0_m1[L1][L2][2 * SB];
static Formura_Comm_Buff send_buf_p1_p1_0[2 * SB][2 * SB][L3], recv_buf_m1_m1_0[2 * SB][2 * SB][L3];
static Formura_Comm_Buff send_buf_p1_0_p1[2 * SB][L2][2 * SB], recv_buf_m1_0_m1[2 * SB][L2][2 * SB];
static Formura_Comm_Buff send_buf_0_p1_p1[L1][2 * SB][2 * SB], recv_buf_0_m1_m1[L1][2 * SB][2 * SB];
static Formura_Comm_Buff send_buf_p1_p1_p1[2 * SB][2 * SB][2 * SB], recv_buf_m1_m1_m1[2 * SB][2 * SB][2 * SB];

Formura_Grid_Struct formura_data;

double to_pos_x(int i, Formura_Navi n) { return n.space_interval_x * ((i + n.offset_x) % n.total_grid_x); }
double to_pos_y(int i, Formura_Navi n) { return n.space_interval_y * ((i + n.offset_y) % n.total_grid_y); }
double to_pos_z(int i, Formura_Navi n) { return n.space_interval_z * ((i + n.offset_z) % n.total_grid_z); }

/* Formura_Setup: the init graph q[i,j,k] = gauss(i*dx,j*dy,k*dz) at sleeve 0
 * (Templates.hs:272-280); LoadIndex k -> i_k + n.offset + block_offset (:481) */
static void Formura_Setup(Formura_Navi n, int block_offset_1, int block_offset_2, int block_offset_3) {
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1) {
    for (int i2 = 0; i2 < L2; i2 += 1) {
      for (int i3 = 0; i3 < L3; i3 += 1) {
        double dx = 1.0 / L1, dy = 1.0 / L2, dz = 1.0 / L3;
        double tx = L1, ty = L2, tz = L3;
        double x = (i1 + n.offset_x + block_offset_1) * dx;
        double y = (i2 + n.offset_y + block_offset_2) * dy;
        double z = (i3 + n.offset_z + block_offset_3) * dz;
        double ax = (x - tx * dx / 2) / (5 * dx);
        double ay = (y - ty * dy / 2) / (5 * dy);
        double az = (z - tz * dz / 2) / (5 * dz);
        double px = pow(ax, 2.0), py = pow(ay, 2.0), pz = pow(az, 2.0);
        double e = -px;
        e = e - py;
        e = e - pz;
        formura_data.q[i1][i2][i3] = exp(e);
      }
    }
  }
}

/* Formura_Step: one loop nest, range [-1,1] -> extent inputSize-2 = output
 * size; buff index i + d + abs(lo) (Templates.hs:473,486) */
static void Formura_Step(Formura_Buff *buff, Formura_Out *rslt, Formura_Navi n, int block_offset_1,
                         int block_offset_2, int block_offset_3) {
  (void)n; (void)block_offset_1; (void)block_offset_2; (void)block_offset_3;
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < OUT1; i1 += 1) {
    for (int i2 = 0; i2 < OUT2; i2 += 1) {
      for (int i3 = 0; i3 < OUT3; i3 += 1) {
        double a0 = 1.0 / L1;                         /* dx */
        double a1 = 1.0 / L2;                         /* dy */
        double a2 = 1.0 / L3;                         /* dz */
        double a3 = buff->q[i1 + 2][i2 + 1][i3 + 1];  /* q[i+1,j,k] */
        double a4 = buff->q[i1 + 0][i2 + 1][i3 + 1];  /* q[i-1,j,k] */
        double a5 = a3 + a4;
        double a6 = 2.0;
        double a7 = buff->q[i1 + 1][i2 + 1][i3 + 1];  /* q[i,j,k] */
        double a8 = a6 * a7;
        double a9 = a5 - a8;
        double a10 = 2.0;
        double a11 = pow(a0, a10);
        double a12 = a9 / a11;                        /* q_xx */
        double a13 = buff->q[i1 + 1][i2 + 2][i3 + 1];
        double a14 = buff->q[i1 + 1][i2 + 0][i3 + 1];
        double a15 = a13 + a14;
        double a16 = 2.0;
        double a17 = a16 * a7;
        double a18 = a15 - a17;
        double a19 = 2.0;
        double a20 = pow(a1, a19);
        double a21 = a18 / a20;                       /* q_yy */
        double a22 = buff->q[i1 + 1][i2 + 1][i3 + 2];
        double a23 = buff->q[i1 + 1][i2 + 1][i3 + 0];
        double a24 = a22 + a23;
        double a25 = 2.0;
        double a26 = a25 * a7;
        double a27 = a24 - a26;
        double a28 = 2.0;
        double a29 = pow(a2, a28);
        double a30 = a27 / a29;                       /* q_zz */
        double a31 = a0 * a0;
        double a32 = 8.0;
        double a33 = a31 / a32;                       /* dt */
        double a34 = a12 + a21;
        double a35 = a34 + a30;
        double a36 = a33 * a35;
        double a37 = a7 + a36;
        rslt->q[i1][i2][i3] = a37;
      }
    }
  }
}

/* isendrecv without MPI for sleeve S (Functions.hs:270-282): pack the high
 * edge of src for every positive basis and copy send -> recv. */
#define PACK(SEND, RECV, E1, E2, E3, O1, O2, O3)                                  \
  _Pragma("omp parallel for collapse(3)") for (int i1 = 0; i1 < (E1); i1 += 1)    \
    for (int i2 = 0; i2 < (E2); i2 += 1)                                          \
      for (int i3 = 0; i3 < (E3); i3 += 1) SEND[i1][i2][i3].q = src->q[i1 + (O1)][i2 + (O2)][i3 + (O3)]; \
  _Pragma("omp parallel for collapse(3)") for (int i1 = 0; i1 < (E1); i1 += 1)    \
    for (int i2 = 0; i2 < (E2); i2 += 1)                                          \
      for (int i3 = 0; i3 < (E3); i3 += 1) RECV[i1][i2][i3].q = SEND[i1][i2][i3].q;

static void isendrecv(const Formura_Grid_Struct *src, int S) {
  PACK(send_buf_p1_0_0, recv_buf_m1_0_0, 2 * S, L2, L3, L1 - 2 * S, 0, 0)
  PACK(send_buf_0_p1_0, recv_buf_0_m1_0, L1, 2 * S, L3, 0, L2 - 2 * S, 0)
  PACK(send_buf_0_0_p1, recv_buf_0_0_m1, L1, L2, 2 * S, 0, 0, L3 - 2 * S)
  PACK(send_buf_p1_p1_0, recv_buf_m1_m1_0, 2 * S, 2 * S, L3, L1 - 2 * S, L2 - 2 * S, 0)
  PACK(send_buf_p1_0_p1, recv_buf_m1_0_m1, 2 * S, L2, 2 * S, L1 - 2 * S, 0, L3 - 2 * S)
  PACK(send_buf_0_p1_p1, recv_buf_0_m1_m1, L1, 2 * S, 2 * S, 0, L2 - 2 * S, L3 - 2 * S)
  PACK(send_buf_p1_p1_p1, recv_buf_m1_m1_m1, 2 * S, 2 * S, 2 * S, L1 - 2 * S, L2 - 2 * S, L3 - 2 * S)
}

/* waitAndCopy (Functions.hs:284-290): recv -> target at offset (0 if b_k else 2S) */
#define UNPACK(RECV, E1, E2, E3, O1, O2, O3)                                      \
  _Pragma("omp parallel for collapse(3)") for (int i1 = 0; i1 < (E1); i1 += 1)    \
    for (int i2 = 0; i2 < (E2); i2 += 1)                                          \
      for (int i3 = 0; i3 < (E3); i3 += 1) tgt->q[i1 + (O1)][i2 + (O2)][i3 + (O3)] = RECV[i1][i2][i3].q;

#define WAIT_AND_COPY(TGT_T, tgt_ptr, S)                                          \
  do {                                                                            \
    TGT_T *tgt = (tgt_ptr);                                                       \
    UNPACK(recv_buf_m1_0_0, 2 * S, L2, L3, 0, 2 * S, 2 * S)                       \
    UNPACK(recv_buf_0_m1_0, L1, 2 * S, L3, 2 * S, 0, 2 * S)                       \
    UNPACK(recv_buf_0_0_m1, L1, L2, 2 * S, 2 * S, 2 * S, 0)                       \
    UNPACK(recv_buf_m1_m1_0, 2 * S, 2 * S, L3, 0, 0, 2 * S)                       \
    UNPACK(recv_buf_m1_0_m1, 2 * S, L2, 2 * S, 0, 2 * S, 0)                       \
    UNPACK(recv_buf_0_m1_m1, L1, 2 * S, 2 * S, 2 * S, 0, 0)                       \
    UNPACK(recv_buf_m1_m1_m1, 2 * S, 2 * S, 2 * S, 0, 0, 0)                       \
  } while (0)

void Formura_Init(int *argc, char ***argv, Formura_Navi *n) {
  (void)argc; (void)argv;
  n->time_step = 0;
  n->lower_x = 0;
  n->lower_y = 0;
  n->lower_z = 0;
  n->upper_x = L1;
  n->upper_y = L2;
  n->upper_z = L3;
  n->space_interval_x = 1.0 / L1; /* double(1.0)/N (NumericalConfig.hs:127) */
  n->space_interval_y = 1.0 / L2;
  n->space_interval_z = 1.0 / L3;
  n->my_rank = 0;
  n->offset_x = 0;
  n->offset_y = 0;
  n->offset_z = 0;
  n->length_x = 1.0;
  n->length_y = 1.0;
  n->length_z = 1.0;
  n->total_grid_x = L1;
  n->total_grid_y = L2;
  n->total_grid_z = L3;
  Formura_Setup(*n, 0, 0, 0);
}

#if !TB
/* noBlocking (Templates.hs:336-346) */
void Formura_Forward(Formura_Navi *n) {
  isendrecv(&formura_data, Ns);
  WAIT_AND_COPY(Formura_Buff, &buff, Ns);
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1)
      for (int i3 = 0; i3 < L3; i3 += 1) buff.q[i1 + 2 * Ns][i2 + 2 * Ns][i3 + 2 * Ns] = formura_data.q[i1][i2][i3];
  Formura_Step(&buff, &formura_data, *n, 0, 0, 0);
  n->offset_x = (n->offset_x - Ns + n->total_grid_x) % n->total_grid_x;
  n->offset_y = (n->offset_y - Ns + n->total_grid_y) % n->total_grid_y;
  n->offset_z = (n->offset_z - Ns + n->total_grid_z) % n->total_grid_z;
  n->time_step += 1;
}
#else
/* updateWithTB (Templates.hs:348-381): sequential nest over blocks j1,j2,j3 in
 * [lo,hi); floor offset gpb*(bpn-1) - gpb*j per axis. */
static void updateWithTB(int lo1, int hi1, int lo2, int hi2, int lo3, int hi3) {
  for (int j1 = lo1; j1 < hi1; j1 += 1) {
    for (int j2 = lo2; j2 < hi2; j2 += 1) {
      for (int j3 = lo3; j3 < hi3; j3 += 1) {
        const int f1 = GB1 * (BP1 - 1) - GB1 * j1;
        const int f2 = GB2 * (BP2 - 1) - GB2 * j2;
        const int f3 = GB3 * (BP3 - 1) - GB3 * j3;
        /* floor -> rslt */
#pragma omp parallel for collapse(3)
        for (int i1 = 0; i1 < GB1; i1 += 1)
          for (int i2 = 0; i2 < GB2; i2 += 1)
            for (int i3 = 0; i3 < GB3; i3 += 1) rslt.q[i1][i2][i3] = tmp_floor.q[i1 + f1][i2 + f2][i3 + f3];
        for (int it = 0; it < NT; it += 1) {
          /* rslt -> buff */
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < GB1; i1 += 1)
            for (int i2 = 0; i2 < GB2; i2 += 1)
              for (int i3 = 0; i3 < GB3; i3 += 1) buff.q[i1][i2][i3] = rslt.q[i1][i2][i3];
          /* walls -> high side of buff */
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < GB2 + 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < GB3 + 2 * Ns; i3 += 1) buff.q[i1 + GB1][i2][i3] = tmp_wall_1.q[j2][j3][it][i1][i2][i3];
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < GB1 + 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < GB3 + 2 * Ns; i3 += 1) buff.q[i1][i2 + GB2][i3] = tmp_wall_2.q[j1][j3][it][i1][i2][i3];
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < GB1 + 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < GB2 + 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < 2 * Ns; i3 += 1) buff.q[i1][i2][i3 + GB3] = tmp_wall_3.q[j1][j2][it][i1][i2][i3];
          Formura_Step(&buff, &rslt, (Formura_Navi){0}, f1, f2, f3);
          /* low side of buff -> walls */
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < GB2 + 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < GB3 + 2 * Ns; i3 += 1) tmp_wall_1.q[j2][j3][it][i1][i2][i3] = buff.q[i1][i2][i3];
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < GB1 + 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < GB3 + 2 * Ns; i3 += 1) tmp_wall_2.q[j1][j3][it][i1][i2][i3] = buff.q[i1][i2][i3];
#pragma omp parallel for collapse(3)
          for (int i1 = 0; i1 < GB1 + 2 * Ns; i1 += 1)
            for (int i2 = 0; i2 < GB2 + 2 * Ns; i2 += 1)
              for (int i3 = 0; i3 < 2 * Ns; i3 += 1) tmp_wall_3.q[j1][j2][it][i1][i2][i3] = buff.q[i1][i2][i3];
        }
        /* rslt -> floor */
#pragma omp parallel for collapse(3)
        for (int i1 = 0; i1 < GB1; i1 += 1)
          for (int i2 = 0; i2 < GB2; i2 += 1)
            for (int i3 = 0; i3 < GB3; i3 += 1) tmp_floor.q[i1 + f1][i2 + f2][i3 + f3] = rslt.q[i1][i2][i3];
      }
    }
  }
}

/* Formura_Forward with temporal blocking (Templates.hs:311-334) */
void Formura_Forward(Formura_Navi *n) {
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1)
      for (int i3 = 0; i3 < L3; i3 += 1)
        tmp_floor.q[i1 + 2 * S_TB][i2 + 2 * S_TB][i3 + 2 * S_TB] = formura_data.q[i1][i2][i3];
  isendrecv(&formura_data, S_TB);
  const int d1 = 2 * S_TB / GB1, d2 = 2 * S_TB / GB2, d3 = 2 * S_TB / GB3;
  /* b0: blocks that do not need the incoming halo */
  updateWithTB(0, BP1 - 1 - d1, 0, BP2 - 1 - d2, 0, BP3 - 1 - d3);
  WAIT_AND_COPY(Formura_Tmp_Floor, &tmp_floor, S_TB);
  /* bs: for axis j the halo-dependent slab, axes before j complete */
  updateWithTB(BP1 - 1 - d1, BP1, 0, BP2 - 1 - d2, 0, BP3 - 1 - d3);
  updateWithTB(0, BP1, BP2 - 1 - d2, BP2, 0, BP3 - 1 - d3);
  updateWithTB(0, BP1, 0, BP2, BP3 - 1 - d3, BP3);
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1)
      for (int i3 = 0; i3 < L3; i3 += 1) formura_data.q[i1][i2][i3] = tmp_floor.q[i1][i2][i3];
  n->offset_x = (n->offset_x - S_TB + n->total_grid_x) % n->total_grid_x;
  n->offset_y = (n->offset_y - S_TB + n->total_grid_y) % n->total_grid_y;
  n->offset_z = (n->offset_z - S_TB + n->total_grid_z) % n->total_grid_z;
  n->time_step += NT;
}
#endif

void Formura_Finalize() {}
