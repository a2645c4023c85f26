/*
 * ORACLE -- test infrastructure only.  Not part of the product; only tests/,
 * the smoke check and bench.py's cpu_baseline leg may build, load or run it.
 *
 * Hand transcription of the C that formura v2.3.2 emits for the reference's
 * sample/1d_hllc/hllc.fmr as it stands (MUSCL-Hancock reconstruction, :167,
 * with the minmod limiter beta = 1, and the HLLC flux), one rank, no MPI, no
 * temporal blocking.  Fields e, rho, rhou, rhov, rhow (Data.Map order,
 * Templates.hs:53-73); sleeve s = 2.
 *
 * Placement, derived from Translate.hs:287-313 (half offsets) and
 * Templates.hs:417-487 (calcRange, mkKernel):
 *   DU = U[+1/2] - U[-1/2] has type offset 1/2 and holds U[j+1] - U[j] at j;
 *   at a cell k, Ob(DU) = DU[-1/2] is DU[k-1] (Shift -1) and Of(DU) = DU[+1/2]
 *   is DU[k].  The temporaries UR_x, UL_x (:153, five components each) read the
 *   state at cursors -1, 0, +1: range [-1,1], extent N+4-2, buff[i+d+1].
 *   DivF (:156) reads them at cursor 0: range [-1,1] merged with [0,0], extent
 *   N+2, UR_x[i].  U_L = Ob(URbar) = URbar at cursor 0 (type offset 1/2),
 *   U_R = Of(ULbar) = ULbar at cursor +1 (Shift +1), so S_L, S_R (:116), S_M
 *   (:117) and the five flux components F (:168) have range
 *   [-1,1] u ([0,0]+[-1,1]) u ([1,1]+[-1,1]) = [-1,2], extent N+1, and read
 *   UR_x[i], DivF[i], UL_x[i+1], DivF[i+1].  The update (:169, :78-80) reads F
 *   at cursors 0 and -1 (F[+1/2], F[-1/2] of a type-offset-1/2 value): range
 *   [-2,2], extent N, F[i+1] and F[i], state buff[i+2].
 *   Halo: buff[j] = data[(j - 2s) mod N] (Functions.hs:265-290), so output i
 *   holds cell i - 2: the offset drops by s = 2 per step.
 * Arithmetic, one C operation per micro instruction, left-associative:
 *   b = U[k] - U[k-1], f = U[k+1] - U[k];
 *   DP = max(max(0.0, min(1.0*b, f)), min(b, 1.0*f));
 *   DN = min(min(0.0, max(1.0*b, f)), max(b, 1.0*f))     (:132-133, min/max :8-11);
 *   Delta = f > 0 ? DP : DN; UL = U - 0.5*Delta; UR = U + 0.5*Delta (:134-136);
 *   DivF = (F(UR) - F(UL))/dx with F the physical flux of a state (:139-141,156);
 *   URbar = UR - (0.5*dt)*DivF; ULbar = UL - (0.5*dt)*DivF (:144-149);
 *   HLLC of (URbar at the left cell, ULbar at the right cell) as in
 *   sample/3d_hllc (:90-128): S_L, S_R from min/max of the velocities and the
 *   larger sound speed, S_M, p*, star states, F_L + min(S_L,0)(U*_L - U_L) if
 *   S_M > 0 else F_R + max(S_R,0)(U*_R - U_R);
 *   U' = U - dt*((F[i+1] - F[i])/dx) (:78-80, :34);
 *   dt = (C_cfl*dx)/max_signal_speed = (0.8*dx)/4 (:19-21), gamma = 7/5,
 *   gamma1 = gamma - 1.
 * Init (:49-58): rho, u, p by x - total_grid_x/2 < 0 (x the cell index),
 * e = p/gamma1 + (0.5*rho)*((u*u + v*v) + w*w), momenta rho*u, rho*v, rho*w.
 *
 * The flux temporaries are external symbols (static in the generated C; the
 * linkage changes no value), for tests.
 * Compile-time parameters: NX (grid_per_node), LX (length_per_node).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include <math.h>

#ifndef NX
#define NX 64
#endif
#ifndef LX
#define LX 4.0
#endif
#define Ns 2
#define L1 NX
#define B1 (NX + 2 * Ns)   /* buff extent */
#define TR (B1 - 2)        /* UR_x, UL_x, DivF: range [-1,1] */
#define TF (B1 - 3)        /* S_L, S_R, S_M, F: range [-1,2] */
#define NC 5               /* conserved components in the program's order: rho, rhou, rhov, rhow, e */

typedef struct {
  double e[L1];
  double rho[L1];
  double rhou[L1];
  double rhov[L1];
  double rhow[L1];
} Formura_Grid_Struct;
typedef struct {
  int time_step;
  int lower_x;
  int upper_x;
  double space_interval_x;
  int my_rank;
  int offset_x;
  double length_x;
  int total_grid_x;
} Formura_Navi;
typedef struct {
  double e[B1];
  double rho[B1];
  double rhou[B1];
  double rhov[B1];
  double rhow[B1];
} Formura_Buff;

static Formura_Buff buff;
Formura_Grid_Struct formura_data;

/* the 23 manifest temporaries (component c of a tuple: 0 rho, 1-3 momenta, 4 e) */
double ur_x[NC][TR], ul_x[NC][TR], divf[NC][TR];
double s_l[TF], s_r[TF], s_m[TF], flux[NC][TF];

double to_pos_x(int i, Formura_Navi n) { return n.space_interval_x * ((i + n.offset_x) % n.total_grid_x); }

static const double gamma_ = 7.0 / 5.0;
#define GAMMA1 (gamma_ - 1.0)
#define DX (LX / NX)
#define DT ((0.8 * DX) / 4.0)

static inline double fmax_(double a, double b) { return a > b ? a : b; }
static inline double fmin_(double a, double b) { return a > b ? b : a; }

static inline double comp(const Formura_Buff *b, int c, int j) {
  switch (c) {
    case 0: return b->rho[j];
    case 1: return b->rhou[j];
    case 2: return b->rhov[j];
    case 3: return b->rhow[j];
    default: return b->e[j];
  }
}

/* MUSCL reconstruction of component c at cell buff[k]: sign +1 gives UR, -1 UL */
static inline double muscl(const Formura_Buff *bf, int c, int k, int right) {
  const double b = comp(bf, c, k) - comp(bf, c, k - 1);
  const double f = comp(bf, c, k + 1) - comp(bf, c, k);
  const double DP = fmax_(fmax_(0.0, fmin_(1.0 * b, f)), fmin_(b, 1.0 * f));
  const double DN = fmin_(fmin_(0.0, fmax_(1.0 * b, f)), fmax_(b, 1.0 * f));
  const double Delta = f > 0.0 ? DP : DN;
  return right ? comp(bf, c, k) + (0.5 * Delta) : comp(bf, c, k) - (0.5 * Delta);
}

typedef struct {
  double U[NC];          /* conserved: rho, rhou, rhov, rhow, e */
  double u, v1, v2, p;   /* velocity components and pressure */
} Side;

static inline Side side_of(const double *U) {
  Side s;
  for (int c = 0; c < NC; ++c) s.U[c] = U[c];
  s.u = U[1] / U[0];
  s.v1 = U[2] / U[0];
  s.v2 = U[3] / U[0];
  s.p = GAMMA1 * (U[4] - ((0.5 * U[0]) * (((s.u * s.u) + (s.v1 * s.v1)) + (s.v2 * s.v2))));
  return s;
}

/* physical flux of a state (Flux_fromU, :139-141 and :65-70) */
static inline double phys_flux(const Side *s, int c) {
  switch (c) {
    case 0: return s->U[0] * s->u;
    case 1: return ((s->U[0] * s->u) * s->u) + s->p;
    case 2: return ((s->U[0] * s->v1) * s->u) + 0.0;
    case 3: return ((s->U[0] * s->v2) * s->u) + 0.0;
    default: return (s->U[4] + s->p) * s->u;
  }
}

/* the Riemann states of face i (between buff[i+1] and buff[i+2]) from the temporaries */
static inline void face_states(int i, Side *L, Side *R) {
  double UL[NC], UR[NC];
  for (int c = 0; c < NC; ++c) {
    UL[c] = ur_x[c][i] - ((0.5 * DT) * divf[c][i]);          /* URbar of the left cell */
    UR[c] = ul_x[c][i + 1] - ((0.5 * DT) * divf[c][i + 1]);  /* ULbar of the right cell */
  }
  *L = side_of(UL);
  *R = side_of(UR);
}

static inline double sound_speed(const Side *s) { return sqrt((gamma_ * s->p) / s->U[0]); }

static void Formura_Step(Formura_Buff *b, Formura_Grid_Struct *rslt, Formura_Navi n, int block_offset_1) {
  (void)n; (void)block_offset_1;
  /* UR_x, UL_x: the cell at buff[i+1] */
  for (int c = 0; c < NC; ++c)
    for (int i = 0; i < TR; ++i) ur_x[c][i] = muscl(b, c, i + 1, 1);
  for (int c = 0; c < NC; ++c)
    for (int i = 0; i < TR; ++i) ul_x[c][i] = muscl(b, c, i + 1, 0);
  /* DivF = (F(UR) - F(UL))/dx, reads UR_x[i], UL_x[i] */
  for (int c = 0; c < NC; ++c)
    for (int i = 0; i < TR; ++i) {
      double a[NC], q[NC];
      for (int d = 0; d < NC; ++d) {
        a[d] = ur_x[d][i];
        q[d] = ul_x[d][i];
      }
      const Side R = side_of(a), L = side_of(q);
      divf[c][i] = (phys_flux(&R, c) - phys_flux(&L, c)) / DX;
    }
  /* wave speeds */
  for (int i = 0; i < TF; ++i) {
    Side L, R;
    face_states(i, &L, &R);
    const double cL = sound_speed(&L), cR = sound_speed(&R), cf = fmax_(cL, cR);
    s_l[i] = fmin_(L.u, R.u) - cf;
  }
  for (int i = 0; i < TF; ++i) {
    Side L, R;
    face_states(i, &L, &R);
    const double cL = sound_speed(&L), cR = sound_speed(&R), cf = fmax_(cL, cR);
    s_r[i] = fmax_(L.u, R.u) + cf;
  }
  for (int i = 0; i < TF; ++i) {
    Side L, R;
    face_states(i, &L, &R);
    const double SL = s_l[i], SR = s_r[i];
    s_m[i] = (((R.p - L.p) + ((L.U[0] * L.u) * (SL - L.u))) - ((R.U[0] * R.u) * (SR - R.u))) /
             ((L.U[0] * (SL - L.u)) - (R.U[0] * (SR - R.u)));
  }
  /* the five flux components */
  for (int c = 0; c < NC; ++c)
    for (int i = 0; i < TF; ++i) {
      Side L, R;
      face_states(i, &L, &R);
      const double SL = s_l[i], SR = s_r[i], SM = s_m[i];
      const double pst = ((((L.U[0] * R.p) * (SL - L.u)) - ((R.U[0] * L.p) * (SR - R.u))) +
                          ((((L.U[0] * R.U[0]) * (SR - R.u)) * (SL - L.u)) * (L.u - R.u))) /
                         ((L.U[0] * (SL - L.u)) - (R.U[0] * (SR - R.u)));
      const Side *K = SM > 0.0 ? &L : &R;
      const double S = SM > 0.0 ? SL : SR;
      const double w = SM > 0.0 ? (SL > 0.0 ? 0.0 : SL) : (SR > 0.0 ? SR : 0.0);
      const double rs = (K->U[0] * (S - K->u)) / (S - SM);
      double us;
      switch (c) {
        case 0: us = rs; break;
        case 1: us = rs * SM; break;
        case 2: us = rs * K->v1; break;
        case 3: us = rs * K->v2; break;
        default: us = ((((S - K->u) * K->U[4]) - (K->p * K->u)) + (pst * SM)) / (S - SM);
      }
      flux[c][i] = phys_flux(K, c) + (w * (us - K->U[c]));
    }
  /* the five updates: U' = U - dt*((F[i+1] - F[i])/dx) at buff[i+2] */
  for (int i = 0; i < L1; ++i) {
    rslt->rho[i] = b->rho[i + 2] - (DT * ((flux[0][i + 1] - flux[0][i]) / DX));
    rslt->rhou[i] = b->rhou[i + 2] - (DT * ((flux[1][i + 1] - flux[1][i]) / DX));
    rslt->rhov[i] = b->rhov[i + 2] - (DT * ((flux[2][i + 1] - flux[2][i]) / DX));
    rslt->rhow[i] = b->rhow[i + 2] - (DT * ((flux[3][i + 1] - flux[3][i]) / DX));
    rslt->e[i] = b->e[i + 2] - (DT * ((flux[4][i + 1] - flux[4][i]) / DX));
  }
}

void Formura_Init(int *argc, char ***argv, Formura_Navi *n) {
  (void)argc; (void)argv;
  n->time_step = 0;
  n->lower_x = 0;
  n->upper_x = L1;
  n->space_interval_x = LX / L1;
  n->my_rank = 0;
  n->offset_x = 0;
  n->length_x = LX;
  n->total_grid_x = L1;
  /* init graph at offset 0: x is the cell index (i + offset + block offset) */
  for (int i = 0; i < L1; ++i) {
    const double x = (double)(i + n->offset_x + 0);
    const bool left = (x - (double)L1 / 2.0) < 0.0;
    const double rho = left ? 1.0 : 0.125;
    const double u = left ? 0.75 : 0.0, v = left ? 0.0 : 0.0, w = left ? 0.0 : 0.0;
    const double p = left ? 1.0 : 0.1;
    formura_data.rho[i] = rho;
    formura_data.rhou[i] = rho * u;
    formura_data.rhov[i] = rho * v;
    formura_data.rhow[i] = rho * w;
    formura_data.e[i] = (p / GAMMA1) + ((0.5 * rho) * (((u * u) + (v * v)) + (w * w)));
  }
}

void Formura_Forward(Formura_Navi *n) {
  /* halo: the low side of width 2s from the high edge (one positive basis in
   * 1-D: pack, send -> recv, unpack), then the interior */
  for (int i = 0; i < 2 * Ns; ++i) {
    buff.e[i] = formura_data.e[L1 - 2 * Ns + i];
    buff.rho[i] = formura_data.rho[L1 - 2 * Ns + i];
    buff.rhou[i] = formura_data.rhou[L1 - 2 * Ns + i];
    buff.rhov[i] = formura_data.rhov[L1 - 2 * Ns + i];
    buff.rhow[i] = formura_data.rhow[L1 - 2 * Ns + i];
  }
  for (int i = 0; i < L1; ++i) {
    buff.e[i + 2 * Ns] = formura_data.e[i];
    buff.rho[i + 2 * Ns] = formura_data.rho[i];
    buff.rhou[i + 2 * Ns] = formura_data.rhou[i];
    buff.rhov[i + 2 * Ns] = formura_data.rhov[i];
    buff.rhow[i + 2 * Ns] = formura_data.rhow[i];
  }
  Formura_Step(&buff, &formura_data, *n, 0);
  n->offset_x = (n->offset_x - Ns + n->total_grid_x) % n->total_grid_x;
  n->time_step += 1;
}

void Formura_Finalize() {}
