/*
 * ORACLE -- test infrastructure only.  Not part of the product; only tests/,
 * the smoke check and bench.py's cpu_baseline leg may build, load or run it.
 *
 * Hand transcription of the C that formura v2.3.2 emits for the first-order
 * form of sample/3d_hllc/3d_hllc.fmr (the first_order_LR line :205,210 in
 * place of MUSCL_Hancock :211) with dt a fixed literal, one rank, no MPI, no
 * temporal blocking.  Fields e, rho, rhou, rhov, rhow.
 *
 *   struct layout (Data.Map key order) ..................... Templates.hs:53-73
 *   Formura_Step (mkKernel) ................................ Templates.hs:426-487
 *     - one full-box nest per manifest node in ascending OM id.  Per face
 *       direction: S_L, S_R (3d_hllc.fmr:129), S_M (:130), then the five
 *       flux components (:212-214); 24 temporaries.  calcRange gives each of
 *       them the range [0,1] (loads at cursor 0 and +e_axis; the wave-speed
 *       temporaries are read at cursor 0, Templates.hs:417-424), so each is a
 *       static array of extent inputSize - 1 = N+1 per axis, indexed
 *       buff[i+d] and formura_mnX[i] (Templates.hs:442-457,473-474);
 *     - then the five Store nests (range [-1,1], extent N) read buff[i+1]
 *       and the fluxes at [i+1] and [i+1-e_axis];
 *   Formura_Forward (noBlocking) ........................... Templates.hs:300-310, 336-346
 *   halo (pack, send->recv, unpack for 7 bases) ............ Functions.hs:265-290
 *
 * The arithmetic follows the expression trees of the program, one C
 * operation per micro instruction, left-associative:
 *   V = m/rho (:74); p = gamma1*(e - (0.5*rho)*((V0*V0 + V1*V1) + V2*V2)) (:75);
 *   c = sqrt((gamma*p)/rho) (:104); cf = cL > cR ? cL : cR (:8,110);
 *   S_L = (uL > uR ? uR : uL) - cf; S_R = (uL > uR ? uL : uR) + cf (:111-112);
 *   S_M = (((pR-pL) + (rhoL*uL)*(S_L-uL)) - (rhoR*uR)*(S_R-uR))
 *         / (rhoL*(S_L-uL) - rhoR*(S_R-uR)) (:116);
 *   p* = ((((rhoL*pR)*(S_L-uL)) - ((rhoR*pL)*(S_R-uR)))
 *         + ((((rhoL*rhoR)*(S_R-uR))*(S_L-uL))*(uL-uR))) / (same denominator) (:120);
 *   rho*_K = (rhoK*(S_K-uK))/(S_K-S_M); e*_K = ((((S_K-uK)*eK) - pK*uK) + p*S_M)/(S_K-S_M) (:97,100);
 *   U*_K = (rho*_K, rho*_K*S_M, rho*_K*V1_K, rho*_K*V2_K, e*_K) (:136-137);
 *   F_K = (rhoK*uK, (rhoK*uK)*uK + pK, (rhoK*V1)*uK + 0, (rhoK*V2)*uK + 0, (eK+pK)*uK) (:78-83);
 *   flux = S_M > 0 ? F_L + (S_L > 0 ? 0 : S_L)*(U*_L - U_L)
 *                  : F_R + (S_R > 0 ? S_R : 0)*(U*_R - U_R) (:85-89);
 *   q' = q - dt*(((Fq[i]-Fq[i-1])/dx + (Gq[j]-Gq[j-1])/dy) + (Hq[k]-Hq[k-1])/dz) (:91-93,38-42).
 * The normal momentum of a face along y is rhov (tangential rhow, rhou), along
 * z rhow (rhou, rhov): the program's YZX / ZXY cycles (:44-45,143-153).
 *
 * The temporaries are external symbols (the generated C has them static; the
 * linkage changes no value) so that tests can read the face fluxes.
 * Compile-time parameters: NX, NY, NZ (grid_per_node), LX, LY, LZ
 * (length_per_node), DX, DY, DZ (the program's dx = L/N rounded to double),
 * DT (the step's dt).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include <math.h>

#ifndef NX
#define NX 16
#endif
#ifndef NY
#define NY NX
#endif
#ifndef NZ
#define NZ NX
#endif
#ifndef LX
#define LX 1.0
#endif
#ifndef LY
#define LY 1.0
#endif
#ifndef LZ
#define LZ 1.0
#endif
#ifndef DX
#define DX (1.0 / NX)
#endif
#ifndef DY
#define DY (1.0 / NY)
#endif
#ifndef DZ
#define DZ (1.0 / NZ)
#endif
#ifndef DT
#define DT 3.25431185102799e-04
#endif
#define Ns 1
#define L1 NX
#define L2 NY
#define L3 NZ
#define T1 (NX + 1)
#define T2 (NY + 1)
#define T3 (NZ + 1)

typedef struct {
  double e[L1][L2][L3];
  double rho[L1][L2][L3];
  double rhou[L1][L2][L3];
  double rhov[L1][L2][L3];
  double rhow[L1][L2][L3];
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
  double e[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns];
  double rho[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns];
  double rhou[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns];
  double rhov[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns];
  double rhow[L1 + 2 * Ns][L2 + 2 * Ns][L3 + 2 * Ns];
} Formura_Buff;
typedef struct {
  double e, rho, rhou, rhov, rhow;
} Formura_Comm_Buff;

static Formura_Buff buff;
Formura_Grid_Struct formura_data;

/* the 24 manifest temporaries, per face direction a = x, y, z: the wave
 * speeds sl_a, sr_a, sm_a and the fluxes flux_a_<field> of the five fields
 * through the face between cells i and i + e_a, stored at index i */
typedef double Tmp[T1][T2][T3];
Tmp sl_x, sr_x, sm_x, flux_x_rho, flux_x_rhou, flux_x_rhov, flux_x_rhow, flux_x_e;
Tmp sl_y, sr_y, sm_y, flux_y_rho, flux_y_rhou, flux_y_rhov, flux_y_rhow, flux_y_e;
Tmp sl_z, sr_z, sm_z, flux_z_rho, flux_z_rhou, flux_z_rhov, flux_z_rhow, flux_z_e;

double to_pos_x(int i, Formura_Navi n) { return n.space_interval_x * ((i + n.offset_x) % n.total_grid_x); }
double to_pos_y(int i, Formura_Navi n) { return n.space_interval_y * ((i + n.offset_y) % n.total_grid_y); }
double to_pos_z(int i, Formura_Navi n) { return n.space_interval_z * ((i + n.offset_z) % n.total_grid_z); }

static const double gamma_ = 7.0 / 5.0;
#define GAMMA1 (gamma_ - 1.0)

/* One cell state seen from a face: density, the normal and two tangential
 * momenta, total energy, and the primitives the program derives from them. */
typedef struct {
  double rho, n, t1, t2, e; /* conserved, rotated */
  double u, v1, v2, p;      /* velocity (normal, tangential) and pressure */
} Side;

static inline Side side_of(double rho, double n, double t1, double t2, double e) {
  Side s;
  s.rho = rho; s.n = n; s.t1 = t1; s.t2 = t2; s.e = e;
  s.u = n / rho;
  s.v1 = t1 / rho;
  s.v2 = t2 / rho;
  s.p = GAMMA1 * (e - ((0.5 * rho) * (((s.u * s.u) + (s.v1 * s.v1)) + (s.v2 * s.v2))));
  return s;
}

static inline double sound_speed(const Side *s) { return sqrt((gamma_ * s->p) / s->rho); }

static inline double max_c(const Side *L, const Side *R) {
  const double cL = sound_speed(L), cR = sound_speed(R);
  return cL > cR ? cL : cR;
}

static inline double wave_left(const Side *L, const Side *R) {
  return (L->u > R->u ? R->u : L->u) - max_c(L, R);
}

static inline double wave_right(const Side *L, const Side *R) {
  return (L->u > R->u ? L->u : R->u) + max_c(L, R);
}

static inline double wave_middle(const Side *L, const Side *R, double S_L, double S_R) {
  return (((R->p - L->p) + ((L->rho * L->u) * (S_L - L->u))) - ((R->rho * R->u) * (S_R - R->u))) /
         ((L->rho * (S_L - L->u)) - (R->rho * (S_R - R->u)));
}

static inline double pressure_star(const Side *L, const Side *R, double S_L, double S_R) {
  return ((((L->rho * R->p) * (S_L - L->u)) - ((R->rho * L->p) * (S_R - R->u))) +
          ((((L->rho * R->rho) * (S_R - R->u)) * (S_L - L->u)) * (L->u - R->u))) /
         ((L->rho * (S_L - L->u)) - (R->rho * (S_R - R->u)));
}

/* component c (0 rho, 1 normal, 2 t1, 3 t2, 4 e) of one side's star-region
 * flux F_K + w*(U*_K - U_K) */
static inline double side_flux(const Side *K, int c, double S, double S_M, double p_st, double w) {
  const double rs = (K->rho * (S - K->u)) / (S - S_M);
  double f, us, uk;
  switch (c) {
    case 0: f = K->rho * K->u; us = rs; uk = K->rho; break;
    case 1: f = ((K->rho * K->u) * K->u) + K->p; us = rs * S_M; uk = K->n; break;
    case 2: f = ((K->rho * K->v1) * K->u) + 0.0; us = rs * K->v1; uk = K->t1; break;
    case 3: f = ((K->rho * K->v2) * K->u) + 0.0; us = rs * K->v2; uk = K->t2; break;
    default:
      f = (K->e + K->p) * K->u;
      us = ((((S - K->u) * K->e) - (K->p * K->u)) + (p_st * S_M)) / (S - S_M);
      uk = K->e;
  }
  return f + (w * (us - uk));
}

/* The cell states of the face between buff[i] and buff[i + e_ax], the
 * momentum rotated so that the first component is normal to the face. */
static inline void face_sides(const Formura_Buff *b, int ax, int i1, int i2, int i3, Side *L, Side *R) {
  const int j1 = i1 + (ax == 0), j2 = i2 + (ax == 1), j3 = i3 + (ax == 2);
  if (ax == 0) {
    *L = side_of(b->rho[i1][i2][i3], b->rhou[i1][i2][i3], b->rhov[i1][i2][i3], b->rhow[i1][i2][i3], b->e[i1][i2][i3]);
    *R = side_of(b->rho[j1][j2][j3], b->rhou[j1][j2][j3], b->rhov[j1][j2][j3], b->rhow[j1][j2][j3], b->e[j1][j2][j3]);
  } else if (ax == 1) {
    *L = side_of(b->rho[i1][i2][i3], b->rhov[i1][i2][i3], b->rhow[i1][i2][i3], b->rhou[i1][i2][i3], b->e[i1][i2][i3]);
    *R = side_of(b->rho[j1][j2][j3], b->rhov[j1][j2][j3], b->rhow[j1][j2][j3], b->rhou[j1][j2][j3], b->e[j1][j2][j3]);
  } else {
    *L = side_of(b->rho[i1][i2][i3], b->rhow[i1][i2][i3], b->rhou[i1][i2][i3], b->rhov[i1][i2][i3], b->e[i1][i2][i3]);
    *R = side_of(b->rho[j1][j2][j3], b->rhow[j1][j2][j3], b->rhou[j1][j2][j3], b->rhov[j1][j2][j3], b->e[j1][j2][j3]);
  }
}

#define TMP_NEST(BODY)                                            \
  _Pragma("omp parallel for collapse(3)")                         \
  for (int i1 = 0; i1 < T1; i1 += 1)                              \
    for (int i2 = 0; i2 < T2; i2 += 1)                            \
      for (int i3 = 0; i3 < T3; i3 += 1) {                        \
        BODY                                                      \
      }

/* S_L, S_R nests: the wave speeds of every face of direction ax */
static void nest_wave_lr(const Formura_Buff *b, int ax, Tmp out, int right) {
  TMP_NEST(Side L; Side R; face_sides(b, ax, i1, i2, i3, &L, &R);
           out[i1][i2][i3] = right ? wave_right(&L, &R) : wave_left(&L, &R);)
}

/* S_M nest: reads the S_L and S_R temporaries at cursor 0 */
static void nest_wave_m(const Formura_Buff *b, int ax, Tmp sl, Tmp sr, Tmp out) {
  TMP_NEST(Side L; Side R; face_sides(b, ax, i1, i2, i3, &L, &R);
           out[i1][i2][i3] = wave_middle(&L, &R, sl[i1][i2][i3], sr[i1][i2][i3]);)
}

/* flux nest of rotated component c: reads S_L, S_R, S_M at cursor 0 */
static void nest_flux(const Formura_Buff *b, int ax, int c, Tmp sl, Tmp sr, Tmp sm, Tmp out) {
  TMP_NEST(Side L; Side R; face_sides(b, ax, i1, i2, i3, &L, &R);
           const double S_L = sl[i1][i2][i3]; const double S_R = sr[i1][i2][i3];
           const double S_M = sm[i1][i2][i3];
           const double p_st = pressure_star(&L, &R, S_L, S_R);
           const double wL = S_L > 0.0 ? 0.0 : S_L;
           const double wR = S_R > 0.0 ? S_R : 0.0;
           out[i1][i2][i3] = S_M > 0.0 ? side_flux(&L, c, S_L, S_M, p_st, wL) : side_flux(&R, c, S_R, S_M, p_st, wR);)
}

/* one Store nest: q'[i] = q - dt*div(F, G, H) at buff[i+1], fluxes at [i+1] and [i+1-e_a] */
#define STORE_NEST(Q, FX, GY, HZ)                                                                  \
  _Pragma("omp parallel for collapse(3)") for (int i1 = 0; i1 < L1; i1 += 1)                        \
    for (int i2 = 0; i2 < L2; i2 += 1)                                                            \
      for (int i3 = 0; i3 < L3; i3 += 1) {                                                        \
        const double a0 = (FX[i1 + 1][i2 + 1][i3 + 1] - FX[i1][i2 + 1][i3 + 1]) / DX;              \
        const double a1 = (GY[i1 + 1][i2 + 1][i3 + 1] - GY[i1 + 1][i2][i3 + 1]) / DY;              \
        const double a2 = (HZ[i1 + 1][i2 + 1][i3 + 1] - HZ[i1 + 1][i2 + 1][i3]) / DZ;              \
        rslt->Q[i1][i2][i3] = buff->Q[i1 + 1][i2 + 1][i3 + 1] - (DT * ((a0 + a1) + a2));           \
      }

static void Formura_Step(Formura_Buff *buff, Formura_Grid_Struct *rslt, Formura_Navi n, int block_offset_1,
                         int block_offset_2, int block_offset_3) {
  (void)n; (void)block_offset_1; (void)block_offset_2; (void)block_offset_3;
  /* x faces: the rotated components (rho, n, t1, t2, e) are (rho, rhou, rhov, rhow, e) */
  nest_wave_lr(buff, 0, sl_x, 0);
  nest_wave_lr(buff, 0, sr_x, 1);
  nest_wave_m(buff, 0, sl_x, sr_x, sm_x);
  nest_flux(buff, 0, 0, sl_x, sr_x, sm_x, flux_x_rho);
  nest_flux(buff, 0, 1, sl_x, sr_x, sm_x, flux_x_rhou);
  nest_flux(buff, 0, 2, sl_x, sr_x, sm_x, flux_x_rhov);
  nest_flux(buff, 0, 3, sl_x, sr_x, sm_x, flux_x_rhow);
  nest_flux(buff, 0, 4, sl_x, sr_x, sm_x, flux_x_e);
  /* y faces: (rho, rhov, rhow, rhou, e) */
  nest_wave_lr(buff, 1, sl_y, 0);
  nest_wave_lr(buff, 1, sr_y, 1);
  nest_wave_m(buff, 1, sl_y, sr_y, sm_y);
  nest_flux(buff, 1, 0, sl_y, sr_y, sm_y, flux_y_rho);
  nest_flux(buff, 1, 1, sl_y, sr_y, sm_y, flux_y_rhov);
  nest_flux(buff, 1, 2, sl_y, sr_y, sm_y, flux_y_rhow);
  nest_flux(buff, 1, 3, sl_y, sr_y, sm_y, flux_y_rhou);
  nest_flux(buff, 1, 4, sl_y, sr_y, sm_y, flux_y_e);
  /* z faces: (rho, rhow, rhou, rhov, e) */
  nest_wave_lr(buff, 2, sl_z, 0);
  nest_wave_lr(buff, 2, sr_z, 1);
  nest_wave_m(buff, 2, sl_z, sr_z, sm_z);
  nest_flux(buff, 2, 0, sl_z, sr_z, sm_z, flux_z_rho);
  nest_flux(buff, 2, 1, sl_z, sr_z, sm_z, flux_z_rhow);
  nest_flux(buff, 2, 2, sl_z, sr_z, sm_z, flux_z_rhou);
  nest_flux(buff, 2, 3, sl_z, sr_z, sm_z, flux_z_rhov);
  nest_flux(buff, 2, 4, sl_z, sr_z, sm_z, flux_z_e);
  STORE_NEST(rho, flux_x_rho, flux_y_rho, flux_z_rho)
  STORE_NEST(rhou, flux_x_rhou, flux_y_rhou, flux_z_rhou)
  STORE_NEST(rhov, flux_x_rhov, flux_y_rhov, flux_z_rhov)
  STORE_NEST(rhow, flux_x_rhow, flux_y_rhow, flux_z_rhow)
  STORE_NEST(e, flux_x_e, flux_y_e, flux_z_e)
}

/* halo: for every positive basis b pack src[.. + (N-2s if b_k)] into an AoS
 * comm buffer, copy send -> recv, unpack into buff[.. + (0 if b_k else 2s)]
 * (Functions.hs:265-290); the buffers are dynamically sized here. */
static void sendrecv_unpack(const Formura_Grid_Struct *src, Formura_Buff *tgt, int b1, int b2, int b3) {
  const int S = Ns;
  const int E1 = b1 ? 2 * S : L1, E2 = b2 ? 2 * S : L2, E3 = b3 ? 2 * S : L3;
  const int O1 = b1 ? L1 - 2 * S : 0, O2 = b2 ? L2 - 2 * S : 0, O3 = b3 ? L3 - 2 * S : 0;
  const int P1 = b1 ? 0 : 2 * S, P2 = b2 ? 0 : 2 * S, P3 = b3 ? 0 : 2 * S;
  Formura_Comm_Buff *send = malloc(sizeof(Formura_Comm_Buff) * E1 * E2 * E3);
  Formura_Comm_Buff *recv = malloc(sizeof(Formura_Comm_Buff) * E1 * E2 * E3);
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < E1; i1 += 1)
    for (int i2 = 0; i2 < E2; i2 += 1)
      for (int i3 = 0; i3 < E3; i3 += 1) {
        Formura_Comm_Buff *c = &send[(i1 * E2 + i2) * E3 + i3];
        c->e = src->e[i1 + O1][i2 + O2][i3 + O3];
        c->rho = src->rho[i1 + O1][i2 + O2][i3 + O3];
        c->rhou = src->rhou[i1 + O1][i2 + O2][i3 + O3];
        c->rhov = src->rhov[i1 + O1][i2 + O2][i3 + O3];
        c->rhow = src->rhow[i1 + O1][i2 + O2][i3 + O3];
      }
  for (long k = 0; k < (long)E1 * E2 * E3; ++k) recv[k] = send[k];
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < E1; i1 += 1)
    for (int i2 = 0; i2 < E2; i2 += 1)
      for (int i3 = 0; i3 < E3; i3 += 1) {
        const Formura_Comm_Buff *c = &recv[(i1 * E2 + i2) * E3 + i3];
        tgt->e[i1 + P1][i2 + P2][i3 + P3] = c->e;
        tgt->rho[i1 + P1][i2 + P2][i3 + P3] = c->rho;
        tgt->rhou[i1 + P1][i2 + P2][i3 + P3] = c->rhou;
        tgt->rhov[i1 + P1][i2 + P2][i3 + P3] = c->rhov;
        tgt->rhow[i1 + P1][i2 + P2][i3 + P3] = c->rhow;
      }
  free(send);
  free(recv);
}

void Formura_Init(int *argc, char ***argv, Formura_Navi *n) {
  (void)argc; (void)argv;
  n->time_step = 0;
  n->lower_x = 0; n->lower_y = 0; n->lower_z = 0;
  n->upper_x = L1; n->upper_y = L2; n->upper_z = L3;
  n->space_interval_x = LX / L1; n->space_interval_y = LY / L2; n->space_interval_z = LZ / L3;
  n->my_rank = 0;
  n->offset_x = 0; n->offset_y = 0; n->offset_z = 0;
  n->length_x = LX; n->length_y = LY; n->length_z = LZ;
  n->total_grid_x = L1; n->total_grid_y = L2; n->total_grid_z = L3;
  /* init graph: a uniform gas at rest, rho = 1, m = 0, e = 2.5 */
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1)
      for (int i3 = 0; i3 < L3; i3 += 1) {
        formura_data.rho[i1][i2][i3] = 1.0;
        formura_data.rhou[i1][i2][i3] = 0.0;
        formura_data.rhov[i1][i2][i3] = 0.0;
        formura_data.rhow[i1][i2][i3] = 0.0;
        formura_data.e[i1][i2][i3] = 2.5;
      }
}

void Formura_Forward(Formura_Navi *n) {
  static const int bases[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int b = 0; b < 7; ++b) sendrecv_unpack(&formura_data, &buff, bases[b][0], bases[b][1], bases[b][2]);
#pragma omp parallel for collapse(3)
  for (int i1 = 0; i1 < L1; i1 += 1)
    for (int i2 = 0; i2 < L2; i2 += 1)
      for (int i3 = 0; i3 < L3; i3 += 1) {
        buff.e[i1 + 2][i2 + 2][i3 + 2] = formura_data.e[i1][i2][i3];
        buff.rho[i1 + 2][i2 + 2][i3 + 2] = formura_data.rho[i1][i2][i3];
        buff.rhou[i1 + 2][i2 + 2][i3 + 2] = formura_data.rhou[i1][i2][i3];
        buff.rhov[i1 + 2][i2 + 2][i3 + 2] = formura_data.rhov[i1][i2][i3];
        buff.rhow[i1 + 2][i2 + 2][i3 + 2] = formura_data.rhow[i1][i2][i3];
      }
  Formura_Step(&buff, &formura_data, *n, 0, 0, 0);
  n->offset_x = (n->offset_x - Ns + n->total_grid_x) % n->total_grid_x;
  n->offset_y = (n->offset_y - Ns + n->total_grid_y) % n->total_grid_y;
  n->offset_z = (n->offset_z - Ns + n->total_grid_z) % n->total_grid_z;
  n->time_step += 1;
}

void Formura_Finalize() {}
