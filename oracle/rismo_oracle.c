/* Oracle: CPU restatement of the reference's UVS2D Newton hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Nothing in the GPU library calls this code.
 *
 * Each routine restates the reference routine cited at its definition, keeping the
 * reference's floating-point operation order (left-to-right evaluation, loop orders,
 * float vs double arithmetic) so that tests can compare it bit for bit against the
 * reference program's own outputs.  Terms that the reference multiplies by quantities
 * that are identically zero for the decks used here (dispersion, actualDisp == 0;
 * Reynolds stresses, constant eddy-viscosity model; point sources, none defined) are
 * left out; dropping an exact zero changes at most the sign of a zero result.
 *
 * Scope: steady, pseudo-time-relaxed or unsteady (theta scheme) Newton, triangles (6 nodes)
 * and quads (8 nodes), QINLET / QTINLET / OUTLET control lines, closed (tangential-flow)
 * walls, Manning friction with Lindner vegetation, constant, algebraic and anisotropic
 * (Elder) eddy viscosity, drying and rewetting by method 2, GPdeg 5 Gauss rules.
 *
 * Parity unpinned: in dry_rewet_method2, the count down of a dry node whose depth has
 * recovered (unreachable in a serial run, a dry node's level is reset to its bed); no
 * golden deck reaches it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------- */
/* reference flag values (Bcon.h:160-190, Node.h, Elem.h) so dumps compare directly      */
enum { kInlet = 1, kOutlet = 2, kOpenBnd = 4, kQInlet = 8, kQTInlet = 16, kSlip = 128, kLoglaw = 256,
       kSetUV = 512, kSetS = 1024, kFixU = 32768, kFixV = 65536, kAutoSlip = 4194304 };
enum { rotTangent = 1, rotNormal = 2 };
/* node flags of this restatement (own bit values) */
enum { fCorn = 1, fMids = 2, fBound = 4, fInlet = 8, fOutlet = 16, fOpen = 32, fRotat = 64,
       fNoMoment = 128, fDry = 256, fMarsh = 512, fDryPrev = 1024, fMarshPrev = 2048 };
/* boundary element flags */
enum { eInlet = 1, eOutlet = 2, eOpen = 4 };
/* region element flags */
enum { elDry = 1, elMarsh = 2 };
/* the reference's node flag bits for dry and marsh nodes (Node.h:100-102) */
enum { refNodeDry = 64, refNodeDryPrev = 128, refNodeMarsh = 256, refNodeMarshPrev = 512 };

#define MAXEQ 24

/* ------------------------------------------------------------------------------------- */
/* shape tables (Shape.cpp:138-395, Triangle.cpp, Square.cpp, Line.cpp), GPdeg = 5       */

typedef struct {
  int ngp, nnd, ncn;
  double f[9][8], dfdx[9][8], dfdy[9][8];    /* quadratic shape at GPs */
  double lf[9][4], ldx[9][4], ldy[9][4];     /* linear shape at GPs */
  double w[9];
  double dndx[8][8];                          /* line only: dN_j/dxi at node i */
} shape_t;

static shape_t g_line, g_tri, g_quad;
static int g_shape_init = 0;

static double lTri(int node, double x, double y)
{
  switch (node) { case 0: return 1.0 - x - y; case 1: return x; case 2: return y; }
  return 0.0;
}
static double lTriDx(int node) { return node == 0 ? -1.0 : (node == 1 ? 1.0 : 0.0); }
static double lTriDy(int node) { return node == 0 ? -1.0 : (node == 2 ? 1.0 : 0.0); }

static double qTri(int node, double x, double y)
{
  if (node < 3) { double Li = lTri(node, x, y); return (2.0 * Li - 1.0) * Li; }
  int i = node % 3, j = (node + 1) % 3;
  return 4.0 * lTri(i, x, y) * lTri(j, x, y);
}
static double qTriDx(int node, double x, double y)
{
  if (node < 3) { double Li = lTri(node, x, y); return (4.0 * Li - 1.0) * lTriDx(node); }
  int i = node % 3, j = (node + 1) % 3;
  double Li = lTri(i, x, y), Lj = lTri(j, x, y);
  return 4.0 * (lTriDx(i) * Lj + Li * lTriDx(j));
}
static double qTriDy(int node, double x, double y)
{
  if (node < 3) { double Li = lTri(node, x, y); return (4.0 * Li - 1.0) * lTriDy(node); }
  int i = node % 3, j = (node + 1) % 3;
  double Li = lTri(i, x, y), Lj = lTri(j, x, y);
  return 4.0 * (lTriDy(i) * Lj + Li * lTriDy(j));
}

static void locSquare(int node, double* xi, double* eta)
{
  static const double L[8][2] = { {-1, -1}, {1, -1}, {1, 1}, {-1, 1},
                                  {0, -1}, {1, 0}, {0, 1}, {-1, 0} };
  *xi = L[node][0]; *eta = L[node][1];
}
static double lSq(int node, double x, double y)
{ double xi, eta; locSquare(node, &xi, &eta); return (1.0 + xi * x) * (1.0 + eta * y) / 4.0; }
static double lSqDx(int node, double y)
{ double xi, eta; locSquare(node, &xi, &eta); return xi * (1.0 + eta * y) / 4.0; }
static double lSqDy(int node, double x)
{ double xi, eta; locSquare(node, &xi, &eta); return eta * (1.0 + xi * x) / 4.0; }
static double qSq(int node, double x, double y)
{
  double xi, eta; locSquare(node, &xi, &eta);
  if (node < 4) return (1.0 + xi * x) * (1.0 + eta * y) * (xi * x + eta * y - 1.0) / 4.0;
  if (node == 4 || node == 6) return (1.0 - x * x) * (1.0 + eta * y) / 2.0;
  return (1.0 + xi * x) * (1.0 - y * y) / 2.0;
}
static double qSqDx(int node, double x, double y)
{
  double xi, eta; locSquare(node, &xi, &eta);
  if (node < 4) return (1.0 + eta * y) * (2.0 * x + xi * eta * y) / 4.0;
  if (node == 4 || node == 6) return -x * (1.0 + eta * y);
  return xi * (1.0 - y * y) / 2.0;
}
static double qSqDy(int node, double x, double y)
{
  double xi, eta; locSquare(node, &xi, &eta);
  if (node < 4) return (1.0 + xi * x) * (2.0 * y + xi * eta * x) / 4.0;
  if (node == 4 || node == 6) return eta * (1.0 - x * x) / 2.0;
  return -y * (1.0 + xi * x);
}
static double qLine(int node, double x)
{ return node == 0 ? x * (x - 1.0) / 2.0 : (node == 1 ? x * (x + 1.0) / 2.0 : 1.0 - x * x); }
static double qLineDx(int node, double x)
{ return node == 0 ? x - 0.5 : (node == 1 ? x + 0.5 : -2.0 * x); }

static void shapes_init(void)
{
  if (g_shape_init) return;
  /* 1D Gauss, 3 points (Shape.cpp:gp) */
  double loc1[3] = { -sqrt(0.6), 0.0, sqrt(0.6) };
  double w1[3] = { 5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0 };
  /* triangle 7-point Cowper rule (Shape.cpp:308-322) */
  static const double tloc[7][2] = {
    { 1.0 / 3.0, 1.0 / 3.0 },
    { 0.0597158717897698, 0.470142064105115 },
    { 0.470142064105115, 0.0597158717897698 },
    { 0.470142064105115, 0.470142064105115 },
    { 0.797426985353087, 0.101286507323456 },
    { 0.101286507323456, 0.797426985353087 },
    { 0.101286507323456, 0.101286507323456 } };
  static const double tw[7] = { 0.1125, 0.0661970763942531, 0.0661970763942531,
                                0.0661970763942531, 0.0629695902724136, 0.0629695902724136,
                                0.0629695902724136 };
  shape_t* s = &g_tri;
  s->ngp = 7; s->nnd = 6; s->ncn = 3;
  for (int g = 0; g < 7; g++) {
    double x = tloc[g][0], y = tloc[g][1];
    s->w[g] = tw[g];
    for (int i = 0; i < 6; i++) {
      s->f[g][i] = qTri(i, x, y); s->dfdx[g][i] = qTriDx(i, x, y); s->dfdy[g][i] = qTriDy(i, x, y);
    }
    for (int i = 0; i < 3; i++) {
      s->lf[g][i] = lTri(i, x, y); s->ldx[g][i] = lTriDx(i); s->ldy[g][i] = lTriDy(i);
    }
  }
  s = &g_quad;
  s->ngp = 9; s->nnd = 8; s->ncn = 4;
  for (int g = 0; g < 9; g++) {
    int ix = g % 3, iy = g / 3;
    double x = loc1[ix], y = loc1[iy];
    s->w[g] = w1[ix] * w1[iy];
    for (int i = 0; i < 8; i++) {
      s->f[g][i] = qSq(i, x, y); s->dfdx[g][i] = qSqDx(i, x, y); s->dfdy[g][i] = qSqDy(i, x, y);
    }
    for (int i = 0; i < 4; i++) {
      s->lf[g][i] = lSq(i, x, y); s->ldx[g][i] = lSqDx(i, y); s->ldy[g][i] = lSqDy(i, x);
    }
  }
  s = &g_line;
  s->ngp = 3; s->nnd = 3; s->ncn = 2;
  for (int g = 0; g < 3; g++) {
    double x = loc1[g];
    s->w[g] = w1[g];
    for (int i = 0; i < 3; i++) { s->f[g][i] = qLine(i, x); s->dfdx[g][i] = qLineDx(i, x); }
    s->lf[g][0] = (1.0 - x) / 2.0; s->lf[g][1] = (1.0 + x) / 2.0;
    s->ldx[g][0] = -0.5; s->ldx[g][1] = 0.5;
  }
  static const double lloc[3] = { -1.0, 1.0, 0.0 };
  for (int n = 0; n < 3; n++)
    for (int j = 0; j < 3; j++) g_line.dndx[n][j] = qLineDx(j, lloc[n]);
  g_shape_init = 1;
}

int orc_shape(int kind, double* f, double* dfdx, double* dfdy, double* lf, double* ldfdx,
              double* ldfdy, double* w)
{
  shapes_init();
  const shape_t* s = kind == 0 ? &g_line : (kind == 1 ? &g_tri : &g_quad);
  for (int g = 0; g < s->ngp; g++) {
    w[g] = s->w[g];
    for (int i = 0; i < s->nnd; i++) {
      f[g * s->nnd + i] = s->f[g][i]; dfdx[g * s->nnd + i] = s->dfdx[g][i];
      dfdy[g * s->nnd + i] = s->dfdy[g][i];
    }
    for (int i = 0; i < s->ncn; i++) {
      lf[g * s->ncn + i] = s->lf[g][i]; ldfdx[g * s->ncn + i] = s->ldx[g][i];
      ldfdy[g * s->ncn + i] = s->ldy[g][i];
    }
  }
  return s->ngp;
}

/* ------------------------------------------------------------------------------------- */
/* model                                                                                  */

struct orc_model {
  int np, ne;
  double *x, *y, *z, *zor, *U, *V, *S, *dUdt, *dVdt, *dSdt, *cf;
  int *elem, *nnd;
  double *manning, *vt, *areaFact;
  double *veg_dp, *veg_sp;          /* emergent vegetation per element (0: none) */
  /* turbulence model: code of $TM_TURBULENCE, st * KDest per element, k-epsilon constants,
   * nodal eddy viscosity and normal Reynolds stress (uu = vv, uv = 0) of the algebraic model */
  int turb, turb_stale;
  double cm, cd;
  double *ste_e, *vtn, *uun;
  /* anisotropic (Elder) model: estx / KDest, esty / KDest per element; nodal exx, exy, eyy
   * in global axes (the local tensor turned into the flow direction) */
  double *exi_e, *eyi_e, *exx, *exy, *eyy;
  /* time integration: state of the previous time step (NODE::vo) and the theta scheme */
  double *Uo, *Vo, *So, *dUdto, *dVdto, *dSdto;
  int unsteady;
  double dt, theta;
  int inflow_qt;                    /* the inlet line is a hydrograph (QTINLET) */
  int n_in, n_out;
  int *in_segs, *out_segs;
  double Q, S_in, S_out;
  /* connectivity */
  int *noel;
  unsigned *flag;
  int nb;              /* boundary line elements */
  int *bnd;            /* [nb*3] corner, corner, midside */
  unsigned *bflag;
  int nme;             /* model element list: region elements with their boundary elements */
  int *me;             /* >= 0 region element, < 0: -(b+1) boundary element b */
  /* boundary conditions */
  unsigned *kind;
  int *rot;
  double *nx, *ny, *niox, *nioy, *valU, *valS;
  /* numbering and pattern */
  int init_done;
  int neq;
  int *eqno;           /* [3*np] */
  int *rowptr, *cols;
  long nnz;
  /* parameters */
  double g, vk, hmin, dryLimit, maxUs, maxDeltaUV, maxDeltaS, convUV, convS;
  double relaxThdt_UV, relaxThdt_H;
  int relaxMethod, dryMethod;
  /* relaxation methods 3 and 4: the minimum (initial) relaxed time step relaxTimeFlow and
   * the stationary cycle's time interval; the relaxed time steps of the cycle and the last
   * one before the latest update (EqsUVS2D.cpp:97-101, 186, 540-627) */
  double relaxDt, interval, dt_UV, dt_H, last_dt;
  /* drying and rewetting (DRYREW, Grid.h): per node the count down and whether it was dry
   * before the last call, per region element its flags and whether it was dry before */
  double rewetLimit;
  int countDown, firstDryRew;
  int bcon_init;       /* boundary conditions generated at least once */
  int *cdown;
  unsigned char *nmark, *emark;
  unsigned *eflag;
};

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); if (!p) abort(); return p; }

/* GRID::AreaFactors (ArFact.cpp:41-85) */
static void area_factors(orc_model* m)
{
  for (int e = 0; e < m->ne; e++) {
    const int* nd = m->elem + 8 * e;
    int ncn = m->nnd[e] == 6 ? 3 : 4;
    double x[4], y[4], z[4];
    x[0] = m->x[nd[0]]; y[0] = m->y[nd[0]]; z[0] = m->z[nd[0]];
    for (int i = 1; i < ncn; i++) {
      x[i] = m->x[nd[i]] - x[0]; y[i] = m->y[nd[i]] - y[0]; z[i] = m->z[nd[i]] - z[0];
    }
    double t1 = y[1] * z[2] - y[2] * z[1];
    double t2 = z[1] * x[2] - z[2] * x[1];
    double t3 = x[1] * y[2] - x[2] * y[1];
    double areaElem = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    double areaPro = t3;
    if (ncn == 4) {
      t1 = y[2] * z[3] - y[3] * z[2];
      t2 = z[2] * x[3] - z[3] * x[2];
      t3 = x[2] * y[3] - x[3] * y[2];
      areaElem += sqrt(t1 * t1 + t2 * t2 + t3 * t3);
      areaPro += t3;
    }
    m->areaFact[e] = areaElem / areaPro;
  }
}

/* GRID::Connection (Connect.cpp:43), MODEL::Boundary (Bound.cpp:50-170),
 * MODEL::Initialize (Model.cpp:121-215): connectivity, boundary line elements for every
 * midside node with one connected element, and the model element list with each region
 * element followed by its boundary elements. */
static void topology(orc_model* m)
{
  /* GRID::Connection(ELEM::kDry) counts wet elements only; MODEL::Boundary clears the
   * boundary flags and the automatic slip conditions of every node first (Bound.cpp:64-80) */
  free(m->bnd); free(m->bflag); free(m->me);
  for (int n = 0; n < m->np; n++) {
    m->noel[n] = 0;
    m->flag[n] &= ~(unsigned)fBound;
    if (m->kind[n] & kAutoSlip) m->kind[n] &= ~(unsigned)(kAutoSlip | kFixU | kFixV);
  }
  for (int e = 0; e < m->ne; e++) {
    int k = m->nnd[e], ncn = k == 6 ? 3 : 4;
    for (int i = 0; i < k; i++) {
      int n = m->elem[8 * e + i];
      m->flag[n] |= (i < ncn) ? fCorn : fMids;
      if (!(m->eflag[e] & elDry)) m->noel[n]++;
    }
  }
  int nb = 0;
  for (int n = 0; n < m->np; n++)
    if (m->noel[n] == 1 && (m->flag[n] & fMids)) { m->flag[n] |= fBound; nb++; }
  m->nb = nb;
  m->bnd = xcalloc((size_t)3 * nb, sizeof(int));
  m->bflag = xcalloc((size_t)nb, sizeof(unsigned));
  m->me = xcalloc((size_t)(m->ne + nb), sizeof(int));
  int b = 0, k = 0;
  for (int e = 0; e < m->ne; e++) {
    if (m->eflag[e] & elDry) continue;
    int nnd = m->nnd[e], ncn = nnd == 6 ? 3 : 4;
    m->me[k++] = e;
    for (int i = ncn; i < nnd; i++) {
      int mid = m->elem[8 * e + i];
      if (m->flag[mid] & fBound) {
        int left = i - ncn, rght = (left + 1) % ncn;
        m->bnd[3 * b + 0] = m->elem[8 * e + left];
        m->bnd[3 * b + 1] = m->elem[8 * e + rght];
        m->bnd[3 * b + 2] = mid;
        m->flag[m->elem[8 * e + left]] |= fBound;
        m->flag[m->elem[8 * e + rght]] |= fBound;
        m->me[k++] = -(b + 1);
        b++;
      }
    }
  }
  m->nme = k;
}

orc_model* orc_create(int np, const double* x, const double* y, const double* z,
                      int ne, const int* elems, const int* nnd,
                      const double* manning, const double* vt,
                      int n_in, const int* in_segs, int n_out, const int* out_segs,
                      double Q, double S_in, double S_out)
{
  shapes_init();
  orc_model* m = xcalloc(1, sizeof(orc_model));
  m->np = np; m->ne = ne;
#define ALLOCD(p) m->p = xcalloc((size_t)np, sizeof(double))
  ALLOCD(x); ALLOCD(y); ALLOCD(z); ALLOCD(zor); ALLOCD(U); ALLOCD(V); ALLOCD(S);
  ALLOCD(dUdt); ALLOCD(dVdt); ALLOCD(dSdt); ALLOCD(cf); ALLOCD(nx); ALLOCD(ny);
  ALLOCD(niox); ALLOCD(nioy); ALLOCD(valU); ALLOCD(valS);
#undef ALLOCD
  memcpy(m->x, x, sizeof(double) * np); memcpy(m->y, y, sizeof(double) * np);
  memcpy(m->z, z, sizeof(double) * np); memcpy(m->zor, z, sizeof(double) * np);
  m->elem = xcalloc((size_t)8 * ne, sizeof(int)); memcpy(m->elem, elems, sizeof(int) * 8 * ne);
  m->nnd = xcalloc((size_t)ne, sizeof(int)); memcpy(m->nnd, nnd, sizeof(int) * ne);
  m->manning = xcalloc((size_t)ne, sizeof(double)); memcpy(m->manning, manning, sizeof(double) * ne);
  m->vt = xcalloc((size_t)ne, sizeof(double)); memcpy(m->vt, vt, sizeof(double) * ne);
  m->areaFact = xcalloc((size_t)ne, sizeof(double));
  m->n_in = n_in; m->n_out = n_out;
  m->in_segs = xcalloc((size_t)3 * n_in, sizeof(int)); memcpy(m->in_segs, in_segs, sizeof(int) * 3 * n_in);
  m->out_segs = xcalloc((size_t)3 * n_out, sizeof(int)); memcpy(m->out_segs, out_segs, sizeof(int) * 3 * n_out);
  m->Q = Q; m->S_in = S_in; m->S_out = S_out;
  m->noel = xcalloc((size_t)np, sizeof(int));
  m->flag = xcalloc((size_t)np, sizeof(unsigned));
  m->kind = xcalloc((size_t)np, sizeof(unsigned));
  m->rot = xcalloc((size_t)np, sizeof(int));
  m->eqno = xcalloc((size_t)3 * np, sizeof(int));
  /* PROJECT defaults (Project.cpp:60-110) and the decks' $DRYREW / $RELAX / $CONVERGENCE */
  m->g = 9.81;
  m->vk = 1.78e-6 / (1.0 + 0.0337 * 10.0 + 0.00022 * 10.0 * 10.0);
  orc_set_params(m, 2, 2, 0.001, 10.0, 10.0, 1.0e-7, 1.0e-7, 100.0);
  for (int i = 0; i < np; i++) m->S[i] = m->z[i];
  m->turb = 1; m->turb_stale = 1;
  m->cm = 0.09; m->cd = 1.00;                   /* KDCONST defaults (Project.cpp:117-118) */
  m->ste_e = xcalloc((size_t)ne, sizeof(double));
  m->vtn = xcalloc((size_t)np, sizeof(double));
  m->uun = xcalloc((size_t)np, sizeof(double));
  m->exi_e = xcalloc((size_t)ne, sizeof(double)); m->eyi_e = xcalloc((size_t)ne, sizeof(double));
  m->exx = xcalloc((size_t)np, sizeof(double)); m->exy = xcalloc((size_t)np, sizeof(double));
  m->eyy = xcalloc((size_t)np, sizeof(double));
  m->cdown = xcalloc((size_t)np, sizeof(int));
  m->nmark = xcalloc((size_t)np, 1); m->emark = xcalloc((size_t)ne, 1);
  m->eflag = xcalloc((size_t)ne, sizeof(unsigned));
  m->rewetLimit = 0.005; m->countDown = 5; m->firstDryRew = 1;   /* the decks' $DRYREW */
  m->Uo = xcalloc((size_t)np, sizeof(double)); m->Vo = xcalloc((size_t)np, sizeof(double));
  m->So = xcalloc((size_t)np, sizeof(double)); m->dUdto = xcalloc((size_t)np, sizeof(double));
  m->dVdto = xcalloc((size_t)np, sizeof(double)); m->dSdto = xcalloc((size_t)np, sizeof(double));
  area_factors(m);
  topology(m);
  return m;
}

void orc_destroy(orc_model* m)
{
  if (!m) return;
  void* ps[] = { m->x, m->y, m->z, m->zor, m->U, m->V, m->S, m->dUdt, m->dVdt, m->dSdt, m->cf,
                 m->nx, m->ny, m->niox, m->nioy, m->valU, m->valS, m->elem, m->nnd, m->manning,
                 m->vt, m->areaFact, m->in_segs, m->out_segs, m->noel, m->flag, m->kind, m->rot,
                 m->eqno, m->bnd, m->bflag, m->me, m->rowptr, m->cols, m->veg_dp, m->veg_sp,
                 m->ste_e, m->vtn, m->uun, m->Uo, m->Vo, m->So, m->dUdto, m->dVdto, m->dSdto,
                 m->exi_e, m->eyi_e, m->exx, m->exy, m->eyy, m->cdown, m->nmark, m->emark,
                 m->eflag };
  for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
  free(m);
}

/* The time terms' diagonal at the start of a flow cycle (EQS_UVS2D::Execute, EqsUVS2D.cpp:
 * 97-147): relaxation methods 3 and 4 start from the relaxed time step relaxTimeFlow (the
 * stationary cycle: 1 / dt_r; unsteady: 1 / dt_r / theta), the others use no time term
 * (stationary) or the time step (1 / dt / theta). */
static void cycle_start(orc_model* m)
{
  const double dt = m->unsteady ? m->dt : m->interval;
  m->dt_UV = m->dt_H = m->last_dt = m->relaxDt > 0.0 && m->relaxDt <= dt ? m->relaxDt : dt;
  if (m->relaxMethod >= 3) {
    const double th = m->unsteady ? m->theta : 1.0;
    m->relaxThdt_UV = 1.0 / m->dt_UV / th;
    m->relaxThdt_H = 1.0 / m->dt_H / th;
  } else if (m->unsteady) {
    m->relaxThdt_UV = m->relaxThdt_H = 1.0 / m->dt / m->theta;
  } else {
    m->relaxThdt_UV = m->relaxThdt_H = 0.0;
  }
}

void orc_set_params(orc_model* m, int relaxMethod, int dryMethod, double dryLimit,
                    double maxDeltaUV, double maxDeltaS, double convUV, double convS,
                    double maxUs)
{
  m->relaxMethod = relaxMethod; m->dryMethod = dryMethod; m->dryLimit = dryLimit;
  m->hmin = dryLimit / 2.0;                     /* Project.cpp:1125 */
  m->maxDeltaUV = maxDeltaUV; m->maxDeltaS = maxDeltaS;
  m->convUV = convUV; m->convS = convS; m->maxUs = maxUs;
  cycle_start(m);
}

void orc_set_relax_time(orc_model* m, double relaxTimeFlow, double interval)
{
  m->relaxDt = relaxTimeFlow;
  m->interval = interval;
  cycle_start(m);
}

void orc_get_relax_time(const orc_model* m, double* dt_UV, double* dt_H, double* last_dt)
{
  if (dt_UV) *dt_UV = m->dt_UV;
  if (dt_H) *dt_H = m->dt_H;
  if (last_dt) *last_dt = m->last_dt;
}

void orc_set_state(orc_model* m, const double* U, const double* V, const double* S)
{
  /* an initial condition: the previous time step's state is the same, with zero time
   * derivatives (GRID::InputInit, Grid.cpp:1119-1129) */
  const size_t b = sizeof(double) * (size_t)m->np;
  memcpy(m->U, U, b); memcpy(m->V, V, b); memcpy(m->S, S, b);
  memcpy(m->Uo, U, b); memcpy(m->Vo, V, b); memcpy(m->So, S, b);
  memset(m->dUdt, 0, b); memset(m->dVdt, 0, b); memset(m->dSdt, 0, b);
  memset(m->dUdto, 0, b); memset(m->dVdto, 0, b); memset(m->dSdto, 0, b);
  m->turb_stale = 1;
}

void orc_set_time_state(orc_model* m, const double* Uo, const double* Vo, const double* So,
                        const double* dUdto, const double* dVdto, const double* dSdto,
                        const double* dUdt, const double* dVdt, const double* dSdt)
{
  const size_t b = sizeof(double) * (size_t)m->np;
  if (Uo) memcpy(m->Uo, Uo, b);
  if (Vo) memcpy(m->Vo, Vo, b);
  if (So) memcpy(m->So, So, b);
  if (dUdto) memcpy(m->dUdto, dUdto, b);
  if (dVdto) memcpy(m->dVdto, dVdto, b);
  if (dSdto) memcpy(m->dSdto, dSdto, b);
  if (dUdt) memcpy(m->dUdt, dUdt, b);
  if (dVdt) memcpy(m->dVdt, dVdt, b);
  if (dSdt) memcpy(m->dSdt, dSdt, b);
}

void orc_get_time_state(const orc_model* m, double* Uo, double* Vo, double* So,
                        double* dUdto, double* dVdto, double* dSdto,
                        double* dUdt, double* dVdt, double* dSdt)
{
  const size_t b = sizeof(double) * (size_t)m->np;
  if (Uo) memcpy(Uo, m->Uo, b);
  if (Vo) memcpy(Vo, m->Vo, b);
  if (So) memcpy(So, m->So, b);
  if (dUdto) memcpy(dUdto, m->dUdto, b);
  if (dVdto) memcpy(dVdto, m->dVdto, b);
  if (dSdto) memcpy(dSdto, m->dSdto, b);
  if (dUdt) memcpy(dUdt, m->dUdt, b);
  if (dVdt) memcpy(dVdt, m->dVdt, b);
  if (dSdt) memcpy(dSdt, m->dSdt, b);
}

/* Start of a flow time step.  dt <= 0: stationary flow, no time terms (EqsUVS2D.cpp:
 * 105-127 with relaxMethod < 3).  dt > 0: the theta scheme, relaxThdt = 1 / dt / theta in
 * the Jacobian (:130-147), then EQS_UVS2D::Predict (:1101-1180): current time derivatives
 * zero; at nodes whose previous level lies below zor + dryLimit the previous time
 * derivatives are zero and, at the first time step or where the node was dry at the
 * previous time step (:1153), the previous state is the current one.  The algebraic eddy viscosity becomes due (first Newton step of the cycle).
 * Returns -1 for theta or dt too small (:91-94). */
int orc_time_step(orc_model* m, double dt, double theta, int first)
{
  m->turb_stale = 1;
  if (dt <= 0.0) {
    m->unsteady = 0;
    cycle_start(m);
    return 0;
  }
  if (fabs(theta) < 1.0e-10 || fabs(dt) < 1.0e-10) return -1;
  m->unsteady = 1; m->dt = dt; m->theta = theta;
  cycle_start(m);
  for (int n = 0; n < m->np; n++) m->dUdt[n] = m->dVdt[n] = m->dSdt[n] = 0.0;
  for (int n = 0; n < m->np; n++) {
    if (m->So[n] < m->zor[n] + m->dryLimit) {
      if (first || (m->flag[n] & fDryPrev)) { m->Uo[n] = m->U[n]; m->Vo[n] = m->V[n]; m->So[n] = m->S[n]; }
      m->dUdto[n] = m->dVdto[n] = m->dSdto[n] = 0.0;
    }
  }
  return 0;
}

/* EQS_UVS2D::Timegrad (EqsUVS2D.cpp:1183-1230): zero at nodes dry or marsh now or at the
 * previous time step, else d/dt = (1 - 1/theta) d/dt_old + (v - v_old) / (dt theta) */
static void timegrad(orc_model* m)
{
  const double iTheta = 1.0 - 1.0 / m->theta;
  const double thdt = 1.0 / m->dt / m->theta;
  for (int n = 0; n < m->np; n++) {
    if (m->flag[n] & (fDry | fDryPrev | fMarsh | fMarshPrev)) {
      m->dUdt[n] = m->dVdt[n] = m->dSdt[n] = 0.0;
      continue;
    }
    m->dUdt[n] = iTheta * m->dUdto[n] + thdt * (m->U[n] - m->Uo[n]);
    m->dVdt[n] = iTheta * m->dVdto[n] + thdt * (m->V[n] - m->Vo[n]);
    m->dSdt[n] = iTheta * m->dSdto[n] + thdt * (m->S[n] - m->So[n]);
  }
}

/* QINLET discharge and reference level for the following Newton steps (a hydrograph gives
 * them per time step: BCONLINE::Inlet, BconLine.cpp:263-276; its line carries the QTINLET
 * kind as well) */
void orc_set_inflow(orc_model* m, double Q, double S_in, int hydrograph)
{
  m->Q = Q;
  m->S_in = S_in;
  m->inflow_qt = hydrograph != 0;
}

/* BCONLINE::FindInterval (BconLine.cpp:167-226) without a second coordinate, then the
 * interpolation of BconLine.cpp:268-275; Q[0], S[0] when no interval holds the time */
void orc_hydrograph(int n, const double* t, const double* Q, const double* S, double time,
                    double* Qout, double* Sout)
{
  const double kEps = 1.0e-20;           /* Defs.h:85 */
  *Qout = Q[0];
  *Sout = S[0];
  for (int i = 0; i < n - 1; i++) {
    double dx = t[i + 1] - t[i];
    double lx = time - t[i];
    if (fabs(dx) > kEps) lx /= dx;
    else lx = dx = 0.0;
    double l;
    if (fabs(dx) < kEps && fabs(lx) < kEps) {
      l = 0.0;                           /* ly = 0 */
    } else {
      if (!(lx > -0.001 && lx < 1.001)) continue;
      l = lx;
    }
    *Qout = Q[i] + l * (Q[i + 1] - Q[i]);
    *Sout = S[i] + l * (S[i + 1] - S[i]);
    return;
  }
}

/* End of a time step: the state and the dry and marsh flags become the previous ones
 * (Compute.cpp:773-796) */
void orc_time_commit(orc_model* m)
{
  const size_t b = sizeof(double) * (size_t)m->np;
  for (int n = 0; n < m->np; n++) {
    m->flag[n] &= ~(unsigned)(fDryPrev | fMarshPrev);
    if (m->flag[n] & fDry) m->flag[n] |= fDryPrev;
    if (m->flag[n] & fMarsh) m->flag[n] |= fMarshPrev;
  }
  memcpy(m->Uo, m->U, b); memcpy(m->Vo, m->V, b); memcpy(m->So, m->S, b);
  memcpy(m->dUdto, m->dUdt, b); memcpy(m->dVdto, m->dVdt, b); memcpy(m->dSdto, m->dSdt, b);
}

void orc_get_state(const orc_model* m, double* U, double* V, double* S, double* z)
{
  if (U) memcpy(U, m->U, sizeof(double) * m->np);
  if (V) memcpy(V, m->V, sizeof(double) * m->np);
  if (S) memcpy(S, m->S, sizeof(double) * m->np);
  if (z) memcpy(z, m->z, sizeof(double) * m->np);
}

static double getrot(const orc_model* m, int n, int r, int c)
{
  /* BCON::Getrot (Bcon.cpp:210-238) */
  switch (m->rot[n]) {
  case rotTangent:
    if (r == 0 && c == 0) return m->ny[n];
    if (r == 0 && c == 1) return m->nx[n];
    if (r == 1 && c == 0) return -m->nx[n];
    return m->ny[n];
  case rotNormal:
    if (r + c == 0) return m->niox[n];
    if (r + c == 1) return m->nioy[n];
    return -m->niox[n];
  }
  return 0.0;
}

void orc_get_nodal(const orc_model* m, double* cf, int* bckind, int* rot, double* rotm,
                   double* bcU, double* bcS)
{
  for (int n = 0; n < m->np; n++) {
    if (cf) cf[n] = m->cf[n];
    if (bckind) bckind[n] = (int)m->kind[n];
    if (rot) rot[n] = m->rot[n];
    if (rotm) {
      rotm[4 * n + 0] = getrot(m, n, 0, 0); rotm[4 * n + 1] = getrot(m, n, 0, 1);
      rotm[4 * n + 2] = getrot(m, n, 1, 0); rotm[4 * n + 3] = getrot(m, n, 1, 1);
    }
    if (bcU) bcU[n] = m->valU[n];
    if (bcS) bcS[n] = m->valS[n];
  }
}

/* ------------------------------------------------------------------------------------- */
/* nodal coefficients                                                                     */

/* Drag coefficient of one cylinder from its Reynolds number (TYPE::dragCoeff,
 * Lindner.cpp:264-278). */
static double drag_cylinder(double v, double d, double vk)
{
  double re = v * d / vk;
  if (re <= 800.0) return 3.07 / pow(re, 0.168);
  if (re <= 6000.0) return 1.0;
  if (re <= 11000.0) return 1.0 + 0.2 * (re - 6000.0) / 5000.0;
  return 1.2;
}

/* Real roots of a x^3 + b x^2 + c x + d (Cardano / trigonometric form, as
 * TYPE::cubeEquation, Lindner.cpp:281-335); returns their number. */
static int cubic_roots(double a, double b, double c, double d, double x[3])
{
  const double pi = 3.141592653589793;
  double ba = b / a / 3.0, ca = c / a;
  double p = ca / 3.0 - ba * ba;
  double q = ba * ba * ba - ba * ca / 2.0 + d / a / 2.0;
  double disc = q * q + p * p * p;
  if (disc > 0.0) {
    double t = -q + sqrt(disc);
    double u = (t / fabs(t)) * pow(fabs(t), 1.0 / 3.0);
    t = -q - sqrt(disc);
    double v = (t / fabs(t)) * pow(fabs(t), 1.0 / 3.0);
    x[0] = (u + v) - ba;
    return 1;
  }
  double t = -q / pow(-p, 1.5);
  double phi = t >= 1.0 ? 0.0 : (t <= -1.0 ? pi : acos(t));
  x[0] = 2.0 * sqrt(-p) * cos(phi / 3.0) - ba;
  x[1] = 2.0 * sqrt(-p) * cos((phi + 2.0 * pi) / 3.0) - ba;
  x[2] = 2.0 * sqrt(-p) * cos((phi + 4.0 * pi) / 3.0) - ba;
  return 3;
}

/* Friction coefficient with emergent vegetation after Lindner (TYPE::lindner,
 * Lindner.cpp:72-262): the drag coefficient cWR of the stems is iterated (Pegasus
 * updates after the third step, at most 200 steps, relative tolerance 1e-3), with the wake
 * length aNL from its own fixed-point iteration (damped by 1/2) and the depth ratio from
 * the cubic of the momentum balance; returns lambda / 8 with
 * lambda = 8 cb + 4 cWR h dp / sp^2, cb without vegetation when dp, sp, v or h are below
 * the thresholds, or -1 when cWR does not converge. */
static double lindner(double dp, double sp, double va, double ha, double cb, double vk, double g)
{
  if (dp < 1.0e-4 || sp < 1.0e-4 || va < 1.0e-3 || ha < 1.0e-3) return cb;
  double cwr = 1.0, anl = sp / 2.0;
  double c1 = 1.0, c2 = 1.0, r1 = 0.0, r2 = 0.0, lambda = 0.0;
  int it, done = 0;
  for (it = 0; it < 200; it++) {
    lambda = 8.0 * cb + 4.0 * cwr * ha * dp / sp / sp;
    double cw = drag_cylinder(va, dp, vk);
    int jt, wake = 0;
    for (jt = 0; jt < 200; jt++) {
      double t1 = 1.0 + anl * lambda / 4.0 / ha;
      double t2 = 30.0 / pow(fabs(t1), 1.5);
      double ra = cw * dp * pow(fabs(t2), 1.429);
      if (fabs((ra - anl) / ra) < 1.0e-3) { anl = ra; wake = 1; break; }
      anl = 0.5 * (ra + anl);
    }
    if (!wake) anl = sp / 2.0;
    double anb = 0.24 * pow(fabs(anl), 0.59) * pow(fabs(cw * dp), 0.41);
    double vr = 1.151 * pow(fabs(anl / sp), -0.483) + 0.5 * pow(fabs(anb / sp), 1.1);
    double fr = va / sqrt(g * ha);
    double al = dp / sp;
    double ca = fr * fr * (1.0 - al * cwr / 2.0);
    double cbb = -fr * fr - (1.0 - al) / 2.0;
    double cd = (1.0 - al) / 2.0;
    double hr = 1.0;
    if (fabs(ca) < 1.0e-10) {
      hr = sqrt(-cd / cbb);
    } else {
      double x[3];
      int nr = cubic_roots(ca, cbb, 0.0, cd, x);
      for (int i = 0; i < nr; i++)
        if (x[i] > 0.0 && x[i] < 1.0) { hr = x[i]; break; }
    }
    double rc = 1.3124 * cw * vr + 2.0 * (1.0 - hr) / fr / fr;
    if (fabs((rc - cwr) / rc) < 1.0e-3) { done = 1; break; }
    double dc = rc - cwr;
    if (it >= 3 && r1 * r2 < 0.0) {
      if (r2 * dc < 0.0) r1 *= r2 / (r2 + dc);
      else { c1 = c2; r1 = r2; }
      c2 = cwr; r2 = dc;
      cwr = c2 - r2 * (c2 - c1) / (r2 - r1);
    } else {
      c1 = c2; r1 = r2;
      c2 = cwr; r2 = dc;
      cwr = (it >= 2 && r1 * r2 < 0.0) ? c2 - r2 * (c2 - c1) / (r2 - r1) : rc;
    }
  }
  if (!done) return -1.0;
  return lambda / 8.0;
}

/* MODEL::DoFriction (DoFriction.cpp:57-205) with TYPE::bottom (Friction.cpp:49-120) for
 * Manning's law (Friction.cpp:238-240, exponent literally 0.333333): cb is Manning plus a
 * zero Nikuradse form-roughness term; emergent vegetation after Lindner
 * (Friction.cpp:101-108: cb + 0.75 h dp / sp^2 where cWR does not converge); no submerged
 * vegetation (TYPE::velzen returns 0 without it, Velzen.cpp:77-78). */
static void do_friction(orc_model* m)
{
  int* counter = xcalloc((size_t)m->np, sizeof(int));
  for (int i = 0; i < m->np; i++) m->cf[i] = 0.0;
  for (int e = 0; e < m->ne; e++) {
    int nnd = m->nnd[e];
    double n = m->manning[e];
    double dp = m->veg_dp ? m->veg_dp[e] : 0.0, sp = m->veg_sp ? m->veg_sp[e] : 0.0;
    for (int i = 0; i < nnd; i++) {
      int nd = m->elem[8 * e + i];
      double h = m->S[nd] - m->z[nd];
      if (h < m->hmin) h = m->hmin;
      double vres = sqrt(m->U[nd] * m->U[nd] + m->V[nd] * m->V[nd]);
      double cb = m->g * (n * n) / pow(h, 0.333333);
      cb += 0.0;      /* Nikuradse term of zero form roughness: 0.41/log(inf) squared */
      double cp = lindner(dp, sp, vres, h, cb, m->vk, m->g);
      if (cp < -0.9) cp = cb + 0.75 * h * dp / sp / sp;
      double cf = cp + 0.0;
      m->cf[nd] += m->areaFact[e] * cf;
      counter[nd]++;
    }
  }
  for (int i = 0; i < m->np; i++) if (counter[i]) m->cf[i] /= counter[i];
  free(counter);
}

void orc_set_vegetation(orc_model* m, const double* dp, const double* sp)
{
  free(m->veg_dp); free(m->veg_sp);
  m->veg_dp = m->veg_sp = NULL;
  if (!dp || !sp) return;
  m->veg_dp = xcalloc((size_t)m->ne, sizeof(double));
  m->veg_sp = xcalloc((size_t)m->ne, sizeof(double));
  memcpy(m->veg_dp, dp, sizeof(double) * m->ne);
  memcpy(m->veg_sp, sp, sizeof(double) * m->ne);
}

/* Turbulence model of the flow cycle.  code is the $TM_TURBULENCE value as Timeint.cpp:
 * 1333-1360 decodes it: last digit 1 = constant eddy viscosity of the element type
 * (kVtConstant), 2 = algebraic model (kVtAlgebraic); tens digit 1 = anisotropic eddy
 * viscosity after Elder (kVtAnisotrop, the flow cycle then runs EQS_UVS2D_AI, Compute.cpp:
 * 451-458); hundreds digit bit 1 = the type's vt is a lower bound (kVtMin), bit 2 = the
 * nodal quantities are recomputed at every Newton step (kVtIterat) instead of only at the
 * first one (EqsUVS2D.cpp:191, 311).
 * st, kdest, estx, esty: per element ($TURBULENCE no vt st estx esty KDest, Type.cpp:
 * 681-683; null: the TYPE defaults 1.0, 0.6, 0, 0); cm, cd: KDCONST.
 * Returns -1 for a code outside these. */
int orc_set_turbulence(orc_model* m, int code, const double* st, const double* kdest,
                       const double* estx, const double* esty, double cm, double cd)
{
  const int model = code % 10, aniso = (code / 10) % 10, hund = (code / 100) % 10;
  if (code < 0 || (model != 1 && model != 2) || aniso > 1 || hund > 3 || code >= 1000)
    return -1;
  m->turb = code;
  for (int e = 0; e < m->ne; e++) {
    const double kd = kdest ? kdest[e] : 0.6;
    m->ste_e[e] = (st ? st[e] : 1.0) * kd;                         /* Type.cpp:73-74 */
    m->exi_e[e] = (estx ? estx[e] : 0.0) / kd;                     /* Turbulence.cpp:1024-1025 */
    m->eyi_e[e] = (esty ? esty[e] : 0.0) / kd;
  }
  m->cm = cm; m->cd = cd;
  m->turb_stale = 1;
  return 0;
}

/* GRID::Turbulence, constant branch (Turbulence.cpp:110-165): nodal vt = the mean of the
 * element types' vt over the incident elements (every element is wet here), uu = 0.  The
 * isotropic flow cycle does not use it (its elements take the type's vt); the anisotropic
 * one interpolates it. */
static void turbulence_constant(orc_model* m)
{
  int* cnt = xcalloc((size_t)m->np, sizeof(int));
  for (int n = 0; n < m->np; n++) { m->vtn[n] = 0.0; m->uun[n] = 0.0; }
  for (int e = 0; e < m->ne; e++)
    for (int i = 0; i < m->nnd[e]; i++) {
      const int no = m->elem[8 * e + i];
      m->vtn[no] += m->vt[e];
      cnt[no]++;
    }
  for (int n = 0; n < m->np; n++)
    if (cnt[n]) m->vtn[n] /= cnt[n];
  free(cnt);
}

/* GRID::Turbulence, anisotropic part (Turbulence.cpp:996-1090): per node the mean of
 * estx / KDest and esty / KDest over the incident elements, turned into the direction of
 * the nodal velocity (angle to the x axis; none below |U| = 1e-6):
 * exx = ex ca^2 + ey sa^2, exy = (ex - ey) sa ca, eyy = ex sa^2 + ey ca^2. */
static void turbulence_elder(orc_model* m)
{
  int* cnt = xcalloc((size_t)m->np, sizeof(int));
  for (int n = 0; n < m->np; n++) { m->exx[n] = 0.0; m->eyy[n] = 0.0; }
  for (int e = 0; e < m->ne; e++)
    for (int i = 0; i < m->nnd[e]; i++) {
      const int no = m->elem[8 * e + i];
      m->exx[no] += m->exi_e[e];
      m->eyy[no] += m->eyi_e[e];
      cnt[no]++;
    }
  for (int n = 0; n < m->np; n++) {
    if (cnt[n]) { m->exx[n] /= cnt[n]; m->eyy[n] /= cnt[n]; }
    const double U = m->U[n], V = m->V[n], Us = sqrt(U * U + V * V);
    double sa = 0.0, ca = 1.0;
    if (Us > 1.0e-6) { sa = V / Us; ca = U / Us; }
    const double ex = m->exx[n], ey = m->eyy[n];
    m->exx[n] = ex * ca * ca + ey * sa * sa;
    m->exy[n] = (ex - ey) * sa * ca;
    m->eyy[n] = ex * sa * sa + ey * ca * ca;
  }
  free(cnt);
}

void orc_get_elder(const orc_model* m, double* exx, double* exy, double* eyy)
{
  if (exx) memcpy(exx, m->exx, sizeof(double) * m->np);
  if (exy) memcpy(exy, m->exy, sizeof(double) * m->np);
  if (eyy) memcpy(eyy, m->eyy, sizeof(double) * m->np);
}

/* GRID::Turbulence, algebraic branch (Turbulence.cpp:292-352): per node the mean of
 * st * KDest over the incident elements (all wet here), then vt = ste H Utau with
 * Utau = sqrt(cf (U^2 + V^2)), H = S - z unclipped, and uu = vv = -2/3 Utau^2
 * sqrt(cK ste / cm / cd), cK = 1 / sqrt(cf); uv = 0.  Nodes of no element keep zeros.
 * Uses the nodal friction coefficient of the current state (do_friction first). */
static void turbulence_algebraic(orc_model* m)
{
  int* cnt = xcalloc((size_t)m->np, sizeof(int));
  double* ste = xcalloc((size_t)m->np, sizeof(double));
  for (int n = 0; n < m->np; n++) { m->vtn[n] = 0.0; m->uun[n] = 0.0; }
  for (int e = 0; e < m->ne; e++)
    for (int i = 0; i < m->nnd[e]; i++) {
      const int no = m->elem[8 * e + i];
      ste[no] += m->ste_e[e];
      cnt[no]++;
    }
  for (int n = 0; n < m->np; n++) {
    if (!cnt[n]) continue;
    ste[n] /= cnt[n];
    const double U = m->U[n], V = m->V[n], H = m->S[n] - m->z[n];
    const double Utau = sqrt(m->cf[n] * (U * U + V * V));
    const double cK = 1.0 / sqrt(m->cf[n]);
    m->vtn[n] = ste[n] * H * Utau;
    m->uun[n] = -2.0 / 3.0 * Utau * Utau * sqrt(cK * ste[n] / m->cm / m->cd);
  }
  free(cnt); free(ste);
}

void orc_get_turbulence(const orc_model* m, double* vt, double* uu)
{
  if (vt) memcpy(vt, m->vtn, sizeof(double) * m->np);
  if (uu) memcpy(uu, m->uun, sizeof(double) * m->np);
}

void orc_turbulence(orc_model* m, double* vt, double* uu)
{
  do_friction(m);
  if (m->turb % 10 == 2) turbulence_algebraic(m);
  else turbulence_constant(m);
  if ((m->turb / 10) % 10 == 1) turbulence_elder(m);
  m->turb_stale = 0;
  if (vt) memcpy(vt, m->vtn, sizeof(double) * m->np);
  if (uu) memcpy(uu, m->uun, sizeof(double) * m->np);
}

double orc_lindner(double dp, double sp, double va, double ha, double cb, double vk, double g)
{
  return lindner(dp, sp, va, ha, cb, vk, g);
}

/* BCONSET::InitBcon (BconSet.cpp:92-480) with BCONLINE::GenBcon / Inlet / Outlet
 * (BconLine.cpp:62-150, 249-640, 855-900) for one QINLET line and one OUTLET line with a
 * single water level, integrated with the quadratic line rule (order 2, Bcon.h:303). */
static void init_bcon(orc_model* m)
{
  int np = m->np;
  char* mark = xcalloc((size_t)np, 1);
  /* bc entries in creation order: node, kind, U, S */
  int nbc = 0;
  int* bno = xcalloc((size_t)np, sizeof(int));
  unsigned* bkind = xcalloc((size_t)np, sizeof(unsigned));
  double* bU = xcalloc((size_t)np, sizeof(double));
  double* bS = xcalloc((size_t)np, sizeof(double));
  /* --- QINLET ----------------------------------------------------------------------- */
  {
    unsigned kind = kQInlet | kInlet | (m->inflow_qt ? kQTInlet : 0u);  /* BconLine.cpp:86-87 */
    double Qspec = m->Q, Sspec = m->S_in;
    double Q = 0.0;
    int any = 0;
    memset(mark, 0, (size_t)np);
    for (int c = 0; c < m->n_in; c++) {
      const int* nd = m->in_segs + 3 * c;
      double q[3];
      for (int i = 0; i < 3; i++) {
        int n = nd[i];
        if (m->S[n] - m->z[n] < 1.0e-9 || (m->flag[n] & (fDry | fMarsh))) {
          q[i] = 0.0;
        } else {
          double h = Sspec - m->z[n];
          if (h > 0.0) q[i] = sqrt(h * h * h); else q[i] = 0.0;
          double cf = m->cf[n];
          if (cf > 0.0) q[i] /= sqrt(cf);
          any = 1;
        }
        mark[n] = 1;
      }
      for (int g = 0; g < g_line.ngp; g++) {
        const double* N = g_line.f[g];
        const double* dN = g_line.dfdx[g];
        double dx = dN[0] * m->x[nd[0]] + dN[1] * m->x[nd[1]] + dN[2] * m->x[nd[2]];
        double dy = dN[0] * m->y[nd[0]] + dN[1] * m->y[nd[1]] + dN[2] * m->y[nd[2]];
        double qg = N[0] * q[0] + N[1] * q[1] + N[2] * q[2];
        double weight = g_line.w[g] * sqrt(dx * dx + dy * dy);
        Q += weight * qg;
      }
    }
    double Qratio = any ? ((fabs(Q) > 0.0) ? (-Qspec / Q) : 0.0) : 0.0;
    for (int n = 0; n < np; n++) {
      if (!mark[n]) continue;
      double h = Sspec - m->z[n];
      double q = (h > 0.0) ? sqrt(h * h * h) : 0.0;
      double cf = m->cf[n];
      if (cf > 0.0) q /= sqrt(cf);
      if (m->flag[n] & (fDry | fMarsh)) q = 0.0;
      bno[nbc] = n; bkind[nbc] = kind; bU[nbc] = q * Qratio; bS[nbc] = Sspec; nbc++;
    }
  }
  /* --- OUTLET ------------------------------------------------------------------------ */
  {
    memset(mark, 0, (size_t)np);
    for (int c = 0; c < m->n_out; c++) {
      const int* nd = m->out_segs + 3 * c;
      for (int i = 0; i < 3; i++) {
        int n = nd[i];
        if (mark[n]) continue;
        mark[n] = 1;
        int b;
        for (b = 0; b < nbc; b++) if (bno[b] == n) break;
        if (b == nbc) {
          bno[nbc] = n; bkind[nbc] = kOutlet; bU[nbc] = 0.0; bS[nbc] = m->S_out; nbc++;
        } else if (!(bkind[b] & kInlet)) {
          bkind[b] |= kOutlet; bS[b] = m->S_out;
        }
      }
    }
  }
  for (int n = 0; n < np; n++) { m->kind[n] = 0; m->valU[n] = 0.0; m->valS[n] = 0.0; }
  for (int b = 0; b < nbc; b++) {
    unsigned k = bkind[b];
    if (k & kInlet) k &= ~(unsigned)(kOutlet | kSlip | kSetUV);
    m->kind[bno[b]] = k; m->valU[bno[b]] = bU[b]; m->valS[bno[b]] = bS[b];
  }
  free(mark); free(bno); free(bkind); free(bU); free(bS);
  m->bcon_init = 1;
}

/* MODEL::SetLocation (Locate.cpp:47-85) */
static void set_location(orc_model* m)
{
  for (int b = 0; b < m->nb; b++) {
    const int* nd = m->bnd + 3 * b;
    int inlet = 1, outlet = 1, open = 1;
    for (int i = 0; i < 2; i++) {
      unsigned k = m->kind[nd[i]];
      if (!(k & kInlet)) inlet = 0;
      if (!(k & kOutlet)) outlet = 0;
      if (!(k & kOpenBnd)) open = 0;
    }
    if (inlet) m->bflag[b] |= eInlet;
    if (outlet) m->bflag[b] |= eOutlet;
    if (open) m->bflag[b] |= eOpen;
    for (int i = 0; i < 3; i++) {
      if (inlet) m->flag[nd[i]] |= fInlet;
      if (outlet) m->flag[nd[i]] |= fOutlet;
      if (open) m->flag[nd[i]] |= fOpen;
    }
  }
}

/* MODEL::SetNormal (Transform.cpp:45-200) */
static void set_normal(orc_model* m)
{
  for (int n = 0; n < m->np; n++) m->nx[n] = m->ny[n] = m->niox[n] = m->nioy[n] = 0.0;
  for (int b = 0; b < m->nb; b++) {
    const int* nd = m->bnd + 3 * b;
    for (int i = 0; i < 3; i++) {
      const double* dndx = g_line.dndx[i];
      double nx = 0.0, ny = 0.0;
      for (int j = 0; j < 3; j++) {
        nx += 2.0 * (dndx[j] * m->y[nd[j]]);
        ny -= 2.0 * (dndx[j] * m->x[nd[j]]);
      }
      int n = nd[i];
      if (m->bflag[b] & (eInlet | eOutlet | eOpen)) { m->niox[n] += nx; m->nioy[n] += ny; }
      else { m->nx[n] += nx; m->ny[n] += ny; }
    }
  }
  for (int n = 0; n < m->np; n++) {
    if (!(m->flag[n] & fBound)) continue;
    double len = sqrt(m->nx[n] * m->nx[n] + m->ny[n] * m->ny[n]);
    if (fabs(len) < 1.0e-12) { m->nx[n] = 0.0; m->ny[n] = 0.0; }
    else { m->nx[n] /= len; m->ny[n] /= len; }
    len = sqrt(m->niox[n] * m->niox[n] + m->nioy[n] * m->nioy[n]);
    if (fabs(len) < 1.0e-12) { m->niox[n] = 0.0; m->nioy[n] = 0.0; }
    else { m->niox[n] /= len; m->nioy[n] /= len; }
  }
}

/* MODEL::SetRotation (Transform.cpp:203-420) */
static void set_rotation(orc_model* m)
{
  for (int n = 0; n < m->np; n++) {
    m->flag[n] &= ~(unsigned)fRotat;
    m->rot[n] = 0;
    unsigned* k = &m->kind[n];
    if (*k & kInlet) *k &= ~(unsigned)(kFixU | kFixV);
    if (*k & kAutoSlip) *k &= ~(unsigned)(kAutoSlip | kFixU | kFixV);
    int bound = (m->flag[n] & fBound) != 0;
    if (bound && !(*k & (kInlet | kOutlet | kOpenBnd | kSlip | kSetUV))) {
      m->rot[n] = rotTangent;
      *k |= kAutoSlip | kFixV;
    }
  }
  for (int b = 0; b < m->nb; b++) {
    if (m->bflag[b] & (eOutlet | eOpen)) continue;
    for (int i = 0; i < 2; i++) {
      int n = m->bnd[3 * b + i];
      unsigned* k = &m->kind[n];
      if ((*k & kOutlet) || (*k & kOpenBnd)) {
        if (!(*k & kSlip) && !(*k & kSetUV)) { m->rot[n] = rotNormal; *k |= kAutoSlip | kFixV; }
      }
    }
  }
  for (int n = 0; n < m->np; n++) {
    if (m->kind[n] & kAutoSlip) {
      double len = sqrt(m->nx[n] * m->nx[n] + m->ny[n] * m->ny[n]);
      if (len < 1.0e-12) m->flag[n] |= fNoMoment;
    }
  }
  for (int n = 0; n < m->np; n++) {
    unsigned* k = &m->kind[n];
    if (*k & kInlet) { m->flag[n] |= fRotat; *k |= kFixV; m->rot[n] = rotNormal; }
    if ((m->flag[n] & fNoMoment) && !(*k & kSetUV) && !(*k & kOutlet) && !(*k & kOpenBnd)) {
      *k |= kAutoSlip | kFixU | kFixV;
      m->U[n] = m->V[n] = 0.0;
    } else if ((*k & kFixV) && (*k & kAutoSlip)) {
      m->flag[n] |= fRotat;
      m->rot[n] = ((*k & kOutlet) || (*k & kOpenBnd)) ? rotNormal : rotTangent;
    }
  }
}

/* model element k -> node list */
static int me_nodes(const orc_model* m, int k, const int** nodes, int* is_bound)
{
  int v = m->me[k];
  if (v >= 0) { *nodes = m->elem + 8 * v; *is_bound = 0; return m->nnd[v]; }
  int b = -v - 1;
  *nodes = m->bnd + 3 * b; *is_bound = 1; return 3;
}

/* EQS::SetEqno / LastEquation / ResetEqOrder (SetEqno.cpp:52-487), serial build:
 * provisional numbers in node order (corner U,V,H; midside U,V; fixed dofs removed),
 * then renumbered in the order of each dof's last occurrence over the model element list. */
static void set_eqno(orc_model* m)
{
  static const unsigned fix[3] = { kFixU, kFixV, kSetS };
  int np = m->np;
  int* eq = m->eqno;
  for (int i = 0; i < 3 * np; i++) eq[i] = -1;
  int neq = 0;
  for (int n = 0; n < np; n++) {
    if (m->flag[n] & fDry) continue;
    int ndf = (m->flag[n] & fCorn) ? 3 : 2;
    for (int j = 0; j < ndf; j++) {
      if (m->kind[n] & fix[j]) continue;
      eq[3 * n + j] = neq++;
    }
  }
  int* counter = xcalloc((size_t)neq, sizeof(int));
  for (int k = 0; k < m->nme; k++) {
    const int* nd; int isb; int nnd = me_nodes(m, k, &nd, &isb);
    for (int i = 0; i < nnd; i++)
      for (int j = 0; j < 3; j++) { int e = eq[3 * nd[i] + j]; if (e >= 0) counter[e]++; }
  }
  int* ord = xcalloc((size_t)neq, sizeof(int));
  int nn = 0;
  for (int k = 0; k < m->nme; k++) {
    const int* nd; int isb; int nnd = me_nodes(m, k, &nd, &isb);
    for (int i = 0; i < nnd; i++)
      for (int j = 0; j < 3; j++) {
        int e = eq[3 * nd[i] + j];
        if (e >= 0) { counter[e]--; if (counter[e] == 0) ord[e] = nn++; }
      }
  }
  for (int i = 0; i < 3 * np; i++) if (eq[i] >= 0) eq[i] = ord[eq[i]];
  m->neq = neq;
  free(counter); free(ord);
}

/* EQS::SetIndexMat (IndexMat.cpp:91-342): each row starts with its diagonal, then the
 * columns in order of first appearance over the model element list. */
static void set_index_mat(orc_model* m)
{
  int neq = m->neq, mceq = 128;
  int* width = xcalloc((size_t)neq, sizeof(int));
  int* index = xcalloc((size_t)neq * mceq, sizeof(int));
  for (int k = 0; k < m->nme; k++) {
    const int* nd; int isb; int nnd = me_nodes(m, k, &nd, &isb);
    for (int i = 0; i < nnd; i++)
      for (int j = 0; j < 3; j++) {
        int row = m->eqno[3 * nd[i] + j];
        if (row < 0) continue;
        int* ip = index + (size_t)row * mceq;
        if (!width[row]) { width[row] = 1; ip[0] = row; }
        for (int kk = 0; kk < nnd; kk++)
          for (int l = 0; l < 3; l++) {
            int col = m->eqno[3 * nd[kk] + l];
            if (col < 0) continue;
            int n;
            for (n = 0; n < width[row]; n++) if (ip[n] == col) break;
            if (n == width[row]) { if (width[row] >= mceq) abort(); ip[width[row]++] = col; }
          }
      }
  }
  free(m->rowptr); free(m->cols);
  m->rowptr = xcalloc((size_t)neq + 1, sizeof(int));
  long nnz = 0;
  for (int r = 0; r < neq; r++) { m->rowptr[r] = (int)nnz; nnz += width[r]; }
  m->rowptr[neq] = (int)nnz;
  m->cols = xcalloc((size_t)nnz, sizeof(int));
  for (int r = 0; r < neq; r++)
    memcpy(m->cols + m->rowptr[r], index + (size_t)r * mceq, sizeof(int) * width[r]);
  m->nnz = nnz;
  free(width); free(index);
}

/* GRID::DryRewet, method 2 (DryRewet.cpp:696-900): nodes with H = S - zor below dryLimit
 * are dry (a count down keeps them dry for countDown further calls); elements with all
 * corners dry are dry, with some corners dry marsh; every node of a wet element is wet.
 * Marsh nodes that were dry before (or on the first call) take the water level of their
 * wet neighbours, clipped to zor + rewetLimit, and the bed is lowered to S - dryLimit
 * where the level lies below zor + dryLimit; midside nodes interpolate S and z; dry
 * nodes are reset.  Counts the elements that got dry and got wet. */
static void dry_rewet_method2(orc_model* m, int* dried, int* wetted)
{
  const double dryLimit = m->dryLimit, rewetLimit = m->rewetLimit;
  for (int n = 0; n < m->np; n++) {
    m->nmark[n] = (m->flag[n] & fDry) != 0;
    m->flag[n] &= ~(unsigned)fMarsh;
    double H = m->S[n] - m->zor[n];
    if (H < dryLimit) {
      m->flag[n] |= fDry;
      m->cdown[n] = m->countDown;
    } else {
      if (m->cdown[n] > 0) m->cdown[n]--;
      else m->flag[n] &= ~(unsigned)fDry;
    }
    m->z[n] = m->zor[n];
  }
  for (int e = 0; e < m->ne; e++) {
    const int* nd = m->elem + 8 * e;
    int ncn = m->nnd[e] == 6 ? 3 : 4;
    m->emark[e] = (m->eflag[e] & elDry) != 0;
    m->eflag[e] &= ~(unsigned)(elDry | elMarsh);
    int ndry = 0;
    for (int i = 0; i < ncn; i++) if (m->flag[nd[i]] & fDry) ndry++;
    if (ndry == ncn) {
      m->eflag[e] |= elDry;
    } else if (ndry) {
      m->eflag[e] |= elMarsh;
      for (int i = 0; i < ncn; i++) if (m->flag[nd[i]] & fDry) m->flag[nd[i]] |= fMarsh;
    }
  }
  for (int n = 0; n < m->np; n++) m->flag[n] |= fDry;
  for (int e = 0; e < m->ne; e++) {
    if (m->eflag[e] & elDry) continue;
    const int* nd = m->elem + 8 * e;
    int nnd = m->nnd[e], ncn = nnd == 6 ? 3 : 4;
    for (int i = 0; i < nnd; i++) m->flag[nd[i]] &= ~(unsigned)fDry;
    for (int i = 0; i < ncn; i++) {
      int n = nd[i];
      if (!(m->flag[n] & fMarsh)) continue;
      if (m->nmark[n] || m->firstDryRew) {
        int j = (i + 1) % ncn, k = (i + ncn - 1) % ncn;
        int mj = (m->flag[nd[j]] & fMarsh) != 0, mk = (m->flag[nd[k]] & fMarsh) != 0;
        double Si;
        if (!mj && !mk) Si = (m->S[nd[j]] + m->S[nd[k]]) / 2.0;
        else if (!mj) Si = m->S[nd[j]];
        else if (!mk) Si = m->S[nd[k]];
        else Si = m->S[nd[(i + 2) % ncn]];     /* the reference reports an error if marsh */
        m->S[n] = Si;
        if (m->S[n] > m->zor[n] + rewetLimit) m->S[n] = m->zor[n] + rewetLimit;
        if (m->S[n] < m->zor[n] + dryLimit) m->z[n] = m->S[n] - dryLimit;
      } else if (m->S[n] < m->zor[n] + dryLimit) {
        m->z[n] = m->S[n] - dryLimit;
      }
    }
  }
  for (int e = 0; e < m->ne; e++) {
    const int* nd = m->elem + 8 * e;
    int nnd = m->nnd[e], ncn = nnd == 6 ? 3 : 4;
    for (int i = ncn; i < nnd; i++) {
      int l = nd[i - ncn], r = nd[(i - ncn + 1) % ncn], n = nd[i];
      m->S[n] = 0.5 * (m->S[l] + m->S[r]);
      m->z[n] = 0.5 * (m->z[l] + m->z[r]);
      if ((m->eflag[e] & elMarsh) && m->z[n] + dryLimit / 10.0 < m->zor[n]) m->flag[n] |= fMarsh;
    }
  }
  for (int n = 0; n < m->np; n++) {
    if (!(m->flag[n] & fDry)) continue;
    m->U[n] = m->V[n] = 0.0;
    m->S[n] = m->z[n] = m->zor[n];
    m->dUdt[n] = m->dVdt[n] = m->dSdt[n] = 0.0;
  }
  int del = 0, wel = 0;
  for (int e = 0; e < m->ne; e++) {
    if ((m->eflag[e] & elDry) && !m->emark[e]) del++;
    if (!(m->eflag[e] & elDry) && m->emark[e]) wel++;
  }
  *dried = del; *wetted = wel;
}

/* GRID::SetSlipFlow (SlipFlow.cpp:43-78): at rotated nodes the velocity and its time
 * derivative keep only their component along the first column of the rotation; applied
 * after a new set-up of the wet model and at the end of the flow cycle
 * (EqsUVS2D.cpp:1078) */
void orc_slip_flow(orc_model* m)
{
  for (int n = 0; n < m->np; n++) {
    if (!(m->flag[n] & fRotat)) continue;
    double r00 = getrot(m, n, 0, 0), r10 = getrot(m, n, 1, 0);
    double Ut = r00 * m->U[n] + r10 * m->V[n];
    m->U[n] = r00 * Ut; m->V[n] = r10 * Ut;
    Ut = r00 * m->dUdt[n] + r10 * m->dVdt[n];
    m->dUdt[n] = r00 * Ut; m->dVdt[n] = r10 * Ut;
  }
}

/* MODEL::DoDryRewet (DoDryRew.cpp:50-238), serial build, method 2.  When elements got
 * dry or wet the model is set up again on the wet elements (MODEL::Initialize, SetNormal,
 * SetRotation: Model.cpp:121-215, Transform.cpp) with the slip flow condition applied to
 * the rotated nodes (GRID::SetSlipFlow, SlipFlow.cpp:43-78), and the equations are
 * numbered anew at the next Newton step.  Returns the number of elements that got wet
 * (-1 for a method other than 2); *dried the number that got dry. */
int orc_dry_rewet(orc_model* m, int* dried)
{
  int del = 0, wel = 0;
  if (m->dryMethod != 2) return -1;        /* methods 1 and 3 are not restated */
  /* the time step's boundary conditions (BCONSET::GenBcon at the start of the time step,
   * Compute.cpp) precede the first call */
  if (!m->bcon_init) init_bcon(m);
  dry_rewet_method2(m, &del, &wel);
  if (del + wel) {
    topology(m);
    set_location(m);
    set_normal(m);
    set_rotation(m);
    orc_slip_flow(m);
    m->init_done = 0;
  } else {
    /* GRID::Connection(ELEM::kDry) only */
    for (int n = 0; n < m->np; n++) m->noel[n] = 0;
    for (int e = 0; e < m->ne; e++)
      if (!(m->eflag[e] & elDry))
        for (int i = 0; i < m->nnd[e]; i++) m->noel[m->elem[8 * e + i]]++;
  }
  m->firstDryRew = 0;
  m->turb_stale = 1;                       /* eddy = true (EqsUVS2D.cpp:216) */
  if (dried) *dried = del;
  return wel;
}

void orc_set_dryrew(orc_model* m, double rewetLimit, int countDown)
{
  m->rewetLimit = rewetLimit; m->countDown = countDown;
}

/* per node the reference's dry and marsh bits and those of the previous time step
 * (Node.h:100-103), per region element 1 dry, 2 marsh */
void orc_get_wetdry(const orc_model* m, int* nflag, int* eflag)
{
  for (int n = 0; n < m->np; n++)
    nflag[n] = ((m->flag[n] & fDry) ? refNodeDry : 0) | ((m->flag[n] & fMarsh) ? refNodeMarsh : 0) |
               ((m->flag[n] & fDryPrev) ? refNodeDryPrev : 0) |
               ((m->flag[n] & fMarshPrev) ? refNodeMarshPrev : 0);
  for (int e = 0; e < m->ne; e++) eflag[e] = (int)m->eflag[e];
}

int orc_prepare(orc_model* m)
{
  /* EqsUVS2D.cpp:300-360 (DoFriction; Turbulence at the first step after a state change
   * or every step with kVtIterat, the constant model needs no nodal data; InitBcon,
   * SetLocation, SetNormal, SetRotation, SetEqno on first use) */
  do_friction(m);
  const int aniso = (m->turb / 10) % 10 == 1;
  if ((m->turb % 10 == 2 || aniso) && (m->turb_stale || ((m->turb / 100) % 10 & 2))) {
    if (m->turb % 10 == 2) turbulence_algebraic(m);
    else turbulence_constant(m);
    if (aniso) turbulence_elder(m);
    m->turb_stale = 0;
  }
  init_bcon(m);
  set_location(m);
  set_normal(m);
  set_rotation(m);
  if (!m->init_done) {
    set_eqno(m);
    set_index_mat(m);
    m->init_done = 1;
  }
  return m->neq;
}

int orc_neq(const orc_model* m) { return m->neq; }
long orc_nnz(const orc_model* m) { return m->nnz; }
void orc_get_numbering(const orc_model* m, int* eqno) { memcpy(eqno, m->eqno, sizeof(int) * 3 * m->np); }
void orc_get_pattern(const orc_model* m, int* rowptr, int* cols)
{
  memcpy(rowptr, m->rowptr, sizeof(int) * (m->neq + 1));
  memcpy(cols, m->cols, sizeof(int) * m->nnz);
}
int orc_n_model_elems(const orc_model* m) { return m->nme; }
int orc_model_elem(const orc_model* m, int k, int* nodes, int* is_bound)
{
  const int* nd; int nnd = me_nodes(m, k, &nd, is_bound);
  memcpy(nodes, nd, sizeof(int) * nnd);
  return nnd;
}

/* ------------------------------------------------------------------------------------- */
/* element coefficients                                                                   */

/* EQS::Rotate2D (Rotate.cpp:42-107) on a MAXEQ x MAXEQ element matrix */
static void rotate2d(const orc_model* m, int nnd, int eqid, const int* nd, double* K, double* F)
{
  for (int i = 0; i < nnd; i++) {
    int n = nd[i];
    if (!(m->flag[n] & fRotat)) continue;
    double r00 = getrot(m, n, 0, 0), r01 = getrot(m, n, 0, 1);
    double r10 = getrot(m, n, 1, 0), r11 = getrot(m, n, 1, 1);
    double det = r00 * r11 - r01 * r10;
    if (fabs(det) < 1.0e-12) continue;
    int j = eqid + i;
    for (int k = 0; k < MAXEQ; k++) {
      double* row = K + k * MAXEQ;
      double tmp = row[i] * r00 + row[j] * r10;
      row[j] = row[i] * r01 + row[j] * r11;
      row[i] = tmp;
    }
    for (int k = 0; k < MAXEQ; k++) {
      double tmp = r00 * K[i * MAXEQ + k] + r10 * K[j * MAXEQ + k];
      K[j * MAXEQ + k] = r01 * K[i * MAXEQ + k] + r11 * K[j * MAXEQ + k];
      K[i * MAXEQ + k] = tmp;
    }
    double tmp = r00 * F[i] + r10 * F[j];
    F[j] = r01 * F[i] + r11 * F[j];
    F[i] = tmp;
  }
}

/* EQS_UVS2D::Bound (CoefsUVS2D.cpp:71-360) for walls and open boundaries without log-law
 * wall friction (solidFlag false). */
static void coefs_bound(orc_model* m, const int* nd, unsigned bflag, double* K, double* F)
{
  (void)bflag;
  const int eqidV = 3, eqidS = 6;
  const double gravity = m->g;
  memset(K, 0, sizeof(double) * MAXEQ * MAXEQ);
  memset(F, 0, sizeof(double) * MAXEQ);
  int outletMid = (m->kind[nd[2]] & kOutlet) != 0;
  for (int g = 0; g < g_line.ngp; g++) {
    const double* mm = g_line.lf[g];
    const double* n = g_line.f[g];
    const double* dn = g_line.dfdx[g];
    double H;
    if (outletMid) {
      double H0 = m->valS[nd[0]] - m->z[nd[0]];
      double H1 = m->valS[nd[1]] - m->z[nd[1]];
      if (H0 <= 0.0) H0 = m->hmin;
      if (H1 <= 0.0) H1 = m->hmin;
      H = mm[0] * H0 + mm[1] * H1;
    } else {
      double H0 = m->S[nd[0]] - m->z[nd[0]];
      double H1 = m->S[nd[1]] - m->z[nd[1]];
      if (H0 <= 0.0) H0 = m->hmin;
      if (H1 <= 0.0) H1 = m->hmin;
      H = mm[0] * H0 + mm[1] * H1;
    }
    double nx = dn[0] * m->y[nd[0]] + dn[1] * m->y[nd[1]] + dn[2] * m->y[nd[2]];
    double ny = -dn[0] * m->x[nd[0]] - dn[1] * m->x[nd[1]] - dn[2] * m->x[nd[2]];
    double weight = g_line.w[g];
    double fU = weight * H * (gravity * H / 2.0 * nx);
    double fV = weight * H * (gravity * H / 2.0 * ny);
    for (int i = 0; i < 3; i++) { F[i] -= n[i] * fU; F[eqidV + i] -= n[i] * fV; }
    double dfUdH = 0.0, dfVdH = 0.0;
    if (!outletMid) {
      dfUdH += weight * (gravity * H * nx);
      dfVdH += weight * (gravity * H * ny);
    }
    for (int i = 0; i < 3; i++) {
      K[i * MAXEQ + eqidS + 0] += n[i] * dfUdH * mm[0];
      K[(eqidV + i) * MAXEQ + eqidS + 0] += n[i] * dfVdH * mm[0];
    }
    for (int i = 0; i < 3; i++) {
      K[i * MAXEQ + eqidS + 1] += n[i] * dfUdH * mm[1];
      K[(eqidV + i) * MAXEQ + eqidS + 1] += n[i] * dfVdH * mm[1];
    }
  }
  rotate2d(m, 3, eqidV, nd, K, F);
  for (int i = 0; i < 3; i++) {
    if (m->kind[nd[i]] & kInlet) {
      for (int j = 0; j < MAXEQ; j++) K[i * MAXEQ + j] = 0.0;
      F[i] = 0.0;
    }
  }
}

/* SHAPE::jacobi2D (Shape.cpp:395-440) */
static double jacobi2d(int nnd, const double* dfdx, const double* dfdy, const double* x,
                       const double* y, double t[2][2])
{
  double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
  for (int n = 0; n < nnd; n++) {
    j00 += dfdx[n] * x[n]; j01 += dfdx[n] * y[n];
    j10 += dfdy[n] * x[n]; j11 += dfdy[n] * y[n];
  }
  double det = j00 * j11 - j01 * j10;
  t[0][0] = j11 / det; t[0][1] = -j01 / det;
  t[1][0] = -j10 / det; t[1][1] = j00 / det;
  return det;
}

/* EQS_UVS2D::Region (CoefsUVS2D.cpp:365-1181) for one triangle or quad */
static void coefs_region(orc_model* m, int e, double* K, double* F)
{
  const int* nd = m->elem + 8 * e;
  const shape_t* s = m->nnd[e] == 6 ? &g_tri : &g_quad;
  const int ngp = s->ngp, nnd = s->nnd, ncn = s->ncn;
  const int eqidV = nnd, eqidS = 2 * nnd;
  const double gravity = m->g;
  const double vtype = m->vt[e];
  const int alg = m->turb % 10 == 2, vtmin = (m->turb / 100) % 10 & 1;
  /* tens digit 1: EQS_UVS2D_AI::Region (CoefsUVS2D_AI.cpp:49-620), the same terms with the
   * eddy viscosity tensor vt (exx, exy, eyy) interpolated from the nodes */
  const int aniso = (m->turb / 10) % 10 == 1;
  memset(K, 0, sizeof(double) * MAXEQ * MAXEQ);
  memset(F, 0, sizeof(double) * MAXEQ);
  double area = 0.0;
  double x[8], y[8];
  x[0] = m->x[nd[0]]; y[0] = m->y[nd[0]];
  for (int i = 1; i < nnd; i++) { x[i] = m->x[nd[i]] - x[0]; y[i] = m->y[nd[i]] - y[0]; }
  x[0] = y[0] = 0.0;

  for (int g = 0; g < ngp; g++) {
    double trafo[2][2];
    double detj = jacobi2d(nnd, s->dfdx[g], s->dfdy[g], x, y, trafo);
    double weight = detj * s->w[g];
    area += weight;
    const double* n = s->f[g];
    double dndx[8], dndy[8], dmdx[4], dmdy[4];
    for (int j = 0; j < nnd; j++) {
      dndx[j] = trafo[0][0] * s->dfdx[g][j] + trafo[0][1] * s->dfdy[g][j];
      dndy[j] = trafo[1][0] * s->dfdx[g][j] + trafo[1][1] * s->dfdy[g][j];
    }
    const double* mm = s->lf[g];
    for (int j = 0; j < ncn; j++) {
      dmdx[j] = trafo[0][0] * s->ldx[g][j] + trafo[0][1] * s->ldy[g][j];
      dmdy[j] = trafo[1][0] * s->ldx[g][j] + trafo[1][1] * s->ldy[g][j];
    }
    double H = 0.0, dHdt = 0.0, dHdx = 0.0, dHdy = 0.0, dadx = 0.0, dady = 0.0, cf = 0.0;
    for (int j = 0; j < ncn; j++) {
      int k = nd[j];
      double ndZ = m->z[k];
      double ndH = m->S[k] - ndZ;
      if (ndH <= 0.0) ndH = m->hmin;
      H += mm[j] * ndH;
      dHdx += dmdx[j] * ndH;
      dHdy += dmdy[j] * ndH;
      dadx += dmdx[j] * ndZ;
      dady += dmdy[j] * ndZ;
      dHdt += mm[j] * m->dSdt[k];
      cf += mm[j] * m->cf[k];
    }
    double U = 0.0, dUdt = 0.0, dUdx = 0.0, dUdy = 0.0;
    double V = 0.0, dVdt = 0.0, dVdx = 0.0, dVdy = 0.0;
    double uu = 0.0;              /* Reynolds stresses: vv = uu, uv = 0 */
    for (int j = 0; j < nnd; j++) {
      int k = nd[j];
      double ndU = m->U[k], ndV = m->V[k];
      U += n[j] * ndU; dUdt += n[j] * m->dUdt[k];
      dUdx += dndx[j] * ndU; dUdy += dndy[j] * ndU;
      V += n[j] * ndV; dVdt += n[j] * m->dVdt[k];
      dVdx += dndx[j] * ndV; dVdy += dndy[j] * ndV;
      if (alg) uu += n[j] * m->uun[k];
    }
    const double vv = uu;
    double vt = 0.0, vtxx = 0.0, vtxy = 0.0, vtyy = 0.0;
    if (aniso) {                  /* CoefsUVS2D_AI.cpp:288-310 */
      for (int j = 0; j < nnd; j++) {
        const int k = nd[j];
        vtxx += n[j] * m->exx[k] * m->vtn[k];
        vtxy += n[j] * m->exy[k] * m->vtn[k];
        vtyy += n[j] * m->eyy[k] * m->vtn[k];
      }
      if (vtmin) {
        if (vtxx < vtype) vtxx = vtype;
        if (vtyy < vtype) vtyy = vtype;
      }
      vtxx += m->vk; vtxy += m->vk; vtyy += m->vk;
    } else {
      if (!alg) vt = vtype;       /* constant eddy viscosity of the element type */
      else
        for (int j = 0; j < nnd; j++) vt += n[j] * m->vtn[nd[j]];
      if (vtmin && vt < vtype) vt = vtype;
      vt += m->vk;
    }
    double Ures = sqrt(U * U + V * V);

    /* residual ------------------------------------------------------------------------ */
    {
      double f, fx, fy;
      f = H * dUdt;
      f += H * (U * dUdx + V * dUdy);
      if (aniso) {
        fx = H * (vtxx * dUdx + vtxy * dUdy);
        fy = H * (vtxy * dUdx + vtyy * dUdy);
      } else {
        fx = H * vt * (dUdx + dUdx);
        fy = H * vt * (dUdy + dVdx);
      }
      fx -= H * uu;
      f += H * gravity * dadx;
      fx -= H * H * gravity / 2.0;
      f += cf * Ures * U;
      f *= weight; fx *= weight; fy *= weight;
      for (int j = 0; j < nnd; j++) F[j] -= n[j] * f + dndx[j] * fx + dndy[j] * fy;

      f = H * dVdt;
      f += H * (U * dVdx + V * dVdy);
      if (aniso) {
        fx = H * (vtxx * dVdx + vtxy * dVdy);
        fy = H * (vtxy * dVdx + vtyy * dVdy);
      } else {
        fx = H * vt * (dVdx + dUdy);
        fy = H * vt * (dVdy + dVdy);
      }
      fy -= H * vv;
      f += H * gravity * dady;
      fy -= H * H * gravity / 2.0;
      f += cf * Ures * V;
      f *= weight; fx *= weight; fy *= weight;
      for (int j = 0; j < nnd; j++) F[eqidV + j] -= n[j] * f + dndx[j] * fx + dndy[j] * fy;

      f = dHdt + H * (dUdx + dVdy) + U * dHdx + V * dHdy;
      f *= weight;
      for (int j = 0; j < ncn; j++) F[eqidS + j] -= mm[j] * f;
    }

    /* Jacobian ------------------------------------------------------------------------ */
    {
      double t[8], tx[8], ty[8];
      double df__, df_x, df_y, dfx_, dfy_, dfxx, dfxy, dfyx, dfyy;
      /* the anisotropic routine guards 1/|U| at 1e-8 (CoefsUVS2D_AI.cpp:379) */
      double iUres = (Ures > (aniso ? 1.0e-8 : 1.0e-9)) ? 1.0 / Ures : 0.0;

      /* U-derivative of x-momentum */
      df__ = weight * H * m->relaxThdt_UV;
      df__ += weight * H * dUdx;
      df_x = weight * H * U;
      df_y = weight * H * V;
      if (aniso) {
        dfxx = weight * H * vtxx; dfxy = weight * H * vtxy;
        dfyx = weight * H * vtxy; dfyy = weight * H * vtyy;
      } else {
        dfxx = weight * H * vt * 2.0;
        dfyy = weight * H * vt;
        dfxy = dfyx = 0.0;
      }
      df__ += weight * cf * (iUres * U * U + Ures);
      for (int j = 0; j < nnd; j++) {
        t[j] = df__ * n[j] + df_x * dndx[j] + df_y * dndy[j];
        tx[j] = aniso ? dfxx * dndx[j] + dfxy * dndy[j] : dfxx * dndx[j];
        ty[j] = aniso ? dfyx * dndx[j] + dfyy * dndy[j] : dfyy * dndy[j];
      }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < nnd; k++)
          K[j * MAXEQ + k] += n[j] * t[k] + dndx[j] * tx[k] + dndy[j] * ty[k];

      /* V-derivative of x-momentum (no viscous coupling in the anisotropic routine) */
      df__ = weight * H * dUdy;
      dfyx = aniso ? 0.0 : weight * H * vt;
      df__ += weight * cf * iUres * U * V;
      for (int j = 0; j < nnd; j++) { t[j] = df__ * n[j]; ty[j] = dfyx * dndx[j]; }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < nnd; k++)
          K[j * MAXEQ + eqidV + k] += n[j] * t[k] + dndy[j] * ty[k];

      /* H-derivative of x-momentum */
      df__ = weight * dUdt;
      df__ += weight * (U * dUdx + V * dUdy);
      if (aniso) {
        dfx_ = weight * (vtxx * dUdx + vtxy * dUdy);
        dfy_ = weight * (vtxy * dUdx + vtyy * dUdy);
      } else {
        dfx_ = weight * vt * (dUdx + dUdx);
        dfy_ = weight * vt * (dUdy + dVdx);
      }
      dfx_ -= weight * uu;
      df__ += weight * gravity * dadx;
      dfx_ -= weight * gravity * H;
      for (int j = 0; j < ncn; j++) { t[j] = df__ * mm[j]; tx[j] = dfx_ * mm[j]; ty[j] = dfy_ * mm[j]; }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < ncn; k++)
          K[j * MAXEQ + eqidS + k] += n[j] * t[k] + dndx[j] * tx[k] + dndy[j] * ty[k];

      /* U-derivative of y-momentum (no viscous coupling in the anisotropic routine) */
      df__ = weight * H * dVdx;
      dfxy = aniso ? 0.0 : weight * H * vt;
      df__ += weight * cf * iUres * U * V;
      for (int j = 0; j < nnd; j++) { t[j] = df__ * n[j]; tx[j] = dfxy * dndy[j]; }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < nnd; k++)
          K[(j + eqidV) * MAXEQ + k] += n[j] * t[k] + dndx[j] * tx[k];

      /* V-derivative of y-momentum */
      df__ = weight * H * m->relaxThdt_UV;
      df__ += weight * H * dVdy;
      df_x = weight * H * U;
      df_y = weight * H * V;
      if (aniso) {
        dfxx = weight * H * vtxx; dfxy = weight * H * vtxy;
        dfyx = weight * H * vtxy; dfyy = weight * H * vtyy;
      } else {
        dfxx = weight * H * vt;
        dfyy = weight * H * vt * 2.0;
        dfxy = dfyx = 0.0;
      }
      df__ += weight * cf * (iUres * V * V + Ures);
      for (int j = 0; j < nnd; j++) {
        t[j] = df__ * n[j] + df_x * dndx[j] + df_y * dndy[j];
        tx[j] = aniso ? dfxx * dndx[j] + dfxy * dndy[j] : dfxx * dndx[j];
        ty[j] = aniso ? dfyx * dndx[j] + dfyy * dndy[j] : dfyy * dndy[j];
      }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < nnd; k++)
          K[(j + eqidV) * MAXEQ + eqidV + k] += n[j] * t[k] + dndx[j] * tx[k] + dndy[j] * ty[k];

      /* H-derivative of y-momentum */
      df__ = weight * dVdt;
      df__ += weight * (U * dVdx + V * dVdy);
      if (aniso) {
        dfx_ = weight * (vtxx * dVdx + vtxy * dVdy);
        dfy_ = weight * (vtxy * dVdx + vtyy * dVdy);
      } else {
        dfx_ = weight * vt * (dVdx + dUdy);
        dfy_ = weight * vt * (dVdy + dVdy);
      }
      dfy_ -= weight * vv;
      df__ += weight * gravity * dady;
      dfy_ -= weight * gravity * H;
      for (int j = 0; j < ncn; j++) { t[j] = df__ * mm[j]; tx[j] = dfx_ * mm[j]; ty[j] = dfy_ * mm[j]; }
      for (int j = 0; j < nnd; j++)
        for (int k = 0; k < ncn; k++)
          K[(j + eqidV) * MAXEQ + eqidS + k] += n[j] * t[k] + dndx[j] * tx[k] + dndy[j] * ty[k];

      /* U-derivative of continuity */
      df__ = weight * dHdx;
      df_x = weight * H;
      for (int j = 0; j < nnd; j++) t[j] = df__ * n[j] + df_x * dndx[j];
      for (int j = 0; j < ncn; j++)
        for (int k = 0; k < nnd; k++) K[(j + eqidS) * MAXEQ + k] += mm[j] * t[k];

      /* V-derivative of continuity */
      df__ = weight * dHdy;
      df_y = weight * H;
      for (int j = 0; j < nnd; j++) t[j] = df__ * n[j] + df_y * dndy[j];
      for (int j = 0; j < ncn; j++)
        for (int k = 0; k < nnd; k++) K[(j + eqidS) * MAXEQ + eqidV + k] += mm[j] * t[k];

      /* H-derivative of continuity */
      df__ = weight * m->relaxThdt_H;
      df__ += weight * (dUdx + dVdy);
      df_x = weight * U;
      df_y = weight * V;
      for (int j = 0; j < ncn; j++) t[j] = df__ * mm[j] + df_x * dmdx[j] + df_y * dmdy[j];
      for (int j = 0; j < ncn; j++)
        for (int k = 0; k < ncn; k++) K[(j + eqidS) * MAXEQ + eqidS + k] += mm[j] * t[k];
    }
  }

  rotate2d(m, nnd, eqidV, nd, K, F);

  /* inlet rows: algebraic discharge equation (CoefsUVS2D.cpp:1071-1180) */
  for (int i = 0; i < ncn; i++) {
    int k = nd[i];
    if (!(m->kind[k] & kInlet)) continue;
    for (int j = 0; j < MAXEQ; j++) K[i * MAXEQ + j] = 0.0;
    double Un = m->U[k] * m->niox[k] + m->V[k] * m->nioy[k];
    double H = m->S[k] - m->z[k];
    double specQ = m->valU[k];
    if (H <= 0.0) { H = m->hmin; specQ = 0.0; }
    F[i] = area * (specQ - Un * H);
    K[i * MAXEQ + i] = area * H;
    K[i * MAXEQ + i + eqidS] = area * Un;
  }
  for (int i = ncn; i < nnd; i++) {
    int k = nd[i];
    if (!(m->kind[k] & kInlet)) continue;
    int l = i - ncn, r = (l + 1) % ncn;
    int kl = nd[l], kr = nd[r];
    for (int j = 0; j < MAXEQ; j++) K[i * MAXEQ + j] = 0.0;
    double Un = m->U[k] * m->niox[k] + m->V[k] * m->nioy[k];
    double Hl = m->S[kl] - m->z[kl];
    double Hr = m->S[kr] - m->z[kr];
    double H = (Hl + Hr) / 2.0;
    double specQ = m->valU[k];
    if (H <= 0.0) { H = m->hmin; specQ = 0.0; }
    F[i] = area * (specQ - Un * H);
    K[i * MAXEQ + i] = area * H;
    K[i * MAXEQ + eqidS + l] = area * Un / 2.0;
    K[i * MAXEQ + eqidS + r] = area * Un / 2.0;
  }
}

int orc_coefs(orc_model* m, int k, double* K, double* F)
{
  int v = m->me[k];
  if (v >= 0) coefs_region(m, v, K, F);
  else coefs_bound(m, m->bnd + 3 * (-v - 1), m->bflag[-v - 1], K, F);
  return 1;
}

/* ------------------------------------------------------------------------------------- */
/* REALPR-dependent routines: compiled twice                                               */

#define REALPR double
#define SFX(name) name##_f64
#include "realpr.inc"
#undef REALPR
#undef SFX
#define REALPR float
#define SFX(name) name##_f32
#include "realpr.inc"
#undef REALPR
#undef SFX

/* ------------------------------------------------------------------------------------- */
/* Newton step (EqsUVS2D.cpp:189-760)                                                     */

int orc_newton_step(orc_model* m, int prec, int solver, int maxIter, double eps, int mkyrl,
                    orc_step_stats* st, double* X_out)
{
  orc_prepare(m);
  int n = m->neq;
  long nnz = m->nnz;
  double* B = xcalloc((size_t)n, sizeof(double));
  double* X = xcalloc((size_t)n, sizeof(double));
  int iters = 0, conv = 0;
  double acc = 1.0, scale;
  if (prec == 32) {
    float* A = xcalloc((size_t)nnz, sizeof(float));
    float* L = xcalloc((size_t)nnz, sizeof(float));
    orc_assemble_f32(m, A, B);
    scale = orc_scale_f32(n, m->rowptr, A, B);
    orc_ilu0_factor_f32(n, m->rowptr, m->cols, A, L);
    if (solver == 2) conv = orc_fgmres_f32(n, m->rowptr, m->cols, A, L, B, X, mkyrl, maxIter, eps, &iters, &acc);
    else conv = orc_bicgstab_f32(n, m->rowptr, m->cols, A, L, B, X, maxIter, eps, &iters, &acc);
    free(A); free(L);
  } else {
    double* A = xcalloc((size_t)nnz, sizeof(double));
    double* L = xcalloc((size_t)nnz, sizeof(double));
    orc_assemble_f64(m, A, B);
    scale = orc_scale_f64(n, m->rowptr, A, B);
    orc_ilu0_factor_f64(n, m->rowptr, m->cols, A, L);
    if (solver == 2) conv = orc_fgmres_f64(n, m->rowptr, m->cols, A, L, B, X, mkyrl, maxIter, eps, &iters, &acc);
    else conv = orc_bicgstab_f64(n, m->rowptr, m->cols, A, L, B, X, maxIter, eps, &iters, &acc);
    free(A); free(L);
  }
  if (X_out) memcpy(X_out, X, sizeof(double) * n);

  /* EQS::Update statistics (Update.cpp:42-187) and the convergence test */
  orc_step_stats s;
  memset(&s, 0, sizeof(s));
  s.iters = iters; s.accuracy = acc; s.converged = conv; s.scale = scale;
  s.nrconv = 1;
  for (int i = 0; i < 3; i++) {
    double mx = 0.0, av = 0.0; int cnt = 0;
    for (int nd = 0; nd < m->np; nd++) {
      int e = m->eqno[3 * nd + i];
      if (e < 0) continue;
      double ch = fabs(X[e]);
      av += ch; cnt++;
      if (ch > mx) mx = ch;
    }
    s.maxAbs[i] = mx; s.avAbs[i] = cnt ? av / cnt : 0.0;
    if (fabs(mx) > (i < 2 ? m->convUV : m->convS)) s.nrconv = 0;
  }
  /* relaxation (EqsUVS2D.cpp:455-627) */
  double relaxUs = 1.0, relaxS = 1.0;
  if (m->relaxMethod >= 3) {
    /* the relaxed time steps: no convergence before they reach the time step; each grows
     * by the ratio of the maximum change to the largest correction, within [relaxTimeFlow,
     * dt], the two kept equal (:540-627) */
    const double dt = m->unsteady ? m->dt : m->interval;
    if (m->dt_H < dt || m->dt_UV < dt) s.nrconv = 0;
    double maxUs = 0.0, maxS = 0.0;
    for (int nd = 0; nd < m->np; nd++) {
      double dU = 0.0, dV = 0.0, dS = 0.0;
      int eU = m->eqno[3 * nd], eV = m->eqno[3 * nd + 1], eS = m->eqno[3 * nd + 2];
      if (eU >= 0) dU = fabs(X[eU]);
      if (eV >= 0) dV = fabs(X[eV]);
      if (eS >= 0) dS = fabs(X[eS]);
      double dUs = sqrt(dU * dU + dV * dV);
      if (dUs > maxUs) maxUs = dUs;
      if (dS > maxS) maxS = dS;
    }
    relaxUs = m->maxDeltaUV / maxUs;
    relaxS = m->maxDeltaS / maxS;
    const double rdt = m->relaxDt > 0.0 && m->relaxDt <= dt ? m->relaxDt : dt;
    m->last_dt = m->dt_H;
    m->dt_UV *= relaxUs;
    m->dt_H *= relaxS;
    if (m->dt_UV > dt) m->dt_UV = dt;
    if (m->dt_UV < rdt) m->dt_UV = rdt;
    if (m->dt_H > dt) m->dt_H = dt;
    if (m->dt_H < rdt) m->dt_H = rdt;
    if (m->dt_H < m->dt_UV) m->dt_UV = m->dt_H; else m->dt_H = m->dt_UV;
    const double th = m->unsteady ? m->theta : 1.0;
    m->relaxThdt_UV = 1.0 / m->dt_UV / th;
    m->relaxThdt_H = 1.0 / m->dt_H / th;
    if (relaxUs > 1.0) relaxUs = 1.0;
    if (relaxS > 1.0) relaxS = 1.0;
  } else if (m->relaxMethod == 2) {
    double maxUs = 0.0, maxS = 0.0;
    for (int nd = 0; nd < m->np; nd++) {
      double dU = 0.0, dV = 0.0, dS = 0.0;
      int eU = m->eqno[3 * nd], eV = m->eqno[3 * nd + 1], eS = m->eqno[3 * nd + 2];
      if (eU >= 0) dU = fabs(X[eU]);
      if (eV >= 0) dV = fabs(X[eV]);
      if (eS >= 0) dS = fabs(X[eS]);
      double dUs = sqrt(dU * dU + dV * dV);
      if (dUs > maxUs) maxUs = dUs;
      if (dS > maxS) maxS = dS;
    }
    relaxUs = m->maxDeltaUV / maxUs;
    relaxS = m->maxDeltaS / maxS;
    if (relaxUs > 1.0) relaxUs = 1.0;
    if (relaxS > 1.0) relaxS = 1.0;
  }
  s.relaxUs = relaxUs; s.relaxS = relaxS;
  /* apply the correction (EqsUVS2D.cpp:636-718) */
  for (int nd = 0; nd < m->np; nd++) {
    double dU = 0.0, dV = 0.0, dS = 0.0;
    int eU = m->eqno[3 * nd], eV = m->eqno[3 * nd + 1], eS = m->eqno[3 * nd + 2];
    if (eU >= 0) dU = X[eU] * relaxUs;
    if (eV >= 0) dV = X[eV] * relaxUs;
    if (eS >= 0) dS = X[eS] * relaxS;
    if (m->flag[nd] & fRotat) {
      m->U[nd] += getrot(m, nd, 0, 0) * dU;
      m->V[nd] += getrot(m, nd, 1, 0) * dU;
      m->U[nd] += getrot(m, nd, 0, 1) * dV;
      m->V[nd] += getrot(m, nd, 1, 1) * dV;
    } else {
      m->U[nd] += dU;
      m->V[nd] += dV;
    }
    double U = m->U[nd], V = m->V[nd];
    double Us = sqrt(U * U + V * V);
    if (Us > m->maxUs) { m->U[nd] = U * m->maxUs / Us; m->V[nd] = V * m->maxUs / Us; }
    if (eS >= 0) {
      m->S[nd] += dS;
      if (m->dryMethod == 0 || m->dryMethod == 2 || m->dryMethod == 3) {
        double H = m->S[nd] - m->zor[nd];
        if (H < m->dryLimit) m->z[nd] = m->S[nd] - m->dryLimit;
        else m->z[nd] = m->zor[nd];
      } else {
        double H = m->S[nd] - m->z[nd];
        if (H <= 0.0) m->S[nd] = m->z[nd] + m->hmin;
      }
    }
  }
  /* water level and bed at midside nodes by linear interpolation (EqsUVS2D.cpp:733-757) */
  for (int e = 0; e < m->ne; e++) {
    int nnd = m->nnd[e], ncn = nnd == 6 ? 3 : 4;
    const int* nd = m->elem + 8 * e;
    for (int i = ncn; i < nnd; i++) {
      int l = i - ncn, r = (l + 1) % ncn;
      m->S[nd[i]] = 0.5 * (m->S[nd[l]] + m->S[nd[r]]);
      m->z[nd[i]] = 0.5 * (m->z[nd[l]] + m->z[nd[r]]);
    }
  }
  /* time derivatives of the corrected state (EqsUVS2D.cpp:811) */
  if (m->unsteady) timegrad(m);
  if (st) *st = s;
  free(B); free(X);
  return n;
}
