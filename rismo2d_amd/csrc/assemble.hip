// Element Jacobian / residual of the Taylor-Hood UVS2D system and its assembly into the
// CSR matrix on gfx950.
//
// One thread owns one (element, node i) pair and computes the element-matrix rows of the
// node's dofs: U_i, V_i and, at corners, H_i (the continuity row), i.e. 2-3 rows of 15
// (triangle), 20 (quad) or 8 (boundary line) columns held in registers.  Every thread of
// an element re-evaluates the Gauss-point interpolation (cheap next to the row work) so
// no inter-thread exchange is needed; the rotation of rotated nodes mixes only the U_i /
// V_i rows of one node, which the same thread owns.  Gauss tables live in constant
// memory.  Elements are coloured so that the elements of one launch share no node, so
// the scatter into A and B needs no atomics; each thread adds its rows through a
// precomputed element-to-CSR slot map.
//
// Terms and their order follow EQS_UVS2D::Region / Bound (CoefsUVS2D.cpp:71-1181),
// EQS::Rotate2D (Rotate.cpp:42) and the inlet discharge rows (CoefsUVS2D.cpp:1071-1180).
#include <cmath>
#include <cstring>

#include "rmx_internal.h"

namespace rmx {

struct ShapeTab {
  double f[9][8], dfdx[9][8], dfdy[9][8];
  double lf[9][4], ldx[9][4], ldy[9][4];
  double w[9];
};

__constant__ ShapeTab c_shape[3];   // ET_TRI, ET_QUAD, ET_LINE

// ---- host: Gauss rules and shape functions at GPdeg 5 (Shape.cpp, Triangle.cpp,
// Square.cpp, Line.cpp): 7-point Cowper triangle, 3x3 Gauss quad, 3-point line ----
static void make_shapes(ShapeTab t[3])
{
  std::memset(t, 0, sizeof(ShapeTab) * 3);
  auto L = [](int k, double x, double y) { return k == 0 ? 1.0 - x - y : (k == 1 ? x : y); };
  const double Ldx[3] = { -1.0, 1.0, 0.0 }, Ldy[3] = { -1.0, 0.0, 1.0 };
  static const double tloc[7][2] = {
    { 1.0 / 3.0, 1.0 / 3.0 }, { 0.0597158717897698, 0.470142064105115 },
    { 0.470142064105115, 0.0597158717897698 }, { 0.470142064105115, 0.470142064105115 },
    { 0.797426985353087, 0.101286507323456 }, { 0.101286507323456, 0.797426985353087 },
    { 0.101286507323456, 0.101286507323456 } };
  static const double tw[7] = { 0.1125, 0.0661970763942531, 0.0661970763942531,
                                0.0661970763942531, 0.0629695902724136, 0.0629695902724136,
                                0.0629695902724136 };
  for (int g = 0; g < 7; g++) {
    double x = tloc[g][0], y = tloc[g][1];
    ShapeTab& s = t[ET_TRI];
    s.w[g] = tw[g];
    for (int k = 0; k < 6; k++) {
      if (k < 3) {
        double Li = L(k, x, y);
        s.f[g][k] = (2.0 * Li - 1.0) * Li;
        s.dfdx[g][k] = (4.0 * Li - 1.0) * Ldx[k];
        s.dfdy[g][k] = (4.0 * Li - 1.0) * Ldy[k];
      } else {
        int i = k % 3, j = (k + 1) % 3;
        double Li = L(i, x, y), Lj = L(j, x, y);
        s.f[g][k] = 4.0 * Li * Lj;
        s.dfdx[g][k] = 4.0 * (Ldx[i] * Lj + Li * Ldx[j]);
        s.dfdy[g][k] = 4.0 * (Ldy[i] * Lj + Li * Ldy[j]);
      }
    }
    for (int k = 0; k < 3; k++) { s.lf[g][k] = L(k, x, y); s.ldx[g][k] = Ldx[k]; s.ldy[g][k] = Ldy[k]; }
  }
  const double loc1[3] = { -std::sqrt(0.6), 0.0, std::sqrt(0.6) };
  const double w1[3] = { 5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0 };
  const double qxi[8] = { -1, 1, 1, -1, 0, 1, 0, -1 }, qeta[8] = { -1, -1, 1, 1, -1, 0, 1, 0 };
  for (int g = 0; g < 9; g++) {
    int ix = g % 3, iy = g / 3;
    double x = loc1[ix], y = loc1[iy];
    ShapeTab& s = t[ET_QUAD];
    s.w[g] = w1[ix] * w1[iy];
    for (int k = 0; k < 8; k++) {
      double xi = qxi[k], eta = qeta[k];
      if (k < 4) {
        s.f[g][k] = (1.0 + xi * x) * (1.0 + eta * y) * (xi * x + eta * y - 1.0) / 4.0;
        s.dfdx[g][k] = (1.0 + eta * y) * (2.0 * x + xi * eta * y) / 4.0;
        s.dfdy[g][k] = (1.0 + xi * x) * (2.0 * y + xi * eta * x) / 4.0;
      } else if (k == 4 || k == 6) {
        s.f[g][k] = (1.0 - x * x) * (1.0 + eta * y) / 2.0;
        s.dfdx[g][k] = -x * (1.0 + eta * y);
        s.dfdy[g][k] = eta * (1.0 - x * x) / 2.0;
      } else {
        s.f[g][k] = (1.0 + xi * x) * (1.0 - y * y) / 2.0;
        s.dfdx[g][k] = xi * (1.0 - y * y) / 2.0;
        s.dfdy[g][k] = -y * (1.0 + xi * x);
      }
    }
    for (int k = 0; k < 4; k++) {
      double xi = qxi[k], eta = qeta[k];
      s.lf[g][k] = (1.0 + xi * x) * (1.0 + eta * y) / 4.0;
      s.ldx[g][k] = xi * (1.0 + eta * y) / 4.0;
      s.ldy[g][k] = eta * (1.0 + xi * x) / 4.0;
    }
  }
  for (int g = 0; g < 3; g++) {
    double x = loc1[g];
    ShapeTab& s = t[ET_LINE];
    s.w[g] = w1[g];
    s.f[g][0] = x * (x - 1.0) / 2.0; s.f[g][1] = x * (x + 1.0) / 2.0; s.f[g][2] = 1.0 - x * x;
    s.dfdx[g][0] = x - 0.5; s.dfdx[g][1] = x + 0.5; s.dfdx[g][2] = -2.0 * x;
    s.lf[g][0] = (1.0 - x) / 2.0; s.lf[g][1] = (1.0 + x) / 2.0;
    s.ldx[g][0] = -0.5; s.ldx[g][1] = 0.5;
  }
}

void upload_shapes(hipStream_t s)
{
  static bool done = false;
  if (done) return;
  ShapeTab t[3];
  make_shapes(t);
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_shape), t, sizeof(t)));
  done = true;
  (void)s;
}

struct NodeView {
  const double *x, *y, *z, *S, *U, *V, *cf, *niox, *nioy, *valU, *valS, *rotm;
  const double *vtn, *uun;      // algebraic model: nodal eddy viscosity, uu (= vv; uv = 0)
  const double *exx, *exy, *eyy;   // anisotropic model: nodal Elder tensor (global axes)
  const double *dUdt, *dVdt, *dSdt;   // time derivatives (zero for stationary flow)
  const uint32_t* flag;
  const int* eqno;
};

struct AsmParams {
  double g, vk, hmin, relaxUV, relaxH;
  int turbAlg, turbMin;         // vt from the nodes (kVtAlgebraic); element vt as floor (kVtMin)
  int turbAniso;                // viscosity tensor exx, exy, eyy times the nodal vt (Elder)
};

// Rows of one region element node: U_i, V_i, H_i (corner only).
template <int NND, int NCN, int NGP>
__device__ __forceinline__ void region_rows(const NodeView& nv, const int* nd, int i, double vte,
                                            const AsmParams& P, double* KU, double* KV,
                                            double* KH, double& FU, double& FV, double& FH,
                                            double& area)
{
  constexpr int NL = 2 * NND + NCN;
  constexpr int T = (NND == 6) ? ET_TRI : ET_QUAD;
  const ShapeTab& s = c_shape[T];
  double x[NND], y[NND], zc[NCN], Hc[NCN], cfc[NCN], Un[NND], Vn[NND], St[NCN];
  double Ut[NND], Vt[NND];
  const double x0 = nv.x[nd[0]], y0 = nv.y[nd[0]];
#pragma unroll
  for (int k = 0; k < NND; k++) {
    x[k] = nv.x[nd[k]] - x0;
    y[k] = nv.y[nd[k]] - y0;
    Un[k] = nv.U[nd[k]];
    Vn[k] = nv.V[nd[k]];
    Ut[k] = nv.dUdt[nd[k]];
    Vt[k] = nv.dVdt[nd[k]];
  }
  x[0] = 0.0; y[0] = 0.0;
#pragma unroll
  for (int k = 0; k < NCN; k++) {
    zc[k] = nv.z[nd[k]];
    double h = nv.S[nd[k]] - zc[k];
    Hc[k] = (h <= 0.0) ? P.hmin : h;
    cfc[k] = nv.cf[nd[k]];
    St[k] = nv.dSdt[nd[k]];
  }
#pragma unroll
  for (int c = 0; c < NL; c++) { KU[c] = 0.0; KV[c] = 0.0; KH[c] = 0.0; }
  FU = FV = FH = 0.0;
  area = 0.0;
  const double grav = P.g;

  for (int g = 0; g < NGP; g++) {
    // isoparametric Jacobian (SHAPE::jacobi2D)
    double j00 = 0.0, j01 = 0.0, j10 = 0.0, j11 = 0.0;
#pragma unroll
    for (int k = 0; k < NND; k++) {
      j00 += s.dfdx[g][k] * x[k]; j01 += s.dfdx[g][k] * y[k];
      j10 += s.dfdy[g][k] * x[k]; j11 += s.dfdy[g][k] * y[k];
    }
    const double det = j00 * j11 - j01 * j10;
    const double idet = 1.0 / det;
    const double t00 = j11 * idet, t01 = -j01 * idet, t10 = -j10 * idet, t11 = j00 * idet;
    const double wgt = det * s.w[g];
    area += wgt;
    double dndx[NND], dndy[NND], dmdx[NCN], dmdy[NCN];
#pragma unroll
    for (int k = 0; k < NND; k++) {
      dndx[k] = t00 * s.dfdx[g][k] + t01 * s.dfdy[g][k];
      dndy[k] = t10 * s.dfdx[g][k] + t11 * s.dfdy[g][k];
    }
#pragma unroll
    for (int k = 0; k < NCN; k++) {
      dmdx[k] = t00 * s.ldx[g][k] + t01 * s.ldy[g][k];
      dmdy[k] = t10 * s.ldx[g][k] + t11 * s.ldy[g][k];
    }
    double H = 0, dHdx = 0, dHdy = 0, dadx = 0, dady = 0, cf = 0, dHdt = 0;
#pragma unroll
    for (int k = 0; k < NCN; k++) {
      const double m = s.lf[g][k];
      H += m * Hc[k]; dHdx += dmdx[k] * Hc[k]; dHdy += dmdy[k] * Hc[k];
      dadx += dmdx[k] * zc[k]; dady += dmdy[k] * zc[k];
      dHdt += m * St[k];
      cf += m * cfc[k];
    }
    double U = 0, V = 0, dUdx = 0, dUdy = 0, dVdx = 0, dVdy = 0, dUdt = 0, dVdt = 0;
#pragma unroll
    for (int k = 0; k < NND; k++) {
      const double n = s.f[g][k];
      U += n * Un[k]; V += n * Vn[k];
      dUdt += n * Ut[k]; dVdt += n * Vt[k];
      dUdx += dndx[k] * Un[k]; dUdy += dndy[k] * Un[k];
      dVdx += dndx[k] * Vn[k]; dVdy += dndy[k] * Vn[k];
    }
    // eddy viscosity and Reynolds stresses (vv = uu, uv = 0 in the algebraic model)
    double vt = vte, uu = 0.0;
    if (P.turbAlg) {
      vt = 0.0;
#pragma unroll
      for (int k = 0; k < NND; k++) {
        vt += s.f[g][k] * nv.vtn[nd[k]];
        uu += s.f[g][k] * nv.uun[nd[k]];
      }
    }
    // anisotropic model (EQS_UVS2D_AI::Region, CoefsUVS2D_AI.cpp:288-310): the tensor
    // sum_k n_k e_k vt_k; the floor applies to its diagonal, vk is added to all three
    double vxx = 0.0, vxy = 0.0, vyy = 0.0;
    if (P.turbAniso) {
#pragma unroll
      for (int k = 0; k < NND; k++) {
        const double n = s.f[g][k];
        const int q = nd[k];
        vxx += n * nv.exx[q] * nv.vtn[q];
        vxy += n * nv.exy[q] * nv.vtn[q];
        vyy += n * nv.eyy[q] * nv.vtn[q];
      }
      if (P.turbMin) {
        if (vxx < vte) vxx = vte;
        if (vyy < vte) vyy = vte;
      }
      vxx += P.vk; vxy += P.vk; vyy += P.vk;
    } else {
      if (P.turbMin && vt < vte) vt = vte;
      vt += P.vk;
    }
    const double vv = uu;
    const double Ures = sqrt(U * U + V * V);
    const double iUres = (Ures > (P.turbAniso ? 1.0e-8 : 1.0e-9)) ? 1.0 / Ures : 0.0;
    // anisotropic viscous fluxes per unit H: x-momentum (ax, ay), y-momentum (bx, by); the
    // isotropic terms below keep their own expressions (and rounding)
    const double ax = vxx * dUdx + vxy * dUdy, ay = vxy * dUdx + vyy * dUdy;
    const double bx = vxx * dVdx + vxy * dVdy, by = vxy * dVdx + vyy * dVdy;
    const double ni = s.f[g][i], nxi = dndx[i], nyi = dndy[i];
    const double wH = wgt * H;

    // residual rows (force = -R)
    {
      double f = H * dUdt + H * (U * dUdx + V * dUdy) + H * grav * dadx + cf * Ures * U;
      double fx = P.turbAniso ? (H * ax - H * uu) - H * H * grav / 2.0
                              : (H * vt * (dUdx + dUdx) - H * uu) - H * H * grav / 2.0;
      double fy = P.turbAniso ? H * ay : H * vt * (dUdy + dVdx);
      FU -= wgt * (ni * f + nxi * fx + nyi * fy);
      f = H * dVdt + H * (U * dVdx + V * dVdy) + H * grav * dady + cf * Ures * V;
      fx = P.turbAniso ? H * bx : H * vt * (dVdx + dUdy);
      fy = P.turbAniso ? (H * by - H * vv) - H * H * grav / 2.0
                       : (H * vt * (dVdy + dVdy) - H * vv) - H * H * grav / 2.0;
      FV -= wgt * (ni * f + nxi * fx + nyi * fy);
      if (i < NCN) {
        const double fc = dHdt + H * (dUdx + dVdy) + U * dHdx + V * dHdy;
        FH -= s.lf[g][i] * (wgt * fc);
      }
    }
    // x-momentum row U_i
    {
      const double dUU = wH * P.relaxUV + wH * dUdx + wgt * cf * (iUres * U * U + Ures);
      const double dUV = wH * dUdy + wgt * cf * iUres * U * V;
      if (P.turbAniso) {
        // d(flux)/d(U_k): x-flux cxx dndx + cxy dndy, y-flux cxy dndx + cyy dndy; no V
        const double cxx = wH * vxx, cxy = wH * vxy, cyy = wH * vyy;
#pragma unroll
        for (int k = 0; k < NND; k++) {
          const double t = dUU * s.f[g][k] + wH * U * dndx[k] + wH * V * dndy[k];
          KU[k] += ni * t + nxi * (cxx * dndx[k] + cxy * dndy[k]) +
                   nyi * (cxy * dndx[k] + cyy * dndy[k]);
          KU[NND + k] += ni * (dUV * s.f[g][k]);
        }
      } else {
        const double dxx = wH * vt * 2.0, dyy = wH * vt;
        const double dyx = wH * vt;
#pragma unroll
        for (int k = 0; k < NND; k++) {
          const double t = dUU * s.f[g][k] + wH * U * dndx[k] + wH * V * dndy[k];
          KU[k] += ni * t + nxi * (dxx * dndx[k]) + nyi * (dyy * dndy[k]);
          KU[NND + k] += ni * (dUV * s.f[g][k]) + nyi * (dyx * dndx[k]);
        }
      }
      const double dH = wgt * dUdt + wgt * (U * dUdx + V * dUdy) + wgt * grav * dadx;
      const double dHx = P.turbAniso ? (wgt * ax - wgt * uu) - wgt * grav * H
                                     : (wgt * vt * (dUdx + dUdx) - wgt * uu) - wgt * grav * H;
      const double dHy = P.turbAniso ? wgt * ay : wgt * vt * (dUdy + dVdx);
#pragma unroll
      for (int k = 0; k < NCN; k++) {
        const double m = s.lf[g][k];
        KU[2 * NND + k] += ni * (dH * m) + nxi * (dHx * m) + nyi * (dHy * m);
      }
    }
    // y-momentum row V_i
    {
      const double dVU = wH * dVdx + wgt * cf * iUres * U * V;
      const double dVV = wH * P.relaxUV + wH * dVdy + wgt * cf * (iUres * V * V + Ures);
      if (P.turbAniso) {
        const double cxx = wH * vxx, cxy = wH * vxy, cyy = wH * vyy;
#pragma unroll
        for (int k = 0; k < NND; k++) {
          KV[k] += ni * (dVU * s.f[g][k]);
          const double t = dVV * s.f[g][k] + wH * U * dndx[k] + wH * V * dndy[k];
          KV[NND + k] += ni * t + nxi * (cxx * dndx[k] + cxy * dndy[k]) +
                         nyi * (cxy * dndx[k] + cyy * dndy[k]);
        }
      } else {
        const double dxy = wH * vt;
        const double dxx = wH * vt, dyy = wH * vt * 2.0;
#pragma unroll
        for (int k = 0; k < NND; k++) {
          KV[k] += ni * (dVU * s.f[g][k]) + nxi * (dxy * dndy[k]);
          const double t = dVV * s.f[g][k] + wH * U * dndx[k] + wH * V * dndy[k];
          KV[NND + k] += ni * t + nxi * (dxx * dndx[k]) + nyi * (dyy * dndy[k]);
        }
      }
      const double dH = wgt * dVdt + wgt * (U * dVdx + V * dVdy) + wgt * grav * dady;
      const double dHx = P.turbAniso ? wgt * bx : wgt * vt * (dVdx + dUdy);
      const double dHy = P.turbAniso ? (wgt * by - wgt * vv) - wgt * grav * H
                                     : (wgt * vt * (dVdy + dVdy) - wgt * vv) - wgt * grav * H;
#pragma unroll
      for (int k = 0; k < NCN; k++) {
        const double m = s.lf[g][k];
        KV[2 * NND + k] += ni * (dH * m) + nxi * (dHx * m) + nyi * (dHy * m);
      }
    }
    // continuity row H_i (corners)
    if (i < NCN) {
      const double mi = s.lf[g][i];
      const double dHU = wgt * dHdx, dHV = wgt * dHdy;
      const double dHH = wgt * P.relaxH + wgt * (dUdx + dVdy);
#pragma unroll
      for (int k = 0; k < NND; k++) {
        KH[k] += mi * (dHU * s.f[g][k] + wH * dndx[k]);
        KH[NND + k] += mi * (dHV * s.f[g][k] + wH * dndy[k]);
      }
#pragma unroll
      for (int k = 0; k < NCN; k++)
        KH[2 * NND + k] += mi * (dHH * s.lf[g][k] + wgt * U * dmdx[k] + wgt * V * dmdy[k]);
    }
  }
}

// Rows U_i, V_i of one boundary line element node (EQS_UVS2D::Bound, no log-law walls).
__device__ __forceinline__ void bound_rows(const NodeView& nv, const int* nd, int i,
                                           const AsmParams& P, double* KU, double* KV,
                                           double& FU, double& FV)
{
  const ShapeTab& s = c_shape[ET_LINE];
#pragma unroll
  for (int c = 0; c < 8; c++) { KU[c] = 0.0; KV[c] = 0.0; }
  FU = FV = 0.0;
  const bool outMid = (nv.flag[nd[2]] & dOutlet) != 0;
  double H0, H1;
  if (outMid) { H0 = nv.valS[nd[0]] - nv.z[nd[0]]; H1 = nv.valS[nd[1]] - nv.z[nd[1]]; }
  else { H0 = nv.S[nd[0]] - nv.z[nd[0]]; H1 = nv.S[nd[1]] - nv.z[nd[1]]; }
  if (H0 <= 0.0) H0 = P.hmin;
  if (H1 <= 0.0) H1 = P.hmin;
  const double x0 = nv.x[nd[0]], x1 = nv.x[nd[1]], x2 = nv.x[nd[2]];
  const double y0 = nv.y[nd[0]], y1 = nv.y[nd[1]], y2 = nv.y[nd[2]];
  for (int g = 0; g < 3; g++) {
    const double* m = s.lf[g];
    const double* dn = s.dfdx[g];
    const double H = m[0] * H0 + m[1] * H1;
    const double nx = dn[0] * y0 + dn[1] * y1 + dn[2] * y2;
    const double ny = -dn[0] * x0 - dn[1] * x1 - dn[2] * x2;
    const double w = s.w[g];
    const double ni = s.f[g][i];
    FU -= ni * (w * H * (P.g * H / 2.0 * nx));
    FV -= ni * (w * H * (P.g * H / 2.0 * ny));
    if (!outMid) {
      const double dUH = w * (P.g * H * nx), dVH = w * (P.g * H * ny);
      KU[6] += ni * dUH * m[0]; KU[7] += ni * dUH * m[1];
      KV[6] += ni * dVH * m[0]; KV[7] += ni * dVH * m[1];
    }
  }
}

// Rotate2D restricted to the rows one thread owns: column rotation of every rotated node
// of the element on rows KU / KV / KH, then the row rotation of node i itself.
template <int NND, int NL>
__device__ __forceinline__ void rotate_rows(const NodeView& nv, const int* nd, int i, int eqid,
                                            double* KU, double* KV, double* KH, bool hasH,
                                            double& FU, double& FV)
{
#pragma unroll
  for (int k = 0; k < NND; k++) {
    if (!(nv.flag[nd[k]] & dRotat)) continue;
    const double* R = nv.rotm + 4 * (size_t)nd[k];
    const double r00 = R[0], r01 = R[1], r10 = R[2], r11 = R[3];
    if (fabs(r00 * r11 - r01 * r10) < 1.0e-12) continue;
    const int j = eqid + k;
    double a = KU[k], b = KU[j];
    KU[k] = a * r00 + b * r10; KU[j] = a * r01 + b * r11;
    a = KV[k]; b = KV[j];
    KV[k] = a * r00 + b * r10; KV[j] = a * r01 + b * r11;
    if (hasH) { a = KH[k]; b = KH[j]; KH[k] = a * r00 + b * r10; KH[j] = a * r01 + b * r11; }
  }
  if (nv.flag[nd[i]] & dRotat) {
    const double* R = nv.rotm + 4 * (size_t)nd[i];
    const double r00 = R[0], r01 = R[1], r10 = R[2], r11 = R[3];
    if (fabs(r00 * r11 - r01 * r10) >= 1.0e-12) {
#pragma unroll
      for (int c = 0; c < NL; c++) {
        const double a = KU[c], b = KV[c];
        KU[c] = r00 * a + r10 * b;
        KV[c] = r01 * a + r11 * b;
      }
      const double a = FU, b = FV;
      FU = r00 * a + r10 * b;
      FV = r01 * a + r11 * b;
    }
  }
}

template <class VT>
__device__ __forceinline__ void add_slot(VT* A, int slot, double v)
{
  A[slot] += (VT)v;
}

// MODE 0: scatter into A (VT) and B; MODE 1: write dense element rows (24x24) to Kout/Fout.
template <int TYPE, class VT, int MODE>
__global__ void __launch_bounds__(256) k_element(NodeView nv, const int* __restrict__ elems,
                                                 int nelem, const int* __restrict__ me_nodes,
                                                 const double* __restrict__ me_vt,
                                                 const long long* __restrict__ slot_off,
                                                 const int* __restrict__ slots, AsmParams P,
                                                 VT* __restrict__ A, double* __restrict__ B,
                                                 double* __restrict__ Kout,
                                                 double* __restrict__ Fout, int first)
{
  constexpr int NND = TYPE == ET_TRI ? 6 : (TYPE == ET_QUAD ? 8 : 3);
  constexpr int NCN = TYPE == ET_TRI ? 3 : (TYPE == ET_QUAD ? 4 : 2);
  constexpr int NGP = TYPE == ET_TRI ? 7 : (TYPE == ET_QUAD ? 9 : 3);
  constexpr int NL = 2 * NND + NCN;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long el = tid / NND;
  const int i = (int)(tid % NND);
  if (el >= nelem) return;
  const int e = MODE == 0 ? elems[el] : first + (int)el;
  int nd[NND];
#pragma unroll
  for (int k = 0; k < NND; k++) nd[k] = me_nodes[8 * (size_t)e + k];
  double KU[NL], KV[NL], KH[NL];
  double FU, FV, FH = 0.0, area = 0.0;
  const bool hasH = TYPE != ET_LINE && i < NCN;
  if constexpr (TYPE == ET_LINE) {
    bound_rows(nv, nd, i, P, KU, KV, FU, FV);
#pragma unroll
    for (int c = 0; c < NL; c++) KH[c] = 0.0;
  } else {
    region_rows<NND, NCN, NGP>(nv, nd, i, me_vt[e], P, KU, KV, KH, FU, FV, FH, area);
  }
  rotate_rows<NND, NL>(nv, nd, i, NND, KU, KV, KH, hasH, FU, FV);
  // inlet rows
  if (nv.flag[nd[i]] & dInlet) {
#pragma unroll
    for (int c = 0; c < NL; c++) KU[c] = 0.0;
    if constexpr (TYPE == ET_LINE) {
      FU = 0.0;
    } else {
      const int k = nd[i];
      const double Un = nv.U[k] * nv.niox[k] + nv.V[k] * nv.nioy[k];
      double specQ = nv.valU[k];
      if (i < NCN) {
        double H = nv.S[k] - nv.z[k];
        if (H <= 0.0) { H = P.hmin; specQ = 0.0; }
        FU = area * (specQ - Un * H);
        KU[i] = area * H;
        KU[2 * NND + i] = area * Un;
      } else {
        const int l = i - NCN, r = (l + 1) % NCN;
        const double Hl = nv.S[nd[l]] - nv.z[nd[l]], Hr = nv.S[nd[r]] - nv.z[nd[r]];
        double H = (Hl + Hr) / 2.0;
        if (H <= 0.0) { H = P.hmin; specQ = 0.0; }
        FU = area * (specQ - Un * H);
        KU[i] = area * H;
        KU[2 * NND + l] = area * Un / 2.0;
        KU[2 * NND + r] = area * Un / 2.0;
      }
    }
  }
  if constexpr (MODE == 1) {
    double* K = Kout + (size_t)el * 576;
    double* F = Fout + (size_t)el * 24;
    const int rows[3] = { i, NND + i, 2 * NND + i };
    const double* src[3] = { KU, KV, KH };
    const double fr[3] = { FU, FV, FH };
    for (int q = 0; q < (hasH ? 3 : 2); q++) {
      for (int c = 0; c < NL; c++) K[rows[q] * 24 + c] = src[q][c];
      F[rows[q]] = fr[q];
    }
  } else {
    const int* sl = slots + slot_off[e];
    const int ldof[3] = { i, NND + i, 2 * NND + i };
    const double fr[3] = { FU, FV, FH };
    const int nrow = hasH ? 3 : 2;
    for (int q = 0; q < nrow; q++) {
      const int eq = nv.eqno[3 * (size_t)nd[i] + q];
      if (eq < 0) continue;
      B[eq] += fr[q];
      const double* Kr = q == 0 ? KU : (q == 1 ? KV : KH);
      const int* srow = sl + ldof[q] * NL;
#pragma unroll
      for (int c = 0; c < NL; c++) {
        const int slot = srow[c];
        if (slot >= 0) add_slot(A, slot, Kr[c]);
      }
    }
  }
}

// ---- emergent vegetation (Lindner) ----
// Drag coefficient of one cylinder from its Reynolds number (TYPE::dragCoeff,
// Lindner.cpp:264-278).
__device__ double drag_cylinder(double v, double d, double vk)
{
  const double re = v * d / vk;
  if (re <= 800.0) return 3.07 / pow(re, 0.168);
  if (re <= 6000.0) return 1.0;
  if (re <= 11000.0) return 1.0 + 0.2 * (re - 6000.0) / 5000.0;
  return 1.2;
}

// Real roots of a x^3 + b x^2 + d (no linear term) in Cardano or trigonometric form
// (TYPE::cubeEquation, Lindner.cpp:281-335); returns their number.
__device__ int cubic_roots(double a, double b, double d, double x[3])
{
  const double pi = 3.141592653589793;
  const double ba = b / a / 3.0, ca = 0.0 / a;
  const double p = ca / 3.0 - ba * ba;
  const double q = ba * ba * ba - ba * ca / 2.0 + d / a / 2.0;
  const double disc = q * q + p * p * p;
  if (disc > 0.0) {
    double t = -q + sqrt(disc);
    const double u = (t / fabs(t)) * pow(fabs(t), 1.0 / 3.0);
    t = -q - sqrt(disc);
    const double v = (t / fabs(t)) * pow(fabs(t), 1.0 / 3.0);
    x[0] = (u + v) - ba;
    return 1;
  }
  const double t = -q / pow(-p, 1.5);
  const double phi = t >= 1.0 ? 0.0 : (t <= -1.0 ? pi : acos(t));
  x[0] = 2.0 * sqrt(-p) * cos(phi / 3.0) - ba;
  x[1] = 2.0 * sqrt(-p) * cos((phi + 2.0 * pi) / 3.0) - ba;
  x[2] = 2.0 * sqrt(-p) * cos((phi + 4.0 * pi) / 3.0) - ba;
  return 3;
}

// Friction coefficient with emergent stems of diameter dp and spacing sp (TYPE::lindner,
// Lindner.cpp:72-262): the stems' drag coefficient cWR is iterated to 1e-3 (Pegasus steps
// after the third iteration, at most 200), each step with the wake length from a damped
// fixed-point iteration and the depth ratio from the momentum-balance cubic.  Returns
// lambda / 8 = cb + cWR h dp / (2 sp^2), cb itself below the thresholds, -1 without
// convergence.
__device__ double lindner(double dp, double sp, double va, double ha, double cb, double vk, double g)
{
  if (dp < 1.0e-4 || sp < 1.0e-4 || va < 1.0e-3 || ha < 1.0e-3) return cb;
  double cwr = 1.0, anl = sp / 2.0;
  double c1 = 1.0, c2 = 1.0, r1 = 0.0, r2 = 0.0, lambda = 0.0;
  bool done = false;
  for (int it = 0; it < 200 && !done; it++) {
    lambda = 8.0 * cb + 4.0 * cwr * ha * dp / sp / sp;
    const double cw = drag_cylinder(va, dp, vk);
    bool wake = false;
    for (int jt = 0; jt < 200; jt++) {
      const double t1 = 1.0 + anl * lambda / 4.0 / ha;
      const double t2 = 30.0 / pow(fabs(t1), 1.5);
      const double ra = cw * dp * pow(fabs(t2), 1.429);
      if (fabs((ra - anl) / ra) < 1.0e-3) { anl = ra; wake = true; break; }
      anl = 0.5 * (ra + anl);
    }
    if (!wake) anl = sp / 2.0;
    const double anb = 0.24 * pow(fabs(anl), 0.59) * pow(fabs(cw * dp), 0.41);
    const double vr = 1.151 * pow(fabs(anl / sp), -0.483) + 0.5 * pow(fabs(anb / sp), 1.1);
    const double fr = va / sqrt(g * ha);
    const double al = dp / sp;
    const double ca = fr * fr * (1.0 - al * cwr / 2.0);
    const double cbb = -fr * fr - (1.0 - al) / 2.0;
    const double cd = (1.0 - al) / 2.0;
    double hr = 1.0;
    if (fabs(ca) < 1.0e-10) {
      hr = sqrt(-cd / cbb);
    } else {
      double x[3];
      const int nr = cubic_roots(ca, cbb, cd, x);
      for (int i = 0; i < nr; i++)
        if (x[i] > 0.0 && x[i] < 1.0) { hr = x[i]; break; }
    }
    const double rc = 1.3124 * cw * vr + 2.0 * (1.0 - hr) / fr / fr;
    if (fabs((rc - cwr) / rc) < 1.0e-3) { done = true; break; }
    const double dc = rc - cwr;
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
  return done ? lambda / 8.0 : -1.0;
}

// ---- nodal friction (MODEL::DoFriction + TYPE::bottom: Manning, emergent vegetation) ----
// Per node the mean over its region elements of areaFact * cf_e, cf_e = Lindner's
// coefficient over the Manning coefficient cb (cb + 0.75 h dp / sp^2 where it does not
// converge, Friction.cpp:101-108), evaluated at the node's depth and speed.
__global__ void k_friction(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                           const double* __restrict__ manning,
                           const double* __restrict__ veg_dp, const double* __restrict__ veg_sp,
                           const double* __restrict__ areaFact, const double* __restrict__ S,
                           const double* __restrict__ z, const double* __restrict__ U,
                           const double* __restrict__ V, double g, double vk, double hmin,
                           double* __restrict__ cf)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  double h = S[n] - z[n];
  if (h < hmin) h = hmin;
  const double ph = pow(h, 0.333333);
  const double vres = sqrt(U[n] * U[n] + V[n] * V[n]);
  double sum = 0.0;
  const int a = n2e_ptr[n], b = n2e_ptr[n + 1];
  for (int q = a; q < b; q++) {
    const int e = n2e[q];
    const double nm = manning[e];
    const double cb = g * (nm * nm) / ph;
    const double dp = veg_dp[e], sp = veg_sp[e];
    double cp = cb;
    if (dp >= 1.0e-4 && sp >= 1.0e-4) {
      cp = lindner(dp, sp, vres, h, cb, vk, g);
      if (cp < -0.9) cp = cb + 0.75 * h * dp / sp / sp;
    }
    sum += areaFact[e] * cp;
  }
  cf[n] = (b > a) ? sum / (double)(b - a) : 0.0;
}

// ---- algebraic eddy viscosity (GRID::Turbulence, kVtAlgebraic, Turbulence.cpp:292-352) ----
// Per node: ste = mean of st*KDest over the incident elements (ascending element order),
// H = S - z unclipped, Utau = sqrt(cf (U^2 + V^2)), vt = ste H Utau,
// uu = vv = -2/3 Utau^2 sqrt(cK ste / cm / cd) with cK = 1/sqrt(cf); uv = 0.
__global__ void k_turb_alg(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                           const double* __restrict__ ste_e, const double* __restrict__ S,
                           const double* __restrict__ z, const double* __restrict__ U,
                           const double* __restrict__ V, const double* __restrict__ cf,
                           double cm, double cd, double* __restrict__ vt, double* __restrict__ uu)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int a = n2e_ptr[n], b = n2e_ptr[n + 1];
  double v = 0.0, r = 0.0;
  if (b > a) {
    double ste = 0.0;
    for (int q = a; q < b; q++) ste += ste_e[n2e[q]];
    ste /= (double)(b - a);
    const double u = U[n], w = V[n], H = S[n] - z[n];
    const double Utau = sqrt(cf[n] * (u * u + w * w));
    const double cK = 1.0 / sqrt(cf[n]);
    v = ste * H * Utau;
    r = -2.0 / 3.0 * Utau * Utau * sqrt(cK * ste / cm / cd);
  }
  vt[n] = v;
  uu[n] = r;
}

// ---- constant eddy viscosity at the nodes (GRID::Turbulence, kVtConstant,
// Turbulence.cpp:110-165): the mean of the incident elements' vt (all elements wet here);
// uu = 0.  Only the anisotropic model reads it (the isotropic elements take their own vt).
__global__ void k_turb_const(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                             const double* __restrict__ vt_e, double* __restrict__ vt,
                             double* __restrict__ uu)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int a = n2e_ptr[n], b = n2e_ptr[n + 1];
  double v = 0.0;
  for (int q = a; q < b; q++) v += vt_e[n2e[q]];
  vt[n] = b > a ? v / (double)(b - a) : 0.0;
  uu[n] = 0.0;
}

// ---- anisotropic eddy viscosity after Elder (Turbulence.cpp:996-1090): per node the mean
// of estx / KDest and esty / KDest over the incident elements, turned into the direction
// of the nodal velocity (angle to the x axis; none below |U| = 1e-6) ----
__global__ void k_turb_elder(int np, const int* __restrict__ n2e_ptr, const int* __restrict__ n2e,
                             const double* __restrict__ exi, const double* __restrict__ eyi,
                             const double* __restrict__ U, const double* __restrict__ V,
                             double* __restrict__ exx, double* __restrict__ exy,
                             double* __restrict__ eyy)
{
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const int a = n2e_ptr[n], b = n2e_ptr[n + 1];
  double ex = 0.0, ey = 0.0;
  for (int q = a; q < b; q++) { ex += exi[n2e[q]]; ey += eyi[n2e[q]]; }
  if (b > a) { ex /= (double)(b - a); ey /= (double)(b - a); }
  const double u = U[n], w = V[n], Us = sqrt(u * u + w * w);
  double sa = 0.0, ca = 1.0;
  if (Us > 1.0e-6) { sa = w / Us; ca = u / Us; }
  exx[n] = ex * ca * ca + ey * sa * sa;
  exy[n] = (ex - ey) * sa * ca;
  eyy[n] = ex * sa * sa + ey * ca * ca;
}

// ---- inlet discharge distribution (BCONLINE::Inlet, QINLET, quadratic line rule) ----
// Dry and marsh nodes of the line take no discharge.
__global__ void k_inlet(int nseg, const int* __restrict__ segs, int nin,
                        const int* __restrict__ inodes, const uint8_t* __restrict__ nwd,
                        const double* __restrict__ x,
                        const double* __restrict__ y, const double* __restrict__ z,
                        const double* __restrict__ S, const double* __restrict__ cf,
                        double Qspec, double Sspec, double* __restrict__ valU)
{
  __shared__ double s_ratio;
  if (threadIdx.x == 0) {
    const ShapeTab& s = c_shape[ET_LINE];
    double Q = 0.0;
    bool any = false;
    for (int c = 0; c < nseg; c++) {
      const int* nd = segs + 3 * c;
      double q[3];
      for (int k = 0; k < 3; k++) {
        const int n = nd[k];
        if (S[n] - z[n] < 1.0e-9 || (nwd[n] & (WD_DRY | WD_MARSH))) { q[k] = 0.0; continue; }
        const double h = Sspec - z[n];
        q[k] = h > 0.0 ? sqrt(h * h * h) : 0.0;
        if (cf[n] > 0.0) q[k] /= sqrt(cf[n]);
        any = true;
      }
      for (int g = 0; g < 3; g++) {
        const double* N = s.f[g];
        const double* dN = s.dfdx[g];
        const double dx = dN[0] * x[nd[0]] + dN[1] * x[nd[1]] + dN[2] * x[nd[2]];
        const double dy = dN[0] * y[nd[0]] + dN[1] * y[nd[1]] + dN[2] * y[nd[2]];
        Q += s.w[g] * sqrt(dx * dx + dy * dy) * (N[0] * q[0] + N[1] * q[1] + N[2] * q[2]);
      }
    }
    s_ratio = (any && fabs(Q) > 0.0) ? -Qspec / Q : 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nin; t += blockDim.x) {
    const int n = inodes[t];
    const double h = Sspec - z[n];
    double q = h > 0.0 ? sqrt(h * h * h) : 0.0;
    if (cf[n] > 0.0) q /= sqrt(cf[n]);
    if (nwd[n] & (WD_DRY | WD_MARSH)) q = 0.0;
    valU[n] = q * s_ratio;
  }
}

static NodeView node_view(Ctx& c)
{
  NodeView v;
  v.x = c.x.p; v.y = c.y.p; v.z = c.z.p; v.S = c.S.p; v.U = c.U.p; v.V = c.V.p; v.cf = c.cf.p;
  v.niox = c.niox.p; v.nioy = c.nioy.p; v.valU = c.valU.p; v.valS = c.valS.p; v.rotm = c.rotm.p;
  v.flag = c.dflag.p; v.eqno = c.eqno.p;
  v.vtn = c.vtn.p; v.uun = c.uun.p;
  v.exx = c.exx.p; v.exy = c.exy.p; v.eyy = c.eyy.p;
  v.dUdt = c.dUdt.p; v.dVdt = c.dVdt.p; v.dSdt = c.dSdt.p;
  return v;
}

static AsmParams asm_params(Ctx& c)
{
  AsmParams p;
  p.g = c.par.g; p.vk = c.par.vk; p.hmin = c.par.hmin();
  // the time terms' diagonal: 1 / (dt theta) for unsteady flow, 1 / (dt_r theta) of the
  // relaxed time steps with relaxation methods 3 and 4 (EqsUVS2D.cpp:97-147, 613-622)
  p.relaxUV = c.thdt_uv;
  p.relaxH = c.thdt_h;
  p.turbAlg = c.par.turb_alg() ? 1 : 0; p.turbMin = c.par.turb_min() ? 1 : 0;
  p.turbAniso = c.par.turb_aniso() ? 1 : 0;
  return p;
}

static void launch_friction(Ctx& c)
{
  const int np = c.hm.np;
  k_friction<<<(np + 255) / 256, 256, 0, c.stream>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_manning.p,
                                                     c.d_veg_dp.p, c.d_veg_sp.p, c.d_areaFact.p,
                                                     c.S.p, c.z.p, c.U.p, c.V.p, c.par.g,
                                                     c.par.vk, c.par.hmin(), c.cf.p);
  HIPCHECK(hipGetLastError());
}

static void launch_turbulence(Ctx& c)
{
  const int np = c.hm.np, g = (np + 255) / 256;
  if (c.par.turb_alg())
    k_turb_alg<<<g, 256, 0, c.stream>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_ste.p, c.S.p, c.z.p, c.U.p,
                                        c.V.p, c.cf.p, c.par.cm, c.par.cd, c.vtn.p, c.uun.p);
  else
    k_turb_const<<<g, 256, 0, c.stream>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_vte.p, c.vtn.p, c.uun.p);
  if (c.par.turb_aniso())
    k_turb_elder<<<g, 256, 0, c.stream>>>(np, c.n2e_ptr.p, c.n2e.p, c.d_exi.p, c.d_eyi.p, c.U.p,
                                          c.V.p, c.exx.p, c.exy.p, c.eyy.p);
  HIPCHECK(hipGetLastError());
  c.turb_stale = false;
}

void dev_turbulence(Ctx& c)
{
  launch_friction(c);
  launch_turbulence(c);
}

void dev_prepare_step(Ctx& c)
{
  upload_shapes(c.stream);
  launch_friction(c);
  // the flow cycle evaluates the eddy viscosity at its first Newton step only, unless
  // kVtIterat asks for every step (EqsUVS2D.cpp:191, 311-315)
  if ((c.par.turb_alg() || c.par.turb_aniso()) && (c.turb_stale || c.par.turb_iterat()))
    launch_turbulence(c);
  const int nseg = (int)(c.hm.in_segs.size() / 3), nin = (int)c.hm.inlet_nodes.size();
  if (nseg > 0) {
    k_inlet<<<1, 256, 0, c.stream>>>(nseg, c.d_in_segs.p, nin, c.d_inlet_nodes.p, c.d_nwd.p,
                                     c.x.p, c.y.p, c.z.p, c.S.p, c.cf.p, c.hm.Q, c.hm.S_in,
                                     c.valU.p);
    HIPCHECK(hipGetLastError());
  }
}

template <class VT>
static void launch_groups(Ctx& c, VT* A)
{
  NodeView nv = node_view(c);
  AsmParams P = asm_params(c);
  for (size_t gI = 0; gI + 1 < c.col_ptr.size(); gI++) {
    const int a = c.col_ptr[gI], b = c.col_ptr[gI + 1];
    const int cnt = b - a;
    const int t = c.col_type[gI];
    const int nnd = t == ET_TRI ? 6 : (t == ET_QUAD ? 8 : 3);
    const long long nthreads = (long long)cnt * nnd;
    const int blocks = (int)((nthreads + 255) / 256);
    const int* el = c.col_elems.p + a;
    if (t == ET_TRI)
      k_element<ET_TRI, VT, 0><<<blocks, 256, 0, c.stream>>>(nv, el, cnt, c.me_nodes.p, c.me_vt.p,
          c.me_slot_off.p, c.slots.p, P, A, c.B.p, nullptr, nullptr, 0);
    else if (t == ET_QUAD)
      k_element<ET_QUAD, VT, 0><<<blocks, 256, 0, c.stream>>>(nv, el, cnt, c.me_nodes.p, c.me_vt.p,
          c.me_slot_off.p, c.slots.p, P, A, c.B.p, nullptr, nullptr, 0);
    else
      k_element<ET_LINE, VT, 0><<<blocks, 256, 0, c.stream>>>(nv, el, cnt, c.me_nodes.p, c.me_vt.p,
          c.me_slot_off.p, c.slots.p, P, A, c.B.p, nullptr, nullptr, 0);
    HIPCHECK(hipGetLastError());
  }
}

void dev_assemble(Ctx& c, int prec, bool zero)
{
  upload_shapes(c.stream);
  if (zero) HIPCHECK(hipMemsetAsync(c.B.p, 0, sizeof(double) * c.n, c.stream));
  if (prec == RMX_F32) {
    c.A32.alloc(c.nnz);
    if (zero) HIPCHECK(hipMemsetAsync(c.A32.p, 0, sizeof(float) * c.nnz, c.stream));
    launch_groups<float>(c, c.A32.p);
    c.have32 = true; c.have64 = false;
    c.a32_version++;
  } else {
    c.A64.alloc(c.nnz);
    if (zero) HIPCHECK(hipMemsetAsync(c.A64.p, 0, sizeof(double) * c.nnz, c.stream));
    launch_groups<double>(c, c.A64.p);
    c.have64 = true; c.have32 = false;
  }
  c.ilu.fac32 = c.ilu.fac64 = c.silu.fac32 = c.silu.fac64 = false;
  c.mat_version++;
}

void dev_element_coefs(Ctx& c, int first, int count, double* K, double* F)
{
  upload_shapes(c.stream);
  NodeView nv = node_view(c);
  AsmParams P = asm_params(c);
  std::vector<int> type(count);
  HIPCHECK(hipMemcpyAsync(type.data(), c.me_type.p + first, sizeof(int) * count,
                          hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  DBuf<double> dK, dF;
  dK.alloc((size_t)count * 576);
  dF.alloc((size_t)count * 24);
  HIPCHECK(hipMemsetAsync(dK.p, 0, sizeof(double) * 576 * count, c.stream));
  HIPCHECK(hipMemsetAsync(dF.p, 0, sizeof(double) * 24 * count, c.stream));
  // one launch per run of equal element types
  int a = 0;
  while (a < count) {
    int b = a;
    while (b < count && type[b] == type[a]) b++;
    const int t = type[a], cnt = b - a;
    const int nnd = t == ET_TRI ? 6 : (t == ET_QUAD ? 8 : 3);
    const int blocks = (int)(((long long)cnt * nnd + 255) / 256);
    double* Ko = dK.p + (size_t)a * 576;
    double* Fo = dF.p + (size_t)a * 24;
    if (t == ET_TRI)
      k_element<ET_TRI, double, 1><<<blocks, 256, 0, c.stream>>>(nv, nullptr, cnt, c.me_nodes.p,
          c.me_vt.p, nullptr, nullptr, P, nullptr, nullptr, Ko, Fo, first + a);
    else if (t == ET_QUAD)
      k_element<ET_QUAD, double, 1><<<blocks, 256, 0, c.stream>>>(nv, nullptr, cnt, c.me_nodes.p,
          c.me_vt.p, nullptr, nullptr, P, nullptr, nullptr, Ko, Fo, first + a);
    else
      k_element<ET_LINE, double, 1><<<blocks, 256, 0, c.stream>>>(nv, nullptr, cnt, c.me_nodes.p,
          c.me_vt.p, nullptr, nullptr, P, nullptr, nullptr, Ko, Fo, first + a);
    HIPCHECK(hipGetLastError());
    a = b;
  }
  dK.download(K, (size_t)count * 576, c.stream);
  dF.download(F, (size_t)count * 24, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
}

}  // namespace rmx
