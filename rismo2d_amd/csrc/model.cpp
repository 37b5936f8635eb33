// Host-side model set-up of the UVS2D solve: topology, boundary line elements, boundary
// kinds / normals / rotations, equation numbering, CSR pattern, element colouring, ILU
// level sets and the element-to-CSR slot map.  Integer work done once per structure;
// every loop that scales with the mesh is split over host threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "rmx_internal.h"

namespace rmx {

void parallel_for(long long n, const std::function<void(long long, long long)>& f)
{
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  long long nt = std::min<long long>(hw, std::max<long long>(1, n / 4096));
  if (nt <= 1) { f(0, n); return; }
  std::vector<std::thread> th;
  long long chunk = (n + nt - 1) / nt;
  for (long long t = 0; t < nt; t++) {
    long long a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) th.emplace_back(f, a, b);
  }
  for (auto& t : th) t.join();
}

// Inverse incidence: for every node the elements that contain it, ascending.
static void node_incidence(int np, int nelem, const std::function<int(int, const int**)>& nodes_of,
                           std::vector<int>& ptr, std::vector<int>& list)
{
  ptr.assign(np + 1, 0);
  for (int e = 0; e < nelem; e++) {
    const int* nd; int k = nodes_of(e, &nd);
    for (int i = 0; i < k; i++) ptr[nd[i] + 1]++;
  }
  for (int i = 0; i < np; i++) ptr[i + 1] += ptr[i];
  list.assign(ptr[np], 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < nelem; e++) {
    const int* nd; int k = nodes_of(e, &nd);
    for (int i = 0; i < k; i++) list[fill[nd[i]]++] = e;
  }
}

// MODEL::Input / Initialize / Boundary and GRID::AreaFactors (Model.cpp:121-215,
// Bound.cpp:50-170, ArFact.cpp:41-85).
void host_mesh(HostModel& hm, const rmx_mesh_desc& d)
{
  hm = HostModel();
  hm.np = d.np; hm.ne = d.ne;
  hm.x.assign(d.x, d.x + d.np); hm.y.assign(d.y, d.y + d.np); hm.zor.assign(d.z, d.z + d.np);
  hm.elem.assign(d.elems, d.elems + 8LL * d.ne); hm.nnd.assign(d.nnd, d.nnd + d.ne);
  hm.manning.assign(d.manning, d.manning + d.ne); hm.vt.assign(d.vt, d.vt + d.ne);
  hm.veg_dp.assign(d.ne, 0.0); hm.veg_sp.assign(d.ne, 0.0);
  if (d.veg_dp && d.veg_sp) {
    hm.veg_dp.assign(d.veg_dp, d.veg_dp + d.ne);
    hm.veg_sp.assign(d.veg_sp, d.veg_sp + d.ne);
  }
  hm.ste.resize(d.ne); hm.exi.resize(d.ne); hm.eyi.resize(d.ne);
  for (int e = 0; e < d.ne; e++) {             // Type.cpp:73-74 defaults st 1.0, KDest 0.6
    const double kd = d.turb_kdest ? d.turb_kdest[e] : 0.6;
    hm.ste[e] = (d.turb_st ? d.turb_st[e] : 1.0) * kd;
    hm.exi[e] = (d.turb_estx ? d.turb_estx[e] : 0.0) / kd;   // Turbulence.cpp:1024-1025
    hm.eyi[e] = (d.turb_esty ? d.turb_esty[e] : 0.0) / kd;
  }
  hm.in_segs.assign(d.inlet_segs, d.inlet_segs + 3LL * d.n_inlet);
  hm.out_segs.assign(d.outlet_segs, d.outlet_segs + 3LL * d.n_outlet);
  hm.Q = d.Q; hm.S_in = d.S_in; hm.S_out = d.S_out;
  for (int e = 0; e < d.ne; e++) {
    if (d.nnd[e] != 6 && d.nnd[e] != 8) fail(RMX_ERR_ARG, "element with nnd other than 6 or 8");
    for (int i = 0; i < d.nnd[e]; i++) {
      int v = d.elems[8LL * e + i];
      if (v < 0 || v >= d.np) fail(RMX_ERR_ARG, "element node index out of range");
    }
  }
  for (int v : hm.in_segs) if (v < 0 || v >= d.np) fail(RMX_ERR_ARG, "inlet node out of range");
  for (int v : hm.out_segs) if (v < 0 || v >= d.np) fail(RMX_ERR_ARG, "outlet node out of range");
  // area factors of inclined elements
  hm.areaFact.assign(d.ne, 1.0);
  parallel_for(d.ne, [&](long long a, long long b) {
    for (long long e = a; e < b; e++) {
      const int* nd = &hm.elem[8 * e];
      int ncn = hm.nnd[e] == 6 ? 3 : 4;
      double x[4], y[4], z[4];
      x[0] = hm.x[nd[0]]; y[0] = hm.y[nd[0]]; z[0] = hm.zor[nd[0]];
      for (int i = 1; i < ncn; i++) {
        x[i] = hm.x[nd[i]] - x[0]; y[i] = hm.y[nd[i]] - y[0]; z[i] = hm.zor[nd[i]] - z[0];
      }
      double t1 = y[1] * z[2] - y[2] * z[1], t2 = z[1] * x[2] - z[2] * x[1];
      double t3 = x[1] * y[2] - x[2] * y[1];
      double ae = std::sqrt(t1 * t1 + t2 * t2 + t3 * t3), ap = t3;
      if (ncn == 4) {
        t1 = y[2] * z[3] - y[3] * z[2]; t2 = z[2] * x[3] - z[3] * x[2];
        t3 = x[2] * y[3] - x[3] * y[2];
        ae += std::sqrt(t1 * t1 + t2 * t2 + t3 * t3); ap += t3;
      }
      hm.areaFact[e] = ae / ap;
    }
  });
  hm.flag.assign(d.np, 0);
  hm.nwd.assign(d.np, 0);
  hm.ewd.assign(d.ne, 0);
  host_topology(hm);
}

// GRID::Connection over the wet elements and MODEL::Initialize / Boundary (Model.cpp:121-215,
// Bound.cpp:50-170): corner / midside flags, the wet elements connected to each node, a
// boundary line element for every midside node of one wet element, and the model element
// list (each wet region element followed by its boundary line elements).  The boundary
// flags of the last set-up are cleared first (Bound.cpp:64-80); dry elements (EL_DRY) are
// left out, so the wet/dry line becomes a boundary.
void host_topology(HostModel& hm)
{
  const int np = hm.np, ne = hm.ne;
  hm.noel.assign(np, 0);
  for (int n = 0; n < np; n++) hm.flag[n] &= ~nBound;
  for (int e = 0; e < ne; e++) {
    int k = hm.nnd[e], ncn = k == 6 ? 3 : 4;
    const bool wet = !(hm.ewd[e] & EL_DRY);
    for (int i = 0; i < k; i++) {
      int n = hm.elem[8LL * e + i];
      if (wet) hm.noel[n]++;
      hm.flag[n] |= (i < ncn) ? nCorn : nMids;
    }
  }
  for (int n = 0; n < np; n++)
    if (hm.noel[n] == 1 && (hm.flag[n] & nMids)) hm.flag[n] |= nBound;
  hm.bnd.clear();
  hm.me.clear();
  hm.me.reserve(ne + ne / 8);
  for (int e = 0; e < ne; e++) {
    if (hm.ewd[e] & EL_DRY) continue;
    int k = hm.nnd[e], ncn = k == 6 ? 3 : 4;
    hm.me.push_back(e);
    for (int i = ncn; i < k; i++) {
      int mid = hm.elem[8LL * e + i];
      if (!(hm.flag[mid] & nBound)) continue;
      int l = i - ncn, r = (l + 1) % ncn;
      int b = (int)(hm.bnd.size() / 3);
      hm.bnd.push_back(hm.elem[8LL * e + l]);
      hm.bnd.push_back(hm.elem[8LL * e + r]);
      hm.bnd.push_back(mid);
      hm.flag[hm.elem[8LL * e + l]] |= nBound;
      hm.flag[hm.elem[8LL * e + r]] |= nBound;
      hm.me.push_back(-(b + 1));
    }
  }
  hm.nb = (int)(hm.bnd.size() / 3);
  hm.bflag.assign(hm.nb, 0);
}

// Static part of BCONSET::InitBcon (BconSet.cpp:92-480, BconLine.cpp:249-900) for one
// QINLET and one OUTLET line, then MODEL::SetLocation (Locate.cpp:47), SetNormal
// (Transform.cpp:45) and SetRotation (Transform.cpp:203).  The inlet velocities depend
// on the nodal friction and are set per Newton step on the device.
void host_boundary_conditions(HostModel& hm)
{
  const int np = hm.np;
  hm.kind.assign(np, 0);
  hm.valS.assign(np, 0.0);
  hm.rot.assign(np, 0);
  std::vector<char> mark(np, 0);
  for (int v : hm.in_segs) mark[v] = 1;
  hm.inlet_nodes.clear();
  for (int n = 0; n < np; n++)
    if (mark[n]) {
      hm.inlet_nodes.push_back(n);
      hm.kind[n] = kQInlet | kInlet | (hm.inflow_qt ? kQTInlet : 0u);   // BconLine.cpp:86-87
      hm.valS[n] = hm.S_in;
    }
  for (int v : hm.out_segs)
    if (!(hm.kind[v] & kInlet)) { hm.kind[v] |= kOutlet; hm.valS[v] = hm.S_out; }
  // SetLocation: line element flags from its corner kinds, node flags from the element
  std::vector<uint32_t> nodeLoc(np, 0);
  for (int b = 0; b < hm.nb; b++) {
    const int* nd = &hm.bnd[3 * b];
    uint32_t f = 7;
    for (int i = 0; i < 2; i++) {
      uint32_t k = hm.kind[nd[i]];
      if (!(k & kInlet)) f &= ~1u;
      if (!(k & kOutlet)) f &= ~2u;
      if (!(k & kOpenBnd)) f &= ~4u;
    }
    hm.bflag[b] |= f;
    for (int i = 0; i < 3; i++) nodeLoc[nd[i]] |= f;
  }
  // SetNormal: sums of the line elements' nodal normals, unit length
  hm.nx.assign(np, 0.0); hm.ny.assign(np, 0.0); hm.niox.assign(np, 0.0); hm.nioy.assign(np, 0.0);
  // derivative of the quadratic line shape at its nodes (xi = -1, +1, 0), Line.cpp
  const double lloc[3] = { -1.0, 1.0, 0.0 };
  auto qLineDx = [](int j, double x) { return j == 0 ? x - 0.5 : (j == 1 ? x + 0.5 : -2.0 * x); };
  for (int b = 0; b < hm.nb; b++) {
    const int* nd = &hm.bnd[3 * b];
    for (int i = 0; i < 3; i++) {
      double nx = 0.0, ny = 0.0;
      for (int j = 0; j < 3; j++) {
        double dn = qLineDx(j, lloc[i]);
        nx += 2.0 * (dn * hm.y[nd[j]]);
        ny -= 2.0 * (dn * hm.x[nd[j]]);
      }
      if (hm.bflag[b] & 7) { hm.niox[nd[i]] += nx; hm.nioy[nd[i]] += ny; }
      else { hm.nx[nd[i]] += nx; hm.ny[nd[i]] += ny; }
    }
  }
  for (int n = 0; n < np; n++) {
    if (!(hm.flag[n] & nBound)) continue;
    double len = std::sqrt(hm.nx[n] * hm.nx[n] + hm.ny[n] * hm.ny[n]);
    if (std::fabs(len) < 1.0e-12) { hm.nx[n] = 0.0; hm.ny[n] = 0.0; }
    else { hm.nx[n] /= len; hm.ny[n] /= len; }
    len = std::sqrt(hm.niox[n] * hm.niox[n] + hm.nioy[n] * hm.nioy[n]);
    if (std::fabs(len) < 1.0e-12) { hm.niox[n] = 0.0; hm.nioy[n] = 0.0; }
    else { hm.niox[n] /= len; hm.nioy[n] /= len; }
  }
  // SetRotation
  for (int n = 0; n < np; n++) {
    hm.flag[n] &= ~nRotat;
    uint32_t& k = hm.kind[n];
    if (k & kInlet) k &= ~(kFixU | kFixV);
    if (k & kAutoSlip) k &= ~(kAutoSlip | kFixU | kFixV);
    if ((hm.flag[n] & nBound) && !(k & (kInlet | kOutlet | kOpenBnd | kSlip | kSetUV))) {
      hm.rot[n] = rotTangent;
      k |= kAutoSlip | kFixV;
    }
  }
  for (int b = 0; b < hm.nb; b++) {
    if (hm.bflag[b] & 6) continue;             // outlet or open line element
    for (int i = 0; i < 2; i++) {
      int n = hm.bnd[3 * b + i];
      uint32_t& k = hm.kind[n];
      if ((k & kOutlet) || (k & kOpenBnd))
        if (!(k & kSlip) && !(k & kSetUV)) { hm.rot[n] = rotNormal; k |= kAutoSlip | kFixV; }
    }
  }
  for (int n = 0; n < np; n++)
    if (hm.kind[n] & kAutoSlip) {
      double len = std::sqrt(hm.nx[n] * hm.nx[n] + hm.ny[n] * hm.ny[n]);
      if (len < 1.0e-12) hm.flag[n] |= nNoMoment;
    }
  for (int n = 0; n < np; n++) {
    uint32_t& k = hm.kind[n];
    if (k & kInlet) { hm.flag[n] |= nRotat; k |= kFixV; hm.rot[n] = rotNormal; }
    if ((hm.flag[n] & nNoMoment) && !(k & kSetUV) && !(k & kOutlet) && !(k & kOpenBnd)) {
      k |= kAutoSlip | kFixU | kFixV;
    } else if ((k & kFixV) && (k & kAutoSlip)) {
      hm.flag[n] |= nRotat;
      hm.rot[n] = ((k & kOutlet) || (k & kOpenBnd)) ? rotNormal : rotTangent;
    }
  }
}

static int me_nodes(const HostModel& hm, int k, const int** nd)
{
  int v = hm.me[k];
  if (v >= 0) { *nd = &hm.elem[8LL * v]; return hm.nnd[v]; }
  *nd = &hm.bnd[3LL * (-v - 1)];
  return 3;
}

// EQS::SetEqno / LastEquation / ResetEqOrder (SetEqno.cpp:52-487): provisional numbers
// in node order without fixed dofs and dry nodes, then renumbering in order of last
// occurrence over the model element list.
void host_numbering(HostModel& hm)
{
  static const uint32_t fix[3] = { kFixU, kFixV, kSetS };
  const int np = hm.np;
  hm.eqno.assign(3LL * np, -1);
  int neq = 0;
  for (int n = 0; n < np; n++) {
    if (hm.nwd[n] & WD_DRY) continue;
    int ndf = (hm.flag[n] & nCorn) ? 3 : 2;
    for (int j = 0; j < ndf; j++) {
      if (hm.kind[n] & fix[j]) continue;
      hm.eqno[3LL * n + j] = neq++;
    }
  }
  const int nme = (int)hm.me.size();
  // the last model element that contains each node
  std::vector<int> last(np, -1);
  for (int k = 0; k < nme; k++) {
    const int* nd; int c = me_nodes(hm, k, &nd);
    for (int i = 0; i < c; i++) last[nd[i]] = k;
  }
  std::vector<int> ord(neq, -1);
  int nn = 0;
  for (int k = 0; k < nme; k++) {
    const int* nd; int c = me_nodes(hm, k, &nd);
    for (int i = 0; i < c; i++) {
      if (last[nd[i]] != k) continue;
      for (int j = 0; j < 3; j++) {
        int e = hm.eqno[3LL * nd[i] + j];
        if (e >= 0) ord[e] = nn++;
      }
    }
  }
  if (nn != neq) fail(RMX_ERR_ARG, "numbering: some equations never occur in an element");
  for (auto& v : hm.eqno) if (v >= 0) v = ord[v];
  hm.neq = neq;
}

// EQS::SetIndexMat (IndexMat.cpp:91-342): row = diagonal, then columns in order of first
// appearance over the model element list.  Each row only depends on the elements that
// contain its node, visited in list order, so rows are built independently.
void host_pattern(HostModel& hm)
{
  const int np = hm.np, nme = (int)hm.me.size(), neq = hm.neq;
  std::vector<int> ptr, list;
  node_incidence(np, nme, [&](int k, const int** nd) { return me_nodes(hm, k, nd); }, ptr, list);
  // owner node and dof of each equation
  std::vector<int> eq_node(neq), eq_dof(neq);
  for (int n = 0; n < np; n++)
    for (int j = 0; j < 3; j++) {
      int e = hm.eqno[3LL * n + j];
      if (e >= 0) { eq_node[e] = n; eq_dof[e] = j; }
    }
  const int MAXW = 128;
  std::vector<int> width(neq, 0);
  std::vector<int> tmp((size_t)neq * 0);
  // pass 1: widths; pass 2: fill.  Both passes use the same per-row routine.
  auto build_row = [&](int row, int* out) -> int {
    int w = 0;
    int buf[MAXW];
    int* ip = out ? out : buf;
    ip[w++] = row;
    int n = eq_node[row];
    for (int q = ptr[n]; q < ptr[n + 1]; q++) {
      const int* nd; int c = me_nodes(hm, list[q], &nd);
      for (int kk = 0; kk < c; kk++)
        for (int l = 0; l < 3; l++) {
          int col = hm.eqno[3LL * nd[kk] + l];
          if (col < 0) continue;
          int s;
          for (s = 0; s < w; s++) if (ip[s] == col) break;
          if (s == w) {
            if (w >= MAXW) fail(RMX_ERR_RANGE, "CSR row wider than 128 columns");
            ip[w++] = col;
          }
        }
    }
    return w;
  };
  parallel_for(neq, [&](long long a, long long b) {
    for (long long r = a; r < b; r++) width[r] = build_row((int)r, nullptr);
  });
  hm.rowptr.assign(neq + 1, 0);
  for (int r = 0; r < neq; r++) hm.rowptr[r + 1] = hm.rowptr[r] + width[r];
  hm.nnz = hm.rowptr[neq];
  hm.cols.assign(hm.nnz, 0);
  parallel_for(neq, [&](long long a, long long b) {
    for (long long r = a; r < b; r++) build_row((int)r, &hm.cols[hm.rowptr[r]]);
  });
}

// Greedy colouring so that no two elements of one launch share a node: region elements
// and boundary line elements are coloured separately (they are assembled in separate
// launches).  Launch groups are (colour, element type).
void host_colouring(const HostModel& hm, std::vector<int>& col_elems, std::vector<int>& col_ptr,
                    std::vector<int>& col_type)
{
  const int nme = (int)hm.me.size(), np = hm.np;
  std::vector<int> ptr, list;
  node_incidence(np, nme, [&](int k, const int** nd) { return me_nodes(hm, k, nd); }, ptr, list);
  std::vector<int> colour(nme, -1);
  int ncol[2] = { 0, 0 };
  for (int pass = 0; pass < 2; pass++) {          // 0 region, 1 boundary
    for (int k = 0; k < nme; k++) {
      bool isb = hm.me[k] < 0;
      if (isb != (pass == 1)) continue;
      const int* nd; int c = me_nodes(hm, k, &nd);
      uint64_t used[4] = { 0, 0, 0, 0 };
      for (int i = 0; i < c; i++)
        for (int q = ptr[nd[i]]; q < ptr[nd[i] + 1]; q++) {
          int o = list[q];
          if (o == k || colour[o] < 0 || ((hm.me[o] < 0) != isb)) continue;
          used[colour[o] >> 6] |= 1ull << (colour[o] & 63);
        }
      int col = 0;
      while (col < 256 && (used[col >> 6] >> (col & 63)) & 1) col++;
      if (col >= 256) fail(RMX_ERR_RANGE, "element colouring needs more than 256 colours");
      colour[k] = col;
      ncol[pass] = std::max(ncol[pass], col + 1);
    }
  }
  // group: region colours first (tri then quad inside a colour), then boundary colours
  auto type_of = [&](int k) { int v = hm.me[k]; return v < 0 ? ET_LINE : (hm.nnd[v] == 6 ? ET_TRI : ET_QUAD); };
  std::vector<std::vector<int>> groups;
  std::vector<int> gtype;
  for (int pass = 0; pass < 2; pass++)
    for (int cc = 0; cc < ncol[pass]; cc++)
      for (int t = 0; t < 3; t++) {
        groups.emplace_back();
        gtype.push_back(t);
      }
  auto gid = [&](int k) {
    int pass = hm.me[k] < 0 ? 1 : 0;
    int base = pass == 0 ? 0 : ncol[0] * 3;
    return base + colour[k] * 3 + type_of(k);
  };
  for (int k = 0; k < nme; k++) groups[gid(k)].push_back(k);
  col_elems.clear(); col_ptr.assign(1, 0); col_type.clear();
  for (size_t gI = 0; gI < groups.size(); gI++) {
    if (groups[gI].empty()) continue;
    col_elems.insert(col_elems.end(), groups[gI].begin(), groups[gI].end());
    col_ptr.push_back((int)col_elems.size());
    col_type.push_back(gtype[gI]);
  }
}

// Element-to-CSR slot map: for every model element, for every pair of local dofs
// (row ldof, column ldof) with ldof = i + j*nnd (U block, V block, corner H block), the
// position of the column in the CSR row, or -1 if either dof is not an equation.
void host_slots(const HostModel& hm, std::vector<long long>& off, std::vector<int>& slots,
                std::vector<int>& type)
{
  const int nme = (int)hm.me.size();
  off.assign(nme + 1, 0);
  type.assign(nme, 0);
  auto nl_of = [](int t) { return t == ET_TRI ? 15 : (t == ET_QUAD ? 20 : 8); };
  for (int k = 0; k < nme; k++) {
    int v = hm.me[k];
    int t = v < 0 ? ET_LINE : (hm.nnd[v] == 6 ? ET_TRI : ET_QUAD);
    type[k] = t;
    off[k + 1] = off[k] + (long long)nl_of(t) * nl_of(t);
  }
  slots.assign(off[nme], -1);
  parallel_for(nme, [&](long long a, long long b) {
    int eq[24];
    for (long long k = a; k < b; k++) {
      const int* nd; int c = me_nodes(hm, (int)k, &nd);
      int ncn = c == 3 ? 2 : (c == 6 ? 3 : 4);
      int nl = 2 * c + ncn;
      for (int q = 0; q < nl; q++) {
        int j = q / c, i = q % c;
        eq[q] = hm.eqno[3LL * nd[i] + j];
      }
      int* s = &slots[off[k]];
      for (int r = 0; r < nl; r++) {
        if (eq[r] < 0) continue;
        const int* ip = &hm.cols[hm.rowptr[eq[r]]];
        int w = hm.rowptr[eq[r] + 1] - hm.rowptr[eq[r]];
        for (int q = 0; q < nl; q++) {
          if (eq[q] < 0) continue;
          int pos = -1;
          for (int t = 0; t < w; t++) if (ip[t] == eq[q]) { pos = t; break; }
          if (pos < 0) fail(RMX_ERR_ARG, "slot map: column missing from pattern");
          s[r * nl + q] = hm.rowptr[eq[r]] + pos;
        }
      }
    }
  });
}

}  // namespace rmx
