// C ABI of the rmx library (declared in include/rmx.h).
#include <chrono>
#include <cmath>
#include <cstring>

#include "rmx_internal.h"

namespace rmx {

static thread_local std::string g_err;

[[noreturn]] void fail(rmx_status code, const std::string& msg) { throw Error{ code, msg }; }

void ensure_work(Ctx& c);

}  // namespace rmx

using namespace rmx;

#define API_BEGIN try {
#define API_END                                                                          \
  }                                                                                      \
  catch (const rmx::Error& e) {                                                          \
    rmx::g_err = e.msg;                                                                  \
    return e.code;                                                                       \
  }                                                                                      \
  catch (const std::exception& e) {                                                      \
    rmx::g_err = e.what();                                                               \
    return RMX_ERR_ARG;                                                                  \
  }                                                                                      \
  return RMX_OK;

Ctx::~Ctx()
{
  for (hipEvent_t e : ktm.ev) (void)hipEventDestroy(e);
}

int rmx::kt_begin(Ctx& c, int cls, double bytes)
{
  KTimer& T = c.ktm;
  if (!T.on) return -1;
  if (2 * (T.used + 1) > T.ev.size()) {          // pool full: wait, read back, reuse
    HIPCHECK(hipStreamSynchronize(c.stream));
    kt_collect(c);
  }
  const size_t k = T.used++;
  T.cls[k] = cls;
  T.pbytes[k] = bytes;
  HIPCHECK(hipEventRecord(T.ev[2 * k], c.stream));
  return (int)k;
}

void rmx::kt_end(Ctx& c, int k)
{
  if (k < 0) return;
  HIPCHECK(hipEventRecord(c.ktm.ev[2 * (size_t)k + 1], c.stream));
}

void rmx::kt_collect(Ctx& c)
{
  KTimer& T = c.ktm;
  for (size_t k = 0; k < T.used; k++) {
    float ms = 0.0f;
    HIPCHECK(hipEventSynchronize(T.ev[2 * k + 1]));
    HIPCHECK(hipEventElapsedTime(&ms, T.ev[2 * k], T.ev[2 * k + 1]));
    T.sec[T.cls[k]] += ms * 1e-3;
    T.bytes[T.cls[k]] += T.pbytes[k];
    T.cnt[T.cls[k]]++;
  }
  T.used = 0;
}

static Ctx* need(rmx_ctx* p)
{
  if (!p) fail(RMX_ERR_ARG, "null context");
  Ctx* c = reinterpret_cast<Ctx*>(p);
  HIPCHECK(hipSetDevice(c->device));
  return c;
}

extern "C" {

const char* rmx_last_error(void) { return rmx::g_err.c_str(); }

rmx_status rmx_create(int device, rmx_ctx** out)
{
  API_BEGIN
  if (!out) fail(RMX_ERR_ARG, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0)
    fail(RMX_ERR_NODEVICE, "no HIP device " + std::to_string(device));
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    fail(RMX_ERR_NODEVICE, std::string("device is ") + prop.gcnArchName + ", not gfx950");
  HIPCHECK(hipSetDevice(device));
  noalloc_setup(device);                // scratch address of rmx_scale_l2_noalloc on this GPU
  Ctx* c = new Ctx();
  c->device = device;
  HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = reinterpret_cast<rmx_ctx*>(c);
  API_END
}

rmx_status rmx_destroy(rmx_ctx* p)
{
  API_BEGIN
  if (!p) return RMX_OK;
  Ctx* c = need(p);
  (void)hipStreamSynchronize(c->stream);
  if (c->h_red) (void)hipHostFree(c->h_red);
  hipStream_t s = c->stream;
  delete c;
  (void)hipStreamDestroy(s);
  API_END
}

// The model element list of the assembly: nodes, type, eddy viscosity.
static void upload_model_elements(Ctx& c)
{
  const HostModel& hm = c.hm;
  const int nme = (int)hm.me.size();
  std::vector<int> men(8LL * nme, -1), met(nme);
  std::vector<double> mvt(nme, 0.0);
  for (int k = 0; k < nme; k++) {
    const int v = hm.me[k];
    if (v >= 0) {
      for (int i = 0; i < 8; i++) men[8LL * k + i] = hm.elem[8LL * v + i];
      met[k] = hm.nnd[v] == 6 ? ET_TRI : ET_QUAD;
      mvt[k] = hm.vt[v];
    } else {
      const int b = -v - 1;
      for (int i = 0; i < 3; i++) men[8LL * k + i] = hm.bnd[3LL * b + i];
      met[k] = ET_LINE;
    }
  }
  c.nme = nme;
  c.me_nodes.upload(men, c.stream); c.me_type.upload(met, c.stream); c.me_vt.upload(mvt, c.stream);
}

rmx_status rmx_mesh_upload(rmx_ctx* p, const rmx_mesh_desc* d)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!d || d->np <= 0 || d->ne <= 0 || !d->x || !d->y || !d->z || !d->elems || !d->nnd ||
      !d->manning || !d->vt)
    fail(RMX_ERR_ARG, "mesh: missing arrays");
  if ((d->n_inlet && !d->inlet_segs) || (d->n_outlet && !d->outlet_segs))
    fail(RMX_ERR_ARG, "mesh: missing control-line segments");
  host_mesh(c.hm, *d);
  const HostModel& hm = c.hm;
  {
    unsigned long long h = 1469598103934665603ULL;            // FNV-1a over the topology
    auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ULL; };
    mix((unsigned long long)hm.np); mix((unsigned long long)hm.ne);
    for (int v : hm.elem) mix((unsigned long long)(unsigned)v);
    for (int v : hm.in_segs) mix((unsigned long long)(unsigned)v);
    for (int v : hm.out_segs) mix((unsigned long long)(unsigned)v);
    c.topo_hash = h;
  }
  const int np = hm.np;
  hipStream_t s = c.stream;
  c.x.upload(hm.x, s); c.y.upload(hm.y, s); c.zor.upload(hm.zor, s); c.z.upload(hm.zor, s);
  c.U.alloc(np); c.V.alloc(np); c.S.upload(hm.zor, s); c.cf.alloc(np);
  HIPCHECK(hipMemsetAsync(c.U.p, 0, sizeof(double) * np, s));
  HIPCHECK(hipMemsetAsync(c.V.p, 0, sizeof(double) * np, s));
  c.d_elem.upload(hm.elem, s); c.d_nnd.upload(hm.nnd, s);
  c.d_manning.upload(hm.manning, s); c.d_areaFact.upload(hm.areaFact, s);
  c.d_veg_dp.upload(hm.veg_dp, s); c.d_veg_sp.upload(hm.veg_sp, s);
  for (DBuf<double>* b : {&c.dUdt, &c.dVdt, &c.dSdt, &c.Uo, &c.Vo, &c.So, &c.dUdto, &c.dVdto,
                          &c.dSdto})
    b->alloc(np);
  dev_time_init(c);
  c.unsteady = false;
  cycle_start(c);
  c.d_ste.upload(hm.ste, s); c.vtn.alloc(np); c.uun.alloc(np);
  c.d_vte.upload(hm.vt, s); c.d_exi.upload(hm.exi, s); c.d_eyi.upload(hm.eyi, s);
  c.exx.alloc(np); c.exy.alloc(np); c.eyy.alloc(np);
  for (DBuf<double>* b : {&c.exx, &c.exy, &c.eyy})
    HIPCHECK(hipMemsetAsync(b->p, 0, sizeof(double) * np, s));
  HIPCHECK(hipMemsetAsync(c.vtn.p, 0, sizeof(double) * np, s));
  HIPCHECK(hipMemsetAsync(c.uun.p, 0, sizeof(double) * np, s));
  c.turb_stale = true;
  // node -> region elements (ascending) for the friction gather
  std::vector<int> ptr(np + 1, 0), list;
  for (int e = 0; e < hm.ne; e++)
    for (int i = 0; i < hm.nnd[e]; i++) ptr[hm.elem[8LL * e + i] + 1]++;
  for (int i = 0; i < np; i++) ptr[i + 1] += ptr[i];
  list.resize(ptr[np]);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < hm.ne; e++)
    for (int i = 0; i < hm.nnd[e]; i++) list[fill[hm.elem[8LL * e + i]]++] = e;
  c.n2e_ptr.upload(ptr, s); c.n2e.upload(list, s);
  upload_model_elements(c);
  c.d_in_segs.upload(hm.in_segs, s);
  // drying and rewetting: every node and element wet, no count down
  c.d_nwd.upload(hm.nwd, s); c.d_ewd.upload(hm.ewd, s);
  c.d_cdown.alloc(np);
  HIPCHECK(hipMemsetAsync(c.d_cdown.p, 0, sizeof(int) * np, s));
  c.first_dry = true;
  c.wet_sets = 0;
  c.mesh_ready = true;
  c.prepared = false;
  c.csr_only = false;
  HIPCHECK(hipStreamSynchronize(s));
  API_END
}

rmx_status rmx_set_params(rmx_ctx* p, const rmx_params* q)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!q) fail(RMX_ERR_ARG, "null params");
  if (q->relax_method != 0 && (q->relax_method < 2 || q->relax_method > 4))
    fail(RMX_ERR_ARG, "params: relaxation method " + std::to_string(q->relax_method) +
                      " not supported (0 none, 2 cap by max_delta_uv / max_delta_s, 3 and 4 "
                      "time-relaxed)");
  if (q->relax_method >= 3 && !(q->interval > 0.0) && !c.unsteady)
    fail(RMX_ERR_ARG, "params: relaxation methods 3 and 4 of stationary flow need the time "
                      "interval (> 0)");
  if (!(q->dry_limit > 0.0))
    fail(RMX_ERR_ARG, "params: dry_limit must be positive (the minimum depth is dry_limit / 2)");
  c.par.relax_method = q->relax_method; c.par.dry_method = q->dry_method;
  c.par.dry_limit = q->dry_limit; c.par.max_delta_uv = q->max_delta_uv;
  c.par.max_delta_s = q->max_delta_s; c.par.conv_uv = q->conv_uv; c.par.conv_s = q->conv_s;
  c.par.max_us = q->max_us;
  {
    const int t = q->turb == 0 ? 1 : q->turb;
    const int model = t % 10, aniso = (t / 10) % 10, hund = (t / 100) % 10;
    if (t < 0 || t >= 1000 || (model != 1 && model != 2) || aniso > 1 || hund > 3)
      fail(RMX_ERR_ARG, "params: turbulence model " + std::to_string(q->turb) +
                        " not supported (1 constant, 2 algebraic, + 10 anisotropic (Elder), "
                        "+ 100 kVtMin, + 200 kVtIterat)");
    c.par.turb = t;
  }
  c.par.cm = q->cm > 0.0 ? q->cm : 0.09;
  c.par.cd = q->cd > 0.0 ? q->cd : 1.00;
  c.par.relax_time = q->relax_time;
  c.par.interval = q->interval;
  if (q->count_down < 0) fail(RMX_ERR_ARG, "params: count_down must not be negative");
  c.par.rewet_limit = q->rewet_limit;
  c.par.count_down = q->count_down;
  cycle_start(c);
  c.turb_stale = true;
  API_END
}

rmx_status rmx_state_upload(rmx_ctx* p, const double* U, const double* V, const double* S)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "state: upload a mesh first");
  const int np = c.hm.np;
  if (U) c.U.upload(U, np, c.stream);
  if (V) c.V.upload(V, np, c.stream);
  if (S) c.S.upload(S, np, c.stream);
  dev_time_init(c);
  c.turb_stale = true;
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_state_download(rmx_ctx* p, double* U, double* V, double* S, double* z)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "state: no mesh");
  const int np = c.hm.np;
  if (U) c.U.download(U, np, c.stream);
  if (V) c.V.download(V, np, c.stream);
  if (S) c.S.download(S, np, c.stream);
  if (z) c.z.download(z, np, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

static void first_prepare(Ctx& c)
{
  HostModel& hm = c.hm;
  host_boundary_conditions(hm);
  host_numbering(hm);
  host_pattern(hm);
  hipStream_t s = c.stream;
  const int np = hm.np;
  // device node BC data
  std::vector<uint32_t> df(np, 0);
  std::vector<double> rotm(4LL * np, 0.0);
  for (int n = 0; n < np; n++) {
    if (hm.flag[n] & nRotat) df[n] |= dRotat;
    if (hm.kind[n] & kInlet) df[n] |= dInlet;
    if (hm.kind[n] & kOutlet) df[n] |= dOutlet;
    double* R = &rotm[4LL * n];
    if (hm.rot[n] == rotTangent) { R[0] = hm.ny[n]; R[1] = hm.nx[n]; R[2] = -hm.nx[n]; R[3] = hm.ny[n]; }
    else if (hm.rot[n] == rotNormal) { R[0] = hm.niox[n]; R[1] = hm.nioy[n]; R[2] = hm.nioy[n]; R[3] = -hm.niox[n]; }
  }
  c.dflag.upload(df, s); c.rotm.upload(rotm, s);
  c.niox.upload(hm.niox, s); c.nioy.upload(hm.nioy, s); c.valS.upload(hm.valS, s);
  c.valU.alloc(np);
  HIPCHECK(hipMemsetAsync(c.valU.p, 0, sizeof(double) * np, s));
  c.d_inlet_nodes.upload(hm.inlet_nodes, s);
  c.eqno.upload(hm.eqno, s);
  // CSR pattern
  c.n = hm.neq;
  c.nnz = hm.nnz;
  c.rowptr.upload(hm.rowptr, s);
  c.cols.upload(hm.cols, s);
  c.B.alloc(c.n); c.X.alloc(c.n);
  // assembly schedule
  std::vector<int> ce, cp, ct;
  host_colouring(hm, ce, cp, ct);
  c.col_ptr = cp; c.col_type = ct;
  c.col_elems.upload(ce, s);
  std::vector<long long> off;
  std::vector<int> slots, type;
  host_slots(hm, off, slots, type);
  c.me_slot_off.upload(off, s);
  c.slots.upload(slots, s);
  HIPCHECK(hipStreamSynchronize(s));
  ensure_work(c);
  spmv_setup(c, hm.rowptr.data(), hm.cols.data());
  setup_ilu_schedule(c);
  c.have32 = c.have64 = false;
  c.prepared = true;
}

rmx_status rmx_prepare(rmx_ctx* p, int* neq, long long* nnz)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "prepare: upload a mesh first");
  if (!c.prepared) first_prepare(c);
  dev_prepare_step(c);
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (neq) *neq = c.n;
  if (nnz) *nnz = c.nnz;
  API_END
}

rmx_status rmx_dry_rewet(rmx_ctx* p, int* dried, int* wetted)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready || c.csr_only) fail(RMX_ERR_ARG, "drying and rewetting: upload a mesh first");
  if (c.par.dry_method != 2)
    fail(RMX_ERR_ARG, "drying and rewetting: params.dry_method " +
                      std::to_string(c.par.dry_method) + " (only method 2 is supported)");
  dev_dry_rewet(c);
  int del = 0, wel = 0;
  for (uint8_t f : c.hm.ewd) {
    if ((f & EL_DRY) && !(f & EL_WASDRY)) del++;
    if (!(f & EL_DRY) && (f & EL_WASDRY)) wel++;
  }
  if (del + wel) {
    // the wet model set up again: boundary line elements on the wet/dry line, boundary
    // conditions and rotations, numbering without the dry nodes, pattern, assembly schedule,
    // ILU plans and coarse space; then the slip condition at the rotated nodes
    host_topology(c.hm);
    upload_model_elements(c);
    c.wet_sets++;
    c.topo_hash = (c.topo_hash ^ (unsigned long long)c.wet_sets) * 1099511628211ULL;
    first_prepare(c);
    dev_slip_flow(c);
  }
  c.turb_stale = true;                       // eddy viscosity due at the next Newton step
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (dried) *dried = del;
  if (wetted) *wetted = wel;
  API_END
}

rmx_status rmx_get_wetdry(rmx_ctx* p, int* nflag, int* eflag)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "wet/dry flags: upload a mesh first");
  std::vector<uint8_t> nw(c.hm.np);
  c.d_nwd.download(nw.data(), nw.size(), c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (nflag)
    for (int n = 0; n < c.hm.np; n++)
      nflag[n] = ((nw[n] & WD_DRY) ? 64 : 0) | ((nw[n] & WD_DRYPREV) ? 128 : 0) |
                 ((nw[n] & WD_MARSH) ? 256 : 0) | ((nw[n] & WD_MARSHPREV) ? 512 : 0);
  if (eflag)
    for (int e = 0; e < c.hm.ne; e++) eflag[e] = c.hm.ewd[e] & (EL_DRY | EL_MARSH);
  API_END
}

rmx_status rmx_get_relax_time(rmx_ctx* p, double* dt_uv, double* dt_h, double* dt)
{
  API_BEGIN
  Ctx& c = *need(p);
  // below method 3 the relaxed time steps are the time step itself (EqsUVS2D.cpp:115-143)
  const double t = c.unsteady ? c.dt : c.par.interval;
  const bool relaxed = c.par.relax_method >= 3;
  if (dt_uv) *dt_uv = relaxed ? c.dt_uv : t;
  if (dt_h) *dt_h = relaxed ? c.dt_h : t;
  if (dt) *dt = t;
  API_END
}

rmx_status rmx_turbulence(rmx_ctx* p, double* vt, double* uu)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "turbulence: upload a mesh first");
  if (!c.par.turb_alg() && !c.par.turb_aniso())
    fail(RMX_ERR_ARG, "turbulence: params.turb selects no nodal eddy viscosity model");
  dev_turbulence(c);
  const int np = c.hm.np;
  if (vt) HIPCHECK(hipMemcpyAsync(vt, c.vtn.p, sizeof(double) * np, hipMemcpyDeviceToHost, c.stream));
  if (uu) HIPCHECK(hipMemcpyAsync(uu, c.uun.p, sizeof(double) * np, hipMemcpyDeviceToHost, c.stream));
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_time_step(rmx_ctx* p, double dt, double theta, int first)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "time step: upload a mesh first");
  c.turb_stale = true;
  if (dt <= 0.0) {
    if (c.par.relax_method >= 3 && !(c.par.interval > 0.0))
      fail(RMX_ERR_ARG, "time step: relaxation methods 3 and 4 of stationary flow need the "
                        "time interval (params.interval > 0)");
    c.unsteady = false;
    cycle_start(c);
  } else {
    if (std::fabs(theta) < 1.0e-10 || std::fabs(dt) < 1.0e-10)
      fail(RMX_ERR_ARG, "time step: theta or time interval too small");
    c.unsteady = true; c.dt = dt; c.theta = theta;
    cycle_start(c);
    dev_time_step(c, first ? 1 : 0);
    HIPCHECK(hipStreamSynchronize(c.stream));
  }
  API_END
}

rmx_status rmx_set_inflow(rmx_ctx* p, double Q, double S_in, int hydrograph)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "inflow: upload a mesh first");
  if (!std::isfinite(Q) || !std::isfinite(S_in)) fail(RMX_ERR_ARG, "inflow: not finite");
  c.hm.Q = Q;
  c.hm.S_in = S_in;
  c.hm.inflow_qt = hydrograph != 0;
  if (c.prepared && !c.csr_only) {                 // prescribed level and kind of the inlet nodes
    for (int n : c.hm.inlet_nodes) {
      c.hm.valS[n] = S_in;
      c.hm.kind[n] = (c.hm.kind[n] & ~(uint32_t)kQTInlet) | (hydrograph ? kQTInlet : 0u);
    }
    c.valS.upload(c.hm.valS, c.stream);
    HIPCHECK(hipStreamSynchronize(c.stream));
  }
  API_END
}

rmx_status rmx_hydrograph(int n, const double* t, const double* Q, const double* S, double time,
                          double* Qout, double* Sout)
{
  API_BEGIN
  if (n < 1 || !t || !Q || !S || !Qout || !Sout) fail(RMX_ERR_ARG, "hydrograph: bad arguments");
  const double eps = 1.0e-20;                      // kEpsilon (Defs.h:85) of the interval test
  double q = Q[0], s = S[0];
  for (int i = 0; i + 1 < n; i++) {
    double dx = t[i + 1] - t[i];
    double lx = time - t[i];
    if (std::fabs(dx) > eps) lx /= dx;
    else lx = dx = 0.0;
    // with no second coordinate the interval test is on lx alone
    if (std::fabs(dx) < eps && std::fabs(lx) < eps) {
      q = Q[i]; s = S[i];                          // degenerate interval at its own time
      break;
    }
    if (lx > -0.001 && lx < 1.001) {
      q = Q[i] + lx * (Q[i + 1] - Q[i]);
      s = S[i] + lx * (S[i + 1] - S[i]);
      break;
    }
  }
  *Qout = q;
  *Sout = s;
  API_END
}

rmx_status rmx_slip_flow(rmx_ctx* p)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "slip flow: call rmx_prepare first");
  dev_slip_flow(c);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_time_commit(rmx_ctx* p)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "time commit: upload a mesh first");
  dev_time_commit(c);
  dev_wetdry_commit(c);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_time_state(rmx_ctx* p, int set, double* const* arrays)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "time state: upload a mesh first");
  if (!arrays) fail(RMX_ERR_ARG, "time state: null array list");
  DBuf<double>* b[9] = {&c.Uo, &c.Vo, &c.So, &c.dUdto, &c.dVdto, &c.dSdto,
                        &c.dUdt, &c.dVdt, &c.dSdt};
  const size_t np = (size_t)c.hm.np;
  for (int k = 0; k < 9; k++) {
    if (!arrays[k]) continue;
    if (set) HIPCHECK(hipMemcpyAsync(b[k]->p, arrays[k], sizeof(double) * np,
                                     hipMemcpyHostToDevice, c.stream));
    else HIPCHECK(hipMemcpyAsync(arrays[k], b[k]->p, sizeof(double) * np,
                                 hipMemcpyDeviceToHost, c.stream));
  }
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_get_numbering(rmx_ctx* p, int* eqno)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "numbering: call rmx_prepare first");
  std::memcpy(eqno, c.hm.eqno.data(), sizeof(int) * c.hm.eqno.size());
  API_END
}

rmx_status rmx_get_pattern(rmx_ctx* p, int* rowptr, int* cols)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "pattern: call rmx_prepare or rmx_csr_upload first");
  c.rowptr.download(rowptr, c.n + 1, c.stream);
  c.cols.download(cols, c.nnz, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_get_nodal(rmx_ctx* p, double* cf, double* bcU, int* kind, double* rotm)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "nodal: call rmx_prepare first");
  const int np = c.hm.np;
  if (cf) c.cf.download(cf, np, c.stream);
  if (bcU) c.valU.download(bcU, np, c.stream);
  if (rotm) c.rotm.download(rotm, 4LL * np, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  if (kind) for (int n = 0; n < np; n++) kind[n] = (int)c.hm.kind[n];
  API_END
}

rmx_status rmx_n_model_elems(rmx_ctx* p, int* n)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "no mesh");
  *n = c.nme;
  API_END
}

rmx_status rmx_element_coefs(rmx_ctx* p, int first, int count, double* K, double* F)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "coefs: call rmx_prepare first");
  if (first < 0 || count <= 0 || first + count > c.nme) fail(RMX_ERR_ARG, "coefs: range");
  dev_element_coefs(c, first, count, K, F);
  API_END
}

rmx_status rmx_assemble(rmx_ctx* p, int prec)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "assemble: call rmx_prepare first");
  dev_assemble(c, prec == RMX_F32 ? RMX_F32 : RMX_F64, true);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_scale_l2(rmx_ctx* p, double* scale)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.have64 && !c.have32) fail(RMX_ERR_ARG, "scale: no matrix");
  const double s = dev_scale_l2(c);
  if (scale) *scale = s;
  c.ilu.fac32 = c.ilu.fac64 = c.silu.fac32 = c.silu.fac64 = false;
  API_END
}

rmx_status rmx_scale_l2_noalloc(int device, int n, const int* rowptr, int prec, void* A,
                                double* B, double* scale, void* stream)
{
  API_BEGIN
  if (n < 1 || !rowptr || !A || !B || !scale)
    fail(RMX_ERR_ARG, "scale (no allocation): n < 1 or a null array");
  if (prec != RMX_F32 && prec != RMX_F64)
    fail(RMX_ERR_ARG, "scale (no allocation): prec must be RMX_F32 or RMX_F64");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    fail(RMX_ERR_NODEVICE, "no HIP device " + std::to_string(device));
  HIPCHECK(hipSetDevice(device));
  dev_scale_l2_noalloc(device, n, rowptr, prec, A, B, scale, static_cast<hipStream_t>(stream));
  API_END
}

rmx_status rmx_convert_matrix(rmx_ctx* p, int to_prec)
{
  API_BEGIN
  Ctx& c = *need(p);
  dev_convert(c, to_prec == RMX_F32 ? RMX_F32 : RMX_F64);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_ilu0_factor(rmx_ctx* p, int prec)
{
  API_BEGIN
  Ctx& c = *need(p);
  dev_ilu0_factor(c, prec == RMX_F32 ? RMX_F32 : RMX_F64);
  API_END
}

rmx_status rmx_silu0_factor(rmx_ctx* p, int prec)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  ilu_factor(c, c.silu, prec == RMX_F32 ? RMX_F32 : RMX_F64);
  API_END
}

rmx_status rmx_silu0_apply(rmx_ctx* p, int prec, const double* b, double* x)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  c.w[0].upload(b, c.n, c.stream);
  ilu_apply(c, c.silu, prec == RMX_F32 ? RMX_F32 : RMX_F64, c.w[0].p, c.w[1].p, nullptr);
  c.w[1].download(x, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_set_strips(rmx_ctx* p, int strips, int segments)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (strips < 0 || segments < 0 || segments > 1024)
    fail(RMX_ERR_ARG, "strips >= 0, 0 <= segments <= 1024 (0: automatic)");
  c.strips = strips;
  c.segments = segments;
  if (c.prepared) setup_ilu_schedule(c);
  API_END
}

rmx_status rmx_set_coarse(rmx_ctx* p, int segments)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (segments < -1 || segments > 1024)
    fail(RMX_ERR_ARG, "coarse segments must be -1 (automatic) or 0..1024");
  c.coarse_segments = segments;
  if (c.prepared) setup_ilu_schedule(c);
  API_END
}

rmx_status rmx_coarse_factor(rmx_ctx* p, int prec)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  coarse_factor(c, prec == RMX_F32 ? RMX_F32 : RMX_F64);
  API_END
}

rmx_status rmx_get_aggregates(rmx_ctx* p, int* agg, int* nc)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || !c.coarse.ready) fail(RMX_ERR_ARG, "no coarse space");
  if (!agg || !nc) fail(RMX_ERR_ARG, "null output");
  c.coarse.agg.download(agg, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  *nc = c.coarse.nc;
  API_END
}

rmx_status rmx_precond_apply(rmx_ctx* p, int kind, int prec, const double* b, double* x)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  if (!b || !x) fail(RMX_ERR_ARG, "null vector");
  const int pr = prec == RMX_F32 ? RMX_F32 : RMX_F64;
  c.w[0].upload(b, c.n, c.stream);
  dev_precond(c, pr, kind, c.w[0].p, c.w[1].p);
  c.w[1].download(x, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_mpnn_n_params(int* n)
{
  API_BEGIN
  if (!n) fail(RMX_ERR_ARG, "null output");
  *n = 51282;
  API_END
}

rmx_status rmx_mpnn_set_weights(rmx_ctx* p, const float* theta, int n)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!theta) fail(RMX_ERR_ARG, "null weights");
  mpnn_set_weights(c, theta, n);
  API_END
}

rmx_status rmx_mpnn_get_weights(rmx_ctx* p, float* theta, int n)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!theta) fail(RMX_ERR_ARG, "null weights");
  mpnn_get_weights(c, theta, n);
  API_END
}

rmx_status rmx_mpnn_forward(rmx_ctx* p, const double* r, double* z, float* feat)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!r || !z) fail(RMX_ERR_ARG, "null vector");
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  c.w[0].upload(r, c.n, c.stream);
  mpnn_forward(c, c.w[0].p, c.w[1].p, feat);
  c.w[1].download(z, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_mpnn_add_sample(rmx_ctx* p, const double* r, int capacity, int* count)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!r) fail(RMX_ERR_ARG, "null vector");
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  c.w[0].upload(r, c.n, c.stream);
  const int k = mpnn_add_sample(c, c.w[0].p, capacity);
  if (count) *count = k;
  API_END
}

rmx_status rmx_mpnn_loss_grad(rmx_ctx* p, int sample, double* loss, float* grad)
{
  API_BEGIN
  Ctx& c = *need(p);
  const double l = mpnn_loss_grad(c, sample, grad);
  if (loss) *loss = l;
  API_END
}

rmx_status rmx_mpnn_train(rmx_ctx* p, int epochs, double lr, double* loss)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (epochs < 1 || !(lr > 0.0)) fail(RMX_ERR_ARG, "epochs >= 1, lr > 0");
  const double l = mpnn_train(c, epochs, lr);
  if (loss) *loss = l;
  API_END
}

rmx_status rmx_get_sweep_info(rmx_ctx* p, int kind, rmx_sweep_info* out)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!out) fail(RMX_ERR_ARG, "null output");
  if (kind != RMX_PREC_ILU0 && kind != RMX_PREC_SILU0) fail(RMX_ERR_ARG, "kind must be an ILU(0)");
  const Ilu& I = ilu_of(c, kind);
  if (!c.prepared || !I.ready) fail(RMX_ERR_ARG, "no system");
  out->lower_entries = I.lo.ne; out->upper_entries = I.up.ne;
  out->lower_levels = I.lo.nlev_max; out->upper_levels = I.up.nlev_max;
  out->strips = I.nblk; out->segments = I.nseg;
  API_END
}

rmx_status rmx_get_blocks(rmx_ctx* p, int* block)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || !c.silu.ready) fail(RMX_ERR_ARG, "no system");
  if (!block) fail(RMX_ERR_ARG, "null output");
  const Ilu& I = c.silu;
  for (int b = 0; b < I.nblk; b++)
    for (int i = I.bstart[b]; i < I.bstart[b + 1]; i++)
      block[i] = b * I.nseg + (I.seg.empty() ? 0 : I.seg[i]);
  API_END
}

rmx_status rmx_ilu0_apply(rmx_ctx* p, int prec, const double* b, double* x)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  c.w[0].upload(b, c.n, c.stream);
  dev_ilu0_apply(c, prec == RMX_F32 ? RMX_F32 : RMX_F64, c.w[0].p, c.w[1].p);
  c.w[1].download(x, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_spmv(rmx_ctx* p, int prec, const double* x, double* r)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  const int pr = prec == RMX_F32 ? RMX_F32 : RMX_F64;
  if ((pr == RMX_F32 && !c.have32) || (pr == RMX_F64 && !c.have64)) fail(RMX_ERR_ARG, "spmv: no matrix");
  c.w[0].upload(x, c.n, c.stream);
  dev_spmv(c, pr, c.w[0].p, c.w[1].p);
  c.w[1].download(r, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_solve(rmx_ctx* p, const rmx_krylov_opts* o, rmx_krylov_out* out)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!o || !out) fail(RMX_ERR_ARG, "solve: null argument");
  if (!c.prepared) fail(RMX_ERR_ARG, "solve: no system");
  if (o->method == RMX_FGMRES && (o->mkyrl < 1 || o->mkyrl > 30))
    fail(RMX_ERR_ARG, "solve: FGMRES restart length must be 1..30");
  dev_solve(c, *o, *out);
  API_END
}

rmx_status rmx_get_matrix(rmx_ctx* p, int prec, void* vals, double* B)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (prec == RMX_F32) {
    if (!c.have32) fail(RMX_ERR_ARG, "no FP32 matrix");
    if (vals) c.A32.download(static_cast<float*>(vals), c.nnz, c.stream);
  } else {
    if (!c.have64) fail(RMX_ERR_ARG, "no FP64 matrix");
    if (vals) c.A64.download(static_cast<double*>(vals), c.nnz, c.stream);
  }
  if (B) c.B.download(B, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_get_factors(rmx_ctx* p, int prec, void* vals)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (prec == RMX_F32) {
    if (!c.ilu.fac32) fail(RMX_ERR_ARG, "no FP32 factors");
    c.ilu.LU32.download(static_cast<float*>(vals), c.nnz, c.stream);
  } else {
    if (!c.ilu.fac64) fail(RMX_ERR_ARG, "no FP64 factors");
    c.ilu.LU64.download(static_cast<double*>(vals), c.nnz, c.stream);
  }
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_get_solution(rmx_ctx* p, double* X)
{
  API_BEGIN
  Ctx& c = *need(p);
  c.X.download(X, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_set_solution(rmx_ctx* p, const double* X)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  if (!X) fail(RMX_ERR_ARG, "null solution");
  c.X.upload(X, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_set_rhs(rmx_ctx* p, const double* B)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared) fail(RMX_ERR_ARG, "no system");
  c.B.upload(B, c.n, c.stream);
  HIPCHECK(hipStreamSynchronize(c.stream));
  API_END
}

rmx_status rmx_csr_upload(rmx_ctx* p, int n, const int* rowptr, const int* cols,
                          const double* vals, const double* B)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (n <= 0 || !rowptr || !cols || !vals || !B) fail(RMX_ERR_ARG, "csr: missing arrays");
  if (rowptr[0] != 0) fail(RMX_ERR_ARG, "csr: rowptr[0] must be 0");
  for (int i = 0; i < n; i++) {
    if (rowptr[i + 1] <= rowptr[i]) fail(RMX_ERR_ARG, "csr: empty row");
    if (cols[rowptr[i]] != i) fail(RMX_ERR_ARG, "csr: the diagonal must come first in each row");
    for (int q = rowptr[i]; q < rowptr[i + 1]; q++)
      if (cols[q] < 0 || cols[q] >= n) fail(RMX_ERR_ARG, "csr: column out of range");
  }
  c.csr_only = true;
  c.mesh_ready = false;
  c.n = n;
  c.nnz = rowptr[n];
  c.rowptr.upload(rowptr, n + 1, c.stream);
  c.cols.upload(cols, c.nnz, c.stream);
  c.A64.upload(vals, c.nnz, c.stream);
  c.B.upload(B, n, c.stream);
  c.X.alloc(n);
  HIPCHECK(hipStreamSynchronize(c.stream));
  ensure_work(c);
  spmv_setup(c, rowptr, cols);
  setup_ilu_schedule(c);
  c.have64 = true; c.have32 = false;
  c.mat_version++;
  c.prepared = true;
  API_END
}

rmx_status rmx_newton_update(rmx_ctx* p, rmx_step_stats* st)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!c.prepared || c.csr_only) fail(RMX_ERR_ARG, "update: call rmx_prepare first");
  rmx_step_stats tmp{};
  dev_newton_update(c, st ? *st : tmp);
  API_END
}

rmx_status rmx_newton_step(rmx_ctx* p, const rmx_krylov_opts* o, rmx_step_stats* st)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!o || !st) fail(RMX_ERR_ARG, "newton: null argument");
  if (!c.mesh_ready) fail(RMX_ERR_ARG, "newton: upload a mesh first");
  *st = rmx_step_stats{};
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  auto t0 = now();
  if (!c.prepared) first_prepare(c);
  dev_prepare_step(c);
  HIPCHECK(hipStreamSynchronize(c.stream));
  auto t1 = now();
  dev_assemble(c, RMX_F64, true);
  st->scale = dev_scale_l2(c);
  auto t2 = now();
  const int prec = o->prec == RMX_F32 ? RMX_F32 : RMX_F64;
  if (prec == RMX_F32) dev_convert(c, RMX_F32);
  const int pk = o->precond;
  if (pk == RMX_PREC_ILU0 || pk == RMX_PREC_SILU0 || pk == RMX_PREC_SILU0C || pk == RMX_PREC_MPNN)
    ilu_factor(c, ilu_of(c, pk), prec);
  if (pk == RMX_PREC_SILU0C || pk == RMX_PREC_MPNN) coarse_factor(c, prec);
  HIPCHECK(hipStreamSynchronize(c.stream));
  auto t3 = now();
  dev_solve(c, *o, st->solve);
  auto t4 = now();
  dev_newton_update(c, *st);
  auto t5 = now();
  st->t_prepare = sec(t0, t1); st->t_assemble = sec(t1, t2); st->t_factor = sec(t2, t3);
  st->t_solve = sec(t3, t4); st->t_update = sec(t4, t5);
  API_END
}

rmx_status rmx_kernel_stats_enable(rmx_ctx* p, int on)
{
  API_BEGIN
  Ctx& c = *need(p);
  KTimer& T = c.ktm;
  HIPCHECK(hipStreamSynchronize(c.stream));
  kt_collect(c);
  if (on)
    for (int k = 0; k < KT_N; k++) { T.sec[k] = 0.0; T.bytes[k] = 0.0; T.cnt[k] = 0; }
  if (on && T.ev.empty()) {
    const size_t pairs = 8192;
    T.ev.resize(2 * pairs);
    for (auto& e : T.ev) HIPCHECK(hipEventCreate(&e));
    T.cls.resize(pairs);
    T.pbytes.resize(pairs);
  }
  T.on = on != 0;
  API_END
}

rmx_status rmx_kernel_stats_get(rmx_ctx* p, rmx_kernel_stats* out)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!out) fail(RMX_ERR_ARG, "null output");
  HIPCHECK(hipStreamSynchronize(c.stream));
  kt_collect(c);
  for (int k = 0; k < KT_N && k < 8; k++) {
    out->seconds[k] = c.ktm.sec[k];
    out->bytes[k] = c.ktm.bytes[k];
    out->launches[k] = c.ktm.cnt[k];
  }
  for (int k = KT_N; k < 8; k++) { out->seconds[k] = 0.0; out->bytes[k] = 0.0; out->launches[k] = 0; }
  API_END
}

rmx_status rmx_profile_kernels(rmx_ctx* p, int prec, int reps, rmx_kernel_times* out)
{
  API_BEGIN
  Ctx& c = *need(p);
  if (!out || reps <= 0) fail(RMX_ERR_ARG, "profile: bad argument");
  if (!c.prepared) fail(RMX_ERR_ARG, "profile: no system");
  const int pr = prec == RMX_F32 ? RMX_F32 : RMX_F64;
  if ((pr == RMX_F32 && !c.have32) || (pr == RMX_F64 && !c.have64)) fail(RMX_ERR_ARG, "profile: no matrix");
  dev_profile(c, pr, reps, *out);
  API_END
}

}  // extern "C"
