// Internal definitions of the rmx library (context, device buffers, host model).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "rmx.h"

namespace rmx {

// Boundary-condition kind bits: the reference's BCON flag values (Bcon.h:160-190).
enum : uint32_t {
  kInlet = 1, kOutlet = 2, kOpenBnd = 4, kQInlet = 8, kQTInlet = 16, kSlip = 128, kSetUV = 512,
  kSetS = 1024, kFixU = 32768, kFixV = 65536, kAutoSlip = 4194304
};
enum { rotTangent = 1, rotNormal = 2 };
// node flags of this library
enum : uint32_t { nCorn = 1, nMids = 2, nBound = 4, nRotat = 64, nNoMoment = 128 };
// device node flag bits (packed for the kernels)
enum : uint32_t { dRotat = 1, dInlet = 2, dOutlet = 4 };
// drying and rewetting, per node: dry, marsh, the same at the previous time step, dry when
// the current drying and rewetting call started; per region element: dry, marsh, dry when
// the call started
enum : uint8_t { WD_DRY = 1, WD_MARSH = 2, WD_DRYPREV = 4, WD_MARSHPREV = 8, WD_WASDRY = 16 };
enum : uint8_t { EL_DRY = 1, EL_MARSH = 2, EL_WASDRY = 4 };

enum ElemType : int { ET_TRI = 0, ET_QUAD = 1, ET_LINE = 2 };

struct Error {
  rmx_status code;
  std::string msg;
};

[[noreturn]] void fail(rmx_status code, const std::string& msg);

#define HIPCHECK(x)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      ::rmx::fail(RMX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
  } while (0)

// A device buffer that frees itself.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count) HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) HIPCHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Host-side model: mesh, boundary line elements, boundary conditions, numbering, pattern.
struct HostModel {
  int np = 0, ne = 0;
  std::vector<double> x, y, zor, areaFact, manning, vt, veg_dp, veg_sp;
  std::vector<double> ste;                    // [ne] st * KDest of the algebraic model
  std::vector<double> exi, eyi;               // [ne] estx / KDest, esty / KDest (Elder)
  std::vector<int> elem, nnd;                 // [ne*8], [ne]
  std::vector<int> in_segs, out_segs;
  double Q = 0, S_in = 0, S_out = 0;
  bool inflow_qt = false;                     // the inlet line is a hydrograph (QTINLET)
  // topology: wet/dry state of nodes and region elements (WD_*, EL_*), connected wet
  // elements per node, node flags
  std::vector<uint8_t> nwd, ewd;
  std::vector<int> noel;
  std::vector<uint32_t> flag;
  int nb = 0;
  std::vector<int> bnd;                       // [nb*3]
  std::vector<uint32_t> bflag;                // 1 inlet, 2 outlet, 4 open
  std::vector<int> me;                        // model element list (>=0 region, <0 -(b+1))
  // boundary conditions (static part)
  std::vector<uint32_t> kind;
  std::vector<int> rot;
  std::vector<double> nx, ny, niox, nioy, valS;
  std::vector<int> inlet_nodes;               // nodes of the QINLET line, ascending
  // numbering and pattern
  int neq = 0;
  std::vector<int> eqno;                      // [3*np]
  std::vector<int> rowptr, cols;              // CSR, diagonal first, insertion order
  long long nnz = 0;
};

struct Params {
  int relax_method = 2, dry_method = 2;
  double dry_limit = 1e-3, max_delta_uv = 10.0, max_delta_s = 10.0, conv_uv = 1e-7,
         conv_s = 1e-7, max_us = 100.0;
  double g = 9.81;
  double vk = 1.78e-6 / (1.0 + 0.0337 * 10.0 + 0.00022 * 10.0 * 10.0);
  int turb = 1;                               // $TM_TURBULENCE code
  double cm = 0.09, cd = 1.00;                // KDCONST
  double relax_time = 0.0, interval = 0.0;    // methods 3, 4: relaxTimeFlow, stationary dt
  double rewet_limit = 0.005;                 // drying and rewetting ($DRYREW)
  int count_down = 5;
  double hmin() const { return dry_limit / 2.0; }
  bool turb_alg() const { return turb % 10 == 2; }
  bool turb_aniso() const { return (turb / 10) % 10 == 1; }
  bool turb_min() const { return (turb / 100) % 10 & 1; }
  bool turb_iterat() const { return (turb / 100) % 10 & 2; }
};

constexpr int RING = 8192;     // LDS ring of solved x values in the packed sweeps

// Packed layout of one triangular sweep over all strips of an ILU system.
struct SweepSet {
  int nblk = 0, nlev_max = 0, nlev_sum = 0, max_nr = 0;
  long long ne = 0;
  DBuf<int> blev, bpos;         // per strip: first level, first position
  DBuf<int> lev_ptr;            // per strip: local level offsets + sentinel
  DBuf<int> prs, prow, pcol, psrc;
  DBuf<double> xp;
};

// Block sweeps of the strip ILU(0): one wave per block (a segment of a strip), the lower
// and the upper sweep of the block in one launch with the block's values in LDS.  The unit
// of a sweep level is a node (the run of its rows in the block, up to 3): the entries that
// couple a node's rows are solved by one lane in registers, so a node adds one level, not
// three.  Local positions are the rows in lower-sweep order.  A level holds at most 32 rows
// and 64 entry pairs (a row's entries are padded to an even count, so no lane holds two
// rows' entries).
struct BlockSweep {
  bool on = false;
  int nb = 0;                      // blocks
  size_t lds = 0;                  // dynamic LDS bytes of the largest block
  long long ne_lo = 0, ne_up = 0, nlev = 0, nunit = 0;
  DBuf<int4> info;                 // [2 nb] {row base, rows, level table, lower levels},
                                   //        {upper levels, unit base, lower / upper entry base}
  DBuf<int> gidx;                  // [n] equation at every local position
  DBuf<unsigned> lev;              // per level: first unit (<< 21) | first entry in the block
  DBuf<unsigned short> mlo, mup;   // per entry: local column | row in its level << 11
  DBuf<unsigned> ulo, uup;         // per unit: position | rows << 11 | first row in level << 13
  DBuf<int> slo, sup;              // CSR position of every entry value (-1: zero)
  DBuf<int> sulo, suup, sdg;       // per unit: CSR positions of its lower / upper couplings
                                   // and of its diagonals (4 each; -1: none)
  DBuf<float> vlo32, vup32, wlo32, wup32;       // packed values (entries, unit couplings)
  DBuf<double> vlo64, vup64, wlo64, wup64;
  DBuf<double> rdg32, rdg64;       // reciprocal diagonals of the FP32 / FP64 factors
};

// One ILU(0) system: pattern, factorisation schedule, sweeps and factors.
struct Ilu {
  bool ready = false;
  int n = 0, nblk = 1, nseg = 1;
  std::vector<int> seg;         // segment of every row (empty: one segment per strip)
  long long nnz = 0;
  std::vector<int> bstart, h_rp, h_cols, flev;
  DBuf<int> rp, cols, srt_cols, flev_rows, fwd2bwd;
  DBuf<unsigned char> srt_pos;
  DBuf<int4> fdesc;             // factorisation: per level position {row, start, width, lower}
  DBuf<int2> fpiv;              // per sorted lower entry: pivot row start, entries after diag
  DBuf<long long> fpbase;       // per level position: offset of its plan entries
  DBuf<unsigned short> plan;    // per plan entry: pivot index | target position << 8
  SweepSet lo, up;
  BlockSweep bs;
  DBuf<double> LU64, plo64, pup64, pdg64;
  DBuf<float> LU32, plo32, pup32, pdg32;
  bool fac64 = false, fac32 = false;
  // dataflow factorisation: per row the epoch of the factorisation that finished it, and
  // {row counter, error bits, abort flag, unused}
  DBuf<int> done, flow_state;
  int epoch = 0;
  int repaired = 0;             // pivots repaired by the last factorisation
};

// FP32 SpMV layout.  Consecutive rows with one column set (the U, V and H rows of a node)
// form a group of up to 4 rows; a group stores its columns once, ascending, as 16-bit
// offsets from its smallest column, and the FP32 values of each of its rows in that column
// order (row after row, so group g's values are the CSR range of its rows).  Values are
// gathered from A32 again whenever A32 changes (a32_version).
struct SpmvGroups {
  bool on = false;                 // layout built (off: CSR kernel)
  long long ng = 0, noff = 0;      // groups, stored offsets
  int lanes = 8;                   // lanes per group (4, 8, 16, 32)
  DBuf<int> row0;                  // [ng+1] first row of each group
  DBuf<int4> meta;                 // [ng] first value (CSR offset of row0), first offset,
                                   //      smallest column, columns
  DBuf<uint16_t> off;              // [noff] column - base
  DBuf<int> perm;                  // [nnz] CSR position of every grouped value
  DBuf<float> vals;                // [nnz] FP32 values in group order
  long long version = -1;          // a32_version the values were gathered at
};

// Coarse space of the two-level preconditioner: one aggregate per (strip, coarse segment,
// variable U/V/H); the coarse matrix R A P (R = P^T, P piecewise constant on aggregates) is
// banded in aggregate order and factored on the host.
struct Coarse {
  bool ready = false, fac = false;
  int nc = 0, kl = 0, ku = 0, fac_prec = 0;
  long long nslot = 0;
  std::vector<int> slot_row, slot_col;        // coarse entry (I, J) of every slot
  DBuf<int> agg, agg_ptr, agg_rows;           // aggregate of every row; rows of each aggregate
  DBuf<long long> slot_ptr;                   // entries of A summed into each slot
  DBuf<int> perm;
  DBuf<double> cval, rc, zc, t;
  // A P: per row the aggregates of its columns and the summed entries (ap_ptr / ap_agg /
  // apv), summed from the entries listed by ap_src (grouped per (row, aggregate) pair)
  long long nap = 0;
  DBuf<int> ap_ptr, ap_agg, ap_src_ptr, ap_src;
  DBuf<double> apv;
  std::vector<double> band;                   // LU factors, band storage (LAPACK layout)
  std::vector<int> ipiv;
  double* h_buf = nullptr;                    // pinned, max(nc, nslot) doubles
  // Block cyclic reduction of R A P on the device, when it is block tridiagonal in strips
  // (every coarse entry couples aggregates of one strip or of neighbouring strips): block
  // i = the aggregates of strip i, padded to bs; factored on the host per factorisation,
  // applied by level kernels.  Off (or a singular reduced block): the host band LU.
  bool cr_ok = false, cr_fac = false;
  int nb = 0, bs = 0, levels = 0;
  std::vector<int> blk_start;                 // [nb + 1] first aggregate of every block
  std::vector<long long> slot_dst;            // slot -> index in the level-0 block arrays
  std::vector<long long> lev_elim, lev_keep;  // per level: offset of its matrices in mats
  long long top_off = 0, mats_n = 0;
  std::vector<double> h_mats;
  DBuf<double> mats, w;                       // factor matrices; padded work vector
  DBuf<int> d_blk_start;
  DBuf<long long> d_slot_dst;
  DBuf<double> d_blocks, d_inv, d_tmp;        // device factorisation: B, A, C blocks; inverses
  DBuf<int> d_bad;                            // a singular block
  ~Coarse() { if (h_buf) (void)hipHostFree(h_buf); }
};

// One stored Newton system of the training set: its FP64 matrix, the inverse row maxima of
// that matrix and the node features of its state, shared by the samples taken from it.
struct MpnnSys {
  DBuf<double> A, rinv;
  DBuf<float> feat;
};

// Learned preconditioner: mesh node graph, weights (FP32 master + BF16 layer copies),
// activations, gradients, Adam moments, stored training samples (residual vectors, each
// naming the stored system it was taken from).
struct Mpnn {
  static constexpr int C = 64, L = 6, F = 8, O = 18;
  bool graph = false, weights = false;
  int np = 0, step = 0, cap = 0, nsamp = 0, tn = 0;
  unsigned long long topo = 0;                // topology fingerprint of the graph's mesh
  std::vector<double> s_rr;                   // ||r||^2 of every stored sample
  std::vector<int> s_sys;                     // stored system of every sample
  std::vector<std::unique_ptr<MpnnSys>> sys;  // stored systems, in the order they arrived
  long long sys_version = -1;                 // matrix version of the last stored system
  DBuf<int> adj_ptr, adj, tptr, trow, tperm;  // node graph; transpose of the system pattern
  DBuf<float> theta, grad, m, v;
  DBuf<uint16_t> wt, wb, H, A, gp;
  DBuf<float> P, Q, feat, out, gh, ga, gout, partial, part2;
  DBuf<double> lpart, ev, rv, s_r, rinv, q, acc;
  long long rinv_version = -1, trunk_version = -1;
  int use_net = -1;                           // per system: -1 undecided, 0 dropped, 1 used
};

// Launch timing per kernel class (opt-in): a pair of HIP events on the context stream
// around every launch of the class, read back when the stream has synchronised.
enum { KT_FACTOR = 0, KT_SWEEP_LO, KT_SWEEP_UP, KT_SPMV, KT_COARSE, KT_GS, KT_RESID64, KT_SWEEP_BLK,
       KT_N };
struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;                 // event pool: pair k = ev[2k], ev[2k+1]
  std::vector<int> cls;                       // class of every pair in use
  std::vector<double> pbytes;                 // algorithmic bytes of every pair in use
  size_t used = 0;
  double sec[KT_N] = {}, bytes[KT_N] = {};
  long long cnt[KT_N] = {};
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Params par;
  HostModel hm;
  bool mesh_ready = false, prepared = false, csr_only = false;

  // ---- device node data (SoA) ----
  DBuf<double> x, y, zor, z, U, V, S, cf, niox, nioy, valU, valS, rotm;  // rotm [np*4]
  // time integration: time derivatives, previous time step's state and derivatives
  DBuf<double> dUdt, dVdt, dSdt, Uo, Vo, So, dUdto, dVdto, dSdto;
  bool unsteady = false;
  double dt = 0.0, theta = 1.0;
  // time terms' diagonal of the Jacobian (1 / dt / theta; methods 3, 4: of the relaxed time
  // steps dt_uv, dt_h, restarted by cycle_start; last_dt: the one before the latest update)
  double thdt_uv = 0.0, thdt_h = 0.0, dt_uv = 0.0, dt_h = 0.0, last_dt = 0.0;
  DBuf<uint32_t> dflag;
  DBuf<int> eqno;                             // [3*np]
  // friction gather: node -> incident region elements (ascending), areaFact*cf summed
  DBuf<int> n2e_ptr, n2e;
  DBuf<int> d_elem, d_nnd;                    // region elements [ne*8], [ne]
  DBuf<double> d_manning, d_areaFact, d_veg_dp, d_veg_sp;
  DBuf<double> d_ste, vtn, uun;               // algebraic model: st*KDest [ne]; nodal vt, uu
  DBuf<double> d_vte, d_exi, d_eyi;           // Elder: element vt, estx/KDest, esty/KDest [ne]
  DBuf<double> exx, exy, eyy;                 // Elder: nodal tensor in global axes [np]
  bool turb_stale = true;                     // nodal vt due at the next Newton step
  DBuf<int> d_in_segs, d_inlet_nodes;
  // drying and rewetting: node and element states (WD_*, EL_*), count down of dry nodes
  DBuf<uint8_t> d_nwd, d_ewd;
  DBuf<int> d_cdown;
  bool first_dry = true;                      // no drying and rewetting call yet
  int wet_sets = 0;                           // wet element sets set up since the mesh upload
  // ---- model elements for assembly ----
  int nme = 0;
  DBuf<int> me_nodes;                         // [nme*8]
  DBuf<double> me_vt;                         // [nme]
  DBuf<int> me_type;                          // [nme]
  DBuf<long long> me_slot_off;                // offset of each element's slot block
  DBuf<int> slots;                            // per element [nrow_ldof * ncol_ldof]
  std::vector<int> col_ptr;                   // colour classes: launches
  std::vector<int> col_type;
  DBuf<int> col_elems;
  // ---- CSR system ----
  int n = 0;
  long long nnz = 0;
  DBuf<int> rowptr, cols;
  DBuf<double> A64, B, X;
  DBuf<float> A32;
  bool have64 = false, have32 = false;
  int spmv_g = 8;                             // SpMV lanes per row
  SpmvGroups sg;                              // FP32 SpMV layout
  long long a32_version = 0;                  // bumped whenever A32 changes
  long long mat_version = 0;                  // bumped whenever A64 changes
  unsigned long long topo_hash = 0;           // fingerprint of the mesh topology
  Ilu ilu;                                    // reference-order ILU(0)
  Ilu silu;                                   // strip ILU(0)
  int strips = 0;                             // 0: two strips per compute unit
  int segments = 0;                           // cross-strip cuts of the strip sweeps (0: auto)
  Coarse coarse;
  Mpnn mpnn;
  int coarse_segments = -1;                   // cross-strip cuts of the coarse aggregates
                                              // (-1: automatic, 0: no coarse space)
  // Krylov work vectors
  DBuf<double> w[10];
  DBuf<double> red;                           // reduction scratch
  double* h_red = nullptr;                    // pinned host mirror of red
  // FGMRES basis
  DBuf<double> V_basis, Z_basis;
  DBuf<float> V32_basis;                      // FP32 copy for the Gram-Schmidt of FP32 solves
  int basis_m = 0;
  // kernel timing of the last profiled solve
  rmx_kernel_times kt{};
  KTimer ktm;                                 // launch timing per kernel class
  ~Ctx();
};

// launch timing: kt_begin returns the pair to end (-1 when timing is off); kt_collect reads
// the finished pairs (the stream must have synchronised)
int kt_begin(Ctx& c, int cls, double bytes);
void kt_end(Ctx& c, int k);
void kt_collect(Ctx& c);

// host model set-up (model.cpp)
void host_mesh(HostModel& hm, const rmx_mesh_desc& d);
void host_topology(HostModel& hm);
void host_boundary_conditions(HostModel& hm);
void host_numbering(HostModel& hm);
void host_pattern(HostModel& hm);
void host_colouring(const HostModel& hm, std::vector<int>& col_elems, std::vector<int>& col_ptr,
                    std::vector<int>& col_type);
void host_slots(const HostModel& hm, std::vector<long long>& off, std::vector<int>& slots,
                std::vector<int>& type);
void parallel_for(long long n, const std::function<void(long long, long long)>& f);

// device entry points implemented in the .hip files
void dev_prepare_step(Ctx& c);                 // friction (+ nodal vt when due) + inlet
void dev_turbulence(Ctx& c);                   // friction + algebraic nodal vt, uu now
void dev_assemble(Ctx& c, int prec, bool zero);
double dev_scale_l2(Ctx& c);
void noalloc_setup(int device);                // device address of the no-allocation scratch
void dev_scale_l2_noalloc(int device, int n, const int* rowptr, int prec, void* A, double* B,
                          double* scale, hipStream_t s);
void dev_convert(Ctx& c, int to_prec);
void dev_ilu0_factor(Ctx& c, int prec);
void dev_ilu0_apply(Ctx& c, int prec, const double* b, double* x);   // device pointers
void dev_spmv(Ctx& c, int prec, const double* x, double* r);        // device pointers
void dev_solve(Ctx& c, const rmx_krylov_opts& o, rmx_krylov_out& out);
void cycle_start(Ctx& c);
void dev_newton_update(Ctx& c, rmx_step_stats& st);
void dev_time_init(Ctx& c);                    // previous state := state, derivatives 0
void dev_slip_flow(Ctx& c);                    // GRID::SetSlipFlow
void dev_dry_rewet(Ctx& c);                    // GRID::DryRewet, method 2 (device pass)
void dev_wetdry_commit(Ctx& c);                // previous dry / marsh flags := current
void dev_time_step(Ctx& c, int first);         // EQS_UVS2D::Predict
void dev_time_commit(Ctx& c);                  // previous state := state
void dev_element_coefs(Ctx& c, int first, int count, double* K, double* F);
void dev_profile(Ctx& c, int prec, int reps, rmx_kernel_times& out);
void dev_resid(Ctx& c, const double* x, const double* b, double* r);
void dev_resid_p(Ctx& c, int prec, const double* x, const double* b, double* r);
void dev_dots(Ctx& c, int K, const double* const* x, const double* const* y, double* dst);
void setup_ilu_schedule(Ctx& c);
Ilu& ilu_of(Ctx& c, int kind);
void ilu_factor(Ctx& c, Ilu& I, int prec);
// x = M^-1 b; with aidx, x[e] += aval[aidx[e]] folded into the sweep's output when the
// staged sweep runs (returns whether it was folded in)
bool ilu_apply(Ctx& c, Ilu& I, int prec, const double* b, double* x, hipEvent_t* ev,
               const int* aidx = nullptr, const double* aval = nullptr);
void upload_shapes(hipStream_t s);
std::vector<int> cut_segments(const std::vector<int>& bstart, const std::vector<int>& group,
                              const HostModel& hm, int nseg);
void coarse_setup(Ctx& c, const std::vector<int>& rp, const std::vector<int>& cl,
                  const std::vector<int>& group, int nseg);
int coarse_segments_of(const Ctx& c);          // coarse segments per strip in effect
void coarse_factor(Ctx& c, int prec);
void coarse_apply(Ctx& c, int prec, const double* r, double* z);   // device pointers
void dev_precond(Ctx& c, int prec, int kind, const double* in, double* out);
void spmv_setup(Ctx& c, const int* rowptr, const int* cols);   // host pattern
double spmv_f32_bytes(const Ctx& c);          // algorithmic bytes of one FP32 SpMV
bool mpnn_ready(Ctx& c);
void mpnn_apply(Ctx& c, int prec, const double* r, double* z);
void mpnn_forward(Ctx& c, const double* r, double* z, float* feat_out);
void mpnn_set_weights(Ctx& c, const float* th, int n);
void mpnn_get_weights(Ctx& c, float* th, int n);
int mpnn_add_sample(Ctx& c, const double* r, int capacity);
double mpnn_loss_grad(Ctx& c, int s, float* grad);
double mpnn_train(Ctx& c, int epochs, double lr);

}  // namespace rmx
