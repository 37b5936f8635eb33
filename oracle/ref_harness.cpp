// Test infrastructure only: golden-vector harness around the reference program.
//
// This file is compiled against the reference's own headers (-I<reference>/sources) and
// linked with the reference's unmodified objects by oracle/Makefile.  It never ships and
// never runs on a GPU machine.  The reference's main() is compiled as rismo_main
// (-Dmain=rismo_main) and called unchanged; four functions are intercepted with the GNU
// linker's --wrap option so that their inputs and outputs can be written at full
// precision while the reference keeps doing exactly what it does:
//
//   EQS_UVS2D::Coefs      (CoefsUVS2D.cpp:49)  per-element estifm / force
//   EQS::Solve            (Solve.cpp:53)       node state + numbering before, Newton
//                                              correction X after each Newton step
//   CRSMAT::ScaleL2Norm   (CRSMat.cpp:319)     assembled pattern, values and RHS before
//                                              scaling, the scale; then, on the scaled
//                                              system, PRECO_ILU0::Factor/Solve
//                                              (Preco_ilu0.cpp:64,497) and CRSMAT::MulVec
//                                              (CRSMat.cpp:466) applied to the scaled RHS
//   GRID::SetSlipFlow     (SlipFlow.cpp)       final node state at the end of Execute
//
// Output: a stream of records  [char tag[8]][int64 nbytes][payload]  written to the file
// named by $RMX_DUMP.  $RMX_DUMP_ELEM_STEP selects the Newton step (0-based) whose element
// matrices are written (-1: none); $RMX_DUMP_ILU_STEP likewise for the ILU records.

#include "Defs.h"
#include "Report.h"
#include "Vars.h"
#include "Shape.h"
#include "Type.h"
#include "Node.h"
#include "Elem.h"
#include "Grid.h"
#include "Model.h"
#include "Project.h"
#include "CRSMat.h"
#include "Eqs.h"
#include "EqsUVS2D.h"
#include "Precon.h"
#include "Preco_ilu0.h"

#include <stdint.h>
#include <vector>

int rismo_main(int argc, char* argv[]);

static FILE*    g_out       = NULL;
static int      g_step      = -1;     // Newton step counter (number of Solve calls)
static int      g_elemStep  = 0;
static int      g_iluStep   = 0;
static EQS*     g_eqs       = NULL;
static PROJECT* g_project   = NULL;

static void rec(const char* tag, const void* data, int64_t nbytes)
{
  if( !g_out ) return;
  char t[8] = {0,0,0,0,0,0,0,0};
  strncpy( t, tag, 8 );
  fwrite( t, 1, 8, g_out );
  fwrite( &nbytes, sizeof(int64_t), 1, g_out );
  if( nbytes ) fwrite( data, 1, (size_t)nbytes, g_out );
}

extern "C" {
int    __real__ZN9EQS_UVS2D5CoefsEP4ELEMP7PROJECTPPdS4_( EQS_UVS2D*, ELEM*, PROJECT*, double**, double* );
int    __real__ZN3EQS5SolveEP5MODELiPdS2_P7PROJECTP6SOLVERPP6PRECONi( EQS*, MODEL*, int, double*, double*,
                                                                   PROJECT*, SOLVER*, PRECON**, int );
double __real__ZN6CRSMAT11ScaleL2NormEPdP6SUBDOM( CRSMAT*, double*, SUBDOM* );
void   __real__ZN4GRID11SetSlipFlowEv( GRID* );
}

static void dump_state( const char* tag, PROJECT* project, EQS* eqs )
{
  GRID* rg = project->M2D->region;
  int np = rg->Getnp();
  // per node: name, flag, bc.kind, bc.rot, eqno[3] (as doubles for a single array)
  // then x,y,z,zor,U,V,S,cf,vt,nx,ny,niox,nioy,bcU,bcS,rot00,rot01,rot11,rot10,uu,uv,vv,
  // dUdt,dVdt,dSdt, previous time step U,V,S,dUdt,dVdt,dSdt
  const int nf = 38;
  std::vector<double> buf( (size_t)np * nf );
  for( int n=0; n<np; n++ )
  {
    NODE* nd = rg->Getnode(n);
    double* b = &buf[(size_t)n * nf];
    b[0] = nd->Getname();
    b[1] = nd->flag;
    b[2] = nd->bc.kind;
    b[3] = nd->bc.rot;
    for( int j=0; j<3; j++ )  b[4+j] = eqs ? eqs->GetEqno( nd, j ) : -1;
    b[7]  = nd->x;     b[8]  = nd->y;     b[9]  = nd->z;     b[10] = nd->zor;
    b[11] = nd->v.U;   b[12] = nd->v.V;   b[13] = nd->v.S;
    b[14] = nd->cf;    b[15] = nd->vt;
    b[16] = nd->bc.nx; b[17] = nd->bc.ny; b[18] = nd->bc.niox; b[19] = nd->bc.nioy;
    b[20] = nd->bc.val ? nd->bc.val->U : 0.0;
    b[21] = nd->bc.val ? nd->bc.val->S : 0.0;
    b[22] = nd->bc.Getrot(0,0);
    b[23] = nd->bc.Getrot(0,1);
    b[24] = nd->bc.Getrot(1,1);
    b[25] = nd->bc.Getrot(1,0);
    b[26] = nd->uu;    b[27] = nd->uv;    b[28] = nd->vv;
    b[29] = nd->v.dUdt;  b[30] = nd->v.dVdt;  b[31] = nd->v.dSdt;
    b[32] = nd->vo.U;    b[33] = nd->vo.V;    b[34] = nd->vo.S;
    b[35] = nd->vo.dUdt; b[36] = nd->vo.dVdt; b[37] = nd->vo.dSdt;
  }
  int64_t hdr[2] = { np, nf };
  rec( "STATEHDR", hdr, sizeof(hdr) );
  rec( tag, buf.data(), (int64_t)buf.size() * 8 );
}

extern "C" int __wrap__ZN9EQS_UVS2D5CoefsEP4ELEMP7PROJECTPPdS4_( EQS_UVS2D* self, ELEM* elem,
                                                                PROJECT* project,
                                                                double** estifm, double* force )
{
  int ret = __real__ZN9EQS_UVS2D5CoefsEP4ELEMP7PROJECTPPdS4_( self, elem, project, estifm, force );
  if( g_out && g_step == g_elemStep )
  {
    int nnd = elem->Getnnd();
    int ndf = 3 * nnd;
    // header: ret, nnd, isBound, node names
    std::vector<double> b( 3 + nnd + ndf + (size_t)ndf * ndf );
    b[0] = ret;
    b[1] = nnd;
    b[2] = isFS(elem->flag, ELEM::kBound) ? 1 : 0;
    for( int i=0; i<nnd; i++ )  b[3+i] = elem->nd[i]->Getname();
    double* f = &b[3+nnd];
    double* k = f + ndf;
    for( int r=0; r<ndf; r++ )
    {
      f[r] = ret ? force[r] : 0.0;
      for( int c=0; c<ndf; c++ )  k[(size_t)r*ndf + c] = ret ? estifm[r][c] : 0.0;
    }
    rec( "ELEM", b.data(), (int64_t)b.size() * 8 );
  }
  return ret;
}

extern "C" int __wrap__ZN3EQS5SolveEP5MODELiPdS2_P7PROJECTP6SOLVERPP6PRECONi(
    EQS* self, MODEL* model, int neq, double* rhs, double* x, PROJECT* project,
    SOLVER* solver, PRECON** precon, int assemble )
{
  g_step++;
  g_eqs = self;
  g_project = project;
  int64_t s = g_step;
  rec( "STEP", &s, 8 );
  int64_t tm = project->GetTimeStep();
  rec( "TSTEP", &tm, 8 );
  if( g_out ) dump_state( "STATE", project, self );
  int err = __real__ZN3EQS5SolveEP5MODELiPdS2_P7PROJECTP6SOLVERPP6PRECONi(
                self, model, neq, rhs, x, project, solver, precon, assemble );
  int64_t hdr[2] = { neq, err };
  rec( "XHDR", hdr, sizeof(hdr) );
  rec( "XSOL", x, (int64_t)neq * 8 );
  return err;
}

extern "C" double __wrap__ZN6CRSMAT11ScaleL2NormEPdP6SUBDOM( CRSMAT* self, double* b, SUBDOM* subdom )
{
  int n = self->m_neq;
  if( g_out )
  {
    // pattern: widths then the column indices of each row in stored order
    std::vector<int32_t> w( n );
    int64_t nnz = 0;
    for( int i=0; i<n; i++ ) { w[i] = self->m_width[i]; nnz += w[i]; }
    std::vector<int32_t> cols( nnz );
    std::vector<double>  vals( nnz );
    int64_t p = 0;
    for( int i=0; i<n; i++ )
      for( int j=0; j<w[i]; j++, p++ )
      {
        cols[p] = self->m_index[i][j];
        vals[p] = (double) self->m_A[i][j];
      }
    int64_t hdr[3] = { n, nnz, (int64_t)sizeof(REALPR) };
    rec( "PATHDR", hdr, sizeof(hdr) );
    rec( "WIDTH", w.data(), (int64_t)n * 4 );
    rec( "COLS", cols.data(), nnz * 4 );
    rec( "AVAL", vals.data(), nnz * 8 );
    rec( "BVEC", b, (int64_t)n * 8 );
  }
  double scale = __real__ZN6CRSMAT11ScaleL2NormEPdP6SUBDOM( self, b, subdom );
  rec( "SCALE", &scale, 8 );
  if( g_out && g_step == g_iluStep && g_eqs && g_project )
  {
    int64_t nnz = 0;
    for( int i=0; i<n; i++ ) nnz += self->m_width[i];
    std::vector<double> vals( nnz );
    int64_t p = 0;
    for( int i=0; i<n; i++ )
      for( int j=0; j<self->m_width[i]; j++ )  vals[p++] = (double) self->m_A[i][j];
    rec( "ASCALED", vals.data(), nnz * 8 );
    rec( "BSCALED", b, (int64_t)n * 8 );

    PRECO_ILU0* ilu = new PRECO_ILU0();
    ilu->crsi = new CRSMAT( self );
    ilu->Factor( g_project, g_eqs, self );
    p = 0;
    for( int i=0; i<n; i++ )
      for( int j=0; j<self->m_width[i]; j++ )  vals[p++] = (double) ilu->crsi->m_A[i][j];
    rec( "ILU", vals.data(), nnz * 8 );
    std::vector<double> y( n ), r( n );
    ilu->Solve( g_project, g_eqs, b, y.data() );
    rec( "ILUAPPLY", y.data(), (int64_t)n * 8 );
    self->MulVec( b, r.data(), g_project, g_eqs );
    rec( "MULVEC", r.data(), (int64_t)n * 8 );
    delete ilu;
  }
  return scale;
}

extern "C" void __wrap__ZN4GRID11SetSlipFlowEv( GRID* self )
{
  __real__ZN4GRID11SetSlipFlowEv( self );
  if( g_out && g_project ) dump_state( "FINAL", g_project, g_eqs );
}

static void close_out()
{
  if( g_out ) fclose( g_out );
  g_out = NULL;
}

int main( int argc, char* argv[] )
{
  const char* path = getenv( "RMX_DUMP" );
  if( path ) g_out = fopen( path, "wb" );
  atexit( close_out );
  const char* es = getenv( "RMX_DUMP_ELEM_STEP" );
  if( es ) g_elemStep = atoi( es );
  const char* is = getenv( "RMX_DUMP_ILU_STEP" );
  if( is ) g_iluStep = atoi( is );
  int ret = rismo_main( argc, argv );
  close_out();
  return ret;
}
