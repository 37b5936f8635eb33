"""The reference-side binding shown in INTEGRATION.md is real code: every C++ block of that
file, put after the reference's own headers (compiled with -I on the reference's source
folder; nothing is copied) and include/rmx.h, compiles, and links into a program with the
reference's own objects (built from its sources by `make -C oracle ref`) and librmx.so, so
that GPU_ILU0 / GPU_KRYLOV override the reference's PRECON / SOLVER interfaces with their
exact signatures (Precon.h:74-75, Solver.h:161) and every rmx_ entry point they call exists.
Nothing is run: that needs a GPU.  Only in the container that holds the reference."""
import glob
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/sources"
LIB = os.path.join(ROOT, "rismo2d_amd", "librmx.so")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                                reason="needs the reference's sources (present only where the "
                                       "reference is)")

PRELUDE = """#include "Defs.h"
#include "Report.h"
#include "Project.h"
#include "Model.h"
#include "CRSMat.h"
#include "Eqs.h"
#include "Precon.h"
#include "Solver.h"
#include "rmx.h"
#include <vector>
"""

# instantiates every class and takes the address of every function the blocks define
MAIN = """
int main(int argc, char**)
{
  rmx_ctx* ctx = nullptr;
  if (argc < 0 || rmx_create(0, &ctx) != RMX_OK) return 0;   // no GPU here: stops
  GpuCrs crs; crs.ctx = ctx;
  PRECON* p = new GPU_ILU0(ctx, RMX_F32);
  GPU_KRYLOV* k = new GPU_KRYLOV; k->ctx = ctx; k->method = RMX_FGMRES; k->prec = RMX_F32;
  SOLVER* s = k;
  void* fns[] = { (void*)&gpu_newton_cycle, (void*)&gpu_solve_with_fallback,
                  (void*)&gpu_time_loop, (void*)&gpu_load_learned };
  delete p; delete s;
  rmx_destroy(ctx);
  return fns[0] == nullptr;
}
"""


def cpp_blocks():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", txt, re.S)


def test_integration_blocks_are_found():
    names = " ".join(cpp_blocks())
    for sym in ("class GPU_ILU0", "class GPU_KRYLOV", "struct GpuCrs", "gpu_newton_cycle",
                "gpu_solve_with_fallback", "gpu_time_loop", "gpu_load_learned"):
        assert sym in names, sym


def test_binding_compiles_against_reference_headers_and_links(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("librmx.so not built")
    src = tmp_path / "binding.cpp"
    src.write_text(PRELUDE + "\n".join(cpp_blocks()) + MAIN)
    obj = tmp_path / "binding.o"
    inc = ["-I" + REF, "-I" + os.path.join(ROOT, "include")]
    r = subprocess.run(["g++", "-std=c++17", "-w", *inc, "-c", str(src), "-o", str(obj)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    # the reference's objects (its Main.cpp with main renamed, as the harness builds them)
    refobj = os.path.join(ROOT, "oracle", "_ref", "obj")
    if not os.path.exists(os.path.join(refobj, "Main.o")):
        m = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "ref"],
                           capture_output=True, text=True)
        assert m.returncode == 0, m.stderr[-4000:]
    objs = [o for o in sorted(glob.glob(os.path.join(refobj, "*.o")))
            if not o.endswith("ref_harness.o")]
    exe = tmp_path / "binding"
    r = subprocess.run(["g++", "-o", str(exe), str(obj), *objs, "-L" + os.path.dirname(LIB),
                        "-lrmx", "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert exe.exists()
