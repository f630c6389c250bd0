"""The C++ layers compile as shipped: include/dfm_eqnsys.hpp (the equation-system classes) compiles and links
against the library, and include/dfm_delfem_adapter.hpp (the reference-side binding) type-checks against the
reference's own headers with the reference's loops and element routines called through it. No GPU is used."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/include"

EQNSYS_TU = r"""
#include "dfm_eqnsys.hpp"
int main() {
  dfm_ctx ctx = nullptr;
  if (dfm_ctx_create(0, &ctx) != DFM_OK) return 0;  // no device: nothing to run
  dfm::Eqn::Mesh m;
  m.ndim = 2;
  m.coords = {0, 0, 1, 0, 1, 1, 0, 1};
  m.lnods = {0, 1, 2, 0, 2, 3};
  dfm::Eqn::CEqnSystem_Solid2D eq(ctx, m);
  eq.SetYoungPoisson(1.0, 0.3, true);
  dfm::Eqn::CEqn_Solid2D e = eq.GetEquation(0);
  e.SetGravitation(0.0, -1.0);
  eq.SetEquation(e);
  dfm::Eqn::ElemAry left;
  left.nnoel = 2;
  left.lnods = {0, 3};
  eq.AddFixElemAry(left);
  eq.SetStationary(true);
  eq.Solve();
  std::vector<std::pair<unsigned int, double> > itr = eq.GetAry_ItrNormRes();
  dfm::Eqn::CEqn_Solid3D_Linear* s3 = nullptr;
  dfm::Eqn::CEqnSystem_Scalar2D* c2 = nullptr;
  dfm::Eqn::CEqn_Scalar3D* c3 = nullptr;
  (void)s3; (void)c2; (void)c3; (void)itr;
  dfm_ctx_destroy(ctx);
  return 0;
}
"""

ADAPTER_TU = r"""
#include "dfm_delfem_adapter.hpp"
#include "delfem/ls/solver_ls_iter.h"
#include "delfem/femeqn/eqn_poisson.h"
#include "delfem/femeqn/eqn_hyper.h"
#include "delfem/field_world.h"

bool solve_through_the_seams(dfm_ctx ctx, Fem::Field::CFieldWorld& world, unsigned int id_field) {
  // the assembly seam: the reference's element routine into the host system, then to the device
  dfm::delfem::CLinearSystemGPU_Field ls(ctx, 0);
  ls.AddPattern_Field(id_field, world);
  ls.InitializeMarge();
  Fem::Eqn::AddLinSys_Poisson(ls, 1.0, 1.0, world, id_field);
  ls.FinalizeMarge();
  ls.Upload(id_field, world);
  double conv = 1e-6;
  unsigned int iter = 1000;
  bool ok = ls.SolvePCG(conv, iter, DFM_MIXED);
  ls.DownloadUpdate(id_field, world);
  ls.UpdateValueOfField(id_field, world, Fem::Field::VALUE);
  // the solver seam: the reference's own host loop over device vectors
  dfm::delfem::CLinearSystemGPU dev(ls.handle());
  dev.SetValuePrecond();
  conv = 1e-6;
  iter = 1000;
  ok = LsSol::Solve_PCG(conv, iter, dev) && ok;
  ok = dev.SolvePBiCGSTAB(conv, iter, DFM_FP32_ILU) && ok;
  return ok;
}

bool solve_multi_segment(dfm_ctx ctx, Fem::Field::CFieldWorld& world, unsigned int id_disp, unsigned int id_lambda) {
  // a P1b/P1 system of three segments, assembled by the reference's own Mooney-Rivlin routine on the host
  Fem::Ls::CLinearSystem_Field ls;
  ls.AddPattern_Field(id_disp, world);
  ls.AddPattern_Field(id_lambda, id_disp, world);
  ls.InitializeMarge();
  Fem::Eqn::AddLinSys_Hyper2D_Static(ls, 1.0, 0.5, 1.0, 0.0, -0.1, id_disp, id_lambda, world);
  ls.FinalizeMarge();
  dfm::delfem::CLinearSystemGPU_MultiSeg dev(ctx);
  dev.Upload(ls.m_ls);
  double conv = 1e-9;
  unsigned int iter = 500;
  const bool ok = dev.SolvePBiCGSTAB(conv, iter);
  dev.DownloadUpdate(ls.m_ls);
  ls.UpdateValueOfField(id_disp, world, Fem::Field::VALUE);
  return ok;
}
"""


def _compile(tu, args, tmp_path, name):
    src = tmp_path / f"{name}.cpp"
    src.write_text(tu)
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror", "-Wno-unused-variable", f"-I{ROOT}/include", str(src)] + args
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_eqnsys_header_compiles_and_links(tmp_path):
    lib = os.path.join(ROOT, "delfem_amd")
    if not os.path.exists(os.path.join(lib, "libdfm_gpu.so")):
        pytest.skip("library not built")
    out = tmp_path / "eqn"
    _compile(EQNSYS_TU, ["-o", str(out), f"-L{lib}", "-ldfm_gpu", f"-Wl,-rpath,{lib}"], tmp_path, "eqn")
    assert out.exists()


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="the reference's headers are not present")
def test_adapter_type_checks_against_reference_headers(tmp_path):
    # the reference's headers are older C++: its own warnings are not this test's subject
    src = tmp_path / "adapter.cpp"
    src.write_text(ADAPTER_TU)
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", f"-I{ROOT}/include", f"-I{REF_INC}", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
