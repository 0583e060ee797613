"""The C-ABI library loads and exports every entry point include/kaos_hip.h
declares (CPU only: nothing is called that touches a device), and the ctypes
mirror's struct layouts match the header's field order and sizes."""
import ctypes as C
import os
import re
import subprocess

import pytest

from kaos_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kaos_hip.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(kaos_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_python_exports():
    assert _declared() == sorted(_abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libkaos_hip.so not built")
    nm = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True)
    assert nm.returncode == 0, nm.stderr
    syms = {l.split()[-1] for l in nm.stdout.splitlines() if l.strip()}
    missing = [s for s in _declared() if s not in syms]
    assert not missing, missing
    lib = C.CDLL(_abi.LIB_PATH)
    for s in _declared():
        assert getattr(lib, s) is not None


def test_struct_layouts_match_the_header():
    # offsets/sizes as a C compiler lays out include/kaos_hip.h
    src = os.path.join(ROOT, "tests", "_layout.c")
    exe = os.path.join(ROOT, "tests", "_layout.bin")
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "kaos_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(kaos_cfg), offsetof(kaos_cfg, seed), sizeof(kaos_grid), offsetof(kaos_grid, d3v));
  printf("%zu %zu %zu\n", offsetof(kaos_grid, n_inj_lb), offsetof(kaos_grid, p), offsetof(kaos_grid, qGL));
  printf("%zu %zu\n", sizeof(kaos_step_out), offsetof(kaos_step_out, status));
  printf("%zu %zu %zu\n", offsetof(kaos_grid, icr), offsetof(kaos_grid, icr_lambda_perp), offsetof(kaos_grid, icr_chi_perp2));
  printf("%zu %zu\n", offsetof(kaos_grid, filter_pt_m1par), offsetof(kaos_grid, n_inj_ub));
  printf("%zu %zu %zu %zu %zu\n", sizeof(kaos_export_info), offsetof(kaos_export_info, species),
         offsetof(kaos_export_info, ns0), offsetof(kaos_export_info, theta), offsetof(kaos_export_info, Ts));
  printf("%zu %zu %zu %zu\n", offsetof(kaos_grid, nvperp1), offsetof(kaos_grid, vperp1GL),
         offsetof(kaos_grid, dvperp2), offsetof(kaos_grid, d3v3));
  printf("%zu %zu %zu %zu\n", sizeof(kaos_spinup_info), offsetof(kaos_spinup_info, nq_lb),
         offsetof(kaos_spinup_info, nsnorm_cub), offsetof(kaos_spinup_info, ub_replenish));
  printf("%zu %zu %zu %zu\n", sizeof(kaos_gridgen_in), offsetof(kaos_gridgen_in, vperp3d),
         offsetof(kaos_gridgen_in, nvpar_gp), offsetof(kaos_gridgen_in, ns0_neut));
  printf("%zu %zu %zu %zu %zu\n", sizeof(kaos_gridgen_out), offsetof(kaos_gridgen_out, counts),
         offsetof(kaos_gridgen_out, nsnorm_cub), offsetof(kaos_gridgen_out, d3v3), offsetof(kaos_gridgen_out, ena_d33v));
  printf("%zu\n", sizeof(kaos_gridgen_dims));
  return 0;
}
'''
    with open(src, "w") as f:
        f.write(prog)
    try:
        r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    finally:
        for p in (src, exe):
            if os.path.exists(p):
                os.remove(p)
    got = [int(x) for x in out]
    want = [C.sizeof(_abi.KaosCfg), _abi.KaosCfg.seed.offset, C.sizeof(_abi.KaosGrid), _abi.KaosGrid.d3v.offset,
            _abi.KaosGrid.n_inj_lb.offset, _abi.KaosGrid.p.offset, _abi.KaosGrid.qGL.offset,
            C.sizeof(_abi.KaosStepOut), _abi.KaosStepOut.status.offset,
            _abi.KaosGrid.icr.offset, _abi.KaosGrid.icr_lambda_perp.offset, _abi.KaosGrid.icr_chi_perp2.offset,
            _abi.KaosGrid.filter_pt_m1par.offset, _abi.KaosGrid.n_inj_ub.offset,
            C.sizeof(_abi.KaosExportInfo), _abi.KaosExportInfo.species.offset, _abi.KaosExportInfo.ns0.offset,
            _abi.KaosExportInfo.theta.offset, _abi.KaosExportInfo.Ts.offset,
            _abi.KaosGrid.nvperp1.offset, _abi.KaosGrid.vperp1GL.offset, _abi.KaosGrid.dvperp2.offset,
            _abi.KaosGrid.d3v3.offset,
            C.sizeof(_abi.KaosSpinupInfo), _abi.KaosSpinupInfo.nq_lb.offset, _abi.KaosSpinupInfo.nsnorm_cub.offset,
            _abi.KaosSpinupInfo.ub_replenish.offset,
            C.sizeof(_abi.KaosGridgenIn), _abi.KaosGridgenIn.vperp3d.offset, _abi.KaosGridgenIn.nvpar_gp.offset,
            _abi.KaosGridgenIn.ns0_neut.offset,
            C.sizeof(_abi.KaosGridgenOut), _abi.KaosGridgenOut.counts.offset, _abi.KaosGridgenOut.nsnorm_cub.offset,
            _abi.KaosGridgenOut.d3v3.offset, _abi.KaosGridgenOut.ena_d33v.offset,
            C.sizeof(_abi.KaosGridgenDims)]
    assert got == want
