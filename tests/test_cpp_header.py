"""The C++ drop-in header compiles against psdp.h and links with libpsdp.so
(host compiler only; nothing is run on a GPU).  CPU only."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = r'''
#include "psd_projection/psd_projection.hpp"
int main(int argc, char**) {
  if (argc > 5) {  // never taken: only the symbols must resolve
    psdp_handle_t h = nullptr;
    double* m = nullptr; double lo, up;
    psd_projection::composite_FP32(h, m, 10);
    psd_projection::composite_FP32_auto_scale(h, h, m, 10);
    psd_projection::composite_FP16(h, m, 10);
    psd_projection::composite_FP16_auto_scale(h, h, m, 10);
    psd_projection::composite_FP32_emulated(h, m, 10);
    psd_projection::composite_FP32_emulated_auto_scale(h, h, m, 10);
    psd_projection::approximate_two_norm(h, h, m, 10, &lo, &up);
    // the MEX gateway's call shape (MATLAB/psd_projection_MATLAB.cu:158-162,175-179):
    // eigenvalues returned as device memory the caller frees
    double* dW = psd_projection::eig_FP64_psd(h, h, m, 10, true);
    psdp_free(dW);
    double* none = psd_projection::eig_FP32_psd(h, h, m, 10);
    psdp_free(none);
    dW = psd_projection::eig_FP32_psd(h, h, m, (size_t)10, true);
    psdp_free(dW);
    psdp_project_method(h, "eig_FP64", m, 10, nullptr);
    size_t r = 0;
    psd_projection::composite_FP32_auto_scale_deflate(h, h, m, 10);
    psd_projection::composite_FP16_auto_scale_deflate(h, h, m, 10);
    psd_projection::eig_FP64_deflate(h, h, m, 10);
    double b = psd_projection::compute_eigenpairs(h, h, m, 10, 3, &r, m, m);
    psd_projection::lobpcg(h, h, m, m, m, 10, 2);
    (void)b;
  }
  psdp_config_t cfg;
  size_t b = 0;
  psdp_default_config(&cfg, PSDP_METHOD_MIXED);
  return psdp_workspace_size(&cfg, 1000, &b) == PSDP_OK && b > 0 ? 0 : 1;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_dropin_header_compiles_and_links(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(SRC)
    exe = tmp_path / "t"
    lib = ROOT / "psd_projection_amd"
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src),
                        f"-L{lib}", "-lpsdp", f"-Wl,-rpath,{lib}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
