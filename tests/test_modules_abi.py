"""CPU checks of the C++ module classes (include/vvhd_modules.hpp): the
header compiles on its own, the shim library exports the reference's class
methods, and a Space without a usable GPU fails with an exception."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "vvflow_amd")
MODULES = os.path.join(LIBDIR, "libvvhd_modules.so")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "h.cpp"
    src.write_text('#include "vvhd_modules.hpp"\nint main() { return sizeof(TObj) == 48 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{INC}", str(src)], check=True)


@pytest.mark.skipif(not os.path.exists(MODULES), reason="module library not built")
def test_exports_reference_methods():
    out = subprocess.run(["nm", "-DC", MODULES], check=True, capture_output=True, text=True).stdout
    for sym in ["stree::build(bool, bool, bool)", "stree::destroy()", "stree::findNode(TVec) const",
                "MConvectiveFast::process_all_lists()", "MConvectiveFast::velocity(TVec) const",
                "MConvectiveFast::calc_circulation(void const**)", "MEpsilonFast::CalcEpsilonFast(bool)",
                "MDiffusiveFast::process_vort_list()",
                "MFlowmove::move_and_clean(bool, void const**, unsigned long*)", "MFlowmove::vortex_shed()",
                "Space::calc_forces()", "Space::zero_forces()", "Space::save(char const*)",
                "Space::load(char const*)", "Stepdata::Stepdata(Space*, char const*, bool)", "Stepdata::write()",
                "Stepdata::flush()", "MDiffusiveFast::process_heat_list()", "MFlowmove::heat_shed()",
                "MFlowmove::streak_shed()", "MEpsilonFast::eps2h(TSortedNode const&, TVec)"]:
        assert sym in out, sym


@pytest.mark.skipif(not os.path.exists(MODULES), reason="module library not built")
def test_space_without_gpu_throws(tmp_path):
    src = tmp_path / "s.cpp"
    src.write_text('#include "vvhd_modules.hpp"\n#include <cstdio>\n#include <stdexcept>\n'
                   'int main() { try { Space S(0); std::puts("gpu"); } '
                   'catch (const std::runtime_error& e) { std::puts("no gpu"); } return 0; }\n')
    exe = tmp_path / "s"
    subprocess.run(["g++", "-std=c++17", f"-I{INC}", str(src), f"-L{LIBDIR}", "-lvvhd_modules", "-lvvhd_gpu",
                    f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1" if os.environ.get("GRAFT_TEST_HIDE_GPU") else
               os.environ.get("HIP_VISIBLE_DEVICES", ""))
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0
    assert r.stdout.strip() in ("gpu", "no gpu")
