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


TTIME_PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>
#include HEADER
int main() {
    // fixed cases, then a seeded sweep of decimal times and sums
    const double ds[] = {0.005, 0.01, 0.02, 0.125, 0.1, 1.0, 2.5, 0.0, 0.3, 7e-4, 1e-3};
    for (double d : ds) {
        TTime t = TTime::makeWithSecondsDecimal(d);
        std::printf("dec %d %u\n", t.value, t.timescale);
    }
    std::srand(7);
    for (int k = 0; k < 400; k++) {
        const double a = (std::rand() % 2000) * 0.005, b = ((std::rand() % 50) + 1) * 0.005;
        TTime ta = TTime::makeWithSecondsDecimal(a), tb = TTime::makeWithSecondsDecimal(b);
        TTime s = TTime::add(ta, tb);
        std::printf("add %d %u div %d %d %.17g\n", s.value, s.timescale, (int)ta.divisibleBy(tb),
                    (int)s.divisibleBy(TTime(1, 100)), double(s));
    }
    TTime z;
    std::printf("zero %d %d\n", (int)z.divisibleBy(TTime(0, 1)), (int)TTime(3, 8).divisibleBy(TTime(1, 8)));
    TTime m = TTime::add(TTime(3, 8), TTime(5, 1000));
    std::printf("mixed %d %u %s\n", m.value, m.timescale, std::string(m).c_str());
    return 0;
}
"""


def run_ttime_probe(tmp_path, header, extra):
    src = tmp_path / "t.cpp"
    src.write_text(TTIME_PROBE.replace("HEADER", f'"{header}"'))
    exe = tmp_path / ("t_" + os.path.basename(header).split(".")[0])
    subprocess.run(["g++", "-std=c++17", "-O1", *extra, str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_ttime_known_answers(tmp_path):
    """TTime (TTime.hpp): decimal construction, exact sums over the least
    common timescale, divisibility, conversions."""
    out = run_ttime_probe(tmp_path, os.path.join(INC, "vvhd_modules.hpp"), [f"-I{INC}"]).splitlines()
    dec = [tuple(map(int, ln.split()[1:])) for ln in out if ln.startswith("dec")]
    assert dec[:9] == [(5, 1000), (1, 100), (2, 100), (125, 1000), (1, 10), (1, 1), (25, 10), (0, 1), (3, 10)]
    zero = [ln for ln in out if ln.startswith("zero")][0].split()
    assert zero[1:] == ["0", "1"]
    assert [ln for ln in out if ln.startswith("mixed")][0].split()[1:] == ["380", "1000", "380/1000"]
    for ln in (l for l in out if l.startswith("add")):
        p = ln.split()
        assert int(p[1]) / int(p[2]) == float(p[6])


REF_TTIME = "/root/reference/libvvhd/headers/TTime.hpp"


@pytest.mark.skipif(not os.path.exists(REF_TTIME), reason="the reference tree is not here")
def test_ttime_matches_reference_header(tmp_path):
    """The same probe built on the reference's own header-only TTime gives the
    same lines (value / timescale of every construction and sum, every
    divisibility answer)."""
    ours = run_ttime_probe(tmp_path, os.path.join(INC, "vvhd_modules.hpp"), [f"-I{INC}"])
    ref = run_ttime_probe(tmp_path, REF_TTIME, ["-include", "cstdio"])
    assert ours == ref
