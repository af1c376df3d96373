"""C++ callers: the reference header's EXT_C is empty on Linux (digHolo.h:148-156),
so a C++ program built against it references mangled names.  libdigholo.so
exports every entry point of include/digHolo.h under its C name and under the
C++ name (digholo_amd/gen_cxx_aliases.py); a C++ program compiled with
DIGHOLO_CXX_LINKAGE (the reference header's linkage) links and runs against it.
No GPU: the program only creates, configures and destroys a handle."""
import os
import subprocess

import pytest

from digholo_amd import api
from digholo_amd.gen_cxx_aliases import HEADER, argument_names, declarations

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_entry_point_has_a_cxx_name():
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    names = [d[1] for d in declarations(open(HEADER).read())]
    assert len(names) >= 100
    demangled = subprocess.run(["c++filt"], input="\n".join(sorted(s for s in exported if s.startswith("_Z"))),
                               capture_output=True, text=True).stdout.split("\n")
    cxx = {d.split("(")[0] for d in demangled if "(" in d}
    assert set(names) <= cxx, sorted(set(names) - cxx)
    # a few mangled names written out, as a caller of the reference header references them
    for s in ("_Z13digHoloCreatev", "_Z14digHoloDestroyi", "_Z15digHoloSetBatchiiPf",
              "_Z20digHoloBasisGetCoefsiPiS_S_"):
        assert s in exported, s


def test_parameter_names_parse():
    assert argument_names("void") == [] and argument_names("") == []
    assert argument_names("int handleIdx, float* buffer") == ["handleIdx", "buffer"]
    with pytest.raises(ValueError):
        argument_names("int, float*")


def test_cxx_program_links_and_runs(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#define DIGHOLO_CXX_LINKAGE
#include "digHolo.h"
#include <cstdio>
int main() {
  int h = digHoloCreate();
  if (h < 0) return 1;
  if (digHoloConfigSetFrameDimensions(h, 640, 512) != DIGHOLO_ERROR_SUCCESS) return 2;
  int w = 0, hh = 0;
  if (digHoloConfigGetFrameDimensions(h, &w, &hh) != DIGHOLO_ERROR_SUCCESS || w != 640 || hh != 512) return 3;
  if (digHoloConfigSetPolCount(h, 2) != DIGHOLO_ERROR_SUCCESS || digHoloConfigGetPolCount(h) != 2) return 4;
  if (digHoloDestroy(h) != DIGHOLO_ERROR_SUCCESS) return 5;
  if (digHoloDestroy(h) != DIGHOLO_ERROR_INVALIDHANDLE) return 6;
  std::printf("ok\n");
  return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.dirname(api.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + libdir, "-ldigholo", "-Wl,-rpath," + libdir], check=True)
    nm = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True).stdout
    assert "_Z13digHoloCreatev" in nm and " digHoloCreate\n" not in nm
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)
