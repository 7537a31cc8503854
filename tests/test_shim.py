"""The reference-side binding (cother_amd/shim/CuBatchProcessingHIP.*), the
class with CuBatchProcessing's method set that a maintainer builds into the
reference in place of its CUDA batch-processing object: it compiles against
the reference's own headers (src/, read where they lie; no CUDA), and a
driver that constructs it and calls every method links against the
reference's host objects (built from its sources by oracle/Makefile.ref) and
libcother_hip.so, and starts (with no device it only reports the device
count).  Skipped where the reference sources or their build are absent."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")
SHIM = os.path.join(ROOT, "cother_amd", "shim")
LIBDIR = os.path.join(ROOT, "cother_amd")
CXX = ["g++", "-std=c++11", "-O1", "-w", f"-I{REF}", f"-I{REF}/extsp", f"-I{ROOT}/include", f"-I{SHIM}",
       "-DCOTHERSTATEDIR=/nonexistent", "-DCOTHER_INSTALL_DIR=/nonexistent"]

DRIVER = r"""
#include <stdio.h>
#include "CuBatchProcessingHIP.h"

// Every method DevCommThread calls on its batch-processing object
// (src/libmycu/cualn/DevCommThread.cpp:155-165,296-356), reached only with
// a device; without one the driver reports the device count and exits.
int main(int argc, char** argv) {
  int n = cth_device_count();
  printf("devices %d\n", n);
  if (n < 1 || argc < 2) return 0;
  CuBatchProcessingHIP b(0, argv[1], 0, NULL, NULL, NULL, NULL, NULL);
  b.SetDbDetails(0, 0);
  b.SetQueryLen(0);
  b.SetScoreThreshold(0.0f);
  b.SetLogEThreshold(0.0f);
  b.TransferCPMData(NULL, NULL, NULL);
  b.ProcessScoreMatrix(0, 0, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
  b.WaitForIdleChilds();
  return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF) or shutil.which("g++") is None,
                    reason="reference sources (/root/reference) or g++ absent")
def test_shim_compiles_against_reference_headers(tmp_path):
    r = subprocess.run(CXX + ["-Wall", "-c", os.path.join(SHIM, "CuBatchProcessingHIP.cpp"), "-o",
                              str(tmp_path / "shim.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _ref_objects():
    objs = []
    for sub in ("liblib", "libpro", "libHDP", "libseg", "extsp"):
        objs += glob.glob(os.path.join(OBJ, sub, "**", "*.o"), recursive=True)
    objs += glob.glob(os.path.join(OBJ, "libmycu", "cualn", "AlnWriter*.o"))
    return sorted(objs)


@pytest.mark.skipif(not os.path.isdir(REF) or not os.path.isdir(OBJ) or
                    not os.path.exists(os.path.join(LIBDIR, "libcother_hip.so")),
                    reason="reference sources, their host build (oracle/Makefile.ref) or the library absent")
def test_shim_links_with_reference_objects_and_library(tmp_path):
    objs = _ref_objects()
    assert any(o.endswith("AlnWriter.o") for o in objs)
    drv = tmp_path / "driver.cpp"
    drv.write_text(DRIVER)
    exe = str(tmp_path / "driver")
    r = subprocess.run(CXX + [str(drv), os.path.join(SHIM, "CuBatchProcessingHIP.cpp"), *objs, f"-L{LIBDIR}",
                              "-lcother_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,--no-undefined", "-lpthread", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120, env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert run.returncode == 0, run.stdout + run.stderr
    assert "devices" in run.stdout
