"""include/cother_hip.h is a plain C interface: a C99 program (and a C++17
one) using the calls a reference-side shim makes compiles against it with
no HIP, torch or C++ headers (the usage of the integration guide's first
example, plus the serving and results-file calls)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "cother_hip.h"

int run(int device, const char* envdir, cth_pset* queries, cth_pset* database, cth_hit* hits, int maxhits) {
  cth_ctx* ctx = 0;
  if (cth_ctx_create(device, envdir, &ctx)) { fprintf(stderr, "%s\n", cth_last_error()); return -1; }
  if (cth_upload_queries(ctx, queries) || cth_upload_db(ctx, database)) { cth_ctx_destroy(ctx); return -1; }
  cth_search_opts o;
  memset(&o, 0, sizeof o);
  o.eval = 10.0f;
  o.ssemodel = 2;
  o.do_map = 1;
  o.minpp = 0.28f;
  int nh[4];
  cth_search_stats st;
  if (cth_search(ctx, 0, 1, &o, hits, maxhits, nh, &st)) fprintf(stderr, "%s\n", cth_last_error());
  cth_server_opts so;
  cth_server_opts_init(&so);
  cth_server* srv = 0;
  if (cth_server_create(ctx, &so, &srv) == 0) {
    uint64_t ticket = 0;
    int n = 0;
    cth_request_info info;
    if (cth_server_submit(srv, 0, &ticket) == 0) cth_server_wait(srv, ticket, hits, maxhits, &n, 0, 0, &info);
    cth_server_destroy(srv);
  }
  cth_ctx_destroy(ctx);
  return 0;
}
"""


def _compile(cc, std, tmp_path, suffix):
    src = tmp_path / f"shim.{suffix}"
    src.write_text(PROGRAM)
    r = subprocess.run([cc, f"-std={std}", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_header_compiles_as_c99(tmp_path):
    _compile("gcc", "c99", tmp_path, "c")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no C++ compiler")
def test_header_compiles_as_cpp17(tmp_path):
    _compile("g++", "c++17", tmp_path, "cpp")


@pytest.mark.skipif(shutil.which("gcc") is None or
                    not os.path.exists(os.path.join(ROOT, "cother_amd", "libcother_hip.so")),
                    reason="no C compiler or library not built")
def test_c_program_links_and_starts(tmp_path):
    """the C99 program links against libcother_hip.so with every symbol it
    calls resolved, and starts (without a device it only counts devices)"""
    src = tmp_path / "prog.c"
    src.write_text(PROGRAM + r"""
int main(int argc, char** argv) {
  int n = cth_device_count();
  printf("devices %d\n", n);
  if (argc > 1 && n > 0) return run(0, argv[1], 0, 0, 0, 0);
  return 0;
}
""")
    exe = str(tmp_path / "prog")
    libdir = os.path.join(ROOT, "cother_amd")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT}/include", str(src), f"-L{libdir}",
                        "-lcother_hip", f"-Wl,-rpath,{libdir}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert run.returncode == 0 and "devices" in run.stdout, run.stdout + run.stderr
