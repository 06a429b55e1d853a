"""C ABI checks that need no GPU: the library loads, exports every function
include/cvt.h declares, the ctypes structs match the C layout, and argument
validation fails with the documented status before any device work."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cvt.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cvt_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = cvt.lib()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(cvt.EXPORTS) == names


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "cvt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(cvt_params), sizeof(cvt_iter_stats), sizeof(cvt_job_info),
         offsetof(cvt_params, debug), offsetof(cvt_job_info, device_bytes), offsetof(cvt_params, tier_disp_abs2));
  printf("%zu %zu %zu %zu %zu\n", sizeof(cvt_server_config), sizeof(cvt_result), sizeof(cvt_server_info),
         offsetof(cvt_result, t_done), offsetof(cvt_server_info, pending));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert vals == [C.sizeof(cvt.Params), C.sizeof(cvt.IterStats), C.sizeof(cvt.JobInfo),
                    cvt.Params.debug.offset, cvt.JobInfo.device_bytes.offset, cvt.Params.tier_disp_abs2.offset,
                    C.sizeof(cvt.ServerConfig), C.sizeof(cvt.Result), C.sizeof(cvt.ServerInfo),
                    cvt.Result.t_done.offset, cvt.ServerInfo.pending.offset]


def test_default_params():
    p = cvt.default_params()
    assert (p.K, p.kproj, p.freeze_mode, p.nan_policy) == (24, 5, 0, 0)
    assert list(p.tier_streak) == [5, 10, 15, 20, 25, 10]


@pytest.mark.parametrize("kw, code", [
    (dict(K=7), -1),              # unsupported K
    (dict(K=24, kproj=3), -1),    # unsupported kproj
    (dict(K=8, freeze_mode=5), -1),
])
def test_create_rejects_bad_params(kw, code):
    V, F = mg.icosphere(2)
    S = mg.sample_surface(V, F, 100, seed=0)
    with pytest.raises(cvt.CVTError) as e:
        cvt.Job(V, F, S, **kw)
    assert e.value.code == code


def test_create_rejects_bad_inputs():
    V, F = mg.icosphere(2)
    S = mg.sample_surface(V, F, 30, seed=0)
    with pytest.raises(cvt.CVTError) as e:   # nS < K + 1 (the reference silently skips n < 1024)
        cvt.Job(V, F, S, K=32)
    assert e.value.code == -1
    S2 = S.copy()
    S2[3, 1] = np.inf
    with pytest.raises(cvt.CVTError) as e:
        cvt.Job(V, F, S2, K=8)
    assert e.value.code == -5
    F2 = F.copy()
    F2[0, 0] = len(V)
    with pytest.raises(cvt.CVTError) as e:
        cvt.Job(V, F2, S, K=8)
    assert e.value.code == -1
    with pytest.raises(cvt.CVTError):
        cvt.Job(V[:2], F[:0], S, K=8)


def test_server_rejects_bad_requests():
    """Argument checks of the request loop run before any device work."""
    V, F = mg.icosphere(2)
    S = mg.sample_surface(V, F, 50, seed=0)
    L = cvt.lib()
    t = C.c_ulonglong()
    p = cvt.default_params()
    # null server handle
    rc = L.cvt_server_submit(None, V.reshape(-1), len(V), F.astype(np.int32).reshape(-1), len(F), S.reshape(-1),
                             len(S), C.byref(p), 1e-3, 10, C.byref(t))
    assert rc == -1
    cfg = cvt.ServerConfig()
    L.cvt_server_config_default(C.byref(cfg))
    assert cfg.hbm_budget_bytes == 200 << 30 and cfg.epoch_iters == 10
    r = cvt.Result()
    assert L.cvt_server_poll(None, 1, C.byref(r)) == -1


def test_cli_usage_and_missing_file(tmp_path):
    """The gpu_cvt command line (main.cu:371-402 arguments) fails cleanly
    without a mesh, before touching a GPU."""
    exe = os.path.join(ROOT, "gpu_cvt_amd", "gpu_cvt")
    assert os.path.exists(exe), "gpu_cvt not built"
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "none.obj")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "cannot read" in r.stderr
