"""The gpu_cvt command line: same arguments, files and CSV schema as the
reference's driver (main.cu:371-1079), and the same iterates as the library."""
import os
import subprocess

import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpu_cvt_amd", "gpu_cvt")
EVAL_HEADER = ("mesh,mode,iter,Qmin,Qavg,theta_min,theta_min_avg,theta_lt_30_pct,theta_gt_90_pct,dH,iter_remesh_ms,"
               "total_remesh_ms,freeze_cell_num,n_vertices,freeze_pct,knn_sites_ms,knn_site_to_mesh_ms,"
               "uv_from_mesh_ms,centroids_ms,knn_centroid_to_mesh_ms,project_ms,freeze_ms,count_same,"
               "count_low_disp,count_both,count_newly_frozen,blocked_by_same,blocked_by_low_disp")


def _write_obj(path, V, F):
    with open(path, "w") as f:
        for v in V:
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for t in F:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_cli_outputs_and_iterates(tmp_path, mode):
    V, F = mg.displaced_sphere(12, seed=3, crease_seed=4)
    obj = tmp_path / "dsph.obj"
    _write_obj(obj, V, F)
    print(f"[cli] mode={mode} nV={len(V)} nF={len(F)}", flush=True)
    r = subprocess.run([EXE, str(obj), str(mode), "12"], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    root = {0: "gpucvt", 1: "freeze", 2: "freeze_tiered"}[mode]
    out = tmp_path / "experiments" / "output" / root / "dsph"
    lines = (out / "eval_iters.csv").read_text().strip().split("\n")
    assert lines[0] == EVAL_HEADER
    rows = [ln.split(",") for ln in lines[1:]]
    assert [int(x[2]) for x in rows] == [0, 10]
    assert all(x[0] == "dsph" and x[1] == root for x in rows)
    for it in (0, 10):
        for suffix in (".obj", "_sites.xyz", "_frozen.txt"):
            assert (out / f"iter_{it:03d}{suffix}").exists()
    runs = (tmp_path / "experiments" / "runs.csv").read_text().strip().split("\n")
    assert runs[0] == "meshname,mode,final_nv,final_nf,converge_rate,used_iterations,total_remeshing_time"
    assert runs[1].split(",")[:3] == ["dsph", root, str(len(V))] and runs[1].split(",")[5] == "12"
    # iterate 10 (11 steps) equals the library run of the same configuration
    Vf = np.asarray(V, np.float32)
    R = float((Vf.max(0) - Vf.min(0)).max())
    fd = np.float32(0.01) * np.float32(R) * np.float32(0.01) * np.float32(R)
    kw = dict(K=32, clip_R_mode=1, clip_R=R, freeze_disp_abs2=float(fd), freeze_need=5)
    if mode:
        kw.update(freeze_mode=2, tier_streak=[5, 10, 15, 20, 25, 25] if mode == 1 else [5, 10, 15, 20, 25, 10],
                  tier_disp_abs2=[float(fd)] * 6)
    with cvt.Job(V, F, np.asarray(V, np.float64), **kw) as job:
        if mode:
            job.compute_tiers(mode)
        job.step(11)
        ref = job.sites()
    got = np.loadtxt(out / "iter_010_sites.xyz")
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    mesh_lines = (out / "iter_010.obj").read_text().split("\n")
    nf = sum(1 for ln in mesh_lines if ln.startswith("f "))
    assert nf > len(V)
