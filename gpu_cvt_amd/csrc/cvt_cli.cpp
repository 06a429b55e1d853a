// cvt_cli.cpp -- `gpu_cvt <mesh.obj> [mode] [iters]`, the reference's command
// line (main.cu:371-1079) over the C ABI of libcvt_hip.so.
//
// Same arguments, defaults and outputs: mode 0 baseline / 1 five-tier freeze /
// 2 six-tier freeze (default 1), 100 iterations, K = 32, KPROJ = 5, sites
// starting at the mesh vertices, freeze threshold (0.01 R)^2 with R the max
// bbox edge; every 10th iteration (DUMP_STRIDE) writes iter_%03d.obj,
// iter_%03d_sites.xyz and iter_%03d_frozen.txt and appends a row to
// eval_iters.csv under experiments/output/{gpucvt|freeze|freeze_tiered}/<mesh>/;
// the run appends a row to experiments/runs.csv.  Mode 0 drops each site's own
// vertex from its centroid -> mesh list (main.cu:728) and every remesh clips
// by the current sites' bbox (R2, main.cu:899-909), as the reference.
//
// The remeshed surface is the library's own dual-triangle extraction (the
// reference's non-Geogram path, main.cu:928-974).  Beside every
// iter_%03d_sites.xyz the Geogram hand-off iter_%03d_co3ne.sh holds the
// reference's vorpalite command and options (main.cu:314-340); with the
// environment variable GPU_CVT_VORPALITE naming a vorpalite executable it is
// run (outside the timed steps, as the reference) into iter_%03d_co3ne.obj,
// and that surface -- the reference's default output -- is the one measured
// in eval_iters.csv and runs.csv.
//
// Quality (evaluate.cuh:388-462) and the one-sided Hausdorff distance are
// computed on the device; the stage columns come from per-kernel timing
// (cvt_job_profile_stages): knn_sites_ms is the site-grid rebuild,
// centroids_ms the cell kernel (site KNN, tangent frame, cell and centroid)
// plus the FP64 pass, project_ms the projection kernel (kproj search,
// reprojection, freeze test and the new site's nearest vertex, the next
// iteration's frame source), so knn_site_to_mesh_ms, uv_from_mesh_ms,
// knn_centroid_to_mesh_ms and freeze_ms are 0.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "cvt.h"

namespace {

// Wavefront OBJ: `v x y z` and `f a b c ...` (1-based or negative indices,
// v/vt/vn tokens, polygons fanned from the first corner).  Returns the max
// bbox edge over all `v` lines in FP32, as load_obj_triangles (main.cu:170-230).
bool load_obj(const std::string& path, std::vector<double>& V, std::vector<int>& F, float& R) {
    std::ifstream in(path);
    if (!in) return false;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    std::string line;
    std::vector<int> idx;
    while (std::getline(in, line)) {
        if (line.size() < 2) continue;
        if (line[0] == 'v' && line[1] == ' ') {
            const char* s = line.c_str() + 2;
            char* e = nullptr;
            float p[3];
            for (int a = 0; a < 3; ++a) {
                p[a] = std::strtof(s, &e);
                s = e;
            }
            for (int a = 0; a < 3; ++a) {
                V.push_back(p[a]);
                lo[a] = std::min(lo[a], p[a]);
                hi[a] = std::max(hi[a], p[a]);
            }
        } else if (line[0] == 'f' && line[1] == ' ') {
            const char* s = line.c_str() + 2;
            idx.clear();
            while (*s) {
                while (*s == ' ' || *s == '\t') ++s;
                if (!*s) break;
                char* e = nullptr;
                long vi = std::strtol(s, &e, 10);
                if (e == s) break;
                const long nv = (long)(V.size() / 3);
                if (vi < 0) vi = nv + vi + 1;
                idx.push_back((int)(vi - 1));
                s = e;
                while (*s && *s != ' ' && *s != '\t') ++s;
            }
            for (size_t k = 1; k + 1 < idx.size(); ++k) {
                F.push_back(idx[0]);
                F.push_back(idx[k]);
                F.push_back(idx[k + 1]);
            }
        }
    }
    R = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    return true;
}

const char* kEvalHeader =
    "mesh,mode,iter,Qmin,Qavg,theta_min,theta_min_avg,theta_lt_30_pct,theta_gt_90_pct,dH,iter_remesh_ms,"
    "total_remesh_ms,freeze_cell_num,n_vertices,freeze_pct,knn_sites_ms,knn_site_to_mesh_ms,uv_from_mesh_ms,"
    "centroids_ms,knn_centroid_to_mesh_ms,project_ms,freeze_ms,count_same,count_low_disp,count_both,"
    "count_newly_frozen,blocked_by_same,blocked_by_low_disp\n";

// The vorpalite command of run_vorpalite (main.cu:316-340): co3ne
// reconstruction of the site cloud with the reference's options.  quote:
// escape the executable's characters (false: wrap it in double quotes as
// given, so that a shell parameter expansion in it still expands).
std::string co3ne_command(const std::string& exe, const std::string& xyz, const std::string& obj, bool quote = true) {
    auto q = [](const std::string& x) {
        std::string r = "\"";
        for (char c : x) {
            if (c == '"' || c == '\\') r += '\\';
            r += c;
        }
        return r + "\"";
    };
    return (quote ? q(exe) : "\"" + exe + "\"") + " " + q(xyz) + " " + q(obj) +
           " co3ne=true pre=false remesh=false post=true co3ne:repair=true co3ne:radius=5%" +
           " co3ne:nb_neighbors=30 log:quiet=true log:pretty=false";
}

int fail_msg(const char* what) {
    std::fprintf(stderr, "gpu_cvt: %s: %s\n", what, cvt_last_error());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: gpu_cvt <mesh.obj> [mode 0|1|2] [iters]\n");
        return 1;
    }
    const std::string input = argv[1];
    const std::string mesh_name = std::filesystem::path(input).stem().string();
    const int mode = argc >= 3 ? std::atoi(argv[2]) : 1;
    const int total_iter = argc >= 4 ? std::atoi(argv[3]) : 100;
    const char* mode_name = mode == 0 ? "baseline" : (mode == 2 ? "freezing_cvt_tiered" : "freezing_cvt");
    std::printf("mode %d (%s)\n", mode, mode_name);
    constexpr int kDumpStride = 10;
    constexpr int kNeigh = 32;

    std::vector<double> V;
    std::vector<int> F;
    float R = 0.f;
    if (!load_obj(input, V, F, R)) {
        std::fprintf(stderr, "gpu_cvt: cannot read %s\n", input.c_str());
        return 1;
    }
    {
        // drop unreferenced vertices (compact_mesh, main.cu:240-268), on the device
        int kept = 0;
        if (cvt_mesh_compact(V.data(), (int)(V.size() / 3), F.data(), (int)(F.size() / 3), V.data(), F.data(), &kept) !=
            CVT_OK)
            return fail_msg("compact");
        V.resize(3 * (size_t)kept);
    }
    const int nV = (int)(V.size() / 3), nF = (int)(F.size() / 3);
    std::printf("V %d\nF %d\n", nV, nF);
    if (nV == 0 || nF == 0) return 1;
    const float freeze_disp = 0.01f * R * 0.01f * R;  // squared, main.cu:431
    std::printf("freeze_disp %.6g (squared, linear=0.01*R R=%.6g)\n", (double)freeze_disp, (double)R);

    cvt_params p;
    cvt_params_default(&p);
    p.K = kNeigh;
    p.kproj = 5;
    p.clip_R_mode = 1;
    p.clip_R = R;
    p.freeze_disp_abs2 = freeze_disp;
    p.freeze_need = 5;
    p.exclude_self_vertex = mode == 0 ? 1 : 0;  // knn_vertices_bruteforce_k<KPROJ, true> (main.cu:728)
    p.remesh_R_mode = 1;                         // R2 from the current sites (main.cu:899-909)
    if (mode == 0) {
        p.freeze_mode = 0;
    } else {
        p.freeze_mode = 2;
        static const int s1[6] = {5, 10, 15, 20, 25, 25}, s2[6] = {5, 10, 15, 20, 25, 10};
        for (int t = 0; t < 6; ++t) {
            p.tier_streak[t] = mode == 2 ? s2[t] : s1[t];
            p.tier_disp_abs2[t] = freeze_disp;
        }
    }
    cvt_job* job = nullptr;
    if (cvt_job_create(V.data(), nV, F.data(), nF, V.data(), nV, &p, &job) != CVT_OK) return fail_msg("create");
    if (mode == 1 || mode == 2) {
        if (cvt_job_compute_tiers(job, mode, nullptr, nullptr, nullptr) != CVT_OK) return fail_msg("tiers");
        std::printf("mode %d: tier assignment done\n", mode);
    }

    const char* root_dir = mode == 0 ? "gpucvt" : (mode == 2 ? "freeze_tiered" : "freeze");
    const std::string out_dir = std::string("experiments/output/") + root_dir + "/" + mesh_name;
    std::filesystem::create_directories(out_dir);
    const std::string eval_csv = out_dir + "/eval_iters.csv";
    {
        std::ofstream h(eval_csv);
        h << kEvalHeader;
    }

    std::vector<double> S((size_t)3 * nV);
    std::vector<float> Sf((size_t)3 * nV);
    std::vector<unsigned char> frozen((size_t)nV), streak((size_t)nV);
    double total_ms = 0;
    int used_iters = 0, final_nf = 0;
    float converge = 0.f;
    const char* vorpalite = std::getenv("GPU_CVT_VORPALITE");
    for (int iter = 0; iter < total_iter; ++iter) {
        std::printf("\n=== iter %d ===\n", iter);
        cvt_stage_ms sm;
        if (cvt_job_profile_stages(job, 1, &sm) != CVT_OK) return fail_msg("step");
        cvt_iter_stats st;
        if (cvt_job_stats(job, 1, &st) != CVT_OK) return fail_msg("stats");
        total_ms += sm.iter;
        std::printf("knn_sites_ms %.3f\nknn_site_to_mesh_ms %.3f\ncentroids_ms %.3f\nproject_ms %.3f\n"
                    "iter_remesh_ms %.3f (total_remesh_ms %.3f)\n",
                    sm.grid, 0.0, sm.cell + sm.precise, sm.project, sm.iter, total_ms);
        std::printf("same neighbors %d\nlow displacement %d\nboth %d\npass2 %d\n", st.count_same, st.count_low_disp,
                    st.count_both, st.count_newly_frozen);
        converge = (float)st.frozen_total / (float)nV;
        std::printf("total frozen so far %d / %d (%.2f%%)\n", st.frozen_total, nV, converge * 100.0f);
        used_iters = iter + 1;
        if (iter % kDumpStride != 0) continue;

        if (cvt_job_get_sites(job, S.data()) != CVT_OK) return fail_msg("sites");
        if (cvt_job_get_freeze(job, frozen.data(), streak.data()) != CVT_OK) return fail_msg("freeze state");
        for (size_t k = 0; k < S.size(); ++k) Sf[k] = (float)S[k];
        int* Fr = nullptr;
        int nFr = 0;
        if (cvt_job_remesh(job, 1e-6, &Fr, &nFr) != CVT_OK) return fail_msg("remesh");
        std::vector<int> Fn(Fr, Fr + 3L * nFr);
        cvt_free(Fr);
        char path[1024];
        std::snprintf(path, sizeof path, "%s/iter_%03d_sites.xyz", out_dir.c_str(), iter);
        const std::string xyz_path = path;
        {
            std::ofstream os(path);
            for (int i = 0; i < nV; ++i) os << Sf[3 * i] << " " << Sf[3 * i + 1] << " " << Sf[3 * i + 2] << "\n";
        }
        std::snprintf(path, sizeof path, "%s/iter_%03d.obj", out_dir.c_str(), iter);
        {
            std::ofstream os(path);
            for (int i = 0; i < nV; ++i) os << "v " << Sf[3 * i] << " " << Sf[3 * i + 1] << " " << Sf[3 * i + 2] << "\n";
            for (int f = 0; f < nFr; ++f)
                os << "f " << Fn[3 * f] + 1 << " " << Fn[3 * f + 1] + 1 << " " << Fn[3 * f + 2] + 1 << "\n";
        }
        std::printf("iter %d V %d faces %d wrote %s\n", iter, nV, nFr, path);
        std::snprintf(path, sizeof path, "%s/iter_%03d_frozen.txt", out_dir.c_str(), iter);
        {
            std::ofstream os(path);
            for (int i = 0; i < nV; ++i) os << (int)frozen[i] << "\n";
        }
        // the Geogram hand-off: the reference's command and options
        std::snprintf(path, sizeof path, "%s/iter_%03d_co3ne.obj", out_dir.c_str(), iter);
        const std::string co3ne_obj = path;
        const std::string cmd = co3ne_command(vorpalite ? vorpalite : "vorpalite", xyz_path, co3ne_obj);
        std::snprintf(path, sizeof path, "%s/iter_%03d_co3ne.sh", out_dir.c_str(), iter);
        {
            std::ofstream os(path);
            namespace fs = std::filesystem;
            os << "#!/bin/sh\ncd \"$(dirname \"$0\")\" && "
               << co3ne_command("${VORPALITE:-vorpalite}", fs::path(xyz_path).filename().string(),
                                fs::path(co3ne_obj).filename().string(), false)
               << "\n";
        }
        cvt_quality q;
        double dH = 0;
        int nf_eval = nFr;
        if (vorpalite) {
            const int vret = std::system(cmd.c_str());
            if (vret != 0) {
                std::fprintf(stderr, "vorpalite failed (exit %d)\n", vret);
                return 1;
            }
            std::vector<double> Vg;
            std::vector<int> Fg;
            float Rg = 0.f;
            if (!load_obj(co3ne_obj, Vg, Fg, Rg)) return fail_msg("co3ne surface");
            const int nVg = (int)(Vg.size() / 3);
            nf_eval = (int)(Fg.size() / 3);
            if (cvt_mesh_quality(Vg.data(), nVg, Fg.data(), nf_eval, &q) != CVT_OK) return fail_msg("quality");
            // one-sided distance over the probes of evaluate.cuh:252-283
            std::vector<double> probes(Vg);
            for (int f = 0; f < nf_eval; ++f) {
                const double* a = &Vg[3 * (size_t)Fg[3 * f]];
                const double* b = &Vg[3 * (size_t)Fg[3 * f + 1]];
                const double* c = &Vg[3 * (size_t)Fg[3 * f + 2]];
                for (int k = 0; k < 3; ++k) probes.push_back((a[k] + b[k] + c[k]) / 3.0);
                for (int k = 0; k < 3; ++k) probes.push_back(0.5 * (a[k] + b[k]));
                for (int k = 0; k < 3; ++k) probes.push_back(0.5 * (b[k] + c[k]));
                for (int k = 0; k < 3; ++k) probes.push_back(0.5 * (c[k] + a[k]));
            }
            std::vector<double> d2(probes.size() / 3);
            if (!d2.empty() &&
                cvt_job_point_mesh_d2(job, probes.data(), (int)d2.size(), d2.data()) != CVT_OK)
                return fail_msg("hausdorff");
            for (double v : d2) dH = std::max(dH, std::sqrt(v));
            std::printf("iter %d co3ne surface: %d vertices, %d faces\n", iter, nVg, nf_eval);
        } else {
            if (cvt_job_quality(job, Fn.data(), nFr, &q) != CVT_OK) return fail_msg("quality");
            if (cvt_job_hausdorff(job, 1e-6, &dH, nullptr) != CVT_OK) return fail_msg("hausdorff");
        }
        final_nf = nf_eval;
        const int blocked_same = std::max(0, st.count_low_disp - st.count_both);
        const int blocked_low = std::max(0, st.count_same - st.count_both);
        std::ofstream out(eval_csv, std::ios::app);
        out << mesh_name << "," << root_dir << "," << iter << "," << q.Qmin << "," << q.Qavg << "," << q.theta_min
            << "," << q.theta_min_avg << "," << q.theta_lt_30_pct << "," << q.theta_gt_90_pct << "," << (float)dH
            << "," << (float)sm.iter << "," << (float)total_ms << "," << st.frozen_total << "," << nV << ","
            << (nV > 0 ? 100.f * (float)st.frozen_total / (float)nV : 0.f) << "," << (float)sm.grid << ","
            << 0.f << "," << 0.f << "," << (float)(sm.cell + sm.precise) << "," << 0.f << ","
            << (float)sm.project << "," << 0.f << "," << st.count_same << "," << st.count_low_disp << ","
            << st.count_both << "," << st.count_newly_frozen << "," << blocked_same << "," << blocked_low << "\n";
    }
    std::printf("\nTotal remeshing time (mesh rebuilding excluded): %.3f ms\n", total_ms);
    {
        const std::string runs = "experiments/runs.csv";
        std::filesystem::create_directories("experiments");
        const bool exists = std::filesystem::exists(runs);
        std::ofstream out(runs, std::ios::app);
        if (!exists) out << "meshname,mode,final_nv,final_nf,converge_rate,used_iterations,total_remeshing_time\n";
        out << mesh_name << "," << root_dir << "," << nV << "," << final_nf << "," << converge << "," << used_iters << ","
            << (float)total_ms << "\n";
    }
    cvt_job_destroy(job);
    return 0;
}
