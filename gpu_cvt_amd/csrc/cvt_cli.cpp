// cvt_cli.cpp -- `gpu_cvt <mesh.obj> [mode] [iters]`, the reference's command
// line (main.cu:371-1079) over the C ABI of libcvt_hip.so.
//
// Same arguments, defaults and outputs: mode 0 baseline / 1 five-tier freeze /
// 2 six-tier freeze (default 1), 100 iterations, K = 32, KPROJ = 5, sites
// starting at the mesh vertices, freeze threshold (0.01 R)^2 with R the max
// bbox edge; every 10th iteration (DUMP_STRIDE) writes iter_%03d.obj,
// iter_%03d_sites.xyz and iter_%03d_frozen.txt and appends a row to
// eval_iters.csv under experiments/output/{gpucvt|freeze|freeze_tiered}/<mesh>/;
// the run appends a row to experiments/runs.csv.  The remeshed surface is the
// library's own dual-triangle extraction (the reference's non-Geogram path,
// main.cu:928-974), since the external vorpalite binary is not part of this
// build.
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

// Drop vertices no face references and renumber (compact_mesh, main.cu:240-268).
void compact(std::vector<double>& V, std::vector<int>& F) {
    const int nV = (int)(V.size() / 3);
    std::vector<int> remap((size_t)nV, -1);
    for (int v : F)
        if (v >= 0 && v < nV) remap[(size_t)v] = 0;
    int n = 0;
    std::vector<double> V2;
    for (int v = 0; v < nV; ++v)
        if (remap[(size_t)v] == 0) {
            remap[(size_t)v] = n++;
            V2.insert(V2.end(), V.begin() + 3L * v, V.begin() + 3L * v + 3);
        }
    for (int& v : F) v = (v >= 0 && v < nV) ? remap[(size_t)v] : -1;
    V.swap(V2);
}

struct Quality {
    float Qmin, Qavg, theta_min, theta_min_avg, lt30, gt90;
};

// Triangle angles and the shape quality q = (6/sqrt 3) A / (s e_max) in FP32
// (eval_quality_angles_cpu, evaluate.cuh:388-462).
Quality quality(const std::vector<float>& P, const std::vector<int>& F) {
    auto sub = [&](int a, int b, float* o) {
        for (int k = 0; k < 3; ++k) o[k] = P[3 * (size_t)a + k] - P[3 * (size_t)b + k];
    };
    auto dot = [](const float* u, const float* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    auto angle = [&](const float* u, const float* v) {
        const float den = std::fmax(std::sqrt(dot(u, u)) * std::sqrt(dot(v, v)), 1e-30f);
        float cs = dot(u, v) / den;
        cs = std::fmin(1.f, std::fmax(-1.f, cs));
        return std::acos(cs) * (180.0f / 3.14159265358979323846f);
    };
    Quality q{1e30f, 0.f, 1e30f, 0.f, 0.f, 0.f};
    double qsum = 0, thsum = 0, lt = 0, gt = 0, cnt = 0;
    const size_t nF = F.size() / 3;
    for (size_t f = 0; f < nF; ++f) {
        const int a = F[3 * f], b = F[3 * f + 1], c = F[3 * f + 2];
        float ab[3], ac[3], ba[3], bc[3], ca[3], cb[3];
        sub(b, a, ab); sub(c, a, ac); sub(a, b, ba); sub(c, b, bc); sub(a, c, ca); sub(b, c, cb);
        const float A = angle(ab, ac), B = angle(ba, bc), C = angle(ca, cb);
        const float tmin = std::fmin(A, std::fmin(B, C));
        q.theta_min = std::fmin(q.theta_min, tmin);
        thsum += tmin;
        lt += (A < 30.f) + (B < 30.f) + (C < 30.f);
        gt += (A > 90.f) + (B > 90.f) + (C > 90.f);
        cnt += 3;
        const float l0 = std::sqrt(dot(ab, ab)), l1 = std::sqrt(dot(bc, bc)), l2 = std::sqrt(dot(ca, ca));
        const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2],
                             ab[0] * ac[1] - ab[1] * ac[0]};
        const float area = 0.5f * std::sqrt(dot(cr, cr));
        const float s = 0.5f * (l0 + l1 + l2), e = std::fmax(l0, std::fmax(l1, l2));
        const float qq = (6.0f / std::sqrt(3.0f)) * (area / std::fmax(s * e, 1e-30f));
        q.Qmin = std::fmin(q.Qmin, qq);
        qsum += qq;
    }
    q.Qavg = nF ? (float)(qsum / (double)nF) : 0.f;
    q.theta_min_avg = nF ? (float)(thsum / (double)nF) : 0.f;
    q.lt30 = cnt > 0 ? (float)(lt * 100.0 / cnt) : 0.f;
    q.gt90 = cnt > 0 ? (float)(gt * 100.0 / cnt) : 0.f;
    return q;
}

const char* kEvalHeader =
    "mesh,mode,iter,Qmin,Qavg,theta_min,theta_min_avg,theta_lt_30_pct,theta_gt_90_pct,dH,iter_remesh_ms,"
    "total_remesh_ms,freeze_cell_num,n_vertices,freeze_pct,knn_sites_ms,knn_site_to_mesh_ms,uv_from_mesh_ms,"
    "centroids_ms,knn_centroid_to_mesh_ms,project_ms,freeze_ms,count_same,count_low_disp,count_both,"
    "count_newly_frozen,blocked_by_same,blocked_by_low_disp\n";

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
    compact(V, F);
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
    for (int iter = 0; iter < total_iter; ++iter) {
        std::printf("\n=== iter %d ===\n", iter);
        double g = 0, l = 0, pr = 0, t = 0;
        if (cvt_job_profile(job, 1, &g, &l, &pr, &t) != CVT_OK) return fail_msg("step");
        cvt_iter_stats st;
        if (cvt_job_stats(job, 1, &st) != CVT_OK) return fail_msg("stats");
        total_ms += t;
        std::printf("grid_ms %.3f lloyd_ms %.3f precise_ms %.3f iter_remesh_ms %.3f (total_remesh_ms %.3f)\n", g, l,
                    pr, t, total_ms);
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
        final_nf = nFr;
        char path[1024];
        std::snprintf(path, sizeof path, "%s/iter_%03d_sites.xyz", out_dir.c_str(), iter);
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
        const Quality q = quality(Sf, Fn);
        double dH = 0;
        if (cvt_job_hausdorff(job, 1e-6, &dH, nullptr) != CVT_OK) return fail_msg("hausdorff");
        // stage columns: the FP32 pass (frame, cell and projection kernels,
        // which do KNN, frame, cell, projection and freeze) is reported as
        // centroids_ms, the site grid rebuild as knn_sites_ms
        const int blocked_same = std::max(0, st.count_low_disp - st.count_both);
        const int blocked_low = std::max(0, st.count_same - st.count_both);
        std::ofstream out(eval_csv, std::ios::app);
        out << mesh_name << "," << root_dir << "," << iter << "," << q.Qmin << "," << q.Qavg << "," << q.theta_min << ","
            << q.theta_min_avg << "," << q.lt30 << "," << q.gt90 << "," << (float)dH << "," << (float)t << ","
            << (float)total_ms << "," << st.frozen_total << "," << nV << ","
            << (nV > 0 ? 100.f * (float)st.frozen_total / (float)nV : 0.f) << "," << (float)g << "," << 0.f << ","
            << 0.f << "," << (float)(l + pr) << "," << 0.f << "," << 0.f << "," << 0.f << "," << st.count_same << ","
            << st.count_low_disp << "," << st.count_both << "," << st.count_newly_frozen << "," << blocked_same << ","
            << blocked_low << "\n";
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
