// host_check.cpp — the library's host-side C++ (problem layout, seeded
// generators, COLMAP and problem-dump I/O) and the CPU oracle, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile)
// and driven over the workloads' generators, the COLMAP fixture, round trips
// and corrupted COLMAP binaries. Exits non-zero on a failed check; the
// sanitizers abort on the first error they find.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/vba.h"
#include "../../include/vba_io.h"
#include "../../include/vba_synth.h"
#include "../../oracle/ba_oracle.h"
#include "../../visloc-rs_amd/csrc/problem.hpp"

namespace {

int failures = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            ++failures;                                                         \
        }                                                                       \
    } while (0)

void default_options(vba_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->max_iterations = 3;
    o->has_initial_lambda = 1;
    o->initial_lambda = 1e-4;
    o->lambda_increase_factor = 10.0;
    o->lambda_decrease_factor = 0.1;
    o->max_lambda = 1e12;
    o->min_lambda = 1e-9;
    o->step_tolerance = 1e-7;
    o->cost_tolerance = 1e-9;
    o->max_pcg_iterations = 50;
    o->pcg_relative_tolerance = 1e-4;
    o->pcg_absolute_tolerance = 1e-12;
}

// A generated problem through the model build, the support check and the
// oracle's solvers.
void check_problem(uint32_t kind, uint32_t cams, uint64_t lms, uint64_t seed) {
    vba_synth_params sp;
    vba_synth_default(kind, &sp);
    sp.num_cameras = cams;
    sp.num_landmarks = lms;
    sp.seed = seed;
    if (kind == VBA_SYNTH_LOOP || kind == VBA_SYNTH_CITY) sp.scene_scale = 0.2 * cams;
    vba_synth_problem* g = vba_synth_generate(&sp);
    CHECK(g != nullptr);
    if (!g) return;
    vba_problem p;
    vba_synth_view(g, &p);
    vba_options o;
    default_options(&o);
    o.robust_kernel = VBA_KERNEL_HUBER;
    o.robust_scale = 1.0;
    std::string reason;
    CHECK(vba::check_support(p, o, &reason) == VBA_OK);
    vba::Model m;
    uint64_t error_id = 0;
    std::string msg;
    CHECK(vba::build_model(p, &m, &error_id, &msg) == VBA_OK);
    std::vector<double> poses(7 * p.num_poses), landmarks(3 * p.num_landmarks);
    std::vector<vba_iteration_stats> stats(8);
    const uint32_t solvers[] = {VBA_SOLVER_DENSE, VBA_SOLVER_SPARSE, VBA_SOLVER_MATRIX_FREE};
    for (uint32_t ls : solvers) {
        if (ls != VBA_SOLVER_DENSE && kind == VBA_SYNTH_POSE_ONLY) continue;
        for (int32_t threads : {1, 3}) {
            o.linear_solver = ls;
            vba_result r;
            std::memset(&r, 0, sizeof(r));
            const int32_t st = oracle_optimize(&p, &o, poses.data(), landmarks.data(), stats.data(),
                                               (uint32_t)stats.size(), &r, threads);
            CHECK(st == VBA_OK);
            CHECK(r.final_cost <= r.initial_cost);
        }
    }
    vba_synth_free(g);
}

std::vector<char> slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
void spit(const std::string& path, const std::vector<char>& data) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    f.write(data.data(), (std::streamsize)data.size());
}

void check_io(const std::string& fixture, const std::string& tmp) {
    vba_model* text = nullptr;
    CHECK(vba_model_read_colmap(fixture.c_str(), 0, &text) == VBA_OK);
    if (!text) return;
    const std::string bin = tmp + "/bin", txt = tmp + "/txt", dump = tmp + "/problem.txt";
    mkdir(bin.c_str(), 0755);
    mkdir(txt.c_str(), 0755);
    CHECK(vba_model_write_colmap(text, bin.c_str(), 1) == VBA_OK);
    CHECK(vba_model_write_colmap(text, txt.c_str(), 0) == VBA_OK);
    vba_model *b = nullptr, *t = nullptr, *d = nullptr;
    CHECK(vba_model_read_colmap(bin.c_str(), 1, &b) == VBA_OK);
    CHECK(vba_model_read_colmap(txt.c_str(), 0, &t) == VBA_OK);
    vba_problem pa, pb;
    const double* w = nullptr;
    CHECK(vba_model_problem(text, &pa, &w) == VBA_OK);
    if (b) CHECK(vba_model_problem(b, &pb, &w) == VBA_OK && pb.num_observations == pa.num_observations);
    // The fixture has several cameras, which the problem dump refuses; the
    // dump round trip takes a generated single-camera problem instead.
    CHECK(vba_model_write_ba_problem(text, dump.c_str()) != VBA_OK);
    {
        vba_synth_params sp;
        vba_synth_default(VBA_SYNTH_LOOP, &sp);
        sp.num_cameras = 10;
        sp.num_landmarks = 400;
        vba_synth_problem* g = vba_synth_generate(&sp);
        CHECK(g != nullptr);
        vba_problem gp;
        vba_synth_view(g, &gp);
        std::vector<double> weights(gp.num_observations, 0.5);
        vba_model* gm = nullptr;
        CHECK(vba_model_from_problem(&gp, weights.data(), &gm) == VBA_OK);
        CHECK(vba_model_write_ba_problem(gm, dump.c_str()) == VBA_OK);
        if (gm) vba_model_free(gm);
        vba_synth_free(g);
    }
    CHECK(vba_model_read_ba_problem(dump.c_str(), &d) == VBA_OK);
    if (d) {
        CHECK(vba_model_keep_local_window(d, 3) == VBA_OK);
        vba_problem pd;
        CHECK(vba_model_problem(d, &pd, &w) == VBA_OK);
    }
    // corrupted binaries: truncations and flipped bytes are errors, never crashes
    std::mt19937 rng(7);
    for (const char* name : {"cameras.bin", "images.bin", "points3D.bin"}) {
        const std::vector<char> good = slurp(bin + "/" + name);
        CHECK(!good.empty());
        for (int trial = 0; trial < 60; ++trial) {
            std::vector<char> bad = good;
            if (trial % 2 == 0) {
                bad.resize(rng() % (good.size() + 1));
            } else {
                for (int k = 0; k < 4; ++k) bad[rng() % bad.size()] = (char)(rng() & 0xff);
            }
            spit(bin + "/" + name, bad);
            vba_model* x = nullptr;
            const int32_t st = vba_model_read_colmap(bin.c_str(), 1, &x);
            if (st == VBA_OK) {
                vba_problem px;
                (void)vba_model_problem(x, &px, &w);
                vba_model_free(x);
            }
        }
        spit(bin + "/" + name, good);
    }
    // a corrupted dump
    {
        const std::vector<char> good = slurp(dump);
        for (int trial = 0; trial < 40; ++trial) {
            std::vector<char> bad = good;
            bad.resize(rng() % (good.size() + 1));
            if (trial % 3 == 0 && !bad.empty()) bad[rng() % bad.size()] = (char)('0' + rng() % 40);
            spit(dump, bad);
            vba_model* x = nullptr;
            if (vba_model_read_ba_problem(dump.c_str(), &x) == VBA_OK) vba_model_free(x);
        }
    }
    for (vba_model* m : {text, b, t, d})
        if (m) vba_model_free(m);
}

// A small point elimination (oracle_reduced_eliminate) with extras, runs and
// aggregates.
void check_elimination() {
    const uint32_t n = 12, P = 40;
    std::vector<uint64_t> point_off{0};
    std::vector<uint32_t> point_nodes;
    std::mt19937 rng(3);
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t s = rng() % (n - 3), len = 2 + rng() % 3;
        for (uint32_t k = 0; k < len && s + k < n; ++k) point_nodes.push_back(s + k);
        point_off.push_back(point_nodes.size());
    }
    std::vector<uint32_t> edges;
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t bb = a + 1; bb < n && bb <= a + 4; ++bb) edges.insert(edges.end(), {a, bb});
    const uint32_t E = edges.size() / 2;
    std::vector<uint64_t> row_off(n + 1, 0), slot_off(n + 1, 0);
    for (uint32_t e = 0; e < E; ++e) ++row_off[edges[2 * e] + 1];
    for (uint32_t a = 0; a < n; ++a) row_off[a + 1] += row_off[a];
    std::vector<std::vector<uint32_t>> per(n);
    for (uint32_t p = 0; p < P; ++p)
        for (uint64_t k = point_off[p]; k < point_off[p + 1]; ++k) {
            per[point_nodes[k]].push_back(p);
            per[point_nodes[k]].push_back((uint32_t)(k - point_off[p]));
        }
    std::vector<uint32_t> slots;
    for (uint32_t a = 0; a < n; ++a) {
        slots.insert(slots.end(), per[a].begin(), per[a].end());
        slot_off[a + 1] = slots.size() / 2;
    }
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> hcp(18 * point_nodes.size()), inv(9 * P), scale(6 * n, 1.0), extra(36);
    for (double& v : hcp) v = u(rng);
    for (uint32_t p = 0; p < P; ++p)
        for (int i = 0; i < 3; ++i) inv[9 * p + 4 * i] = 1.0 + 0.1 * i;
    for (double& v : extra) v = u(rng);
    std::vector<uint8_t> has(P, 1);
    has[5] = 0;
    const uint32_t extra_edge = 1, run_edges[] = {0, 2, 3};
    vba_point_elimination el;
    std::memset(&el, 0, sizeof(el));
    el.structure_id = 9;
    el.num_nodes = n;
    el.num_edges = E;
    el.edges = edges.data();
    el.row_off = row_off.data();
    el.slot_off = slot_off.data();
    el.slots = slots.data();
    el.num_points = P;
    el.point_off = point_off.data();
    el.point_nodes = point_nodes.data();
    el.hcp = hcp.data();
    el.point_inverse = inv.data();
    el.has_inverse = has.data();
    el.scale = scale.data();
    el.num_extra = 1;
    el.extra_edge = &extra_edge;
    el.extra_blocks = extra.data();
    el.num_run_edges = 3;
    el.run_edges = run_edges;
    el.coarse_group = 4;
    el.coarse_nodes = n;
    std::vector<double> blocks(36 * E), runs(36 * 3), coarse(36 * 9);
    CHECK(oracle_reduced_eliminate(&el, blocks.data(), runs.data(), coarse.data()) == VBA_OK);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: host_check <colmap text fixture dir> <scratch dir>\n");
        return 2;
    }
    check_problem(VBA_SYNTH_CIRCLE, 8, 300, 1);
    check_problem(VBA_SYNTH_LOOP, 30, 1500, 2);
    check_problem(VBA_SYNTH_STEREO_WINDOW, 6, 300, 3);
    check_problem(VBA_SYNTH_POSE_ONLY, 1, 200, 4);
    check_problem(VBA_SYNTH_CITY, 12, 3000, 5);
    check_io(argv[1], argv[2]);
    check_elimination();
    std::printf("host_check: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
