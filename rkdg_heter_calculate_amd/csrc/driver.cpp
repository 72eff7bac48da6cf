// rkdg_solver: the reference's command-line surface over librkdg.
//
// Mirrors main() + CCUDARkdgSolver::run/runNext/runAfter/outputSolution +
// RkdgAdvance::advance (src/main.cpp:24-159, src/cudarkdgsolver.cu:93-368,860-1001,
// src/rkdgadvance.cpp:8-263) on one GPU, without MPI: read main.conf (same keys and
// defaults), the mesh named by its gridconf key and the partition file, set the
// free-stream initial state, advance, write the residual history and the .plt file.
//
//   rkdg_solver [main.conf] [--partition FILE] [--steps N | --to-time] [--degree P]
//               [--dump FILE] [--restart FILE] [--quiet]
//
// Without --steps/--to-time the run stops after 11 steps like the reference's loop
// (src/rkdgadvance.cpp:236-242); --to-time runs to the configured 'time'.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>
#include "../../include/rkdg.h"

namespace {

struct Config {
    // defaults of the CCUDARkdgSolver constructor (src/cudarkdgsolver.cu:3-32)
    std::string title = "Unknown case";
    double alpha_deg = 5.0, gamma = 1.4, mach = 0.4, cfl = 0.18, rhoref = 1.0, pref = 1.0, time = 15;
    char loghistory = 'Y';
    int printinterval = 1000, threadsperblock = 512, reductionthreads = 512;
    std::string gridconf = "input/mesh.conf", solutionfile = "output/solution.dat", logfile = "output/log.dat",
                residualfile = "output/residual.dat";
};

std::string get(const char* path, const char* key) {
    char buf[4096];
    if (rkdg_config_get(path, key, buf, sizeof buf) != RKDG_OK) return "";
    return buf;
}

bool load_config(const char* path, Config& c) {
    std::ifstream probe(path);
    if (!probe) return false;
    std::string v;
    if (!(v = get(path, "title")).empty()) c.title = v;
    if (!(v = get(path, "gamma")).empty()) c.gamma = std::atof(v.c_str());
    if (!(v = get(path, "alpha")).empty()) c.alpha_deg = std::atof(v.c_str());
    if (!(v = get(path, "mach")).empty()) c.mach = std::atof(v.c_str());
    if (!(v = get(path, "cfl")).empty()) c.cfl = std::atof(v.c_str());
    if (!(v = get(path, "rhoref")).empty()) c.rhoref = std::atof(v.c_str());
    if (!(v = get(path, "pref")).empty()) c.pref = std::atof(v.c_str());
    if (!(v = get(path, "time")).empty()) c.time = std::atof(v.c_str());
    if (!(v = get(path, "gridconf")).empty()) c.gridconf = v;
    if (!(v = get(path, "solutionfile")).empty()) c.solutionfile = v;
    if (!(v = get(path, "logfile")).empty()) c.logfile = v;
    if (!(v = get(path, "residualfile")).empty()) c.residualfile = v;
    if (!(v = get(path, "loghistory")).empty()) c.loghistory = (char)std::toupper((unsigned char)v[0]);
    if (!(v = get(path, "threadsperblock")).empty()) c.threadsperblock = std::atoi(v.c_str());
    if (!(v = get(path, "printinterval")).empty()) c.printinterval = std::abs(std::atoi(v.c_str()));
    if (!(v = get(path, "reductionthreads")).empty()) c.reductionthreads = std::atoi(v.c_str());
    return true;
}

std::string now() {
    std::time_t t = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&t));
    return buf;
}

int die(const char* what, int rc, rkdg_ctx* ctx = nullptr) {
    std::fprintf(stderr, "rkdg_solver: %s failed: %s (%s)\n", what, rkdg_strerror(rc),
                 ctx ? rkdg_last_error(ctx) : rkdg_last_error(nullptr));
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    const char* conf_path = "input/main.conf";  // src/main.cpp:56
    std::string partition = "input/ElemDepart.dat";  // src/unstructuredgrid.cpp:38
    std::string dump, restart;
    int steps = 11, degree = 2;
    bool to_time = false, quiet = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--partition" && i + 1 < argc) partition = argv[++i];
        else if (a == "--steps" && i + 1 < argc) steps = std::atoi(argv[++i]);
        else if (a == "--to-time") to_time = true;
        else if (a == "--degree" && i + 1 < argc) degree = std::atoi(argv[++i]);
        else if (a == "--dump" && i + 1 < argc) dump = argv[++i];
        else if (a == "--restart" && i + 1 < argc) restart = argv[++i];
        else if (a == "--quiet") quiet = true;
        else if (a[0] != '-') conf_path = argv[i];
        else { std::fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
    }
    auto t_start = std::chrono::steady_clock::now();
    Config c;
    if (!load_config(conf_path, c)) { std::fprintf(stderr, "rkdg_solver: cannot open %s\n", conf_path); return 1; }
    std::ofstream flog(c.logfile.c_str());
    if (flog) flog << now() << ": programs starts" << std::endl;

    rkdg_grid grid;
    std::ifstream pf(partition.c_str());
    int rc = rkdg_grid_read(c.gridconf.c_str(), pf ? partition.c_str() : nullptr, 15.0, &grid);
    if (rc) return die("reading the mesh", rc);
    if (flog) flog << now() << ": grid read, " << grid.n_elements << " triangles" << std::endl;

    rkdg_params prm;
    rkdg_default_params(&prm);
    prm.degree = degree;
    prm.gamma = c.gamma;
    prm.cfl = c.cfl;
    prm.reduction_threads = c.reductionthreads;
    prm.dt_policy = RKDG_DT_REFERENCE;
    // far-field state of src/rkdgadvance.cpp:31-34 (alpha converted as initConfig does)
    double alpha = c.alpha_deg * std::atan(1.0) * 4 / 180;
    double ut = std::sqrt(c.gamma * c.pref / c.rhoref) * c.mach;
    prm.freestream[0] = c.rhoref;
    prm.freestream[1] = c.rhoref * ut * std::cos(alpha);
    prm.freestream[2] = c.rhoref * ut * std::sin(alpha);
    prm.freestream[3] = 0.5 * c.rhoref * (ut * ut) + c.pref / (c.gamma - 1);

    rkdg_mesh mesh;
    rkdg_grid_as_mesh(&grid, &mesh);
    std::vector<int32_t> order(grid.n_elements);
    if (rkdg_grid_locality_order(&grid, order.data()) == RKDG_OK) mesh.order = order.data();
    rkdg_ctx* ctx = nullptr;
    rc = rkdg_create(&mesh, &prm, &ctx);
    if (rc) return die("rkdg_create", rc);
    const int Np = (degree + 1) * (degree + 2) / 2;
    const int E = grid.n_elements;
    std::vector<double> U((size_t)4 * Np * E);
    double t0 = 0;
    if (!restart.empty()) {
        rc = rkdg_load_state(restart.c_str(), Np, E, &t0, U.data());
        if (rc) return die("reading the restart file", rc);
        rc = rkdg_set_state(ctx, RKDG_BASIS_ELEMENT, U.data());
    } else {
        // CCUDARkdgSolver::initRKDG (src/cudarkdgsolver.cu:326-368)
        double p[5] = {c.gamma, alpha, c.mach, c.rhoref, c.pref};
        rc = rkdg_project_ic(ctx, RKDG_IC_FREESTREAM, p, 5);
    }
    if (rc) return die("initial state", rc, ctx);
    if (!quiet) {
        std::cout << "====================\nProgram configures: \n====================\n"
                  << "gamma:              " << c.gamma << "\nalpha:              " << c.alpha_deg
                  << "\nmach:               " << c.mach << "\ncfl:                " << c.cfl
                  << "\nrhoref:             " << c.rhoref << "\npref:               " << c.pref
                  << "\ntime:               " << c.time << "\n====================\n";
    }

    std::ofstream fres;
    bool hist = c.loghistory == 'Y';
    if (hist) {
        fres.open(c.residualfile.c_str());
        if (fres) fres << "N, rho" << std::endl;
    }
    auto t_loop = std::chrono::steady_clock::now();
    int count = 0;
    double rhoE_inf = prm.freestream[3];
    if (to_time) {
        // run to the configured end time, writing the residual every step
        double t = 0;
        while (t < c.time) {
            int done = 0;
            rc = rkdg_advance(ctx, 1, c.time, &done, &t);
            if (rc) return die("advance", rc, ctx);
            if (!done) break;
            ++count;
            double res[2];
            rkdg_residual(ctx, res);
            if (hist && fres) fres << count << "   " << std::log(res[0] / c.rhoref) / std::log(10.0) << std::endl;
            if (!quiet && (count - 1) % c.printinterval == 0)
                std::cout << "Step: " << count << ", time: " << t << ", rhomax: " << res[0] / c.rhoref
                          << "   E: " << res[1] / rhoE_inf << std::endl;
        }
    } else {
        // the reference loop: step, and log the previous step's residual with the
        // current count (src/rkdgadvance.cpp:91-256)
        double res[2] = {0, 0};
        for (count = 1; count <= steps; ++count) {
            double dt;
            rc = rkdg_step(ctx, &dt);
            if (rc) return die("step", rc, ctx);
            if (hist && count > 1 && fres) fres << count << "   " << std::log(res[0] / c.rhoref) / std::log(10.0) << std::endl;
            rkdg_residual(ctx, res);
            if (!quiet && (count - 1) % c.printinterval == 0)
                std::cout << "Step: " << count << ", time step: " << dt << std::endl;
        }
        --count;
        if (hist && fres) fres << count << "   " << std::log(res[0] / c.rhoref) / std::log(10.0) << std::endl;
    }
    rkdg_sync(ctx);
    auto t_end = std::chrono::steady_clock::now();

    rc = rkdg_get_state(ctx, RKDG_BASIS_ELEMENT, U.data());
    if (rc) return die("get_state", rc, ctx);
    std::vector<int32_t> flags(E);
    rkdg_get_flags(ctx, flags.data());
    double t_now = 0;
    int64_t nsteps = 0;
    rkdg_time(ctx, &t_now, &nsteps);
    rc = rkdg_write_plt(c.solutionfile.c_str(), &grid, Np, U.data(), prm.limiter ? flags.data() : nullptr, c.gamma);
    if (rc) std::fprintf(stderr, "Failed to open solution file: %s and output will be omitted.\n", c.solutionfile.c_str());
    if (!dump.empty()) {
        rc = rkdg_dump_state(dump.c_str(), Np, E, t0 + t_now, U.data());
        if (rc) return die("state dump", rc);
    }
    double loop_s = std::chrono::duration<double>(t_end - t_loop).count();
    double wall_s = std::chrono::duration<double>(t_end - t_start).count();
    if (flog) flog << now() << ": complete solving flow, " << count << " steps, t = " << t_now << std::endl;
    std::cout << "steps: " << count << "  simulated time: " << t_now << "  loop wall time: " << loop_s
              << " s  total wall time: " << wall_s << " s" << std::endl;
    rkdg_destroy(ctx);
    rkdg_grid_free(&grid);
    return 0;
}
