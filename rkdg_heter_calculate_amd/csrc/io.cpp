// Output surface and configuration: the reference's Tecplot writer, a raw state dump,
// and its 'key value # comment' parser.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>
#include "internal.h"

namespace rkdg {

// Restates CConfig::parseConfigLine (src/config.cpp:57-112): leading blanks skipped,
// '#' starts a comment line, the key ends at ' ', '=' or TAB, the value starts after
// any of those and ends at one of ",; #\t".  Lines without a value are ignored.
bool config_parse(const char* path, std::vector<std::pair<std::string, std::string>>& items) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t first = line.find_first_not_of(" \t");
        if (first == std::string::npos || line[first] == '#') continue;
        size_t mid = line.find_first_of(" =\t", first);
        if (mid == std::string::npos) continue;
        std::string key = line.substr(first, mid - first);
        size_t vstart = line.find_first_not_of(" =\t", mid);
        if (vstart == std::string::npos) continue;
        size_t vend = line.find_first_of(",; #\t", vstart);
        if (vend == vstart) continue;
        std::string value = vend == std::string::npos ? line.substr(vstart) : line.substr(vstart, vend - vstart);
        bool replaced = false;
        for (auto& kv : items)
            if (kv.first == key) { kv.second = value; replaced = true; }
        if (!replaced) items.emplace_back(key, value);
    }
    return true;
}

}  // namespace rkdg

using namespace rkdg;

extern "C" {

int rkdg_config_get(const char* path, const char* key, char* out, int32_t outlen) {
    if (!path || !key || !out || outlen < 1) return RKDG_ERR_ARG;
    std::vector<std::pair<std::string, std::string>> items;
    if (!config_parse(path, items)) return RKDG_ERR_IO;
    out[0] = 0;
    for (auto& kv : items)
        if (kv.first == key) {
            if ((int)kv.second.size() + 1 > outlen) return RKDG_ERR_ARG;
            std::memcpy(out, kv.second.c_str(), kv.second.size() + 1);
        }
    return RKDG_OK;
}

// Format of CCUDARkdgSolver::outputSolution (src/cudarkdgsolver.cu:879-998): BLOCK
// data, each value followed by two blanks, a newline after index i with i%6 == 0, the
// default ostream precision, 1-based connectivity separated by four blanks.
int rkdg_write_plt(const char* path, const rkdg_grid* g, int32_t Np, const double* U, const int32_t* flags,
                   double gamma) {
    if (!path || !g || !U || Np < 1) return RKDG_ERR_ARG;
    std::ofstream f(path);
    if (!f) return RKDG_ERR_IO;
    // '\n' rather than std::endl: the same bytes without a flush per line
    const int nv = g->n_vertices, ne = g->n_elements;
    auto mean = [&](int v, int e) { return U[((size_t)v * Np) * ne + e]; };
    f << "TITLE=RKDG" << '\n';
    f << "VARIABLES=X , Y , rho , u , v , p, Ma , FLAG" << '\n';
    f << "ZONE T= T1 N= " << nv << " , E= " << ne << " , ZONETYPE=FETRIANGLE" << '\n';
    f << "DATAPACKING=BLOCK" << '\n';
    f << "VARLOCATION=([1-2]=NODAL,[3-8]=CELLCENTERED)" << '\n';
    f << "DT=(SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE)" << '\n';
    auto block = [&](int n, auto value) {
        for (int i = 0; i < n; ++i) {
            value(i);
            f << "  ";
            if (i % 6 == 0) f << '\n';
        }
    };
    block(nv, [&](int i) { f << g->vx[i]; });
    block(nv, [&](int i) { f << g->vy[i]; });
    f << '\n';
    block(ne, [&](int i) { f << mean(0, i); });
    f << '\n';
    block(ne, [&](int i) { f << mean(1, i) / mean(0, i); });
    f << '\n';
    block(ne, [&](int i) { f << mean(2, i) / mean(0, i); });
    f << '\n';
    auto prim = [&](int i, double& rho, double& u, double& v, double& p) {
        rho = mean(0, i); u = mean(1, i) / rho; v = mean(2, i) / rho;
        p = (gamma - 1) * (mean(3, i) - 0.5 * rho * (u * u + v * v));
    };
    block(ne, [&](int i) { double r, u, v, p; prim(i, r, u, v, p); f << p; });
    f << '\n';
    block(ne, [&](int i) {
        double r, u, v, p;
        prim(i, r, u, v, p);
        double a = std::sqrt(gamma * p / r);
        f << std::sqrt(u * u + v * v) / a;
    });
    f << '\n';
    block(ne, [&](int i) { f << (flags ? flags[i] : 1); });
    f << '\n';
    for (int i = 0; i < ne; ++i)
        f << g->tri[3 * i] + 1 << "    " << g->tri[3 * i + 1] + 1 << "    " << g->tri[3 * i + 2] + 1 << '\n';
    return f ? RKDG_OK : RKDG_ERR_IO;
}

int rkdg_convert_basis(int32_t degree, int32_t n, const double* ex, const double* ey, int32_t from_basis,
                       const double* U_in, double* U_out) {
    if (degree < 1 || degree > 3 || n < 1 || !ex || !ey || !U_in || !U_out) return RKDG_ERR_ARG;
    if (from_basis != RKDG_BASIS_ELEMENT && from_basis != RKDG_BASIS_NATIVE) return RKDG_ERR_ARG;
    const int Np = np_of(degree);
    parallel_for(n, [&](int a, int b) {
        double in[4 * MAXNP], o[4 * MAXNP];
        for (int e = a; e < b; ++e) {
            for (int v = 0; v < 4; ++v)
                for (int i = 0; i < Np; ++i) in[v * Np + i] = U_in[((size_t)v * Np + i) * n + e];
            if (from_basis == RKDG_BASIS_ELEMENT) element_to_native(degree, ex + 3 * e, ey + 3 * e, in, o);
            else native_to_element(degree, ex + 3 * e, ey + 3 * e, in, o);
            for (int v = 0; v < 4; ++v)
                for (int i = 0; i < Np; ++i) U_out[((size_t)v * Np + i) * n + e] = o[v * Np + i];
        }
    });
    return RKDG_OK;
}

static const char kMagic[8] = {'R', 'K', 'D', 'G', 'S', 'T', '0', '1'};

int rkdg_dump_state(const char* path, int32_t Np, int32_t n_elements, double t, const double* U) {
    if (!path || !U || Np < 1 || n_elements < 1) return RKDG_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return RKDG_ERR_IO;
    int32_t hdr[2] = {Np, n_elements};
    size_t n = (size_t)4 * Np * n_elements;
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(&t, sizeof t, 1, f) == 1 && std::fwrite(U, sizeof(double), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RKDG_OK : RKDG_ERR_IO;
}

int rkdg_load_state(const char* path, int32_t Np, int32_t n_elements, double* t, double* U) {
    if (!path || !U || Np < 1 || n_elements < 1) return RKDG_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return RKDG_ERR_IO;
    char mg[8];
    int32_t hdr[2];
    double tt;
    size_t n = (size_t)4 * Np * n_elements;
    bool ok = std::fread(mg, 1, 8, f) == 8 && std::memcmp(mg, kMagic, 8) == 0 && std::fread(hdr, sizeof hdr, 1, f) == 1 &&
              hdr[0] == Np && hdr[1] == n_elements && std::fread(&tt, sizeof tt, 1, f) == 1 &&
              std::fread(U, sizeof(double), n, f) == n;
    std::fclose(f);
    if (!ok) return RKDG_ERR_IO;
    if (t) *t = tt;
    return RKDG_OK;
}

}  // extern "C"
