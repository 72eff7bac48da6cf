// Mesh input: split-quad generator, the reference mesh format, O(N) connectivity and a
// locality order.  Replaces the reference's readers and its O(N^2) neighbour/ghost
// set-up (src/unstructuredgrid.cpp:14-80,88-107,193-305,502-643,729-760,934-976).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "internal.h"

namespace rkdg {

namespace {

// splitmix64 (Steele, Lea, Flood 2014): the generator behind every seeded draw
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

template <class T>
T* dup(const std::vector<T>& v) {
    T* p = (T*)std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

// Half-edge matching by sorting (key = sorted canonical vertex pair).
int match_edges(int E, const int32_t* tri, const std::vector<int64_t>& canon, int32_t* nbr, int32_t* se,
                std::string& err) {
    std::vector<std::pair<uint64_t, int64_t>> h((size_t)3 * E);
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) {
            uint64_t a = (uint64_t)canon[tri[3 * e + (k + 1) % 3]], b = (uint64_t)canon[tri[3 * e + (k + 2) % 3]];
            if (a > b) std::swap(a, b);
            h[(size_t)3 * e + k] = {(a << 32) | b, (int64_t)3 * e + k};
        }
    std::sort(h.begin(), h.end());
    for (size_t i = 0; i < h.size();) {
        size_t j = i + 1;
        while (j < h.size() && h[j].first == h[i].first) ++j;
        if (j - i == 1) {
            nbr[h[i].second] = -1;
            se[h[i].second] = 0;
        } else if (j - i == 2) {
            int64_t p = h[i].second, q = h[i + 1].second;
            nbr[p] = (int32_t)(q / 3); se[p] = (int32_t)(q % 3);
            nbr[q] = (int32_t)(p / 3); se[q] = (int32_t)(p % 3);
        } else {
            err = "non-manifold edge shared by more than two triangles";
            return RKDG_ERR_MESH;
        }
        i = j;
    }
    return RKDG_OK;
}

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// Reads whitespace-separated numbers, skipping any line that does not start with a
// number (the reference's files carry one '#...' header line; ElemNeigh's header is
// not skipped by the reference reader, src/unstructuredgrid.cpp:220).
template <class T>
bool read_numbers(const std::string& path, size_t count, std::vector<T>& out) {
    std::ifstream f(path.c_str());
    if (!f) return false;
    out.clear();
    out.reserve(count);
    std::string line;
    while (out.size() < count && std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty()) continue;
        char c = t[0];
        if (!(std::isdigit((unsigned char)c) || c == '-' || c == '+' || c == '.')) continue;
        std::istringstream is(t);
        T v;
        while (out.size() < count && (is >> v)) out.push_back(v);
    }
    return out.size() == count;
}

}  // namespace

int build_shared_edges(int E, const int32_t* tri, const int32_t* nbr, int32_t* se, std::string& err, const double* vx,
                       const double* vy) {
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) {
            int n = nbr[3 * e + k];
            se[3 * e + k] = 0;
            if (n < 0) continue;
            if (n >= E) { err = "neighbour id out of range"; return RKDG_ERR_MESH; }
            int a = tri[3 * e + (k + 1) % 3], b = tri[3 * e + (k + 2) % 3];
            int found = -1;
            for (int kk = 0; kk < 3; ++kk) {
                int c = tri[3 * n + (kk + 1) % 3], d = tri[3 * n + (kk + 2) % 3];
                if (c == b && d == a) found = kk;
            }
            if (found < 0 && vx && vy) {
                // a periodic seam: the two sides carry different vertex ids; the shared edge
                // is the neighbour's edge that points back to this element and is this edge
                // moved by one translation (its end b maps to the neighbour's start c, a to d)
                const double lx = vx[b] - vx[a], ly = vy[b] - vy[a];
                const double tol = 1e-9 * std::sqrt(lx * lx + ly * ly);
                for (int kk = 0; kk < 3; ++kk) {
                    if (nbr[3 * n + kk] != e) continue;
                    int c = tri[3 * n + (kk + 1) % 3], d = tri[3 * n + (kk + 2) % 3];
                    const double tx = vx[c] - vx[b], ty = vy[c] - vy[b];
                    if (std::fabs(vx[d] - vx[a] - tx) <= tol && std::fabs(vy[d] - vy[a] - ty) <= tol &&
                        std::fabs(vx[d] - vx[c] + lx) <= tol && std::fabs(vy[d] - vy[c] + ly) <= tol)
                        found = kk;
                }
            }
            if (found < 0) {
                char buf[160];
                std::snprintf(buf, sizeof buf, "edge relation error: triangle %d edge %d is not an edge of neighbour %d",
                              e, k, n);
                err = buf;
                return RKDG_ERR_MESH;
            }
            se[3 * e + k] = found;
        }
    return RKDG_OK;
}

}  // namespace rkdg

using namespace rkdg;

static thread_local std::string g_grid_err;

namespace rkdg {

void morton_order(int n, const double* bx, const double* by, int32_t* order) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int i = 0; i < n; ++i) {
        xmin = std::min(xmin, bx[i]); xmax = std::max(xmax, bx[i]);
        ymin = std::min(ymin, by[i]); ymax = std::max(ymax, by[i]);
    }
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int i = 0; i < n; ++i) {
        uint32_t ix = (uint32_t)std::min(65535.0, std::max(0.0, (bx[i] - xmin) / (xmax - xmin + 1e-300) * 65535.0));
        uint32_t iy = (uint32_t)std::min(65535.0, std::max(0.0, (by[i] - ymin) / (ymax - ymin + 1e-300) * 65535.0));
        uint64_t m = 0;
        for (int b = 0; b < 16; ++b) m |= (uint64_t)((ix >> b) & 1) << (2 * b) | (uint64_t)((iy >> b) & 1) << (2 * b + 1);
        key[i] = {m, i};
    }
    std::stable_sort(key.begin(), key.end());
    for (int i = 0; i < n; ++i) order[i] = key[i].second;
}

}  // namespace rkdg

extern "C" {

int rkdg_grid_splitquad(int32_t nx, int32_t ny, double x0, double y0, double lx, double ly, int32_t periodic,
                        int32_t diagonal, double jitter, uint64_t seed, int32_t bc, rkdg_grid* out) {
    if (!out || nx < 1 || ny < 1 || lx <= 0 || ly <= 0 || jitter < 0 || jitter >= 0.25) return RKDG_ERR_ARG;
    if ((int64_t)nx * ny * 2 > (1 << 29)) return RKDG_ERR_ARG;
    std::memset(out, 0, sizeof *out);
    const int nvx = nx + 1, nvy = ny + 1;
    const int NV = nvx * nvy, E = 2 * nx * ny;
    double hx = lx / nx, hy = ly / ny;
    std::vector<double> vx(NV), vy(NV);
    SplitMix64 rj{seed * 2 + 1}, rd{seed * 2 + 2};
    for (int j = 0; j < nvy; ++j)
        for (int i = 0; i < nvx; ++i) {
            int id = j * nvx + i;
            vx[id] = x0 + i * hx;
            vy[id] = y0 + j * hy;
            if (jitter > 0 && i > 0 && i < nx && j > 0 && j < ny) {
                vx[id] += (2 * rj.uniform() - 1) * jitter * hx;
                vy[id] += (2 * rj.uniform() - 1) * jitter * hy;
            }
        }
    std::vector<int32_t> tri((size_t)3 * E);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int q = j * nx + i;
            int v00 = j * nvx + i, v10 = v00 + 1, v01 = v00 + nvx, v11 = v01 + 1;
            int d = diagonal ? (int)(rd.next() >> 63) : 0;
            int32_t* t = &tri[(size_t)6 * q];
            if (d == 0) { t[0] = v00; t[1] = v10; t[2] = v11; t[3] = v00; t[4] = v11; t[5] = v01; }
            else { t[0] = v00; t[1] = v10; t[2] = v01; t[3] = v10; t[4] = v11; t[5] = v01; }
        }
    std::vector<int64_t> canon(NV);
    for (int j = 0; j < nvy; ++j)
        for (int i = 0; i < nvx; ++i)
            canon[j * nvx + i] = periodic ? (int64_t)(j % ny) * nx + (i % nx) : (int64_t)j * nvx + i;
    std::vector<int32_t> nbr((size_t)3 * E), se((size_t)3 * E), bnd((size_t)3 * E, 0);
    int rc = match_edges(E, tri.data(), canon, nbr.data(), se.data(), g_grid_err);
    if (rc) return rc;
    for (size_t h = 0; h < nbr.size(); ++h)
        if (nbr[h] < 0) bnd[h] = bc ? bc : RKDG_BC_OUTFLOW;
    out->n_vertices = NV;
    out->n_elements = E;
    out->vx = dup(vx); out->vy = dup(vy);
    out->tri = dup(tri); out->neighbour = dup(nbr); out->shared_edge = dup(se); out->boundary = dup(bnd);
    out->partition = nullptr;
    return RKDG_OK;
}

int rkdg_grid_read(const char* meshconf, const char* partition_file, double box, rkdg_grid* out) {
    if (!meshconf || !out) return RKDG_ERR_ARG;
    std::memset(out, 0, sizeof *out);
    std::vector<std::pair<std::string, std::string>> items;
    if (!config_parse(meshconf, items)) { g_grid_err = std::string("cannot open ") + meshconf; return RKDG_ERR_IO; }
    auto get = [&](const char* k) -> std::string {
        for (auto& kv : items) if (kv.first == k) return kv.second;
        return "";
    };
    int E = std::atoi(get("trianglenum").c_str()), NV = std::atoi(get("verticenum").c_str());
    if (E <= 0 || NV <= 0) { g_grid_err = "trianglenum/verticenum missing"; return RKDG_ERR_IO; }
    // file names in mesh.conf are relative to the working directory, as in the reference
    std::vector<double> xy;
    std::vector<int32_t> tri, nbr;
    if (!read_numbers(get("vertice_filename"), (size_t)2 * NV, xy)) { g_grid_err = "vertex file"; return RKDG_ERR_IO; }
    if (!read_numbers(get("trivertice_filename"), (size_t)3 * E, tri)) { g_grid_err = "element file"; return RKDG_ERR_IO; }
    if (!read_numbers(get("trineigh_filename"), (size_t)3 * E, nbr)) { g_grid_err = "neighbour file"; return RKDG_ERR_IO; }
    std::vector<double> vx(NV), vy(NV);
    for (int i = 0; i < NV; ++i) { vx[i] = xy[2 * i]; vy[i] = xy[2 * i + 1]; }
    for (int i = 0; i < 3 * E; ++i)
        if (tri[i] < 0 || tri[i] >= NV || nbr[i] < -1 || nbr[i] >= E) { g_grid_err = "id out of range"; return RKDG_ERR_MESH; }
    std::vector<int32_t> se((size_t)3 * E), bnd((size_t)3 * E, 0);
    int rc = build_shared_edges(E, tri.data(), nbr.data(), se.data(), g_grid_err, vx.data(), vy.data());
    if (rc) return rc;
    // boundary tag from the barycentre of the mirrored (ghost) triangle, as
    // markLocalBoundaryTriangles (src/unstructuredgrid.cpp:729-760)
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) {
            if (nbr[3 * e + k] >= 0) continue;
            int a = tri[3 * e + k], b = tri[3 * e + (k + 1) % 3], c = tri[3 * e + (k + 2) % 3];
            double gx = vx[b] + vx[c] - vx[a], gy = vy[b] + vy[c] - vy[a];
            double bx = (gx + vx[b] + vx[c]) / 3.0, by = (gy + vy[b] + vy[c]) / 3.0;
            bool far = bx < -box || bx > box || by < -box || by > box;
            bnd[3 * e + k] = far ? RKDG_BC_FARFIELD : RKDG_BC_REFLECTION;
        }
    std::vector<int32_t> part;
    if (partition_file && partition_file[0]) {
        std::ifstream f(partition_file);
        if (!f) { g_grid_err = "partition file"; return RKDG_ERR_IO; }
        int v;
        while ((int)part.size() < E && (f >> v)) part.push_back(v);
        part.resize(E, part.empty() ? 0 : part.back());  // tolerate short files
    }
    out->n_vertices = NV;
    out->n_elements = E;
    out->vx = dup(vx); out->vy = dup(vy);
    out->tri = dup(tri); out->neighbour = dup(nbr); out->shared_edge = dup(se); out->boundary = dup(bnd);
    out->partition = part.empty() ? nullptr : dup(part);
    return RKDG_OK;
}

int rkdg_grid_write(const char* dir, const rkdg_grid* g) {
    if (!dir || !g || g->n_elements < 1 || g->n_vertices < 1 || !g->vx || !g->vy || !g->tri || !g->neighbour) {
        g_grid_err = "grid_write arguments";
        return RKDG_ERR_ARG;
    }
    const int E = g->n_elements, NV = g->n_vertices;
    // edge k of element e lies opposite vertex k; an edge takes its id where it is first met
    // (ascending element, then edge), and the neighbour across it reuses that id
    std::vector<int32_t> eid((size_t)3 * E, -1);
    std::vector<int32_t> ev;
    std::vector<int32_t> se((size_t)3 * E);
    if (g->shared_edge) std::memcpy(se.data(), g->shared_edge, sizeof(int32_t) * 3 * E);
    else if (int rc = build_shared_edges(E, g->tri, g->neighbour, se.data(), g_grid_err, g->vx, g->vy)) return rc;
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 3; ++k) {
            const int n = g->neighbour[3 * e + k];
            if (n >= E || (n >= 0 && (se[3 * e + k] < 0 || se[3 * e + k] > 2))) {
                g_grid_err = "neighbour or shared edge out of range";
                return RKDG_ERR_MESH;
            }
            if (n >= 0 && n < e) {
                eid[3 * e + k] = eid[3 * n + se[3 * e + k]];
                continue;
            }
            eid[3 * e + k] = (int32_t)(ev.size() / 2);
            ev.push_back(g->tri[3 * e + (k + 1) % 3]);
            ev.push_back(g->tri[3 * e + (k + 2) % 3]);
        }
    const int NE = (int)(ev.size() / 2);
    const std::string d(dir);
    auto path = [&](const char* f) { return d + "/" + f; };
    auto open = [&](const std::string& p, std::FILE*& f) {
        f = std::fopen(p.c_str(), "w");
        if (!f) g_grid_err = "cannot write " + p;
        return f != nullptr;
    };
    std::FILE* f = nullptr;
    // one header line each (the reference's readers skip it), except ElemNeigh, whose
    // reader does not (src/unstructuredgrid.cpp:220)
    if (!open(path("Vertice.dat"), f)) return RKDG_ERR_IO;
    std::fprintf(f, "#vertist data\n");
    for (int i = 0; i < NV; ++i) std::fprintf(f, "%.17g %.17g\n", g->vx[i], g->vy[i]);
    bool ok = std::fclose(f) == 0;
    if (!open(path("ElemVertice.dat"), f)) return RKDG_ERR_IO;
    std::fprintf(f, "#ElemVertex Data\n");
    for (int e = 0; e < E; ++e) std::fprintf(f, "%d %d %d\n", g->tri[3 * e], g->tri[3 * e + 1], g->tri[3 * e + 2]);
    ok = std::fclose(f) == 0 && ok;
    if (!open(path("ElemNeigh.dat"), f)) return RKDG_ERR_IO;
    for (int e = 0; e < E; ++e)
        std::fprintf(f, "%d %d %d\n", g->neighbour[3 * e], g->neighbour[3 * e + 1], g->neighbour[3 * e + 2]);
    ok = std::fclose(f) == 0 && ok;
    if (!open(path("ElemEdge.dat"), f)) return RKDG_ERR_IO;
    std::fprintf(f, "#Elemside data\n");
    for (int e = 0; e < E; ++e) std::fprintf(f, "%d %d %d\n", eid[3 * e], eid[3 * e + 1], eid[3 * e + 2]);
    ok = std::fclose(f) == 0 && ok;
    if (!open(path("EdgeVertice.dat"), f)) return RKDG_ERR_IO;
    std::fprintf(f, "#Sidelist data\n");
    for (int i = 0; i < NE; ++i) std::fprintf(f, "%d %d\n", ev[2 * i], ev[2 * i + 1]);
    ok = std::fclose(f) == 0 && ok;
    if (g->partition) {
        // gridpartition's output: one part id per line, no header (src/unstructuredgrid.cpp:290-305)
        if (!open(path("ElemDepart.dat"), f)) return RKDG_ERR_IO;
        for (int e = 0; e < E; ++e) std::fprintf(f, "%d\n", g->partition[e]);
        ok = std::fclose(f) == 0 && ok;
    }
    // mesh.conf: the keys of CUnstructuredGrid::initializeGrid (src/unstructuredgrid.cpp:19-36);
    // file names as written here (relative to the working directory when dir is relative,
    // like the reference's input/ paths)
    if (!open(path("mesh.conf"), f)) return RKDG_ERR_IO;
    std::fprintf(f, "# mesh written by rkdg_grid_write\n");
    std::fprintf(f, "trianglenum %d\nverticenum %d\nedgenum %d\n", E, NV, NE);
    std::fprintf(f, "trivertice_filename %s\n", path("ElemVertice.dat").c_str());
    std::fprintf(f, "trineigh_filename %s\n", path("ElemNeigh.dat").c_str());
    std::fprintf(f, "triedge_filename %s\n", path("ElemEdge.dat").c_str());
    std::fprintf(f, "vertice_filename %s\n", path("Vertice.dat").c_str());
    std::fprintf(f, "edgevertice_filename %s\n", path("EdgeVertice.dat").c_str());
    std::fprintf(f, "output_filename %s\n", path("mesh.plt").c_str());
    ok = std::fclose(f) == 0 && ok;
    if (!ok) { g_grid_err = "write error in " + d; return RKDG_ERR_IO; }
    return RKDG_OK;
}

int rkdg_grid_locality_order(const rkdg_grid* g, int32_t* order) {
    if (!g || !order) return RKDG_ERR_ARG;
    const int E = g->n_elements;
    if (g->partition) {
        // gridpartition's parts, each in element order
        std::vector<std::pair<uint64_t, int32_t>> key(E);
        for (int e = 0; e < E; ++e) key[e] = {((uint64_t)(uint32_t)g->partition[e] << 32) | (uint64_t)e, e};
        std::stable_sort(key.begin(), key.end());
        for (int i = 0; i < E; ++i) order[i] = key[i].second;
        return RKDG_OK;
    }
    std::vector<double> bx(E), by(E);
    for (int e = 0; e < E; ++e) {
        double sx = 0, sy = 0;
        for (int k = 0; k < 3; ++k) { sx += g->vx[g->tri[3 * e + k]]; sy += g->vy[g->tri[3 * e + k]]; }
        bx[e] = sx / 3;
        by[e] = sy / 3;
    }
    rkdg::morton_order(E, bx.data(), by.data(), order);
    return RKDG_OK;
}

void rkdg_grid_free(rkdg_grid* g) {
    if (!g) return;
    std::free(g->vx); std::free(g->vy); std::free(g->tri); std::free(g->neighbour);
    std::free(g->shared_edge); std::free(g->boundary); std::free(g->partition);
    std::memset(g, 0, sizeof *g);
}

void rkdg_grid_as_mesh(const rkdg_grid* g, rkdg_mesh* m) {
    std::memset(m, 0, sizeof *m);
    m->n_vertices = g->n_vertices;
    m->n_elements = g->n_elements;
    m->vx = g->vx; m->vy = g->vy; m->tri = g->tri;
    m->neighbour = g->neighbour; m->shared_edge = g->shared_edge; m->boundary = g->boundary;
}

const char* rkdg_grid_last_error(void) { return g_grid_err.c_str(); }

}  // extern "C"
