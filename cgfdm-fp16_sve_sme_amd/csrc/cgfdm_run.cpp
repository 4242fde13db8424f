// cgfdm_run.cpp — command-line driver: the drop-in for the reference executable's
// main() -> run() -> propagate() (main.cpp:10-41, run.cpp:5-120, propagate.cpp:24-228) on one GPU.
//
// Reads params.json and station.json from the working directory (getParams.cpp:12-252,
// station.cpp:83-296), builds the model (flat or Gauss-hill grid, homogeneous or layered medium,
// coord.cpp / medium.cpp), runs the time loop through the C ABI and writes the reference's output
// files. The single-GPU run is split into the per-rank files the reference's MPI decomposition
// (PX x PY x PZ from params.json, init_grid.cpp:14-130) would have written, so its post-processing
// scripts read them unchanged:
//   {OUT}/station_mpi_X_Y_Z.bin      int32 XYZ[n][3] then float wave[n][NT][9] (station.cpp:382-405)
//   {OUT}/PGV_Z_mpi_X_Y_Z.bin        float [ny][nx][2] on the top ranks (PGV.cpp:103-111)
//   {OUT}/V?_{it}_{X|Y|Z}_mpi_*.bin  velocity planes every IT_SKIP steps (data_io.cpp:434-495)
//   {OUT}/FreeSurfV?_{it}_Z_mpi_*.bin
//   {OUT}/coord?_*, Vs_*, Vp_*, rho_* model planes (data_io.cpp:497-526)
//
// Usage: cgfdm_run [--params params.json] [--stations station.json] [--precision fp32|fp16|bf16]
//                  [--device N] [--model-only] [--chi 2lambda+mu|lambda+2mu] [--scales reference|auto]
// --model-only writes the model planes and stops before any device call.
// --chi selects the coefficient of the free-surface matrices: the reference's 2*lambda + mu
// (contravariant.cpp:266, the default, for parity) or lambda + 2*mu, the traction-free condition;
// the run prints which one it uses.
// --scales (16-bit storage) keeps the reference's fixed scale factors (the default) or chooses them
// from the source (cgfdm_source_scales, with the medium's largest shear modulus), so a strong
// source stays inside the FP16 range.
// Exit status follows the reference's abort codes where a status can carry them: 100 no
// params.json, 101 unparseable params.json (reference 1001), 110 CFL violation, 180 no source.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/cgfdm.h"

namespace {

// ------------------------------------------------------------------ minimal JSON
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // insertion order kept (station order)

    // case-insensitive lookup, like cJSON_GetObjectItem (cjson.c:39, :500)
    const Json* get(const std::string& key) const {
        for (auto& kv : obj) {
            if (kv.first.size() != key.size()) continue;
            bool eq = true;
            for (size_t q = 0; q < key.size() && eq; ++q)
                eq = std::tolower((unsigned char)kv.first[q]) == std::tolower((unsigned char)key[q]);
            if (eq) return &kv.second;
        }
        return nullptr;
    }
};

struct Parser {
    const std::string& s;
    size_t p = 0;
    std::string err;
    explicit Parser(const std::string& t) : s(t) {}
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    }
    bool parse(Json& out) {
        ws();
        if (p >= s.size()) return fail("unexpected end");
        char c = s[p];
        if (c == '{') {
            out.kind = Json::Obj;
            ++p;
            ws();
            if (p < s.size() && s[p] == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string key;
                if (!str(key)) return false;
                ws();
                if (p >= s.size() || s[p] != ':') return fail("expected ':'");
                ++p;
                Json v;
                if (!parse(v)) return false;
                out.obj.emplace_back(key, std::move(v));
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            out.kind = Json::Arr;
            ++p;
            ws();
            if (p < s.size() && s[p] == ']') { ++p; return true; }
            for (;;) {
                Json v;
                if (!parse(v)) return false;
                out.arr.push_back(std::move(v));
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            out.kind = Json::Str;
            return str(out.str);
        }
        if (s.compare(p, 4, "true") == 0) { out.kind = Json::Bool; out.b = true; p += 4; return true; }
        if (s.compare(p, 5, "false") == 0) { out.kind = Json::Bool; out.b = false; p += 5; return true; }
        if (s.compare(p, 4, "null") == 0) { out.kind = Json::Null; p += 4; return true; }
        const char* b = s.c_str() + p;
        char* e = nullptr;
        errno = 0;
        double v = std::strtod(b, &e);
        if (e == b) return fail("bad value");
        out.kind = Json::Num;
        out.num = v;
        p += (size_t)(e - b);
        return true;
    }
    bool str(std::string& out) {
        if (p >= s.size() || s[p] != '"') return fail("expected string");
        ++p;
        while (p < s.size() && s[p] != '"') {
            if (s[p] == '\\' && p + 1 < s.size()) {
                char e = s[p + 1];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e;
                p += 2;
            } else {
                out += s[p++];
            }
        }
        if (p >= s.size()) return fail("unterminated string");
        ++p;
        return true;
    }
    bool fail(const char* m) {
        err = std::string(m) + " at offset " + std::to_string(p);
        return false;
    }
};

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

double num(const Json& o, const char* k, double def) {
    const Json* v = o.get(k);
    return (v && v->kind == Json::Num) ? v->num : def;
}
int inum(const Json& o, const char* k, int def) {
    const Json* v = o.get(k);
    return (v && v->kind == Json::Num) ? (int)v->num : def;  // cJSON valueint truncates
}
std::string sval(const Json& o, const char* k, const std::string& def) {
    const Json* v = o.get(k);
    return (v && v->kind == Json::Str) ? v->str : def;
}

// ------------------------------------------------------------------ params
struct Params {
    double TMAX = 0, DT = 0, DH = 0;
    int NX = 0, NY = 0, NZ = 0, PX = 1, PY = 1, PZ = 1;
    int centerX = 0, centerY = 0, IT_SKIP = 1;
    int sliceX = 0, sliceY = 0, sliceZ = 0, sliceFreeSurf = 0;
    int nPML = 12;
    int gauss_hill = 0, useTerrain = 0, useMedium = 0, useMultiSource = 0, useSingleSource = 0;
    float rickerfc = 0;
    int LayeredModel = 0, ShenModel = 0, Crust_1Medel = 0;
    std::string LayeredFileName, OUT = ".";
    double Depth = 0;  // km
    int sourceX = 0, sourceY = 0, sourceZ = 0;
    std::string sourceFile, sourceDir = ".";
    double centerLongitude = 0, centerLatitude = 0;
};

// getParams (getParams.cpp:55-247)
// returns 0, or the reference's abort code: 100 missing file, 1001 unparseable
int load_params(const std::string& path, Params& P, std::string& err) {
    std::string text;
    if (!read_file(path, text)) { err = "There is not " + path + " file!"; return 100; }
    Json root;
    Parser ps(text);
    if (!ps.parse(root) || root.kind != Json::Obj) { err = "Can't parse json file! " + ps.err; return 1001; }
    P.TMAX = num(root, "TMAX", 0); P.DT = num(root, "DT", 0); P.DH = num(root, "DH", 0);
    P.NX = inum(root, "NX", 0); P.NY = inum(root, "NY", 0); P.NZ = inum(root, "NZ", 0);
    P.PX = inum(root, "PX", 1); P.PY = inum(root, "PY", 1); P.PZ = inum(root, "PZ", 1);
    P.centerX = inum(root, "centerX", 0); P.centerY = inum(root, "centerY", 0);
    P.IT_SKIP = inum(root, "IT_SKIP", 1);
    P.sliceX = inum(root, "sliceX", 0); P.sliceY = inum(root, "sliceY", 0); P.sliceZ = inum(root, "sliceZ", 0);
    P.sliceFreeSurf = inum(root, "sliceFreeSurf", 0);
    P.nPML = inum(root, "nPML", 12);
    P.gauss_hill = inum(root, "gauss_hill", 0);
    P.useTerrain = inum(root, "useTerrain", 0);
    P.useMedium = inum(root, "useMedium", 0);
    P.useMultiSource = inum(root, "useMultiSource", 0);
    P.LayeredModel = inum(root, "LayeredModel", 0);
    P.LayeredFileName = sval(root, "LayeredFileName", "");
    P.useSingleSource = inum(root, "useSingleSource(ricker)", 0);
    P.rickerfc = (float)num(root, "rickerfc", 0);
    P.ShenModel = inum(root, "ShenModel", 0);
    P.Crust_1Medel = inum(root, "Crust_1Medel", 0);
    P.OUT = sval(root, "OUT", ".");
    P.Depth = num(root, "Depth(km)", 0);
    P.sourceX = inum(root, "sourceX", 0); P.sourceY = inum(root, "sourceY", 0); P.sourceZ = inum(root, "sourceZ", 0);
    P.sourceFile = sval(root, "sourceFile", "");
    P.sourceDir = sval(root, "sourceDir", ".");
    P.centerLongitude = num(root, "centerLongitude", 0);
    P.centerLatitude = num(root, "centerLatitude", 0);
    if (!P.useTerrain) P.Depth = P.DH * P.NZ * 1e-3;
    if (P.IT_SKIP <= 0) P.IT_SKIP = 1;
    return 0;
}

// ------------------------------------------------------------------ rank partition
struct Part { int n, front; };
// init_grid.cpp:38-75: the first N % P ranks get one more point
Part part(int N, int P, int r) {
    int n = N / P, res = N % P;
    if (r < res) return {n + 1, r * (n + 1)};
    return {n, res * (n + 1) + (r - res) * n};
}

bool mkdir_p(const std::string& d) {
    std::string cur;
    for (size_t q = 0; q < d.size(); ++q) {
        cur += d[q];
        if (d[q] == '/' || q + 1 == d.size()) mkdir(cur.c_str(), 0755);
    }
    struct stat st;
    return stat(d.c_str(), &st) == 0;
}

bool write_bin(const std::string& path, const void* a, size_t n1, const void* b = nullptr, size_t n2 = 0) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(a, 1, n1, f) == n1;
    if (b && n2) ok = ok && std::fwrite(b, 1, n2, f) == n2;
    std::fclose(f);
    return ok;
}

std::string rank_name(int X, int Y, int Z) {
    return "_mpi_" + std::to_string(X) + "_" + std::to_string(Y) + "_" + std::to_string(Z) + ".bin";
}

// Writes the X/Y/Z planes of a field given as a full interior-plane sampler, split per rank
// (data2D_output_bin, data_io.cpp:323-431).
template <class F>
bool write_planes(const Params& P, const std::string& name, int sx, int sy, int sz, F&& plane) {
    // plane(axis, index) -> full plane (axis 0: [NZ][NY], 1: [NZ][NX], 2: [NY][NX])
    for (int axis = 0; axis < 3; ++axis) {
        const int idx = axis == 0 ? sx : (axis == 1 ? sy : sz);
        const int lim = axis == 0 ? P.NX : (axis == 1 ? P.NY : P.NZ);
        if (idx < 0 || idx >= lim) continue;
        std::vector<float> full = plane(axis, idx);
        const int nA = axis == 0 ? P.NY : P.NX;  // fast dimension of the plane
        for (int X = 0; X < P.PX; ++X)
            for (int Y = 0; Y < P.PY; ++Y)
                for (int Z = 0; Z < P.PZ; ++Z) {
                    Part px = part(P.NX, P.PX, X), py = part(P.NY, P.PY, Y), pz = part(P.NZ, P.PZ, Z);
                    const Part own = axis == 0 ? px : (axis == 1 ? py : pz);
                    if (idx < own.front || idx >= own.front + own.n) continue;
                    const Part a = axis == 0 ? py : px;  // fast
                    const Part b = axis == 2 ? py : pz;  // slow
                    std::vector<float> out((size_t)a.n * b.n);
                    for (int s = 0; s < b.n; ++s)
                        for (int f = 0; f < a.n; ++f)
                            out[(size_t)s * a.n + f] = full[(size_t)(b.front + s) * nA + a.front + f];
                    const char* ax = axis == 0 ? "_X" : (axis == 1 ? "_Y" : "_Z");
                    const std::string fn = name + ax + rank_name(X, Y, Z);
                    if (!write_bin(fn, out.data(), out.size() * sizeof(float))) return false;
                }
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    std::string paramsPath = "params.json", stationPath = "station.json", precision = "fp32";
    int device = 0;
    bool modelOnly = false;
    int chiMode = 0;
    bool autoScales = false;
    for (int a = 1; a < argc; ++a) {
        std::string s = argv[a];
        auto next = [&]() -> std::string { return a + 1 < argc ? argv[++a] : ""; };
        if (s == "--params") paramsPath = next();
        else if (s == "--stations") stationPath = next();
        else if (s == "--precision") precision = next();
        else if (s == "--device") device = std::atoi(next().c_str());
        else if (s == "--model-only") modelOnly = true;
        else if (s == "--chi") {
            const std::string v = next();
            if (v == "2lambda+mu") chiMode = 0;
            else if (v == "lambda+2mu") chiMode = 1;
            else { std::fprintf(stderr, "--chi takes 2lambda+mu or lambda+2mu\n"); return 2; }
        }
        else if (s == "--scales") {
            const std::string v = next();
            if (v == "reference") autoScales = false;
            else if (v == "auto") autoScales = true;
            else { std::fprintf(stderr, "--scales takes reference or auto\n"); return 2; }
        }
        else { std::fprintf(stderr, "unknown argument %s\n", s.c_str()); return 2; }
    }
    Params P;
    std::string err;
    if (int code = load_params(paramsPath, P, err)) { std::fprintf(stderr, "%s\n", err.c_str()); return code == 1001 ? 101 : code; }
    if (P.useTerrain) { std::fprintf(stderr, "useTerrain (SRTM + PROJ) is not supported; use gauss_hill or a flat grid\n"); return 1; }
    if (P.useMedium && (P.ShenModel || P.Crust_1Medel)) { std::fprintf(stderr, "only homogeneous and layered media are supported\n"); return 1; }
    if (!P.useSingleSource && !P.useMultiSource) { std::fprintf(stderr, "No source configured (modelChecking)\n"); return 180; }
    if (P.useSingleSource && P.useMultiSource) { std::fprintf(stderr, "useSingleSource together with useMultiSource is not supported\n"); return 1; }

    const int NX = P.NX, NY = P.NY, NZ = P.NZ;
    const int nx_ = NX + 6, ny_ = NY + 6;
    const float DTf = (float)P.DT;
    const int NT = (int)(P.TMAX / DTf);  // propagate.cpp:35-36
    cgfdm_desc d{};
    d.NX = NX; d.NY = NY; d.NZ = NZ; d.nPML = P.nPML;
    d.DH = P.DH; d.DT = P.DT; d.NT = NT;
    d.precision = precision == "fp16" ? CGFDM_FP16 : precision == "bf16" ? CGFDM_BF16 : CGFDM_FP32;
    d.free_surface = 1; d.pml = 1;
    d.use_terrain = P.gauss_hill ? 1 : 0;
    d.depth = P.Depth * 1000.0;
    d.centerX = P.centerX; d.centerY = P.centerY;
    d.check_cfl = 1;
    d.chi_mode = chiMode;

    // coordinates and terrain (coord.cpp:106-307)
    std::vector<double> terrain((size_t)nx_ * ny_, 0.0);
    const double D = std::fabs(d.depth);
    if (P.gauss_hill) {
        const double h = 0.2 * D, aa = 0.1 * D;
        for (int j = 0; j < ny_; ++j)
            for (int i = 0; i < nx_; ++i) {
                const double x = (i - 3) * P.DH - P.centerX * P.DH, y = (j - 3) * P.DH - P.centerY * P.DH;
                terrain[i + (size_t)j * nx_] = h * std::exp(-0.5 * (x * x / (aa * aa) + y * y / (aa * aa)));
            }
    }
    auto coordz = [&](int i, int j, int k) {  // padded indices
        if (P.gauss_hill) return -D + (terrain[i + (size_t)j * nx_] + D) / double(NZ - 1) * (k - 3);
        return (k - 3 + 1) * P.DH - NZ * P.DH;
    };
    // medium (medium.cpp:59-216), interior points
    const size_t nin = (size_t)NX * NY * NZ;
    std::vector<float> vs(nin, 3464.0f), vp(nin, 6000.0f), rho(nin, 2670.0f);
    if (P.useMedium && P.LayeredModel) {
        FILE* f = std::fopen(P.LayeredFileName.c_str(), "rb");
        if (!f) { std::fprintf(stderr, "Can't open Layered Model file %s\n", P.LayeredFileName.c_str()); return 1; }
        int layers = 0;
        if (std::fread(&layers, 4, 1, f) != 1 || layers < 1) { std::fclose(f); std::fprintf(stderr, "bad layer file\n"); return 1; }
        std::vector<float> H(layers), VS(layers), VP(layers), RH(layers);
        for (int F = 0; F < layers; ++F) {
            float rec[4];
            if (std::fread(rec, 4, 4, f) != 4) { std::fclose(f); std::fprintf(stderr, "bad layer file\n"); return 1; }
            H[F] = rec[0] * 1000.0f; VS[F] = rec[1] * 1000.f; VP[F] = rec[2] * 1000.f; RH[F] = rec[3] * 1000.f;
        }
        std::fclose(f);
        const float tol = 1.0f;
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j)
                for (int i = 0; i < NX; ++i) {
                    const size_t q = i + (size_t)j * NX + (size_t)k * NX * NY;
                    const double ter = P.gauss_hill ? terrain[(i + 3) + (size_t)(j + 3) * nx_] : 0.0;  // LayeredStructureTerrain
                    const float cz = (float)coordz(i + 3, j + 3, k + 3);
                    for (int F = 1; F < layers; ++F) {
                        const float h0 = (float)(-H[0] + ter), h1 = (float)(-H[F - 1] + ter), h2 = (float)(-H[F] + ter);
                        if (cz > h0 - tol) { vs[q] = VS[0]; vp[q] = VP[0]; rho[q] = RH[0]; }
                        if (h1 - tol > cz && cz > h2 - tol) { vs[q] = VS[F - 1]; vp[q] = VP[F - 1]; rho[q] = RH[F - 1]; }
                        if (h2 - tol > cz) { vs[q] = VS[F]; vp[q] = VP[F]; rho[q] = RH[F]; }
                    }
                }
    }
    // stations (station.cpp:83-296): global indices in file order, kept if inside the grid
    std::vector<int> st;
    {
        std::string text;
        if (read_file(stationPath, text)) {
            Json root;
            Parser ps(text);
            if (!ps.parse(root)) { std::fprintf(stderr, "Can't parse json file %s\n", stationPath.c_str()); return 1001; }
            const Json* arr = root.get("station(point)");
            if (arr) {
                auto add = [&](const Json& e) {
                    if (e.kind != Json::Arr || e.arr.size() != 3) return false;
                    int X = (int)e.arr[0].num, Y = (int)e.arr[1].num, Z = (int)e.arr[2].num;
                    if (X >= 0 && X < NX && Y >= 0 && Y < NY && Z >= 0 && Z < NZ) { st.push_back(X); st.push_back(Y); st.push_back(Z); }
                    return true;
                };
                if (arr->kind == Json::Obj) { for (auto& kv : arr->obj) if (!add(kv.second)) { std::fprintf(stderr, "station index count != 3\n"); return 1; } }
                else if (arr->kind == Json::Arr) { for (auto& e : arr->arr) if (!add(e)) { std::fprintf(stderr, "station index count != 3\n"); return 1; } }
            }
        } else {
            std::printf("There is not %s file!\n", stationPath.c_str());
        }
    }
    const int nst = (int)st.size() / 3;

    mkdir_p(P.OUT);
    // model planes (data2D_Model_out, data_io.cpp:497-526)
    {
        auto cplane = [&](int comp) {
            return [&, comp](int axis, int idx) {
                std::vector<float> out;
                const int n1 = axis == 0 ? NY : NX, n2 = axis == 2 ? NY : NZ;
                out.resize((size_t)n1 * n2);
                for (int s = 0; s < n2; ++s)
                    for (int f = 0; f < n1; ++f) {
                        int i = axis == 0 ? idx : f, j = axis == 0 ? f : (axis == 1 ? idx : s), k = axis == 2 ? idx : s;
                        double c = comp == 0 ? (i - P.centerX) * P.DH : comp == 1 ? (j - P.centerY) * P.DH : coordz(i + 3, j + 3, k + 3);
                        out[(size_t)s * n1 + f] = (float)c;
                    }
                return out;
            };
        };
        auto mplane = [&](const std::vector<float>& a) {
            return [&](int axis, int idx) {
                const int n1 = axis == 0 ? NY : NX, n2 = axis == 2 ? NY : NZ;
                std::vector<float> out((size_t)n1 * n2);
                for (int s = 0; s < n2; ++s)
                    for (int f = 0; f < n1; ++f) {
                        int i = axis == 0 ? idx : f, j = axis == 0 ? f : (axis == 1 ? idx : s), k = axis == 2 ? idx : s;
                        out[(size_t)s * n1 + f] = a[i + (size_t)j * NX + (size_t)k * NX * NY];
                    }
                return out;
            };
        };
        write_planes(P, P.OUT + "/coordX", P.sliceX, P.sliceY, P.sliceZ, cplane(0));
        write_planes(P, P.OUT + "/coordY", P.sliceX, P.sliceY, P.sliceZ, cplane(1));
        write_planes(P, P.OUT + "/coordZ", P.sliceX, P.sliceY, P.sliceZ, cplane(2));
        write_planes(P, P.OUT + "/Vs", P.sliceX, P.sliceY, P.sliceZ, mplane(vs));
        write_planes(P, P.OUT + "/Vp", P.sliceX, P.sliceY, P.sliceZ, mplane(vp));
        write_planes(P, P.OUT + "/rho", P.sliceX, P.sliceY, P.sliceZ, mplane(rho));
    }

    if (modelOnly) {
        std::printf("model written to %s (%d stations)\n", P.OUT.c_str(), nst);
        return 0;
    }
    std::printf("free surface: chi = %s\n", chiMode ? "lambda + 2 mu (traction-free, --chi lambda+2mu)"
                                                   : "2 lambda + mu (the reference's, contravariant.cpp:266)");
    int rc = 0;
    // init_MultiSource (multiSource.cpp:824-922): the finite-fault file mapped onto the grid (host
    // work, before the domain exists, so the scales can follow the source)
    long long nsrc = 0;
    int snt = 0;
    float sdt = 0.f;
    std::vector<int> sxyz;
    std::vector<float> srate;
    if (P.useMultiSource) {
        const std::string path = P.sourceDir + "/" + P.sourceFile;
        const double* tp = P.gauss_hill ? terrain.data() : nullptr;
        rc = cgfdm_finite_fault_map(path.c_str(), &d, tp, P.centerLongitude, P.centerLatitude, nullptr, nullptr, 0,
                                    &nsrc, &snt, &sdt);
        sxyz.resize((size_t)std::max<long long>(nsrc, 1) * 3);
        srate.resize((size_t)std::max<long long>(nsrc, 1) * snt * 6);
        if (!rc)
            rc = cgfdm_finite_fault_map(path.c_str(), &d, tp, P.centerLongitude, P.centerLatitude, sxyz.data(),
                                        srate.data(), nsrc, &nsrc, &snt, &sdt);
        if (rc) { std::fprintf(stderr, "source file: %s\n", cgfdm_last_error()); return 1; }
        std::printf("finite fault %s: %lld grid nodes, nt = %d, dt = %g\n", path.c_str(), nsrc, snt, sdt);
    }
    if (autoScales && d.precision != CGFDM_FP32) {
        double mu = 0.0;
        for (size_t q = 0; q < nin; ++q) mu = std::max(mu, (double)rho[q] * vs[q] * vs[q]);
        const bool table = P.useMultiSource && nsrc > 0;
        if (!table && P.useSingleSource) {
            // the medium at the source cell
            const size_t q = (size_t)P.sourceX + (size_t)P.sourceY * NX + (size_t)P.sourceZ * NX * NY;
            const double muSrc = q < nin ? (double)rho[q] * vs[q] * vs[q] : mu;
            rc = cgfdm_point_source_scales(d.DH, P.rickerfc, muSrc, mu, &d.Cv, &d.Cs);
        } else {
            rc = cgfdm_source_scales(d.DH, table ? sxyz.data() : nullptr, table ? srate.data() : nullptr,
                                     table ? nsrc : 0, table ? snt : 0, table ? sdt : 0.f, mu, &d.Cv, &d.Cs);
        }
        if (rc) {
            std::fprintf(stderr, "scales: %s\n", cgfdm_last_error());
            return 1;
        }
        std::printf("scales from the source: Cv = %.6g, Cs = %.6g\n", d.Cv, d.Cs);
    }
    cgfdm_domain* h = nullptr;
    rc = cgfdm_domain_create(&d, device, &h);
    if (rc) { std::fprintf(stderr, "create: %s\n", cgfdm_last_error()); return 1; }
    auto guard = std::unique_ptr<cgfdm_domain, int (*)(cgfdm_domain*)>(h, cgfdm_domain_destroy);
    rc = cgfdm_set_model(h, P.gauss_hill ? terrain.data() : nullptr, vs.data(), vp.data(), rho.data());
    if (rc == CGFDM_ECFL) { std::fprintf(stderr, "The parameters can't afford the CFL condition! %s\n", cgfdm_last_error()); return 110; }
    if (rc) { std::fprintf(stderr, "set_model: %s\n", cgfdm_last_error()); return 1; }
    if (P.useMultiSource && nsrc > 0 && (rc = cgfdm_set_source_table(h, sxyz.data(), srate.data(), nsrc, snt, sdt))) {
        std::fprintf(stderr, "source: %s\n", cgfdm_last_error());
        return 1;
    }
    if (P.useSingleSource) {
        rc = cgfdm_set_source_point(h, P.sourceX, P.sourceY, P.sourceZ, P.rickerfc);
        if (rc) { std::fprintf(stderr, "source: %s\n", cgfdm_last_error()); return 1; }
    }
    if (nst && (rc = cgfdm_set_stations(h, st.data(), nst))) { std::fprintf(stderr, "stations: %s\n", cgfdm_last_error()); return 1; }

    auto vplane = [&](int var) {
        return [&, var](int axis, int idx) {
            const int n1 = axis == 0 ? NY : NX, n2 = axis == 2 ? NY : NZ;
            std::vector<float> out((size_t)n1 * n2);
            if (cgfdm_get_slice(h, axis, idx, var, out.data())) std::fprintf(stderr, "slice: %s\n", cgfdm_last_error());
            return out;
        };
    };
    for (int it = 0; it < NT; ++it) {
        if ((rc = cgfdm_step(h, 1))) { std::fprintf(stderr, "step: %s\n", cgfdm_last_error()); return 1; }
        if (it % P.IT_SKIP == 0) {  // propagate.cpp:174-179
            const char* names[3] = {"Vx", "Vy", "Vz"};
            for (int v = 0; v < 3; ++v) {
                write_planes(P, P.OUT + "/" + names[v] + "_" + std::to_string(it), P.sliceX, P.sliceY, P.sliceZ, vplane(v));
                if (P.sliceFreeSurf)
                    write_planes(P, P.OUT + "/FreeSurf" + names[v] + "_" + std::to_string(it), -1, -1, NZ - 1, vplane(v));
            }
        }
        if (it % 10 == 0) { std::printf("it = %8d.\n", it); std::fflush(stdout); }
    }
    cgfdm_sync(h);
    // stations, split per rank in file order (stationWrite, station.cpp:382-405)
    if (nst) {
        std::vector<float> wave((size_t)nst * NT * 9);
        cgfdm_get_stations(h, wave.data());
        for (int X = 0; X < P.PX; ++X)
            for (int Y = 0; Y < P.PY; ++Y)
                for (int Z = 0; Z < P.PZ; ++Z) {
                    Part px = part(NX, P.PX, X), py = part(NY, P.PY, Y), pz = part(NZ, P.PZ, Z);
                    std::vector<int> xyz;
                    std::vector<float> w;
                    for (int s = 0; s < nst; ++s) {
                        int sx = st[3 * s], sy = st[3 * s + 1], sz = st[3 * s + 2];
                        if (sx < px.front || sx >= px.front + px.n || sy < py.front || sy >= py.front + py.n ||
                            sz < pz.front || sz >= pz.front + pz.n) continue;
                        xyz.insert(xyz.end(), {sx, sy, sz});
                        w.insert(w.end(), wave.begin() + (size_t)s * NT * 9, wave.begin() + (size_t)(s + 1) * NT * 9);
                    }
                    if (xyz.empty()) continue;
                    write_bin(P.OUT + "/station" + rank_name(X, Y, Z), xyz.data(), xyz.size() * 4, w.data(), w.size() * 4);
                }
    }
    // PGV on the top ranks (outputPgvData, PGV.cpp:103-111)
    {
        std::vector<float> pgv((size_t)NX * NY * 2);
        cgfdm_get_pgv(h, pgv.data());
        const int Z = P.PZ - 1;
        for (int X = 0; X < P.PX; ++X)
            for (int Y = 0; Y < P.PY; ++Y) {
                Part px = part(NX, P.PX, X), py = part(NY, P.PY, Y);
                std::vector<float> out((size_t)px.n * py.n * 2);
                for (int j = 0; j < py.n; ++j)
                    for (int i = 0; i < px.n; ++i)
                        for (int c = 0; c < 2; ++c)
                            out[((size_t)j * px.n + i) * 2 + c] = pgv[((size_t)(py.front + j) * NX + px.front + i) * 2 + c];
                write_bin(P.OUT + "/PGV_Z" + rank_name(X, Y, Z), out.data(), out.size() * 4);
            }
    }
    std::printf("done: %d steps, %d stations, output in %s\n", NT, nst, P.OUT.c_str());
    return 0;
}
