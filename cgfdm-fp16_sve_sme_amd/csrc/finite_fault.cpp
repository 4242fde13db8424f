// finite_fault.cpp — host-side reader of the reference's finite-fault source file and its
// mapping onto the grid (reference multiSource.cpp, init_MultiSource and the functions it calls):
// read the points (readSourceInfo / readSourceLocation, :14-100), project lon/lat onto the grid
// plane (projTrans, :109-173), find each point's grid node (srcCoord2Index, :177-320), read the
// fault parameters of the points that land in the grid (readSourceParams, :365-413), form the
// moment-rate tensors (solveMomentRate, :478-545), merge points that share a node
// (dealDuplicateIndex, :551-607) and order the nodes (changeStorageOrder, :752-796). The result is
// the table cgfdm_set_source_table takes. No device work.
//
// The projection is the azimuthal equidistant projection on the WGS84 ellipsoid centred on
// (lon0, lat0), which the reference obtains from PROJ ("+proj=aeqd ... +ellps=WGS84"): a point
// maps to (s sin(az), s cos(az)) with s the geodesic distance from the centre and az the azimuth
// at the centre. PROJ solves the inverse geodesic problem with Karney's series; here it is solved
// with Vincenty's iteration (T. Vincenty, Survey Review 23(176), 88-93, 1975), which agrees to well
// below a millimetre at the distances of a simulation grid.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/cgfdm.h"

namespace cg {
void set_err(const char* fmt, ...);
}

namespace {

constexpr double kWgs84A = 6378137.0;
constexpr double kWgs84F = 1.0 / 298.257223563;

// Geodesic distance s (m) and forward azimuth az1 (radians, clockwise from north) from (phi1,
// lam1) to (phi2, lam2), radians, on WGS84 (Vincenty's inverse solution). Returns false if the
// iteration does not converge (nearly antipodal points).
bool vincenty_inverse(double phi1, double lam1, double phi2, double lam2, double& s, double& az1) {
    const double a = kWgs84A, f = kWgs84F, b = a * (1.0 - f);
    const double L = lam2 - lam1;
    const double U1 = std::atan((1.0 - f) * std::tan(phi1));
    const double U2 = std::atan((1.0 - f) * std::tan(phi2));
    const double sU1 = std::sin(U1), cU1 = std::cos(U1), sU2 = std::sin(U2), cU2 = std::cos(U2);
    double lam = L, sinS = 0, cosS = 1, sig = 0, cos2a = 1, cos2sm = 0;
    for (int it = 0; it < 200; ++it) {
        const double sl = std::sin(lam), cl = std::cos(lam);
        const double t1 = cU2 * sl, t2 = cU1 * sU2 - sU1 * cU2 * cl;
        sinS = std::sqrt(t1 * t1 + t2 * t2);
        if (sinS == 0.0) {  // coincident points
            s = 0.0;
            az1 = 0.0;
            return true;
        }
        cosS = sU1 * sU2 + cU1 * cU2 * cl;
        sig = std::atan2(sinS, cosS);
        const double sina = cU1 * cU2 * sl / sinS;
        cos2a = 1.0 - sina * sina;
        cos2sm = cos2a != 0.0 ? cosS - 2.0 * sU1 * sU2 / cos2a : 0.0;
        const double C = f / 16.0 * cos2a * (4.0 + f * (4.0 - 3.0 * cos2a));
        const double prev = lam;
        lam = L + (1.0 - C) * f * sina * (sig + C * sinS * (cos2sm + C * cosS * (-1.0 + 2.0 * cos2sm * cos2sm)));
        if (std::fabs(lam - prev) < 1e-13) {
            const double u2 = cos2a * (a * a - b * b) / (b * b);
            const double A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)));
            const double B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)));
            const double dsig =
                B * sinS *
                (cos2sm + B / 4.0 * (cosS * (-1.0 + 2.0 * cos2sm * cos2sm) -
                                     B / 6.0 * cos2sm * (-3.0 + 4.0 * sinS * sinS) * (-3.0 + 4.0 * cos2sm * cos2sm)));
            s = b * A * (sig - dsig);
            const double sl2 = std::sin(lam), cl2 = std::cos(lam);
            az1 = std::atan2(cU2 * sl2, cU1 * sU2 - sU1 * cU2 * cl2);
            return true;
        }
    }
    return false;
}

struct FaultFile {
    FILE* f = nullptr;
    ~FaultFile() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" int cgfdm_aeqd_forward(double lon0, double lat0, long long n, const double* lon, const double* lat,
                                  double* x, double* y) {
    if (n < 0 || (n > 0 && (!lon || !lat || !x || !y))) {
        cg::set_err("bad arguments");
        return CGFDM_EINVAL;
    }
    if (std::fabs(lat0) > 89.999) {
        cg::set_err("polar projection centres are not supported");
        return CGFDM_EINVAL;
    }
    const double d2r = M_PI / 180.0;
    for (long long i = 0; i < n; ++i) {
        double s, az;
        if (!vincenty_inverse(lat0 * d2r, lon0 * d2r, lat[i] * d2r, lon[i] * d2r, s, az)) {
            cg::set_err("geodesic from the centre to point %lld did not converge", i);
            return CGFDM_EINVAL;
        }
        x[i] = s * std::sin(az);
        y[i] = s * std::cos(az);
    }
    return CGFDM_OK;
}

extern "C" int cgfdm_finite_fault_map(const char* path, const cgfdm_desc* desc, const double* terrain, double lon0,
                                      double lat0, int* xyz, float* rate, long long cap, long long* npts_out,
                                      int* nt_out, float* dt_out) {
    using cg::set_err;
    if (!path || !desc || !npts_out) {
        set_err("null argument");
        return CGFDM_EINVAL;
    }
    if (desc->use_terrain && !terrain) {
        set_err("terrain heights required (desc->use_terrain)");
        return CGFDM_EINVAL;
    }
    FaultFile ff;
    ff.f = std::fopen(path, "rb");
    if (!ff.f) {
        set_err("can't open source file %s", path);
        return CGFDM_EIO;
    }
    int npts = 0, nt = 0;
    float dt = 0.f;
    if (std::fread(&npts, 4, 1, ff.f) != 1 || std::fread(&nt, 4, 1, ff.f) != 1 || std::fread(&dt, 4, 1, ff.f) != 1 ||
        npts < 0 || nt <= 0) {
        set_err("bad finite-fault header in %s", path);
        return CGFDM_EIO;
    }
    // per point: lon, lat, z, area, strike, dip, rake[nt], rate[nt] (float32)
    const long long rec = 6 + 2LL * nt;
    std::vector<float> lonf(npts), latf(npts), zf(npts);
    for (int i = 0; i < npts; ++i) {
        float head[3];
        if (std::fseek(ff.f, 12 + (long)(rec * 4) * i, SEEK_SET) != 0 || std::fread(head, 4, 3, ff.f) != 3) {
            set_err("truncated finite-fault file %s at point %d", path, i);
            return CGFDM_EIO;
        }
        lonf[i] = head[0];
        latf[i] = head[1];
        zf[i] = head[2];
    }

    // projTrans: lon/lat converted to radians with the reference's float pi (macro.h:60-63), the
    // centre with PROJ's; coordinates kept as float
    const double ref_d2r = (double)3.141592657f / 180.0;
    const double d2r = M_PI / 180.0;
    std::vector<float> x(npts), y(npts);
    for (int i = 0; i < npts; ++i) {
        double s, az;
        if (!vincenty_inverse(lat0 * d2r, lon0 * d2r, latf[i] * ref_d2r, lonf[i] * ref_d2r, s, az)) {
            set_err("geodesic from the centre to point %d did not converge", i);
            return CGFDM_EINVAL;
        }
        x[i] = (float)(s * std::sin(az));
        y[i] = (float)(s * std::cos(az));
    }

    // srcCoord2Index, the branch the reference compiles (NO_SOURCE_SMOOTH is defined in
    // macro.h:76; multiSource.cpp:243-264): depths relative to the shallowest point; node column by
    // rounding toward zero of x/DH + 0.5 from the projection centre (centerX, centerY), kept only
    // inside the interior columns [HALO, nx+HALO) x [HALO, ny+HALO); node layer = the interior
    // layer k in [HALO, nz+HALO) whose cell [z_k - dz1, z_k + dz2) (half distances to the
    // neighbouring layers) holds the point, the last matching layer if rounding lets two match.
    // Points outside these ranges are dropped without an error, as the reference drops them.
    const int H = 3;  // halo width of the padded grid (reference macro.h HALO)
    const int nx_ = desc->NX + 2 * H, ny_ = desc->NY + 2 * H;
    const float DH = (float)desc->DH;
    const float Depth = (float)desc->depth;
    float maxZ = -Depth;
    for (int i = 0; i < npts; ++i) maxZ = std::max(maxZ, zf[i]);
    for (int i = 0; i < npts; ++i) zf[i] = zf[i] - maxZ;
    auto coord_z = [&](int i, int j, int k) -> float {  // construct_{flat,terrain}_coord (coord.cpp)
        if (desc->use_terrain) {
            const float height = (float)terrain[i + (long long)j * nx_];
            const float DZ = (float)(double(height + std::fabs(Depth)) / double(desc->NZ - 1));
            return -std::fabs(Depth) + DZ * (float)(k - H);
        }
        return (float)(k - H + 1) * DH - (float)desc->NZ * DH;
    };
    std::map<long long, long long> point_index;  // point -> node (padded linear index)
    for (int i = 0; i < npts; ++i) {
        const int sx = int(x[i] / DH + 0.5) + desc->centerX + H;
        const int sy = int(y[i] / DH + 0.5) + desc->centerY + H;
        if (sx < H || sx >= desc->NX + H || sy < H || sy >= desc->NY + H) continue;
        for (int k = H; k < desc->NZ + H; ++k) {
            const float zk = coord_z(sx, sy, k);
            const double dz1 = (zk - coord_z(sx, sy, k - 1)) * 0.5;
            const double dz2 = (coord_z(sx, sy, k + 1) - zk) * 0.5;
            if (zk - dz1 <= zf[i] && zf[i] < dz2 + zk) point_index[i] = ((long long)k * ny_ + sy) * nx_ + sx;
        }
    }

    // readSourceParams + solveMomentRate for the points in the grid, in point order
    const long long m = (long long)point_index.size();
    std::vector<float> mr((size_t)(m * nt * 6), 0.f);
    {
        std::vector<float> buf((size_t)(3 + 2LL * nt));
        long long p = 0;
        for (const auto& kv : point_index) {
            if (std::fseek(ff.f, 12 + (long)(rec * 4) * kv.first + 12, SEEK_SET) != 0 ||
                std::fread(buf.data(), 4, buf.size(), ff.f) != buf.size()) {
                set_err("truncated finite-fault file %s at point %lld", path, kv.first);
                return CGFDM_EIO;
            }
            const float a = buf[0];
            // the reference's argument types: strike, dip and rake in float radians; sin/cos of a
            // float in float, of 2.0*angle in double (solveMomentRate)
            const float s = (float)(buf[1] * ref_d2r), d = (float)(buf[2] * ref_d2r);
            for (int it = 0; it < nt; ++it) {
                const float r = (float)(buf[3 + it] * ref_d2r);
                const float rt = buf[3 + nt + it];
                const float sd = std::sin(d), cd = std::cos(d), sr = std::sin(r), cr = std::cos(r);
                const float ss = std::sin(s), cs = std::cos(s);
                const double s2s = std::sin(2.0 * s), c2s = std::cos(2.0 * s);
                const double s2d = std::sin(2.0 * d), c2d = std::cos(2.0 * d);
                // Aki & Richards moment tensor (x north, y east, z down), stored as (xx, yy, zz,
                // xy, xz, yz) of the east-north-up grid frame
                const float M11 = (float)-((sd * cr) * s2s + s2d * sr * ss * ss);
                const float M22 = (float)((sd * cr) * s2s - s2d * sr * cs * cs);
                const float M33 = -(M11 + M22);
                const float M12 = (float)((sd * cr) * c2s + 0.5 * s2d * sr * s2s);
                const float M13 = (float)-((cd * cr) * cs + c2d * sr * ss);
                const float M23 = (float)-((cd * cr) * ss - c2d * sr * cs);
                float* o = &mr[(size_t)((p * nt + it) * 6)];
                o[0] = M22 * a * rt;
                o[1] = M11 * a * rt;
                o[2] = M33 * a * rt;
                o[3] = M12 * a * rt;
                o[4] = -M23 * a * rt;
                o[5] = -M13 * a * rt;
            }
            ++p;
        }
    }

    // dealDuplicateIndex: each node keeps its last point (in point order); the earlier points of
    // the node are added into it in point order
    std::map<long long, long long> index_point;  // node -> position of its last point
    {
        long long q = 0;
        for (const auto& kv : point_index) index_point[kv.second] = q++;
        q = 0;
        for (const auto& kv : point_index) {
            const long long last = index_point[kv.second];
            if (last > q)
                for (long long e = 0; e < (long long)nt * 6; ++e) mr[(size_t)(last * nt * 6 + e)] += mr[(size_t)(q * nt * 6 + e)];
            ++q;
        }
    }
    const long long nout = (long long)index_point.size();
    *npts_out = nout;
    if (nt_out) *nt_out = nt;
    if (dt_out) *dt_out = dt;
    if (!xyz && !rate) return CGFDM_OK;  // size query
    if (!xyz || !rate || cap < nout) {
        set_err("output arrays for %lld nodes required", nout);
        return CGFDM_EINVAL;
    }
    // changeStorageOrder: nodes in increasing linear index, rates [nt][node][6]
    long long J = 0;
    for (const auto& kv : index_point) {
        const long long node = kv.first, ii = kv.second;
        const int k = (int)(node / ((long long)nx_ * ny_)), rem = (int)(node % ((long long)nx_ * ny_));
        xyz[3 * J + 0] = rem % nx_ - H;
        xyz[3 * J + 1] = rem / nx_ - H;
        xyz[3 * J + 2] = k - H;
        for (int it = 0; it < nt; ++it)
            std::memcpy(&rate[(size_t)(((long long)it * nout + J) * 6)], &mr[(size_t)((ii * nt + it) * 6)], 24);
        ++J;
    }
    return CGFDM_OK;
}
