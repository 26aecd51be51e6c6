// Host-only check of the launch tile table (build_tiles in csrc/mnad_runtime.hip).
// Built with hipcc against the runtime source alone; the kernel launchers are
// stubbed (the tile builder calls none of them) and no GPU is touched.
//
// For each case it prints one line "ok <name>" or "FAIL <name>: <reason>":
//  - every grid point of every shot is covered by exactly one listed tile
//    (one-point tiles of kernels 4/7, and for kernels 6 and 10 one-point
//    tiles off the multi-column interior plus the interior tiles);
//  - the table length is a multiple of 8 (XCD x takes entries b % 8 == x);
//  - each XCD's tiles are a contiguous range of the shot-major tile order;
//  - the estimated work per XCD differs by at most two tile weights, half a
//    tile at either end of each XCD's range (two-step kernels: counts);
//  - with strip_first, each XCD dispatches its CPML-strip tiles first;
//  - kernel 13: a wave (56 columns of its tile) takes the interior path when
//    it passes q_interior_tile, and then every load of its march (four
//    cells around it) lies in the grid and every level-s+1 halo point (two
//    cells) outside the CPML; a tile weighs its interior waves plus
//    kQEdgeWeight per edge wave, and counts as a strip tile when it holds an
//    edge wave;
// and the block heights the resident kernel (9) picks for a few batches;
// and the host pass over a velocity model (scan_model, model_digest):
//  - the first value that is not positive and finite (0, -0, negative, NaN,
//    +inf; a subnormal is valid) is found wherever it sits, lanes and tail;
//  - the range equals the one std::minmax_element gives;
//  - the digest is the same for equal arrays, and a one-ulp change of any
//    single value, or the swap of two neighbours, changes both halves.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <vector>

#include "mnad_internal.h"

extern "C" {
hipError_t mnad_launch_setup_model(const double*, int, int, int, int, double, double, double,
                                   float*, float*, float*, hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_step(const mnad::ShotDev*, const mnad::TileRef*, int, const int*, int, int,
                            int, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_step2q(const mnad::ShotDev*, int, const mnad::TileRef*, int, const int*,
                              int, int, int, int, int, hipStream_t)
{
    return hipErrorNotSupported;
}
hipError_t mnad_launch_advance(int*, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_init_source(mnad::ShotDev*, int, hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_make_radj(const float*, const float*, const float*, int, int, float*,
                                 hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_stack(const float*, const float*, const float*, float, long, float*, float*,
                             float*, float*, hipStream_t) { return hipErrorNotSupported; }
hipError_t mnad_launch_fill(float*, long, float, hipStream_t) { return hipErrorNotSupported; }
int mnad_resident_max_rows() { return 64; }
hipError_t mnad_resident_fit(int, int*) { return hipErrorNotSupported; }
long mnad_resident_edge_records(int) { return 0; }
hipError_t mnad_launch_resident(const mnad::ShotDev*, const mnad::ResBlock*, int,
                                unsigned long long*, int*, int, int, int, hipStream_t)
{
    return hipErrorNotSupported;
}
}

namespace mnad {
int autotune(mnad_ctx*, const Model&, int, TunePick*) { return MNAD_ERR_STATE; }
void server_destroy(mnad_ctx*) {}
}

using namespace mnad;

namespace mnad {
int resident_rows(const std::vector<ShotDev>& h, int ncu, int per_cu, long* nblocks);
}

struct Shape { int nz, nx, np; };

static bool is_strip_tile(int r0, int r1, int c0, int c1, const Shape& s)
{
    return c0 < s.np || c1 > s.nx - s.np || r0 < s.np || r1 > s.nz - s.np;
}

static std::string check(const char* name, int kernel, int rz, int brows, int strip_first,
                         const std::vector<Shape>& shapes)
{
    mnad_ctx c;
    c.kernel = kernel;
    c.rz = rz;
    c.brows = brows;
    c.strip_first = strip_first;
    std::vector<ShotDev> h(shapes.size());
    for (size_t k = 0; k < shapes.size(); ++k) {
        h[k] = ShotDev{};
        h[k].nz = shapes[k].nz;
        h[k].nx = shapes[k].nx;
        h[k].npml = shapes[k].np;
    }
    std::vector<TileRef> table;
    build_tiles(&c, h, table);
    int tw, th, spl;
    tile_shape(&c, &tw, &th, &spl);
    const int nc = interior_columns_per_lane(kernel);
    const bool x4 = nc > 0, two = two_step_kernel(kernel), q13 = kernel == 13;
    char buf[256];
    if (table.size() % 8 != 0) return "table length not a multiple of 8";
    // kernel 13: the tile's interior and edge waves
    auto q_waves = [&](const TileRef& t, int* nin, int* ned) {
        const Shape& s = shapes[t.shot];
        const ShotDev& d = h[t.shot];
        const int tj = t.tile / d.ntx, ti = t.tile - tj * d.ntx;
        *nin = *ned = 0;
        for (int x0 = ti * tw; x0 < std::min(s.nx, (ti + 1) * tw); x0 += 56) {
            const int zA = tj * th;
            if (!q_interior_tile(s.nx, s.nz, s.np, x0, zA, 56, th)) {
                ++*ned;
                continue;
            }
            ++*nin;
            if (x0 - 4 < 0 || x0 + 56 + 4 > s.nx || zA - 4 < 0 || zA + th + 4 > s.nz)
                return "an interior wave's march loads outside the grid";
            if (x0 - 2 < s.np || x0 + 56 + 2 > s.nx - s.np || zA - 2 < s.np ||
                zA + th + 2 > s.nz - s.np)
                return "an interior wave's level-s+1 halo meets the CPML";
        }
        return "";
    };
    if (q13)
        for (const TileRef& t : table) {
            if (t.shot < 0) continue;
            int nin, ned;
            const char* why = q_waves(t, &nin, &ned);
            if (*why) return why;
        }

    // coverage: every point exactly once
    std::vector<std::vector<int>> cover(shapes.size());
    for (size_t k = 0; k < shapes.size(); ++k) cover[k].assign((size_t)shapes[k].nz * shapes[k].nx, 0);
    std::vector<std::pair<int, int>> listed;
    for (const TileRef& t : table) {
        if (t.shot < 0) continue;
        if (t.shot >= (int)shapes.size()) return "shot index out of range";
        listed.emplace_back(t.shot, t.tile);
        const Shape& s = shapes[t.shot];
        const ShotDev& d = h[t.shot];
        if (x4 && t.tile >= d.nt1) {
            const int u = t.tile - d.nt1, tb = u / d.nx4, tx = u - tb * d.nx4;
            const int lo = d.x4lo + tx * 62 * nc;
            for (int j = s.np + tb * th; j < std::min(s.nz - s.np, s.np + (tb + 1) * th); ++j)
                for (int i = std::max(lo, s.np); i < std::min(lo + 62 * nc, d.x4hi); ++i)
                    cover[t.shot][(size_t)j * s.nx + i]++;
        } else {
            const int tj = t.tile / d.ntx, ti = t.tile - tj * d.ntx;
            for (int j = tj * th; j < std::min(s.nz, (tj + 1) * th); ++j)
                for (int i = ti * tw; i < std::min(s.nx, (ti + 1) * tw); ++i) {
                    if (x4 && j >= s.np && j < s.nz - s.np && i >= s.np && i < d.x4hi) continue;
                    cover[t.shot][(size_t)j * s.nx + i]++;
                }
        }
    }
    for (size_t k = 0; k < shapes.size(); ++k)
        for (size_t q = 0; q < cover[k].size(); ++q)
            if (cover[k][q] != 1) {
                snprintf(buf, sizeof buf, "shot %zu point %zu covered %d times", k, q, cover[k][q]);
                return buf;
            }
    std::vector<std::pair<int, int>> sorted = listed;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return "tile listed twice";

    // per XCD: contiguous range, balanced work, strip tiles first
    auto weight = [&](const TileRef& t, bool* strip) {
        const Shape& s = shapes[t.shot];
        const ShotDev& d = h[t.shot];
        if (q13) {
            int nin, ned;
            q_waves(t, &nin, &ned);
            *strip = ned > 0;
            return nin + kQEdgeWeight * ned;
        }
        if (x4 && t.tile >= d.nt1) {
            *strip = false;
            return c.x4_weight * nc / 4.0;
        }
        const int tj = t.tile / d.ntx, ti = t.tile - tj * d.ntx;
        *strip = q13 ? !q_interior_tile(s.nx, s.nz, s.np, ti * tw, tj * th, tw, th)
                     : is_strip_tile(tj * th, std::min(s.nz, (tj + 1) * th), ti * tw,
                                     std::min(s.nx, (ti + 1) * tw), s);
        return two ? 1.0 : (*strip ? c.strip_weight : 1.0);
    };
    const std::pair<size_t, size_t> part{0, table.size()};
    {
    std::vector<std::pair<int, int>> psorted;
    for (size_t b = part.first; b < part.second; ++b)
        if (table[b].shot >= 0) psorted.emplace_back(table[b].shot, table[b].tile);
    std::sort(psorted.begin(), psorted.end());
    double wmax = 0.0, wlo = 1e300, whi = 0.0;
    for (int x = 0; x < 8; ++x) {
        std::vector<std::pair<int, int>> mine;
        double w = 0.0;
        bool seen_interior = false;
        for (size_t b = part.first + x; b < part.second; b += 8) {
            const TileRef& t = table[b];
            if (t.shot < 0) continue;
            bool strip = false;
            const double tw1 = weight(t, &strip);
            w += tw1;
            wmax = std::max(wmax, tw1);
            if (strip_first) {
                if (!strip) seen_interior = true;
                else if (seen_interior) return "a strip tile after an interior tile on one XCD";
            }
            mine.emplace_back(t.shot, t.tile);
        }
        wlo = std::min(wlo, w);
        whi = std::max(whi, w);
        if (mine.empty()) continue;
        std::sort(mine.begin(), mine.end());
        const auto first = std::lower_bound(psorted.begin(), psorted.end(), mine.front());
        if (first + (long)mine.size() > psorted.end() ||
            !std::equal(mine.begin(), mine.end(), first))
            return "an XCD's tiles are not a contiguous range";
    }
    // a run gains or loses at most half a tile at each end of its range
    if (whi - wlo > 2 * wmax + 1e-9) {
        snprintf(buf, sizeof buf, "work per XCD from %.1f to %.1f (tile weight at most %.1f)", wlo,
                 whi, wmax);
        return buf;
    }
    }
    return "";
}

int main()
{
    struct Case {
        const char* name;
        int kernel, rz, brows, strip_first;
        std::vector<Shape> shapes;
    };
    const Shape c2{561, 1061, 30}, c1{241, 241, 20}, tiny{13, 13, 6}, nocpml{70, 300, 0};
    const Shape c5w{1061, 561, 30};
    const std::vector<Case> cases = {
        {"k4_b1_c2", 4, 1, 1, 1, {c2}},
        {"k4_b4_c2x2", 4, 1, 4, 1, {c2, c2}},
        {"k4_b1_mixed", 4, 1, 1, 1, {c2, c1, tiny, nocpml}},
        {"k4_b2_nostripfirst", 4, 1, 2, 0, {c1, c2}},
        {"k0_r2_c1", 0, 2, 4, 1, {c1}},
        {"k6_b2_c2", 6, 1, 2, 1, {c2}},
        {"k6_b4_mixed", 6, 1, 4, 1, {c1, c2, tiny, nocpml}},
        {"k10_b1_c2", 10, 1, 1, 1, {c2}},
        {"k11_r2_mixed", 11, 2, 1, 1, {c2, c1, tiny, nocpml}},
        {"k10_b2_mixed", 10, 1, 2, 1, {c1, c2, tiny, nocpml}},
        {"k7_r5_b8_c2x3", 7, 5, 8, 1, {c2, c2, c2}},
        {"k7_r3_b16_mixed", 7, 3, 16, 1, {c1, tiny, c2}},
        {"k13_r8_b1_c2", 13, 8, 1, 1, {c2}},
        {"k13_r4_b4_mixed", 13, 4, 4, 1, {c1, c2, tiny, nocpml}},
        {"k13_r2_b2_c5w", 13, 2, 2, 0, {c5w, c2}},
        {"k13_r2_b4_mixed", 13, 2, 4, 1, {c2, c5w, c1, tiny, nocpml}},
    };
    int bad = 0;
    // Block heights of the resident kernel (9): the smallest height >= 4 whose
    // 64-column blocks number at most per_cu x ncu; 0 when none up to 64 does.
    struct RCase {
        const char* name;
        std::vector<Shape> shapes;
        int ncu, per_cu, bh;
        long nblocks;
    };
    const Shape c3{811, 3461, 30}, c5{2061, 8061, 30};
    const std::vector<RCase> rcases = {
        {"res_c2_1shot", {c2}, 256, 2, 19, 17L * 30},
        {"res_c2_2shots", {c2, c2}, 256, 2, 38, 2L * 17 * 15},
        {"res_c1_3shots", {c1, c1, c1}, 256, 1, 12, 3L * 4 * 21},
        {"res_c3_1shot", {c3}, 256, 4, 46, 55L * 18},
        {"res_c5_too_big", {c5}, 256, 4, 0, 0},
    };
    for (const RCase& rc : rcases) {
        std::vector<ShotDev> h(rc.shapes.size());
        for (size_t k = 0; k < rc.shapes.size(); ++k) {
            h[k] = ShotDev{};
            h[k].nz = rc.shapes[k].nz;
            h[k].nx = rc.shapes[k].nx;
            h[k].npml = rc.shapes[k].np;
        }
        long n = 0;
        const int bh = resident_rows(h, rc.ncu, rc.per_cu, &n);
        if (bh == rc.bh && (bh == 0 || n == rc.nblocks)) {
            printf("ok %s\n", rc.name);
        } else {
            printf("FAIL %s: block height %d, %ld blocks (expected %d, %ld)\n", rc.name, bh, n,
                   rc.bh, rc.nblocks);
            ++bad;
        }
    }
    {   // model scan
        std::string why;
        for (size_t n : {9u, 15u, 16u, 17u, 1001u}) {
            std::vector<double> v(n);
            for (size_t k = 0; k < n; ++k) v[k] = 1500.0 + 37.0 * (double)((k * 7919) % 101);
            double lo = 0, hi = 0;
            if (scan_model(v.data(), n, &lo, &hi) != n) why += " valid model rejected;";
            const auto mm = std::minmax_element(v.begin(), v.end());
            if (lo != *mm.first || hi != *mm.second) why += " wrong range;";
            const double bads[] = {0.0, -0.0, -1500.0, std::numeric_limits<double>::quiet_NaN(),
                                   std::numeric_limits<double>::infinity()};
            for (size_t at : {(size_t)0, n / 2, n - 1})
                for (double b : bads) {
                    std::vector<double> w = v;
                    w[at] = b;
                    if (at + 1 < n) w[n - 1] = -1.0;   // a later bad value must not be reported
                    if (scan_model(w.data(), n, &lo, &hi) != at) why += " bad value missed;";
                }
            std::vector<double> w = v;
            w[n - 1] = std::numeric_limits<double>::denorm_min();
            if (scan_model(w.data(), n, &lo, &hi) != n || lo != w[n - 1]) why += " subnormal;";
        }
        printf(why.empty() ? "ok %s\n" : "FAIL %s:%s\n", "model scan", why.c_str());
        bad += !why.empty();
    }
    {   // model digest
        std::string why;
        for (size_t n : {9u, 15u, 16u, 17u, 1001u}) {
            std::vector<double> v(n);
            for (size_t k = 0; k < n; ++k) v[k] = 1500.0 + 37.0 * (double)((k * 7919) % 101);
            uint64_t a0, b0, a1, b1;
            model_digest(v.data(), n, &a0, &b0);
            std::vector<double> c = v;
            model_digest(c.data(), n, &a1, &b1);
            if (a1 != a0 || b1 != b0) why += " equal arrays differ;";
            for (size_t k = 0; k < n; ++k) {
                c[k] = std::nextafter(v[k], 1e9);
                model_digest(c.data(), n, &a1, &b1);
                if (a1 == a0 || b1 == b0) why += " one-ulp change at " + std::to_string(k) + ";";
                c[k] = v[k];
            }
            for (size_t k = 0; k + 1 < n; ++k) {
                if (v[k] == v[k + 1]) continue;
                std::swap(c[k], c[k + 1]);
                model_digest(c.data(), n, &a1, &b1);
                if (a1 == a0 || b1 == b0) why += " swap at " + std::to_string(k) + ";";
                std::swap(c[k], c[k + 1]);
            }
        }
        printf(why.empty() ? "ok %s\n" : "FAIL %s:%s\n", "model digest", why.substr(0, 300).c_str());
        bad += !why.empty();
    }
    for (const Case& cs : cases) {
        const std::string r = check(cs.name, cs.kernel, cs.rz, cs.brows, cs.strip_first, cs.shapes);
        if (r.empty()) {
            printf("ok %s\n", cs.name);
        } else {
            printf("FAIL %s: %s\n", cs.name, r.c_str());
            ++bad;
        }
    }
    return bad ? 1 : 0;
}
