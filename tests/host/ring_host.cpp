// Host check of the boundary-ring layout and code (csrc/mnad_ring.h), built
// with g++ alone (no HIP).  One line per case, "ok <name>" or
// "FAIL <name>: <reason>":
//  - ring_point(r) is the point ring_index maps to r, for every cell;
//  - ring_index maps the two-cell band around the reconstruction region onto
//    [0, nring) one to one, and gives -1 inside it and beyond the band;
//  - every 16-bit code decodes to the value an independent formula gives
//    (F16: ldexp of sign, exponent, mantissa; BF16: the top half of the
//    float), and encoding rounds to the nearest representable value, ties to
//    even, for random floats of every magnitude;
//  - closed-loop coding of random band-limited series (predictor orders 2-4,
//    F16 and BF16 codes, amplitudes from 1e-30 to 1e3, zeros, a step): the
//    reverse recurrence run from the last P decoded values reproduces every
//    decoded value of the forward bit for bit, and each decoded value lies
//    within one code rounding of the residual (plus one fixed-point unit) of
//    the true value;
//  - on a 30 Hz-band series sampled at 1 ms, the decoded error falls with the
//    order: relative error of order 3 below order 2's by more than 4x.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mnad_ring.h"

using namespace mnad;

static int failures = 0;

static void report(const std::string& name, bool ok, const std::string& why = "")
{
    if (ok) {
        std::printf("ok %s\n", name.c_str());
    } else {
        std::printf("FAIL %s: %s\n", name.c_str(), why.c_str());
        ++failures;
    }
}

static void check_index(int nz, int nx, int np)
{
    const int n = ring_cells(nz, nx, np);
    std::vector<int> hits(n, 0);
    bool ok = true;
    std::string why;
    for (int j = 0; j < nz && ok; ++j)
        for (int i = 0; i < nx; ++i) {
            const int r = ring_index(j, i, nz, nx, np);
            const bool band = j >= np - 2 && j < nz - np + 2 && i >= np - 2 && i < nx - np + 2 &&
                              !in_rec_region(j, i, nz, nx, np);
            if (band != (r >= 0)) {
                ok = false;
                why = "cell (" + std::to_string(j) + "," + std::to_string(i) + ") misclassified";
                break;
            }
            if (r >= n) {
                ok = false;
                why = "index out of range";
                break;
            }
            if (r >= 0) ++hits[r];
        }
    for (int r = 0; r < n && ok; ++r)
        if (hits[r] != 1) {
            ok = false;
            why = "cell " + std::to_string(r) + " hit " + std::to_string(hits[r]) + " times";
        }
    for (int r = 0; r < n && ok; ++r) {   // ring_point inverts ring_index
        int j = -1, i = -1;
        ring_point(r, nz, nx, np, &j, &i);
        if (ring_index(j, i, nz, nx, np) != r) {
            ok = false;
            why = "ring_point of cell " + std::to_string(r) + " maps back to " +
                  std::to_string(ring_index(j, i, nz, nx, np));
        }
    }
    report("ring_index_" + std::to_string(nz) + "x" + std::to_string(nx) + "_np" + std::to_string(np),
           ok, why);
}

static double f16_ref(uint16_t h)
{
    const int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
    double v;
    if (e == 0) v = std::ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp(1024.0 + m, e - 25);
    return s ? -v : v;
}

static void check_formats()
{
    bool ok = true;
    std::string why;
    for (uint32_t h = 0; h < 65536 && ok; ++h) {
        const double a = f16_ref((uint16_t)h), b = f16_dec((uint16_t)h);
        if (!(a == b || (std::isnan(a) && std::isnan(b)))) {
            ok = false;
            why = "f16 code " + std::to_string(h);
        }
        const float c = bf16_dec((uint16_t)h);
        uint32_t bits;
        std::memcpy(&bits, &c, 4);
        if (bits != (h << 16)) {
            ok = false;
            why = "bf16 code " + std::to_string(h);
        }
    }
    report("decode_every_code", ok, why);
    // encode: nearest representable, ties to even
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ue(-40.0, 18.0), um(1.0, 2.0);
    ok = true;
    for (int k = 0; k < 200000 && ok; ++k) {
        const float x = (float)((k & 1 ? -1 : 1) * um(rng) * std::ldexp(1.0, (int)ue(rng)));
        for (int fmt = 1; fmt <= 2 && ok; ++fmt) {
            const uint16_t c = fmt == 1 ? f16_enc(x) : bf16_enc(x);
            const double v = fmt == 1 ? f16_dec(c) : bf16_dec(c);
            if (fmt == 1 && std::fabs((double)x) >= 65520.0) continue;
            // neighbours of the code toward zero and away from it
            const uint16_t lo = (uint16_t)(c - 1), hi = (uint16_t)(c + 1);
            const double vl = fmt == 1 ? f16_dec(lo) : bf16_dec(lo);
            const double vh = fmt == 1 ? f16_dec(hi) : bf16_dec(hi);
            const double d = std::fabs(v - x);
            if ((c & 0x7fff) != 0 && std::fabs(vl - x) < d) ok = false;
            if ((c & 0x7fff) != 0x7fff && std::isfinite(vh) && std::fabs(vh - x) < d) ok = false;
            if ((std::fabs(vl - x) == d || std::fabs(vh - x) == d) && (c & 1)) ok = false;
            if (!ok) why = "x = " + std::to_string(x) + " fmt " + std::to_string(fmt);
        }
    }
    report("encode_rounds_to_nearest_even", ok, why);
}

// Forward-encode `x`, then decode backward; returns the max relative error
// of the decoded values (against max |x|), or -1 on a bit mismatch / bound
// violation (reason in *why).
static double roundtrip(int P, int fmt, const std::vector<float>& x, int E, std::string* why)
{
    const double rsc = std::ldexp(1.0, E), rsi = std::ldexp(1.0, -E);
    const double rcs = fmt == 1 ? std::ldexp(1.0, -(26 + P)) : 1.0, rci = 1.0 / rcs;
    const int n = (int)x.size();
    std::vector<int64_t> G(n + 1, 0);   // G[L], L = 1..n; G[0] = 0
    std::vector<uint16_t> code(n + 1, 0);
    int64_t st[RING_MAX_ORDER] = {0, 0, 0, 0};   // slot L % P
    double xmax = 0, emax = 0;
    for (int L = 1; L <= n; ++L) {
        int64_t g[RING_MAX_ORDER] = {0, 0, 0, 0};
        for (int i = 1; i <= P; ++i) g[i - 1] = st[((L - i) % P + P) % P];
        int64_t gL;
        code[L] = ring_encode(P, g, x[L - 1], rsc, rcs, rci, fmt, &gL);
        st[L % P] = gL;
        G[L] = gL;
        const double pred = (double)ring_pred(P, g[0], g[1], g[2], g[3]);
        const double res = std::fabs((double)x[L - 1] * rsc - pred);
        const double err = std::fabs((double)gL - (double)x[L - 1] * rsc);
        const double bound = res * (fmt == 1 ? std::ldexp(1.0, -11) : std::ldexp(1.0, -8)) + 1.0 +
                             (fmt == 1 ? std::ldexp(1.0, -24) * rci : 0.0);
        if (err > bound) {
            *why = "level " + std::to_string(L) + ": error " + std::to_string(err) + " > " +
                   std::to_string(bound);
            return -1;
        }
        xmax = std::max(xmax, (double)std::fabs(x[L - 1]));
        emax = std::max(emax, err * rsi);
    }
    // backward from the final window
    for (int L = n; L >= P + 1; --L) {
        int64_t g[RING_MAX_ORDER] = {0, 0, 0, 0};
        for (int i = 1; i < P; ++i) g[i - 1] = st[((L - i) % P + P) % P];
        const int64_t back = ring_decode(P, st[L % P], g, code[L], fmt, rci);
        if (back != G[L - P]) {
            *why = "level " + std::to_string(L - P) + " decoded differently";
            return -1;
        }
        st[(L - P) % P] = back;   // same slot as L
    }
    return xmax > 0 ? emax / xmax : emax;
}

static void check_codec()
{
    std::mt19937_64 rng(11);
    std::normal_distribution<double> nd;
    std::uniform_real_distribution<double> uf(0.0, 1.0);
    const double amps[] = {1e-30, 1e-6, 1.0, 1e3};
    for (int P = 2; P <= 4; ++P)
        for (int fmt = 1; fmt <= 2; ++fmt) {
            bool ok = true;
            std::string why;
            for (double amp : amps) {
                for (int trial = 0; trial < 4 && ok; ++trial) {
                    const int n = 3000;
                    std::vector<float> x(n, 0.0f);
                    // a few damped tones (0.02 .. 0.3 rad per step) + FP32-level noise
                    for (int m = 0; m < 5; ++m) {
                        const double w = 0.02 + 0.28 * uf(rng), ph = 6.283 * uf(rng);
                        const double a = amp * nd(rng), t0 = n * uf(rng), s = 50 + 400 * uf(rng);
                        for (int t = 0; t < n; ++t)
                            x[t] += (float)(a * std::sin(w * t + ph) *
                                            std::exp(-(t - t0) * (t - t0) / (2 * s * s)));
                    }
                    if (trial == 1)
                        for (int t = n / 2; t < n; ++t) x[t] += (float)amp;   // a step
                    if (trial == 2)
                        for (int t = 0; t < n / 3; ++t) x[t] = 0.0f;           // silence
                    // fixed point: 2^-E = 2^-40 of the amplitude, as the runtime picks
                    const int E = 40 - (int)std::ceil(std::log2(amp));
                    if (roundtrip(P, fmt, x, E, &why) < 0) ok = false;
                }
            }
            report("codec_P" + std::to_string(P) + (fmt == 1 ? "_f16" : "_bf16"), ok, why);
        }
    // order helps on a band-limited series (30 Hz Ricker-like band, dt = 1 ms)
    const int n = 4000;
    std::vector<float> x(n);
    for (int t = 0; t < n; ++t) {
        const double tau = 3.14159265358979 * (t * 1e-3 - 0.5) * 30.0;
        const double tau2 = 3.14159265358979 * (t * 1e-3 - 1.7) * 18.0;
        x[t] = (float)((1 - 2 * tau * tau) * std::exp(-tau * tau) +
                       0.4 * (1 - 2 * tau2 * tau2) * std::exp(-tau2 * tau2));
    }
    std::string why;
    const double e2 = roundtrip(2, 2, x, 40, &why), e3 = roundtrip(3, 2, x, 40, &why),
                 e4 = roundtrip(4, 2, x, 40, &why);
    std::printf("# bf16 decoded error / max: P2 %.3g  P3 %.3g  P4 %.3g\n", e2, e3, e4);
    report("order_reduces_error", e2 > 0 && e3 > 0 && e4 > 0 && e3 * 4 < e2 && e4 < e3,
           "errors " + std::to_string(e2) + " " + std::to_string(e3) + " " + std::to_string(e4));
}

int main()
{
    check_index(9, 11, 2);
    check_index(64, 70, 3);
    check_index(121, 97, 20);
    check_index(100, 300, 30);
    check_formats();
    check_codec();
    return failures ? 1 : 0;
}
