// fp32_trust.cpp — how often FP32 deltas can be trusted at the deep-zoom frames.
//
// For seeded pixels of the 1920x1080 frame at the deep-zoom centre, runs the
// oracle's FP32-delta perturbation loop (rebasing; or_perturb_pixel_f's steps)
// while carrying a first-order bound E on the absolute error of z = Z + d:
//   E' = (|2Z + d|_1 + |d|_1) E + u (|2Z + d|_1 |d|_1 + |Z|_1 |d|_1 + |dc|_1 + |d'|_1),
// with u = 2^-24 and |.|_1 = |re| + |im| >= |.| (propagation by |2z| through the
// step's derivative, and one rounding of each term of d' = (2Z + d) d + dc plus
// the orbit entry's own rounding); a rebase adds u |z|_1.  A pixel's count is
// "trusted" when its escape test and the test before it clear the bailout radius
// by more than E (|z_n| - R > E_n and R - |z_{n-1}| > E_{n-1}).  An orbit that
// never escapes is taken as trusted (over 50000 steps the L1 bound grows without
// meaning there, and such pixels rarely change: "caught" shows how rarely).
// Untrusted pixels would be re-rendered with FP64 deltas.
//
// Prints, for the frame: pixels, the share whose FP32 count differs from the
// FP64-delta count, the share flagged untrusted, the share of differing pixels
// that are flagged, and the share still differing after the flagged ones take
// their FP64-delta count.  Also checks that the counts of this loop equal
// or_perturb_pixel_f's (the bound changes nothing in the iteration).
//
// usage: fp32_trust <width> <max_iter> <pixels> <seed>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../oracle/chaos_oracle.h"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: fp32_trust <width> <max_iter> <pixels> <seed>\n");
        return 2;
    }
    const double width = std::atof(argv[1]);
    const uint32_t max_iter = (uint32_t)std::atol(argv[2]);
    const int n_pix = std::atoi(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const uint32_t W = 1920, H = 1080;
    const double cx = -0.743643887037151, cy = 0.131825904205330;
    const double pw = width / W;
    const double bsq = 256.0;
    const float R = 16.0f;
    const float u = std::ldexp(1.0f, -24);

    std::vector<float> orbit_f((size_t)(max_iter + 1) * 4);
    std::vector<double> orbit_d((size_t)(max_iter + 1) * 4);
    const uint32_t Mf = or_reference_orbit_f(cx, 0, cy, 0, max_iter, bsq, orbit_f.data());
    const uint32_t Md = or_reference_orbit_d(cx, 0, cy, 0, max_iter, bsq, orbit_d.data());

    std::mt19937_64 rng(seed);
    int differ = 0, flagged = 0, caught = 0, after = 0, loop_mismatch = 0;
    for (int k = 0; k < n_pix; ++k) {
        const uint32_t px = (uint32_t)(rng() % W), py = (uint32_t)(rng() % H);
        double ox, oy;
        or_deep_offset(px, py, W, H, pw, pw, &ox, &oy);
        const float dcx = (float)ox, dcy = (float)oy;
        const float dcn = std::fabs(dcx) + std::fabs(dcy);

        // the oracle's FP32 loop with the bound carried alongside
        float dx = 0, dy = 0, E = u * dcn, E_prev = 0, mag_prev = 0;
        uint32_t m = 0, n = 0;
        bool escaped = false, trusted = true;
        for (;;) {
            float Zx = orbit_f[4 * m], Zy = orbit_f[4 * m + 1];
            const float h = orbit_f[4 * m + 2];
            const float zx = Zx + dx, zy = Zy + dy;
            const float mag = std::fmaf(zx, zx, zy * zy);
            if (!(n < max_iter)) break;
            if (!(mag < (float)bsq)) {
                escaped = true;
                trusted = std::sqrt(mag) - R > E && R - std::sqrt(mag_prev) > E_prev;
                break;
            }
            if (std::fmaf(Zx, dx, Zy * dy) < -h || m == Mf) {
                dx = zx;
                dy = zy;
                m = 0;
                Zx = 0;
                Zy = 0;
                E += u * (std::fabs(zx) + std::fabs(zy));
            }
            const float ax = std::fmaf(2.0f, Zx, dx), ay = std::fmaf(2.0f, Zy, dy);
            const float ndx = std::fmaf(ax, dx, std::fmaf(-ay, dy, dcx));
            const float ndy = std::fmaf(ax, dy, std::fmaf(ay, dx, dcy));
            const float dn = std::fabs(dx) + std::fabs(dy), an = std::fabs(ax) + std::fabs(ay);
            const float Zn = std::fabs(Zx) + std::fabs(Zy);
            E_prev = E;
            mag_prev = mag;
            E = (an + dn) * E + u * (an * dn + Zn * dn + dcn + std::fabs(ndx) + std::fabs(ndy));
            dx = ndx;
            dy = ndy;
            ++m;
            ++n;
        }
        const uint32_t c32 = escaped ? n : 0;

        float mf;
        int sf;
        const uint32_t nf = or_perturb_pixel_f(orbit_f.data(), Mf, dcx, dcy, max_iter, (float)bsq, 1u,
                                               1e-6f, &mf, &sf);
        if ((sf == 1 ? nf : 0u) != c32) ++loop_mismatch;
        double md;
        int sd;
        const uint32_t nd = or_perturb_pixel_d(orbit_d.data(), Md, ox, oy, max_iter, bsq, 1u, 1e-6,
                                               &md, &sd);
        const uint32_t c64 = sd == 1 ? nd : 0u;

        const bool bad = c32 != c64;
        differ += bad;
        flagged += !trusted;
        caught += bad && !trusted;
        after += (trusted ? c32 : c64) != c64;
    }
    std::printf("pixels %d differ %.4f flagged %.4f caught %.4f after %.4f loop_mismatch %d\n",
                n_pix, (double)differ / n_pix, (double)flagged / n_pix,
                differ ? (double)caught / differ : 1.0, (double)after / n_pix, loop_mismatch);
    return 0;
}
