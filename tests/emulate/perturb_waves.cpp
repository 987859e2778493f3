// A CPU model of the perturbation kernel's wave control flow (perturb_pixel in
// chaos-ultra_amd/csrc/kernels/perturb.hip), checked pixel by pixel against the
// oracle's checked loop (oracle/deep_oracle.cpp, or_perturb_pixel_f/_d).
//
// The kernel does not step each pixel through the oracle's loop: 64 lanes run
// unchecked blocks of K steps together, a block in which a rebase test fires is
// rerun up to the first step where one does (every lane stops there, the lanes
// that fired take the rebase), escape is tested once per block and a block with
// an escape is replayed with checked steps.  This model runs exactly that
// schedule on 8x8-pixel waves of a deep frame, in the same arithmetic (C fma,
// contraction off), and reports every pixel whose count, state or |z|^2 differs
// from the oracle's.  Memory placement (scalar, LDS or global reads of the
// orbit) does not change values and is not modelled.
//
//   perturb_waves <32|64> <centre_x> <centre_y> <width_plane> <max_iter> <waves> <seed>
//
// prints "mismatches N waves W pixels P blocks B reruns R checked C" and exits 1
// when N > 0.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../oracle/chaos_oracle.h"

namespace {

constexpr int K = 32;  // steps per unchecked block (perturbFloat / perturbDouble)
constexpr uint32_t kWidth = 1920, kHeight = 1080;

inline float fma_r(float a, float b, float c) { return fmaf(a, b, c); }
inline double fma_r(double a, double b, double c) { return fma(a, b, c); }

template <class Real>
struct Lane {
    Real dx = 0, dy = 0, dcx = 0, dcy = 0, mag = 0;
    uint32_t m = 0, n = 0;
    int state = 0;  // 0 running, 1 escaped, 2 max_iter, 4 outside the frame
};

template <class Real>
struct Model {
    std::vector<Real> Z;  // {x, y, |Z|^2/2, 0} per entry
    uint32_t M = 0, max_iter = 0;
    Real bsq = 256;
    uint64_t blocks = 0, reruns = 0, checked = 0;

    Real zx(uint32_t m) const { return Z[4 * m]; }
    Real zy(uint32_t m) const { return Z[4 * m + 1]; }
    Real zh(uint32_t m) const { return Z[4 * m + 2]; }

    bool fires(const Lane<Real> &L, uint32_t m) const {
        return fma_r(zx(m), L.dx, zy(m) * L.dy) < -zh(m);
    }
    void step(Lane<Real> &L, uint32_t m) const {
        const Real ax = fma_r((Real)2, zx(m), L.dx), ay = fma_r((Real)2, zy(m), L.dy);
        const Real ndx = fma_r(ax, L.dx, fma_r(-ay, L.dy, L.dcx));
        const Real ndy = fma_r(ax, L.dy, fma_r(ay, L.dx, L.dcy));
        L.dx = ndx;
        L.dy = ndy;
    }
    // checked_step of the kernel (the oracle's loop body, rebasing on)
    int checked_step(Lane<Real> &L) const {
        Real Zx = zx(L.m), Zy = zy(L.m);
        const Real h = zh(L.m);
        const Real x = Zx + L.dx, y = Zy + L.dy;
        const Real mag = fma_r(x, x, y * y);
        if (!(L.n < max_iter)) return 2;
        if (!(mag < bsq)) {
            L.mag = mag;
            return 1;
        }
        if (fma_r(Zx, L.dx, Zy * L.dy) < -h || L.m == M) {
            L.dx = x;
            L.dy = y;
            L.m = 0;
            Zx = 0;
            Zy = 0;
        }
        const Real ax = fma_r((Real)2, Zx, L.dx), ay = fma_r((Real)2, Zy, L.dy);
        const Real ndx = fma_r(ax, L.dx, fma_r(-ay, L.dy, L.dcx));
        const Real ndy = fma_r(ax, L.dy, fma_r(ay, L.dx, L.dcy));
        L.dx = ndx;
        L.dy = ndy;
        ++L.m;
        ++L.n;
        return 0;
    }

    void wave(Lane<Real> (&L)[64]) {
        for (;;) {
            bool any = false, all = true;
            for (auto &l : L) {
                if (l.state != 0) continue;
                any = true;
                if (!(max_iter - l.n >= (uint32_t)K && M - l.m >= (uint32_t)K)) all = false;
            }
            if (!any) return;
            bool replay[64] = {};
            if (all) {  // every running lane takes the block
                ++blocks;
                Lane<Real> start[64];
                std::memcpy(start, L, sizeof L);
                int k0 = K;
                for (int k = 0; k < K && k0 == K; ++k) {
                    for (auto &l : L)
                        if (l.state == 0 && fires(l, l.m + k)) k0 = k;
                    if (k0 == K)
                        for (auto &l : L)
                            if (l.state == 0) step(l, l.m + k);
                }
                if (k0 < K) ++reruns;
                for (int j = 0; j < 64; ++j) {
                    Lane<Real> &l = L[j];
                    if (l.state != 0) continue;
                    const bool fired = k0 < K && fires(l, l.m + k0);
                    const uint32_t me = l.m + k0;
                    const Real x = zx(me) + l.dx, y = zy(me) + l.dy;
                    const Real mag = fma_r(x, x, y * y);
                    if (!(mag < bsq)) {
                        l.dx = start[j].dx;
                        l.dy = start[j].dy;
                        replay[j] = true;
                    } else {
                        l.n += k0;
                        l.m += k0;
                        if (fired) {
                            l.dx = x;
                            l.dy = y;
                            l.m = 0;
                        }
                    }
                }
            }
            int most = 0;
            for (int j = 0; j < 64; ++j) {
                Lane<Real> &l = L[j];
                if (l.state != 0 || (all && !replay[j])) continue;
                const int steps = all ? K : 1;
                int s = 0;
                for (; s < steps && l.state == 0; ++s) l.state = checked_step(l);
                if (s > most) most = s;
            }
            checked += most;
        }
    }
};

template <class Real>
int run(double cx, double cy, double width_plane, uint32_t max_iter, int waves, unsigned seed) {
    Model<Real> mdl;
    mdl.max_iter = max_iter;
    mdl.Z.assign(4 * ((size_t)max_iter + 2), 0);
    if (sizeof(Real) == 4)
        mdl.M = or_reference_orbit_f(cx, 0, cy, 0, max_iter, 256.0, (float *)mdl.Z.data());
    else
        mdl.M = or_reference_orbit_d(cx, 0, cy, 0, max_iter, 256.0, (double *)mdl.Z.data());
    const double pw = width_plane / kWidth;
    srand(seed);
    uint64_t bad = 0, pixels = 0;
    for (int w = 0; w < waves; ++w) {
        const uint32_t wx = (uint32_t)(rand() % (kWidth / 8)) * 8, wy = (uint32_t)(rand() % (kHeight / 8)) * 8;
        Lane<Real> L[64];
        for (int l = 0; l < 64; ++l) {
            double ox, oy;
            or_deep_offset(wx + (l & 7), wy + (l >> 3), kWidth, kHeight, pw, pw, &ox, &oy);
            L[l].dcx = (Real)ox;
            L[l].dcy = (Real)oy;
        }
        mdl.wave(L);
        for (int l = 0; l < 64; ++l) {
            int state = 0;
            uint32_t n;
            Real mag;
            if (sizeof(Real) == 4) {
                float mg;
                n = or_perturb_pixel_f((const float *)mdl.Z.data(), mdl.M, (float)L[l].dcx,
                                       (float)L[l].dcy, max_iter, 256.f, 1u, 1e-6f, &mg, &state);
                mag = (Real)mg;
            } else {
                double mg;
                n = or_perturb_pixel_d((const double *)mdl.Z.data(), mdl.M, (double)L[l].dcx,
                                       (double)L[l].dcy, max_iter, 256.0, 1u, 1e-6, &mg, &state);
                mag = (Real)mg;
            }
            const Real got = L[l].state == 1 ? L[l].mag : (Real)0;
            ++pixels;
            if (n != L[l].n || state != L[l].state || (state == 1 && std::memcmp(&mag, &got, sizeof(Real)) != 0)) {
                if (bad < 5)
                    printf("pixel (%u, %u): oracle n %u state %d, model n %u state %d\n", wx + (l & 7),
                           wy + (l >> 3), n, state, L[l].n, L[l].state);
                ++bad;
            }
        }
    }
    printf("mismatches %llu waves %d pixels %llu M %u blocks %llu reruns %llu checked %llu\n",
           (unsigned long long)bad, waves, (unsigned long long)pixels, mdl.M,
           (unsigned long long)mdl.blocks, (unsigned long long)mdl.reruns,
           (unsigned long long)mdl.checked);
    return bad != 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s <32|64> <cx> <cy> <width_plane> <max_iter> <waves> <seed>\n", argv[0]);
        return 2;
    }
    const int prec = atoi(argv[1]);
    const double cx = atof(argv[2]), cy = atof(argv[3]), width = atof(argv[4]);
    const uint32_t max_iter = (uint32_t)strtoul(argv[5], nullptr, 10);
    const int waves = atoi(argv[6]);
    const unsigned seed = (unsigned)strtoul(argv[7], nullptr, 10);
    return prec == 32 ? run<float>(cx, cy, width, max_iter, waves, seed)
                      : run<double>(cx, cy, width, max_iter, waves, seed);
}
