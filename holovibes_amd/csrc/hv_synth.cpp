// hv_synth.cpp -- seeded synthetic interferogram source (host side).
//
// Stands in for a camera (ICamera::get_frames, Camera/include/icamera.hh:152).
// I = |R + O_t|^2 with R = exp(2 pi i (tilt_x x + tilt_y y)),
// O_t = O_0 exp(i q r^2 / r_max^2) exp(2 pi i f_d(x, y) t / period),
// O_0 = sigma (n1 + i n2) / sqrt(2) fixed per source, f_d an integer drawn per
// pixel inside each Doppler disc (0 outside), then Poisson(K I) with
// K = mean_counts / (1 + sigma^2), clipped to 2^bits - 1.
// Every random draw comes from SplitMix64 of a key built from
// (seed, stream, frame, pixel, draw index), so any frame can be regenerated
// alone and the result does not depend on threading.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/holo_hip.h"

namespace {

inline uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t key;
    uint64_t n = 0;
    Rng(uint64_t seed, uint64_t stream, uint64_t frame, uint64_t pixel)
        : key(splitmix(splitmix(splitmix(seed ^ (stream * 0x632be59bd9b4e019ull)) ^ frame) ^ pixel)) {}
    // uniform in [0, 1)
    double uniform() { return (double)(splitmix(key + 0x9e3779b97f4a7c15ull * (++n)) >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
    // Poisson: multiplication method below 10, PTRS (Hormann 1993,
    // "The transformed rejection method for generating Poisson random
    // variables") above.
    long poisson(double lam) {
        if (lam <= 0) return 0;
        if (lam < 10) {
            const double L = std::exp(-lam);
            long k = 0;
            double p = uniform();
            while (p > L) {
                ++k;
                p *= uniform();
            }
            return k;
        }
        const double slam = std::sqrt(lam), loglam = std::log(lam);
        const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
        const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
        for (;;) {
            const double U = uniform() - 0.5, V = uniform();
            const double us = 0.5 - std::fabs(U);
            const long k = (long)std::floor((2 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr) return k;
            if (k < 0 || (us < 0.013 && V > us)) continue;
            if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
                -lam + k * loglam - std::lgamma((double)k + 1))
                return k;
        }
    }
};

}  // namespace

extern "C" hv_status hv_synth_frames(const hv_synth_params* p, uint64_t first_frame, uint32_t count, void* out) {
    if (!p || !out) return HV_ERR_INVALID_ARG;
    if (p->width == 0 || p->height == 0 || p->period == 0) return HV_ERR_INVALID_ARG;
    if (p->depth != HV_DEPTH_8 && p->depth != HV_DEPTH_16) return HV_ERR_INVALID_ARG;
    if (p->n_discs > 4 || p->bits == 0 || p->bits > 16) return HV_ERR_INVALID_ARG;
    const uint32_t W = p->width, H = p->height;
    const size_t P = (size_t)W * H;
    const long maxv = (1L << p->bits) - 1;
    const double K = p->mean_counts / (1.0 + (double)p->sigma * p->sigma);
    const double cx = (W - 1) / 2.0, cy = (H - 1) / 2.0;
    const double rmax2 = cx * cx + cy * cy;
    // per-pixel source state (fixed per source): O_0 * quadratic phase, f_d
    std::vector<double> ore(P), oim(P);
    std::vector<int> fd(P, 0);
    auto per_pixel = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const double x = (double)(i % W), y = (double)(i / W);
            Rng r(p->seed, 1, 0, i);
            const double n1 = r.normal(), n2 = r.normal();
            double a = p->sigma * n1 / std::sqrt(2.0), b = p->sigma * n2 / std::sqrt(2.0);
            if (p->quad_phase != 0.f) {
                const double ph = p->quad_phase * ((x - cx) * (x - cx) + (y - cy) * (y - cy)) / rmax2;
                const double c = std::cos(ph), s = std::sin(ph);
                const double na = a * c - b * s, nb = a * s + b * c;
                a = na;
                b = nb;
            }
            ore[i] = a;
            oim[i] = b;
            for (uint32_t d = 0; d < p->n_discs; ++d) {
                const double dx = x - p->disc_cx[d], dy = y - p->disc_cy[d];
                if (dx * dx + dy * dy <= (double)p->disc_r[d] * p->disc_r[d]) {
                    Rng rd(p->seed, 2 + d, 0, i);
                    const int span = p->fd_hi[d] - p->fd_lo[d] + 1;
                    fd[i] = p->fd_lo[d] + (int)std::min<double>(span - 1, std::floor(rd.uniform() * span));
                    break;
                }
            }
        }
    };
    auto per_frame = [&](size_t lo, size_t hi) {
        for (size_t idx = lo; idx < hi; ++idx) {
            const uint64_t f = idx / P;
            const size_t i = idx % P;
            const uint64_t t = first_frame + f;
            const double x = (double)(i % W), y = (double)(i / W);
            const double pr = 2.0 * M_PI * (p->tilt_x * x + p->tilt_y * y);
            double re = std::cos(pr), im = std::sin(pr);
            const double pd = 2.0 * M_PI * (double)fd[i] * (double)(t % p->period) / (double)p->period;
            const double c = std::cos(pd), s = std::sin(pd);
            re += ore[i] * c - oim[i] * s;
            im += ore[i] * s + oim[i] * c;
            const double I = re * re + im * im;
            Rng r(p->seed, 7, t, i);
            long v = r.poisson(K * I);
            if (v > maxv) v = maxv;
            if (p->depth == HV_DEPTH_8) ((uint8_t*)out)[idx] = (uint8_t)v;
            else ((uint16_t*)out)[idx] = (uint16_t)v;
        }
    };
    auto parallel = [](size_t n, auto&& fn) {
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        if (n < 65536) nt = 1;
        std::vector<std::thread> th;
        const size_t chunk = (n + nt - 1) / nt;
        for (unsigned k = 0; k < nt; ++k) {
            const size_t lo = k * chunk, hi = std::min(n, lo + chunk);
            if (lo < hi) th.emplace_back(fn, lo, hi);
        }
        for (auto& t : th) t.join();
    };
    parallel(P, per_pixel);
    parallel(P * count, per_frame);
    return HV_OK;
}
