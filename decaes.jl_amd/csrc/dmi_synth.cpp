// Seeded synthetic multi-echo spin-echo volumes: the workload generator for tests and benchmarks.
//
// Per in-mask voxel: s = sum_c f_c * EPG(alpha, T2_c; TE, T1, refocusing 180) plus Rician noise
// sqrt((s + sigma z1)^2 + (sigma z2)^2) with sigma = max(s) / SNR (linear SNR). Modelled on mock_image
// (src/utils.jl:623-658) with the compartments of the benchmark workloads:
//   short  T2 ~ U(15,25) ms, fraction ~ U(0,0.3)
//   long   T2 ~ U(1,2) s,    fraction ~ U(0,0.1)
//   (4th)  T2 ~ U(35,50) ms, fraction ~ U(0,0.2)   [ncomp = 4 only]
//   middle T2 ~ U(60,90) ms, fraction = 1 - the others
//   alpha ~ U(120,180) deg, SNR ~ U(50,200).
// Voxels outside a seeded ellipsoid holding `mask_fraction` of the volume are zero.
// Every random number is a pure function of (seed, voxel, draw index): results never depend on threading.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "decaes_amd.h"

namespace {

inline uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

struct CounterRng {
    uint64_t key;
    uint64_t ctr = 0;
    double uniform() { return double(mix64(key ^ mix64(ctr++)) >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Constant-flip EPG echo train (Prasloski 2012 appendix; same state update as the device kernels):
// per echo, each state (F, Fbar, Z) maps to F_{k+1}, Fbar_{k-1}, Z_k through the refocusing pulse with
// relaxation over TE; the echo is the state-1 F after the transition.
void epg_const(double *dc, int etl, double alpha_deg, double TE, double T2, double T1, std::vector<double> &F,
               std::vector<double> &Fb, std::vector<double> &Z) {
    const int L = etl + 2;
    F.assign(L, 0.0);
    Fb.assign(L, 0.0);
    Z.assign(L, 0.0);
    const double al = alpha_deg * (3.141592653589793 / 180.0);
    const double E1 = std::exp(-(TE / 2) / T1), E2 = std::exp(-(TE / 2) / T2);
    const double a = (E2 * E2) / 2, b = a * std::cos(al), c = E1 * E2 * std::sin(al), d = E1 * E1 * std::cos(al);
    const double cp = -c / 2;
    F[0] = 1.0;
    std::vector<double> nF(L), nFb(L), nZ(L);
    for (int i = 0; i < etl; ++i) {
        for (int s = 0; s < L; ++s) {
            double C = F[s] + Fb[s], S = F[s] - Fb[s], Cp = a * C, Sp = b * S;
            double fr = std::fma(c, Z[s], Cp + Sp), fbr = std::fma(-c, Z[s], Cp - Sp), zr = std::fma(cp, S, d * Z[s]);
            if (s + 1 < L) nF[s + 1] = fr;
            if (s == 0) nF[0] = fbr;
            else nFb[s - 1] = fbr;
            nZ[s] = zr;
        }
        nFb[L - 1] = 0.0;
        F.swap(nF);
        Fb.swap(nFb);
        Z.swap(nZ);
        dc[i] = F[0];
    }
    const double m0 = std::sin(alpha_deg / 2 * (3.141592653589793 / 180.0));
    for (int i = 0; i < etl; ++i) dc[i] = std::fabs(m0 * dc[i]);
}

struct Ellipsoid {
    double c[3], r[3];
    bool inside(int64_t x, int64_t y, int64_t z) const {
        double u = (x + 0.5 - c[0]) / r[0], v = (y + 0.5 - c[1]) / r[1], w = (z + 0.5 - c[2]) / r[2];
        return u * u + v * v + w * w <= 1.0;
    }
};

}  // namespace

extern "C" int64_t dmi_synth_mse(double *image, int64_t nx, int64_t ny, int64_t nz, int32_t nTE, double TE,
                                 double T1, int32_t ncomp, double mask_fraction, uint64_t seed, int32_t nthreads) {
    if (!image || nx < 1 || ny < 1 || nz < 1 || nTE < 1 || !(TE > 0) || (ncomp != 3 && ncomp != 4)) return -1;
    const int64_t V = nx * ny * nz;
    // Seeded ellipsoid: centre jitter and axis ratios from the seed, then a common scale chosen by bisection
    // so that the fraction of voxel centres inside matches mask_fraction.
    CounterRng g{mix64(seed ^ 0x5eed0e11ULL)};
    Ellipsoid e;
    const double n3[3] = {double(nx), double(ny), double(nz)};
    double ax[3];
    for (int k = 0; k < 3; ++k) {
        e.c[k] = n3[k] / 2 * (1.0 + g.uniform(-0.05, 0.05));
        ax[k] = n3[k] / 2 * g.uniform(0.9, 1.1);
    }
    auto frac_inside = [&](double s) {
        Ellipsoid t = e;
        for (int k = 0; k < 3; ++k) t.r[k] = s * ax[k];
        int64_t cnt = 0;
        for (int64_t z = 0; z < nz; ++z)
            for (int64_t y = 0; y < ny; ++y)
                for (int64_t x = 0; x < nx; ++x) cnt += t.inside(x, y, z);
        return double(cnt) / double(V);
    };
    double slo = 0.0, shi = 4.0;
    if (mask_fraction >= 1.0) {
        slo = shi = 1e9;
    } else {
        for (int it = 0; it < 40; ++it) {
            double s = 0.5 * (slo + shi);
            (frac_inside(s) < mask_fraction ? slo : shi) = s;
        }
    }
    for (int k = 0; k < 3; ++k) e.r[k] = shi * ax[k];

    int nt = nthreads > 0 ? nthreads : int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<int64_t> counts(nt, 0);
    auto work = [&](int t) {
        std::vector<double> F, Fb, Z, dc(nTE), s(nTE);
        int64_t cnt = 0;
        for (int64_t I = t; I < V; I += nt) {
            int64_t x = I % nx, y = (I / nx) % ny, z = I / (nx * ny);
            if (!e.inside(x, y, z)) {
                for (int i = 0; i < nTE; ++i) image[I + V * i] = 0.0;
                continue;
            }
            ++cnt;
            CounterRng r{mix64(seed) ^ mix64(uint64_t(I) * 0x9e3779b97f4a7c15ULL + 1)};
            double T2s = r.uniform(15e-3, 25e-3), fs = r.uniform(0.0, 0.3);
            double T2m = r.uniform(60e-3, 90e-3);
            double T2l = r.uniform(1.0, 2.0), fl = r.uniform(0.0, 0.1);
            double T2c = 0, fc = 0;
            if (ncomp == 4) { T2c = r.uniform(35e-3, 50e-3); fc = r.uniform(0.0, 0.2); }
            double fm = 1.0 - fs - fl - fc;
            double alpha = r.uniform(120.0, 180.0);
            double snr = r.uniform(50.0, 200.0);
            for (int i = 0; i < nTE; ++i) s[i] = 0.0;
            const double T2v[4] = {T2s, T2m, T2l, T2c}, fv[4] = {fs, fm, fl, fc};
            for (int k = 0; k < ncomp; ++k) {
                epg_const(dc.data(), nTE, alpha, TE, T2v[k], T1, F, Fb, Z);
                for (int i = 0; i < nTE; ++i) s[i] += fv[k] * dc[i];
            }
            double smax = 0.0;
            for (int i = 0; i < nTE; ++i) smax = std::max(smax, s[i]);
            double sigma = smax / snr;
            for (int i = 0; i < nTE; ++i) {
                double z1 = r.normal(), z2 = r.normal();
                double re = s[i] + sigma * z1, im = sigma * z2;
                image[I + V * i] = std::sqrt(re * re + im * im);
            }
        }
        counts[t] = cnt;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &h : th) h.join();
    int64_t total = 0;
    for (auto c : counts) total += c;
    return total;
}
