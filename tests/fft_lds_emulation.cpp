// CPU run of the LDS FFT convolution in rb_fft_lds.h: the device code's group functions and pass
// schedules (fft_convolve_r16, fft_convolve_r16_br) executed by a host runner that walks the groups
// of each pass in order (on the device the groups of a pass are disjoint and the passes are
// separated by barriers, so the order does not matter), checked against a direct FP64 linear
// convolution of two real paths for every transform length, including ragged N.  Exit status 0
// when every case is within 2e-5 relative of the direct sum and no pad slot was written.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "rb_fft_lds.h"

using namespace rb;
typedef std::complex<double> cd;

struct HostRunner {
    template <int LOG2L, int SB, int R, int KIND>
    void pass(float2* z, const float2* tw, const float2* G) const {
        for (int t = 0; t < ((1 << LOG2L) >> R); ++t) fft_group<LOG2L, SB, R, KIND>(z, tw, G, t);
    }
};

static unsigned brev(unsigned i, int b) {
    unsigned r = 0;
    for (int k = 0; k < b; ++k) r |= ((i >> k) & 1u) << (b - 1 - k);
    return r;
}

// FP64 discrete Fourier transform (recursive radix 2), sign -1 forward
static void dft(std::vector<cd>& a, int sign) {
    const size_t n = a.size();
    if (n == 1) return;
    std::vector<cd> e(n / 2), o(n / 2);
    for (size_t i = 0; i < n / 2; ++i) {
        e[i] = a[2 * i];
        o[i] = a[2 * i + 1];
    }
    dft(e, sign);
    dft(o, sign);
    for (size_t k = 0; k < n / 2; ++k) {
        const cd t = std::polar(1.0, sign * 2.0 * M_PI * double(k) / double(n)) * o[k];
        a[k] = e[k] + t;
        a[k + n / 2] = e[k] - t;
    }
}

static bool pads_clean(const std::vector<float2>& z) {
    for (size_t i = 32; i < z.size(); i += 33)
        if (!std::isnan(z[i].x)) return false;
    return true;
}

template <int LOG2L>
static int check(int N) {
    const int L = 1 << LOG2L;
    std::vector<float2> tw(L / 2);
    for (int k = 0; k < L / 2; ++k) {
        const double a = -2.0 * M_PI * k / L;
        tw[k] = make_float2(float(std::cos(a)), float(std::sin(a)));
    }
    std::mt19937 g(LOG2L * 7 + N);
    std::normal_distribution<double> nd;
    std::vector<double> xa(N), xb(N), gam(N);
    for (int i = 0; i < N; ++i) {
        xa[i] = nd(g);
        xb[i] = nd(g);
        gam[i] = 1.0 / (1.0 + i) + 0.1 * nd(g);
    }
    std::vector<cd> gh(L, 0.0);
    for (int i = 0; i < N; ++i) gh[i] = gam[i];
    dft(gh, -1);
    // direct linear convolution (the circular one of length L >= 2N - 1 equals it below N)
    std::vector<double> ra(N, 0.0), rb(N, 0.0);
    double scale = 0.0;
    for (int i = 0; i < N; ++i) {
        for (int k = 0; k <= i; ++k) {
            ra[i] += gam[k] * xa[i - k];
            rb[i] += gam[k] * xb[i - k];
        }
        scale = std::max(scale, std::max(std::abs(ra[i]), std::abs(rb[i])));
    }
    int bad = 0;
    // natural order: forward top stage while writing, spectrum bit-reversed, inverse top stage
    // while reading back
    {
        std::vector<float2> G(L);
        for (int p = 0; p < L; ++p) {
            const cd v = gh[brev(p, LOG2L)];
            G[p] = make_float2(float(v.real()), float(v.imag()));
        }
        std::vector<float2> z(fft_lds_slots(LOG2L), make_float2(NAN, NAN));
        for (int i = 0; i < L / 2; ++i) {
            const float2 x = i < N ? make_float2(float(xa[i]), float(xb[i])) : make_float2(0.f, 0.f);
            z[fpad(i)] = x;
            z[fpad(i + L / 2)] = cmulf(x, tw[i]);
        }
        fft_convolve_r16<LOG2L>(z.data(), tw.data(), G.data(), HostRunner{});
        double e = 0.0;
        for (int i = 0; i < N; ++i) {
            const float2 lo = z[fpad(i)], hi = z[fpad(i + L / 2)];
            const float2 p = cmulf(make_float2(tw[i].x, -tw[i].y), hi);
            e = std::max(e, std::max(std::abs(double(lo.x + p.x) / L - ra[i]), std::abs(double(lo.y + p.y) / L - rb[i])));
        }
        const bool ok = e <= 2e-5 * scale && pads_clean(z);
        std::printf("natural      L=2^%d N=%5d max err %.2e (scale %.2f) %s\n", LOG2L, N, e, scale, ok ? "ok" : "FAIL");
        bad += !ok;
    }
    // bit-reversed order: input in bit-reversed slots, spectrum natural, output bit-reversed
    if constexpr (LOG2L >= 12) {
        std::vector<float2> G(L);
        for (int p = 0; p < L; ++p) G[p] = make_float2(float(gh[p].real()), float(gh[p].imag()));
        std::vector<float2> z(fft_lds_slots(LOG2L), make_float2(NAN, NAN));
        for (int i = 0; i < L; ++i)
            z[fpad(brev(i, LOG2L))] = i < N ? make_float2(float(xa[i]), float(xb[i])) : make_float2(0.f, 0.f);
        fft_convolve_r16_br<LOG2L>(z.data(), tw.data(), G.data(), HostRunner{});
        double e = 0.0;
        for (int i = 0; i < N; ++i) {
            const float2 o = z[fpad(brev(i, LOG2L))];
            e = std::max(e, std::max(std::abs(double(o.x) / L - ra[i]), std::abs(double(o.y) / L - rb[i])));
        }
        const bool ok = e <= 2e-5 * scale && pads_clean(z);
        std::printf("bit-reversed L=2^%d N=%5d max err %.2e (scale %.2f) %s\n", LOG2L, N, e, scale, ok ? "ok" : "FAIL");
        bad += !ok;
    }
    return bad;
}

int main() {
    int bad = 0;
    bad += check<10>(1) + check<10>(33) + check<10>(512);
    bad += check<11>(513) + check<11>(1024);
    bad += check<12>(1025) + check<12>(2048);
    bad += check<13>(2049) + check<13>(4096);
    bad += check<14>(4097) + check<14>(8192);
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
