// Host: impulse response of the 16-line Hadamard feedback delay network that
// renders the late reverberation tail (acoustic_agent/steam_rt.py:2235-2328).
//
// Per sample n, every line l reads the value at its delay position, runs it
// through its cascade of num_bands biquads (direct form I, b0 b1 b2 a1 a2),
// and the wave of line outputs is mixed by the unnormalised Sylvester
// Hadamard matrix (in-place butterflies, stride 1, 2, 4, ...); each line then
// stores (n == 0 ? 1 : 0) + 0.25 * mixed[l] at its position, which advances
// modulo its delay.  The output sample is the mean of the line outputs, run
// through the Schroeder all-pass stages (w = x + 0.5 w[n-m]; y = w[n-m] -
// 0.5 w) and the tone biquad cascade.  Everything is binary64 in this order,
// built without contraction, so it rounds as the reference recursion does.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/aa_tracer.h"
#include "aa_internal.h"

namespace {

struct Biquad {
    double b0, b1, b2, a1, a2;
    double x1 = 0.0, x2 = 0.0, y1 = 0.0, y2 = 0.0;
    double step(double x) {
        const double y = b0 * x + b1 * x1 + b2 * x2 - a1 * y1 - a2 * y2;
        x2 = x1;
        x1 = x;
        y2 = y1;
        y1 = y;
        return y;
    }
};

}  // namespace

extern "C" int aa_fdn_impulse(const int64_t* delays, int32_t num_delays, const double* absorptive,
                              int32_t num_bands, int64_t num_samples, const int64_t* allpass_delays,
                              int32_t num_allpass, const double* tone, double* out) {
    if (!delays || !absorptive || !tone || !out || (num_allpass > 0 && !allpass_delays))
        return aa::set_error(AA_ERR_INVALID, "aa_fdn_impulse: NULL array");
    if (num_delays < 1 || (num_delays & (num_delays - 1)) != 0 || num_bands < 0 || num_samples < 0 ||
        num_allpass < 0)
        return aa::set_error(AA_ERR_INVALID, "aa_fdn_impulse: need a power-of-two line count and non-negative sizes");
    for (int32_t l = 0; l < num_delays; ++l)
        if (delays[l] < 1) return aa::set_error(AA_ERR_INVALID, "aa_fdn_impulse: delay %d is < 1", (int)l);
    for (int32_t a = 0; a < num_allpass; ++a)
        if (allpass_delays[a] < 1) return aa::set_error(AA_ERR_INVALID, "aa_fdn_impulse: all-pass delay %d < 1", (int)a);

    const int L = num_delays, NB = num_bands;
    std::vector<std::vector<double>> line(L);
    std::vector<int64_t> pos(L, 0);
    for (int l = 0; l < L; ++l) line[l].assign((size_t)delays[l], 0.0);
    std::vector<Biquad> filt((size_t)L * NB);
    for (int l = 0; l < L; ++l)
        for (int b = 0; b < NB; ++b) {
            const double* c = absorptive + ((size_t)l * NB + b) * 5;
            filt[(size_t)l * NB + b] = Biquad{c[0], c[1], c[2], c[3], c[4]};
        }
    std::vector<Biquad> tone_f((size_t)NB);
    for (int b = 0; b < NB; ++b) tone_f[b] = Biquad{tone[b * 5], tone[b * 5 + 1], tone[b * 5 + 2], tone[b * 5 + 3],
                                                    tone[b * 5 + 4]};
    std::vector<std::vector<double>> ap(num_allpass);
    std::vector<int64_t> ap_pos(num_allpass, 0);
    for (int a = 0; a < num_allpass; ++a) ap[a].assign((size_t)allpass_delays[a], 0.0);
    std::vector<double> delayed(L), mixed(L);

    for (int64_t n = 0; n < num_samples; ++n) {
        for (int l = 0; l < L; ++l) {
            double v = line[l][(size_t)pos[l]];
            for (int b = 0; b < NB; ++b) v = filt[(size_t)l * NB + b].step(v);
            delayed[l] = v;
            mixed[l] = v;
        }
        for (int stride = 1; stride < L; stride *= 2)
            for (int base = 0; base < L; base += 2 * stride)
                for (int o = 0; o < stride; ++o) {
                    const double a = mixed[base + o], b = mixed[base + o + stride];
                    mixed[base + o] = a + b;
                    mixed[base + o + stride] = a - b;
                }
        const double injection = n == 0 ? 1.0 : 0.0;
        for (int l = 0; l < L; ++l) {
            line[l][(size_t)pos[l]] = injection + 0.25 * mixed[l];
            if (++pos[l] >= delays[l]) pos[l] = 0;
        }
        double v = 0.0;
        for (int l = 0; l < L; ++l) v += delayed[l];
        v /= L;
        for (int a = 0; a < num_allpass; ++a) {
            const double prev = ap[a][(size_t)ap_pos[a]];
            const double w = v + 0.5 * prev;
            ap[a][(size_t)ap_pos[a]] = w;
            v = prev - 0.5 * w;
            if (++ap_pos[a] >= allpass_delays[a]) ap_pos[a] = 0;
        }
        for (int b = 0; b < NB; ++b) v = tone_f[b].step(v);
        out[n] = v;
    }
    return AA_OK;
}
