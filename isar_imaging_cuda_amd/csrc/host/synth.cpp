// synth.cpp — seeded synthetic STRETCH echo generator and .wbd writer (see isar_synth.h).
#include "isar_synth.h"
#include "synth_model.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 3e8;

// SplitMix64 (Steele, Lea, Flood 2014): a 64-bit counter passed through a fixed mixer.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
    double gauss() {                                                            // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

uint64_t stream_seed(uint64_t seed, uint64_t purpose, uint64_t index) {
    Rng r(seed ^ (purpose * 0xD1B54A32D192ED03ULL) ^ (index * 0xABC98388FB8FAC03ULL));
    r.next();
    return r.next();
}

enum : uint64_t { kScene = 1, kPhase = 2, kWalk = 3, kNoise = 4 };

}  // namespace

void isar_synth_model(const isar_scene_config* c, isar_synth_model_t* out) {
    const int M = c->echo_num, S = c->n_scatterers;
    const double dr = kC / (2.0 * c->band_width);
    const double T = M / c->prf;
    out->sx.assign(S, 0.0);
    out->sy.assign(S, 0.0);
    out->sz.assign(S, 0.0);
    out->amp.assign(S, cd(0, 0));
    // scene
    {
        Rng r(stream_seed(c->seed, kScene, 0));
        for (int i = 0; i < S; ++i) {
            out->sx[i] = (r.uniform() - 0.5) * c->box_x;
            out->sy[i] = (r.uniform() - 0.5) * c->box_y;
            out->sz[i] = (r.uniform() - 0.5) * c->box_z;
            double g1 = r.gauss(), g2 = r.gauss();   // Rayleigh magnitude (sigma = 1), uniform phase
            out->amp[i] = cd(g1, g2);
        }
    }
    // per-pulse phase error
    out->perr.assign(M, 0.0);
    {
        Rng r(stream_seed(c->seed, kPhase, 0));
        if (c->phase_error_model == ISAR_PHASE_UNIFORM) {
            for (int m = 0; m < M; ++m) out->perr[m] = (2.0 * r.uniform() - 1.0) * kPi;
        } else if (c->phase_error_model == ISAR_PHASE_POLY) {
            double ck[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int k = 2; k <= 6; ++k) ck[k] = c->poly_sigma * r.gauss();
            for (int m = 0; m < M; ++m) {
                double tau = M > 1 ? -1.0 + 2.0 * m / (M - 1) : 0.0, acc = 0, tp = tau * tau;
                for (int k = 2; k <= 6; ++k) { acc += ck[k] * tp; tp *= tau; }
                out->perr[m] = acc + c->white_sigma * r.gauss();
            }
        }
    }
    // residual range walk after coarse tracking
    out->walk.assign(M, 0.0);
    {
        Rng r(stream_seed(c->seed, kWalk, 0));
        double psi = 2.0 * kPi * r.uniform();
        for (int m = 0; m < M; ++m) {
            double t = m / c->prf;
            if (c->walk_cells < 0) {
                out->walk[m] = c->v * t + 0.5 * c->a * t * t;
            } else if (c->walk_cells > 0) {
                double u = T > 0 ? t / T : 0;
                out->walk[m] = dr * c->walk_cells * (0.6 * (2 * u - 1) + 0.35 * std::sin(2 * kPi * 1.3 * u + psi)) +
                               dr * 0.3 * r.gauss();
            }
        }
    }
}

uint64_t isar_synth_noise_seed(uint64_t seed, int row) { return stream_seed(seed, kNoise, static_cast<uint64_t>(row)); }


extern "C" {

int isar_synth_preset(int id, isar_scene_config* c) {
    if (!c) return -1;
    std::memset(c, 0, sizeof(*c));
    c->fc = 10e9;
    c->band_width = 1e9;
    c->prf = 200.0;
    c->fs = 64e6;
    c->peak_adc = 12000.0;
    c->quantize = 1;
    c->seed = 0x1234ABCDULL + static_cast<uint64_t>(id);
    switch (id) {
    case 6:  // IFDS raw echoes: 128 pulses of 6000 samples at 32 MHz, 20 MHz chirp of 40 us
        c->echo_num = 128; c->range_num = 6000; c->n_scatterers = 40;
        c->phase_error_model = ISAR_PHASE_UNIFORM;
        c->box_x = 60; c->box_y = 60; c->box_z = 0;
        c->band_width = 20e6; c->fs = 32e6; c->tp = 40e-6;
        c->omega = 0.02; c->v = 20; c->a = 0; c->walk_cells = 3; c->snr_db = 20;
        return 0;
    case 1:  // 128 x 1024, 40 scatterers in 20 x 20 m, v = 20 m/s untracked, iid phase, 20 dB
        c->echo_num = 128; c->range_num = 1024; c->n_scatterers = 40;
        c->box_x = 20; c->box_y = 20; c->box_z = 0;
        c->omega = 0.01; c->v = 20; c->a = 0; c->walk_cells = -1;
        c->phase_error_model = ISAR_PHASE_UNIFORM; c->snr_db = 20;
        break;
    case 2:  // 256 x 4096, 200 scatterers in a 60 x 15 x 10 m ship box, high speed
        c->echo_num = 256; c->range_num = 4096; c->n_scatterers = 200;
        c->box_x = 60; c->box_y = 15; c->box_z = 10; c->elevation = 5.0 * kPi / 180.0;
        c->prf = 400.0;   // keeps the +-100 Hz Doppler spread of the 60 m hull unaliased
        c->omega = 0.05; c->v = 3000; c->a = 50; c->walk_cells = 48;
        c->phase_error_model = ISAR_PHASE_POLY; c->poly_sigma = std::sqrt(2.0); c->white_sigma = 0.5;
        c->snr_db = 15;
        break;
    case 3:  // one 8192-pulse dwell x 4096 cells with the config-1 scene and motion (tracked)
        c->echo_num = 8192; c->range_num = 4096; c->n_scatterers = 40;
        c->box_x = 20; c->box_y = 20; c->box_z = 0;
        c->omega = 0.01; c->v = 20; c->a = 0; c->walk_cells = 48;
        c->phase_error_model = ISAR_PHASE_UNIFORM; c->snr_db = 20;
        break;
    case 4:  // 256 x 2048 range-aligned frames, iid phase error, 15 dB
        c->echo_num = 256; c->range_num = 2048; c->n_scatterers = 40;
        c->box_x = 20; c->box_y = 20; c->box_z = 0;
        c->omega = 0.01; c->v = 20; c->a = 0; c->walk_cells = 0;
        c->phase_error_model = ISAR_PHASE_UNIFORM; c->snr_db = 15;
        break;
    case 5:  // 1024 x 16384, 1000 scatterers, v = 7000 m/s
        c->echo_num = 1024; c->range_num = 16384; c->n_scatterers = 1000;
        c->box_x = 30; c->box_y = 15; c->box_z = 10; c->elevation = 5.0 * kPi / 180.0;
        c->omega = 0.08; c->v = 7000; c->a = 100; c->walk_cells = 48;
        c->phase_error_model = ISAR_PHASE_POLY; c->poly_sigma = std::sqrt(2.0); c->white_sigma = 0.5;
        c->snr_db = 15;
        break;
    default:
        return -1;
    }
    c->tp = c->range_num / c->fs;
    return 0;
}

void isar_synth_radar_header(const isar_scene_config* c, int64_t* B, int64_t* fc, double* tp, int64_t* fs) {
    // the .wbd header stores MHz and microseconds as integers (common.cu:816-826)
    if (B) *B = static_cast<int64_t>(std::llround(c->band_width / 1e6)) * 1000000LL;
    if (fc) *fc = static_cast<int64_t>(std::llround(c->fc / 1e6)) * 1000000LL;
    if (tp) *tp = std::llround(c->tp * 1e6) / 1e6;
    if (fs) *fs = static_cast<int64_t>(std::llround(c->fs / 1e6)) * 1000000LL;
}

int isar_synth_frame(const isar_scene_config* c, float* echo, double* velocity, double* truth_phase,
                     double* truth_walk) {
    if (!c || !echo || !velocity || c->echo_num < 1 || c->range_num < 1 || c->n_scatterers < 0) return -1;
    const int M = c->echo_num, N = c->range_num, S = c->n_scatterers;
    const double K = c->band_width / c->tp;
    const double f_hi = c->fc + c->band_width / 2;   // down-chirp starts at the top of the band
    const double T = M / c->prf;
    isar_synth_model_t md;
    isar_synth_model(c, &md);
    const std::vector<double>&sx = md.sx, &sy = md.sy, &sz = md.sz, &perr = md.perr, &walk = md.walk;
    const std::vector<cd>& amp = md.amp;
    const double cel = std::cos(c->elevation), sel = std::sin(c->elevation);
    std::vector<double> pw(M, 0.0);

#pragma omp parallel
    {
        std::vector<cd> acc(N);
#pragma omp for schedule(dynamic, 4)
        for (int m = 0; m < M; ++m) {
            const double t = m / c->prf;
            const double vm = c->v + c->a * t;
            const double th = c->omega * (t - T / 2);
            const double ct = std::cos(th), st = std::sin(th);
            std::fill(acc.begin(), acc.end(), cd(0, 0));
            for (int i = 0; i < S; ++i) {
                const double dR = walk[m] + (sy[i] * ct - sx[i] * st) * cel + sz[i] * sel;
                const double phi0 = -(4 * kPi / kC) * f_hi * dR + perr[m];
                const double phi1 = (-(4 * kPi / kC) * f_hi * vm + (4 * kPi * K / kC) * dR) / c->fs;
                const cd rot(std::cos(phi1), std::sin(phi1));
                for (int n0 = 0; n0 < N; n0 += 256) {   // re-anchor the recurrence every 256 samples
                    const double a0 = phi0 + phi1 * n0;
                    cd z = amp[i] * cd(std::cos(a0), std::sin(a0));
                    const int n1 = std::min(N, n0 + 256);
                    for (int n = n0; n < n1; ++n) { acc[n] += z; z *= rot; }
                }
            }
            const double phi2 = (4 * kPi * K / kC) * vm / (c->fs * c->fs);
            double p = 0;
            for (int n = 0; n < N; ++n) {
                const double a2 = phi2 * static_cast<double>(n) * n;
                acc[n] *= cd(std::cos(a2), std::sin(a2));
                p += std::norm(acc[n]);
                echo[2 * (static_cast<size_t>(m) * N + n)] = static_cast<float>(acc[n].real());
                echo[2 * (static_cast<size_t>(m) * N + n) + 1] = static_cast<float>(acc[n].imag());
            }
            pw[m] = p;
            velocity[m] = vm;
        }
    }
    double psig = 0;
    for (double p : pw) psig += p;
    psig /= static_cast<double>(M) * N;
    const bool noisy = c->snr_db < 200 && psig > 0;
    const double nsig = noisy ? std::sqrt(psig / std::pow(10.0, c->snr_db / 10.0) / 2.0) : 0.0;
    const size_t tot = 2 * static_cast<size_t>(M) * N;
    if (noisy) {
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) {
            Rng r(stream_seed(c->seed, kNoise, static_cast<uint64_t>(m)));
            for (int q = 0; q < 2 * N; ++q) echo[2 * static_cast<size_t>(m) * N + q] += static_cast<float>(nsig * r.gauss());
        }
    }
    if (c->peak_adc > 0) {
        float pk = 0;
        for (size_t i = 0; i < tot; ++i) pk = std::max(pk, std::fabs(echo[i]));
        const double g = pk > 0 ? c->peak_adc / pk : 1.0;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < static_cast<long long>(tot); ++i) {
            double v = echo[i] * g;
            if (c->quantize) v = std::max(-32768.0, std::min(32767.0, std::nearbyint(v)));
            echo[i] = static_cast<float>(v);
        }
    }
    if (truth_phase) std::copy(perr.begin(), perr.end(), truth_phase);
    if (truth_walk) std::copy(walk.begin(), walk.end(), truth_walk);
    return 0;
}

int isar_synth_ifds(const isar_scene_config* c, int len, float* iq, double* velocity) {
    if (!c || !iq || !velocity || c->echo_num < 1 || len < 2 || c->n_scatterers < 0) return -1;
    const int M = c->echo_num, Sc = c->n_scatterers;
    const int S = static_cast<int>(c->fs * c->tp);   // reference chirp samples (common.cu:671)
    if (S < 2 || S > len) return -1;
    const double T = M / c->prf;
    const double cell = kC / (2.0 * c->fs);           // metres per sample of two-way delay
    std::vector<double> sx(Sc), sy(Sc), sz(Sc);
    std::vector<cd> amp(Sc);
    {
        Rng r(stream_seed(c->seed, kScene, 0));
        for (int i = 0; i < Sc; ++i) {
            sx[i] = (r.uniform() - 0.5) * c->box_x;
            sy[i] = (r.uniform() - 0.5) * c->box_y;
            sz[i] = (r.uniform() - 0.5) * c->box_z;
            double g1 = r.gauss(), g2 = r.gauss();
            amp[i] = cd(g1, g2);
        }
    }
    std::vector<double> perr(M, 0.0);
    if (c->phase_error_model == ISAR_PHASE_UNIFORM) {
        Rng r(stream_seed(c->seed, kPhase, 0));
        for (int m = 0; m < M; ++m) perr[m] = (2.0 * r.uniform() - 1.0) * kPi;
    }
    const double cel = std::cos(c->elevation), sel = std::sin(c->elevation);
    const int d0 = (len - S) / 2;
    const double c2 = kPi * c->band_width / c->tp;
    std::vector<double> pw(M, 0.0);
#pragma omp parallel
    {
        std::vector<cd> acc(len), u(S);
#pragma omp for schedule(dynamic, 2)
        for (int m = 0; m < M; ++m) {
            const double t = m / c->prf;
            const double vm = c->v + c->a * t;
            const double v2 = 2.0 * vm / kC, v2_ = 1.0 - v2;
            const double c1 = -2.0 * kPi * c->fc * v2 / v2_;
            for (int k = 0; k < S; ++k) {   // the transmitted chirp as the receiver models it
                const double tk = v2_ * (-c->tp / 2.0 + k * c->tp / (static_cast<double>(S) - 1.0));
                const double ph = (c1 + c2 * tk) * tk;
                u[k] = cd(std::cos(ph), std::sin(ph));
            }
            const double walk = c->walk_cells * cell * (2.0 * t / (T > 0 ? T : 1.0) - 1.0);
            const double th = c->omega * (t - T / 2);
            const double ct = std::cos(th), st = std::sin(th);
            std::fill(acc.begin(), acc.end(), cd(0, 0));
            for (int i = 0; i < Sc; ++i) {
                const double dR = walk + (sy[i] * ct - sx[i] * st) * cel + sz[i] * sel;
                const int d = d0 + static_cast<int>(std::lround(dR / cell));
                const double phi = -(4 * kPi / kC) * c->fc * dR + perr[m];
                const cd z = amp[i] * cd(std::cos(phi), std::sin(phi));
                for (int k = std::max(0, -d); k < S && d + k < len; ++k) acc[d + k] += z * u[k];
            }
            double p = 0;
            for (int k = 0; k < len; ++k) {
                p += std::norm(acc[k]);
                iq[2 * (static_cast<size_t>(m) * len + k)] = static_cast<float>(acc[k].real());
                iq[2 * (static_cast<size_t>(m) * len + k) + 1] = static_cast<float>(acc[k].imag());
            }
            pw[m] = p;
            velocity[m] = vm;
        }
    }
    double psig = 0;
    for (double p : pw) psig += p;
    psig /= static_cast<double>(M) * len;
    const bool noisy = c->snr_db < 200 && psig > 0;
    const double nsig = noisy ? std::sqrt(psig / std::pow(10.0, c->snr_db / 10.0) / 2.0) : 0.0;
    const size_t tot = 2 * static_cast<size_t>(M) * len;
    if (noisy) {
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) {
            Rng r(stream_seed(c->seed, kNoise, static_cast<uint64_t>(m)));
            for (int q = 0; q < 2 * len; ++q) iq[2 * static_cast<size_t>(m) * len + q] += static_cast<float>(nsig * r.gauss());
        }
    }
    if (c->peak_adc > 0) {
        float pk = 0;
        for (size_t i = 0; i < tot; ++i) pk = std::max(pk, std::fabs(iq[i]));
        const double g = pk > 0 ? c->peak_adc / pk : 1.0;
        for (size_t i = 0; i < tot; ++i) {
            double v = iq[i] * g;
            if (c->quantize) v = std::max(-32768.0, std::min(32767.0, std::nearbyint(v)));
            iq[i] = static_cast<float>(v);
        }
    }
    return 0;
}

int isar_write_wbd(const char* path, const float* echo, int frames, int N, int64_t fc_hz, int64_t B_hz,
                   double tp_s, int64_t fs_hz, const double* range_m, const double* vel,
                   const double* az, const double* pitch) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    const uint32_t frame_len = 256u + 4u * static_cast<uint32_t>(N);
    std::vector<uint8_t> buf(frame_len);
    auto put32 = [&](int word, uint32_t v) { std::memcpy(&buf[4 * word], &v, 4); };
    auto angle_word = [](double deg) {   // int32 in units of 360 / 2^24 degrees
        double u = deg * (16777216.0 / 360.0);
        return static_cast<uint32_t>(static_cast<int32_t>(std::llround(u)));
    };
    for (int k = 0; k < frames; ++k) {
        std::fill(buf.begin(), buf.end(), 0);
        put32(4, frame_len / 4);                                              // frame length / 4
        put32(12, static_cast<uint32_t>(fc_hz / 1000000));                    // fc, MHz
        put32(13, static_cast<uint32_t>(B_hz / 1000000));                     // bandwidth, MHz
        put32(15, static_cast<uint32_t>(std::llround(tp_s * 1e6)));           // pulse width, us
        put32(17, static_cast<uint32_t>(fs_hz / 1000000) & 0xFFFFu);          // Fs, MHz (low 16)
        put32(19, static_cast<uint32_t>(std::llround((range_m ? range_m[k] : 0.0) * 10.0)));
        put32(20, static_cast<uint32_t>(static_cast<int32_t>(std::llround((vel ? vel[k] : 0.0) * 10.0))));
        put32(21, angle_word(az ? az[k] : 0.0));
        put32(22, angle_word(pitch ? pitch[k] : 0.0));
        int16_t* iq = reinterpret_cast<int16_t*>(&buf[256]);
        for (int n = 0; n < 2 * N; ++n) {
            double v = std::nearbyint(echo[2 * static_cast<size_t>(k) * N + n]);
            iq[n] = static_cast<int16_t>(std::max(-32768.0, std::min(32767.0, v)));
        }
        if (std::fwrite(buf.data(), 1, frame_len, f) != frame_len) { std::fclose(f); return -1; }
    }
    std::fclose(f);
    return 0;
}

}  // extern "C"
