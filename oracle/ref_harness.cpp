// ref_harness.cpp -- exports the reference's own PRNG, uniform mapping,
// sampling helpers and common.hpp currents as a C ABI, so golden vectors can
// be generated from the reference code itself.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) against
// the headers where they lie (src/includes/xoshiro.hpp, src/includes/common.hpp
// of the reference); output goes to oracle/_ref/ and is never committed.
// Only these two headers are used because they compile standalone; the rest
// of the reference path needs Eigen, which this image lacks.
#include <cstdint>
#include <cstring>
#include <random>

#include "xoshiro.hpp"
#include "common.hpp"

namespace {

// A generator that replays caller-supplied raw 64-bit outputs. It is passed
// as the Generator template argument of urand/sample_weights/sample_binomial,
// which otherwise seed a thread_local engine from std::random_device.
const uint64_t *g_replay = nullptr;
size_t g_pos = 0;

struct ReplayGen {
    using result_type = std::uint64_t;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~0ULL; }
    explicit ReplayGen(unsigned) {}
    result_type operator()() { return g_replay[g_pos++]; }
};

template <class G>
void fill_raw(uint64_t seed, int n, uint64_t *out) {
    G g(seed);
    for (int i = 0; i < n; ++i) out[i] = g();
}

template <class G>
void fill_uniform(uint64_t seed, int n, double *out) {
    G g(seed);
    std::uniform_real_distribution<double> dist(0.0, 1.0);  // as common.hpp:27
    for (int i = 0; i < n; ++i) out[i] = dist(g);
}

}  // namespace

extern "C" {

int ref_splitmix64(uint64_t seed, int n, uint64_t *out) {
    XoshiroCpp::SplitMix64 sm(seed);
    for (int i = 0; i < n; ++i) out[i] = sm();
    return 0;
}

// prng ids follow include/gw_capi.h GW_PRNG_*
int ref_raw(int prng, uint64_t seed, int n, uint64_t *out) {
    using namespace XoshiroCpp;
    switch (prng) {
    case 0: fill_raw<Xoshiro256StarStar>(seed, n, out); break;
    case 1: fill_raw<Xoshiro256PlusPlus>(seed, n, out); break;
    case 2: fill_raw<Xoshiro256Plus>(seed, n, out); break;
    case 3: fill_raw<Xoroshiro128StarStar>(seed, n, out); break;
    case 4: fill_raw<Xoroshiro128PlusPlus>(seed, n, out); break;
    case 5: fill_raw<Xoroshiro128Plus>(seed, n, out); break;
    default: return -1;
    }
    return 0;
}

int ref_uniform(int prng, uint64_t seed, int n, double *out) {
    using namespace XoshiroCpp;
    switch (prng) {
    case 0: fill_uniform<Xoshiro256StarStar>(seed, n, out); break;
    case 1: fill_uniform<Xoshiro256PlusPlus>(seed, n, out); break;
    default: return -1;
    }
    return 0;
}

// urand over replayed raw values: out[i] = urand() after feeding raw[i].
int ref_urand_replay(const uint64_t *raw, int n, double *out) {
    g_replay = raw;
    g_pos = 0;
    for (int i = 0; i < n; ++i) out[i] = urand<double, ReplayGen>();
    return 0;
}

// sample_weights over replayed raw values (one raw value per call).
int ref_sample_weights(const double *w, double total, int size, const uint64_t *raw, int n, int *out) {
    g_replay = raw;
    g_pos = 0;
    for (int i = 0; i < n; ++i) out[i] = sample_weights<double, int, ReplayGen>(w, total, size);
    return 0;
}

// sample_binomial(p, N) over replayed raw values; returns draws consumed.
int ref_sample_binomial(double p, int N, const uint64_t *raw, int *out) {
    g_replay = raw;
    g_pos = 0;
    *out = sample_binomial<double, ReplayGen>(p, N);
    return (int)g_pos;
}

// common.hpp point functions
double ref_Nernst(double Xi, double Xo, double RT_F, double z) { return common::Nernst<double>(Xi, Xo, RT_F, z); }
double ref_INa(double V, double m, double h, double j, double GNa, double ENa) { return common::INa<double>(V, m, h, j, GNa, ENa); }
double ref_Ib(double V, double G, double E) { return common::Ib<double>(V, G, E); }
double ref_IpCa(double Cai, double IpCamax, double Km) { return common::IpCa<double>(Cai, IpCamax, Km); }
double ref_fNaK2(double VF_RT, double sigma) { return common::fNaK<double>(VF_RT, sigma); }
double ref_fNaK3(double VF_RT, double e, double sigma) { return common::fNaK<double>(VF_RT, e, sigma); }
double ref_INaK5(double VF_RT, double Nai, double sigma, double KmNai, double mult) { return common::INaK<double>(VF_RT, Nai, sigma, KmNai, mult); }
double ref_INaK6(double VF_RT, double e, double Nai, double sigma, double KmNai, double mult) { return common::INaK<double>(VF_RT, e, Nai, sigma, KmNai, mult); }
double ref_INaCa8(double VF_RT, double Nai, double Cai, double Nao3, double Cao, double eta, double ksat, double mult) {
    return common::INaCa<double>(VF_RT, Nai, Cai, Nao3, Cao, eta, ksat, mult);
}
double ref_INaCa9(double VF_RT, double e, double Nai, double Cai, double Nao3, double Cao, double eta, double ksat, double mult) {
    return common::INaCa<double>(VF_RT, e, Nai, Cai, Nao3, Cao, eta, ksat, mult);
}
double ref_alpha_m(double V) { return common::alpha_m<double>(V); }
double ref_beta_m(double V) { return common::beta_m<double>(V); }
double ref_alpha_h(double V) { return common::alpha_h<double>(V); }
double ref_beta_h(double V) { return common::beta_h<double>(V); }
double ref_alpha_j(double V) { return common::alpha_j<double>(V); }
double ref_beta_j(double V) { return common::beta_j<double>(V); }

}  // extern "C"
