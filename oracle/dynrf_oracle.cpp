// Oracle (TEST INFRASTRUCTURE ONLY) of DynamicRFKickMap::__calcModulation
// (src/SM/DynamicRFKickMap.cpp:80-95): the queue of (phase, amplitude) RF
// modulations, one per step.  The reference draws the noise from the C++
// standard library (std::mt19937 seeded by std::random_device,
// std::normal_distribution<meshaxis_t>); this restatement uses the same
// generators with a given seed, and the reference's float arithmetic:
//   _phasenoise = phasespread / sqrt(revolutionpart)   (double, held as float)
//   _amplnoise  = amplspread / sqrt(revolutionpart)
//   _modtimedelta = 2 pi modtimeincrement               (held as float)
//   step i: phase = _syncphase + N(0,1) _phasenoise + modampl sin(_modtimedelta i)
//           ampl  = 1 + N(0,1) _amplnoise               (phase drawn first)
#include <cmath>
#include <cstdint>
#include <random>

extern "C" void or_dynrf_modulation(uint32_t steps, uint32_t seed,
                                    float syncphase, double revolutionpart,
                                    float phasespread, float amplspread,
                                    float modampl, double modtimeincrement,
                                    float* out)
{
    const double two_pi = 6.283185307179586476925286766559005768;
    const float phasenoise = phasespread / std::sqrt(revolutionpart);
    const float amplnoise = amplspread / std::sqrt(revolutionpart);
    const float modtimedelta = two_pi * modtimeincrement;
    std::mt19937 prng(seed);
    std::normal_distribution<float> dist(0, 1);
    for (uint32_t i = 0; i < steps; i++) {
        float pn = dist(prng) * phasenoise;
        float an = dist(prng) * amplnoise;
        float pm = modampl * std::sin(modtimedelta * (i));
        out[2 * i] = syncphase + pn + pm;
        out[2 * i + 1] = 1 + an;
    }
}

// the 10000th output of a default-constructed std::mt19937 (the C++
// standard fixes it at 4123659995), for pinning the generator
extern "C" uint32_t or_mt19937_10000th(void)
{
    std::mt19937 g;
    g.discard(9999);
    return (uint32_t)g();
}
