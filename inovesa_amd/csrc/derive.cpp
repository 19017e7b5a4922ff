// Ring parameters -> mesh, map and field parameters, restating the
// derivations of src/main.cpp:179-355 (single bunch).  Types follow the
// reference: options held as meshaxis_t/frequency_t are float, the rest
// double (inc/IO/ProgramOptions.hpp:237-287).
#include <algorithm>
#include <cmath>

#include "vfps_gpu.h"

namespace {
constexpr double C_LIGHT = 2.99792458e8;
constexpr double EPS0 = 8.854187817e-12;
constexpr double E_CHARGE = 1.602e-19;
constexpr double I_ALFVEN = 17045;
constexpr double ME = 510998.9;
constexpr double TWO_PI = 6.283185307179586476925286766559005768;

uint64_t upper_power_of_two(uint64_t v)
{
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    v |= v >> 32;
    v++;
    return v;
}
}  // namespace

extern "C" int vfps_derive(const vfps_ring_params* p, vfps_derived* d)
{
    if (!p || !d || p->grid_size < 3) {
        return VFPS_ERR_INVALID;
    }
    const uint32_t ps_bins = p->grid_size;
    const float pqsize = (float)p->pq_size;
    const float shiftx = (float)p->shift_x, shifty = (float)p->shift_y;
    // src/main.cpp:179-187
    const float qcenter = -shiftx * pqsize / (float)(ps_bins - 1);
    const float pcenter = -shifty * pqsize / (float)(ps_bins - 1);
    const float pqhalf = pqsize / 2;
    d->qmax = qcenter + pqhalf;
    d->qmin = qcenter - pqhalf;
    d->pmax = pcenter + pqhalf;
    d->pmin = pcenter - pqhalf;
    const double sE = p->energy_spread;
    const double E0 = p->beam_energy;
    const double dE = sE * E0;
    // src/main.cpp:203-211
    const double f_rev = (double)(float)p->f_rev;
    const double R_bend = (p->bending_radius > 0)
                              ? p->bending_radius
                              : C_LIGHT / (TWO_PI * f_rev);
    const double harmonic_number = (double)(float)p->harmonic_number;
    const double f_RF = f_rev * harmonic_number;
    const double bunchspacing = 1.0 / f_RF;
    const double gap = p->vacuum_gap;
    const double V_RF = p->rf_voltage;
    // src/main.cpp:218-238
    const double lorentzgamma = E0 / ME;
    const double V0 = E_CHARGE * std::pow(lorentzgamma, 4) / (3 * EPS0 * R_bend);
    const double W0 = V0 * E_CHARGE;
    const double V_eff = std::sqrt(V_RF * V_RF - V0 * V0);
    double fs = (double)(float)p->f_s;
    double alpha0 = (double)(float)p->alpha0;
    if (std::fpclassify(fs) == FP_ZERO) {
        fs = f_rev * std::sqrt(alpha0 * harmonic_number * V_eff / (TWO_PI * E0));
    } else {
        alpha0 = (fs > 0 ? 1.0 : -1.0) * TWO_PI * E0 / (harmonic_number * V_eff)
                 * std::pow(fs / f_rev, 2);
    }
    const float alpha[3] = {(float)alpha0, (float)p->alpha1, (float)p->alpha2};
    // src/main.cpp:245-246
    const double bl = C_LIGHT * dE / harmonic_number / std::pow(f_rev, 2.0)
                      / V_eff * fs;
    const double Ib = (double)(float)p->bunch_current;
    // src/main.cpp:284-297
    const double steps = (p->steps_per_trev > 0)
                             ? p->steps_per_trev * f_rev / fs
                             : std::max(p->steps_per_ts, 1u);
    const float rotations = (float)p->rotations;
    const double calc_damp = E0 * E_CHARGE / W0 / f_rev;
    const double t_damp = (p->damping_time < 0) ? calc_damp : p->damping_time;
    const double dt = 1.0 / (fs * steps);
    const double revolutionpart = f_rev * dt;
    // src/main.cpp:304-321
    const double spacing_ps = bunchspacing * C_LIGHT / bl / (double)pqsize;
    const uint32_t spacing_bins = (uint32_t)std::round(ps_bins * spacing_ps);
    const float fmax = (float)(ps_bins * C_LIGHT / ((double)pqsize * bl));
    const double padding = std::max(p->padding, 1.0);
    size_t padded_bins = (size_t)std::ceil(ps_bins * padding);
    if (p->round_padding) {
        padded_bins = upper_power_of_two(padded_bins);
    }
    // src/main.cpp:353-355
    const float angle = (float)(TWO_PI / steps);
    d->laststep = (uint32_t)std::ceil(steps * rotations);
    // src/main.cpp:376-390 (scaling law)
    double shield = 0;
    if (gap > 0) {
        shield = bl * std::sqrt(R_bend) * std::pow(gap, -3. / 2.);
    }
    const double Inorm = I_ALFVEN / ME * TWO_PI * std::pow(dE * fs / f_rev, 2)
                         / V_eff / harmonic_number
                         * std::pow(bl / R_bend, 1. / 3.);
    // src/main.cpp:663-664 and 682
    d->slip[0] = angle;
    d->slip[1] = alpha[1] / alpha[0] * angle;
    d->slip[2] = alpha[2] / alpha[0] * angle;
    d->e1 = (t_damp > 0) ? (float)(2.0 / (fs * t_damp * steps)) : 0.0f;
    d->t_damp = t_damp;
    d->qscale = bl;
    d->pscale = dE;
    d->R_bend = R_bend;
    d->f0 = C_LIGHT / (TWO_PI * R_bend);
    d->V0 = V0;
    d->V_eff = V_eff;
    d->f_s = fs;
    d->alpha0 = alpha0;
    d->f_RF = f_RF;
    d->steps = steps;
    d->dt = dt;
    d->revolutionpart = revolutionpart;
    d->angle = angle;
    d->padded_bins = (uint32_t)padded_bins;
    d->spacing_bins = spacing_bins;
    d->fmax = fmax;
    d->Ib = Ib;
    d->shield = shield;
    d->Inorm = Inorm;
    d->Ith = Inorm * (0.5 + 0.34 * shield);
    // ElectricField scaling argument (inc/PS/ElectricField.hpp:81-103);
    // delta of the energy axis from its Ruler
    const float d1 = (d->pmax - d->pmin) / (float)(ps_bins - 1);
    d->wakescaling = (float)(Ib * dt * C_LIGHT / (double)(float)bl
                             / ((double)d1 * sE * E0));
    return VFPS_OK;
}
