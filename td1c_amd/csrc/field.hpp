// Laser field of one pulse (clfield1, src/field/clfield1.cpp:14-500), usable
// on the host (td1c_field_value) and on the device (the propagation loop
// evaluates the stage fields itself).  Envelopes: 0 zero, 1 sin^2, 2 flat;
// gauges: 0 length1 (E explicit), 1 velocity1 (A explicit, E = -dA/dt by
// the 6-point central difference of clfield1::get_evalue_numder).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/td1c_hprod.h"

namespace td1c {

struct Pulse {
  int env_type, gauge;
  double famp, freq, period, cep, tau, tleft, tright, numcyc;
};

inline Pulse make_pulse(const td1c_field* f) {
  constexpr double pi = 3.14159265358979323846264338327950288;
  Pulse p{};
  p.env_type = f->env_type;
  p.gauge = f->gauge;
  p.cep = f->cep / 180.0 * pi;
  p.famp = std::sqrt(f->fint / 3.50944506E+16);
  p.freq = 1239.84190 / (f->wlen * 27.2113845);
  p.period = 2.0 * pi / p.freq;
  p.numcyc = f->numcyc;
  p.tau = f->env_type == 0 ? 0.0 : p.numcyc * p.period * 0.5;
  const double delay = f->delay / 24.188843265;
  // clfield::gen: tshift = -(delay - tau) for a single pulse
  const double tshift = -(delay - p.tau < 0.0 ? delay - p.tau : 0.0);
  const double tcenter = tshift + delay;
  p.tleft = tcenter - p.tau;
  p.tright = tcenter + p.tau;
  return p;
}

// amplitude factor of E (scale 1) or A (scale 1/freq) at time t
__host__ __device__ inline double pulse_wave(const Pulse& p, double t, double scale) {
  constexpr double pi = 3.14159265358979323846264338327950288;
  const double trel = t - p.tleft;
  if (p.env_type == 2) return p.famp * scale * sin(p.freq * trel + p.cep);
  if (p.env_type == 1) {
    if (t < p.tleft || t > p.tright) return 0.0;
    const double env = sin(pi * trel / (p.numcyc * p.period));
    return env * env * p.famp * scale * sin(p.freq * trel + p.cep);
  }
  return 0.0;
}

__host__ __device__ inline double pulse_e_from_a(const Pulse& p, double t) {
  constexpr double pi = 3.14159265358979323846264338327950288;
  const double h = 2.0 * pi / (1239.84190 / (800.0 * 27.2113845)) / 100000.0;
  const double sa = 1.0 / p.freq;
  double s = 0.0;
  s += pulse_wave(p, t + 3 * h, sa) / 60.0;
  s -= pulse_wave(p, t + 2 * h, sa) / 20.0 * 3.0;
  s += pulse_wave(p, t + h, sa) / 4.0 * 3.0;
  s -= pulse_wave(p, t - h, sa) / 4.0 * 3.0;
  s += pulse_wave(p, t - 2 * h, sa) / 20.0 * 3.0;
  s -= pulse_wave(p, t - 3 * h, sa) / 60.0;
  return -s / h;
}

// lfield[0..2] chosen gauge, [3..5] E, [6..8] A (z components used)
__host__ __device__ inline void pulse_lfield(const Pulse& p, double t, double* lf) {
  for (int i = 0; i < 9; ++i) lf[i] = 0.0;
  if (p.env_type == 0) return;
  if (!(t > p.tleft && t < p.tright)) return;
  if (p.gauge == 0) {
    lf[2] = pulse_wave(p, t, 1.0);
    lf[5] = lf[2];
    // A in the length gauge is a running integral in the reference
    // (get_avalue_numint); the hprod does not read it and it stays 0 here.
  } else {
    lf[2] = pulse_wave(p, t, 1.0 / p.freq);
    lf[5] = pulse_e_from_a(p, t);
    lf[8] = lf[2];
  }
}

// d lfield / dt in lfield[0..2] (clfield1::get_der, src/field/clfield1.cpp:
// 193-211): length1 the analytic derivative of E (get_eder_explicit,
// :324-378), velocity1 dA/dt = -E of the 6-point difference
// (get_ader_explicit, :228-238); zero outside (tleft, tright).
__host__ __device__ inline void pulse_dlfield(const Pulse& p, double t, double* df) {
  constexpr double pi = 3.14159265358979323846264338327950288;
  for (int i = 0; i < 9; ++i) df[i] = 0.0;
  if (p.env_type == 0) return;
  if (!(t > p.tleft && t < p.tright)) return;
  if (p.gauge == 0) {
    const double trel = t - p.tleft;
    if (p.env_type == 2) {
      df[2] = p.famp * p.freq * cos(p.freq * trel + p.cep);
    } else {
      const double fac = pi / (p.numcyc * p.period);
      double env = sin(fac * trel);
      env = env * env;
      const double denv = 2.0 * fac * sin(fac * trel) * cos(fac * trel);
      df[2] = p.famp * (p.freq * env * cos(p.freq * trel + p.cep) + denv * sin(p.freq * trel + p.cep));
    }
  } else {
    df[2] = -pulse_e_from_a(p, t);
  }
}

}  // namespace td1c
