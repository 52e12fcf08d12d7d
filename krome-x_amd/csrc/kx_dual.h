// Forward-mode dual numbers over the generated rhs: a value and K tangent
// components, with the arithmetic and the kx:: math the rates, the rhs and
// the thermal terms are written in (they are templates on the real type R).
// Evaluated at a state whose tangents are seeded with unit directions, the
// rhs returns its exact directional derivatives: the columns of the
// Jacobian, the temperature column and the density-dependent rates
// included, to rounding (no difference quotients).  The adjoint kernel
// (kx_adjoint.h) builds the Jacobians of its backward steps from them.
//
// Values follow the same kx:: functions as the double path, so the value
// part of a dual evaluation equals the double evaluation.  Derivatives:
// d exp(x) = exp(x) dx, d log(x) = dx / x, d x^y = x^y (y (dx / x) + log x dy)
// (each term only where its tangent is nonzero, and none at a zero base:
// x = 0, an underflow, or a constant exponent adds no NaN), d sqrt(x) = dx / (2 sqrt x), d tanh = (1 - t^2) dx;
// fmin, fmax, fabs and comparisons act on the values (the derivative of
// the branch taken).
//
// Included by kx_physics.h after the scalar kx:: math (exp, log, log10,
// sqrt, pow, tanh) that it builds on, before the templates that use it.
#pragma once

namespace kx {

template <int K>
struct Dual {
  double v;
  double d[K];
  KX_HD Dual() : v(0.0) {
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  KX_HD Dual(double x) : v(x) {
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  KX_HD explicit operator double() const { return v; }
  KX_HD Dual& operator+=(const Dual& b) {
    v += b.v;
    for (int k = 0; k < K; ++k) d[k] += b.d[k];
    return *this;
  }
  KX_HD Dual& operator-=(const Dual& b) {
    v -= b.v;
    for (int k = 0; k < K; ++k) d[k] -= b.d[k];
    return *this;
  }
  KX_HD Dual& operator*=(const Dual& b) { return *this = *this * b; }
  KX_HD Dual& operator/=(const Dual& b) { return *this = *this / b; }
  friend KX_HD Dual operator+(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v + b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
  }
  friend KX_HD Dual operator-(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v - b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
  }
  friend KX_HD Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
  }
  // (a factor exactly 0, an underflowed fit, leaves no term: its
  // partner's derivative may have overflowed where the fit underflows)
  friend KX_HD Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
    for (int k = 0; k < K; ++k)
      r.d[k] = (b.v != 0.0 ? a.d[k] * b.v : 0.0) +
               (a.v != 0.0 ? a.v * b.d[k] : 0.0);
    return r;
  }
  // (an infinite denominator, an overflowed power in a fit, leaves the
  // quotient at 0 whatever its tangents: no derivative there)
  friend KX_HD Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
    const bool fin = __builtin_isfinite(b.v);
    for (int k = 0; k < K; ++k)
      r.d[k] = fin ? (a.d[k] - r.v * b.d[k]) / b.v : 0.0;
    return r;
  }
  friend KX_HD bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend KX_HD bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend KX_HD bool operator<=(const Dual& a, const Dual& b) {
    return a.v <= b.v;
  }
  friend KX_HD bool operator>=(const Dual& a, const Dual& b) {
    return a.v >= b.v;
  }
  friend KX_HD bool operator==(const Dual& a, const Dual& b) {
    return a.v == b.v;
  }
  friend KX_HD bool operator!=(const Dual& a, const Dual& b) {
    return a.v != b.v;
  }
};

// a finite value (kx::finite of the double and float forms)
template <int K>
KX_HD inline bool finite(const Dual<K>& x) {
  return __builtin_isfinite(x.v);
}

// the value's function and its derivative f'(v) applied to every tangent
template <int K>
KX_HD inline Dual<K> chain(const Dual<K>& x, double fv, double dfdx) {
  Dual<K> r;
  r.v = fv;
  for (int k = 0; k < K; ++k) r.d[k] = dfdx * x.d[k];
  return r;
}

template <int K>
KX_HD inline Dual<K> exp(const Dual<K>& x) {
  const double e = kx::exp(x.v);
  return chain(x, e, e);
}
// (the tangent divided by the value, not multiplied by its reciprocal:
// 1 / x overflows for the smallest arguments, dx / x does not)
template <int K>
KX_HD inline Dual<K> log(const Dual<K>& x) {
  Dual<K> r;
  r.v = kx::log(x.v);
  for (int k = 0; k < K; ++k) r.d[k] = x.d[k] / x.v;
  return r;
}
template <int K>
KX_HD inline Dual<K> log10(const Dual<K>& x) {
  Dual<K> r;
  r.v = kx::log10(x.v);
  for (int k = 0; k < K; ++k)
    r.d[k] = 0.43429448190325182765 * (x.d[k] / x.v);
  return r;
}
template <int K>
KX_HD inline Dual<K> sqrt(const Dual<K>& x) {
  Dual<K> r;
  r.v = kx::sqrt(x.v);
  for (int k = 0; k < K; ++k) r.d[k] = x.d[k] / (2.0 * r.v);
  return r;
}
template <int K>
KX_HD inline Dual<K> tanh(const Dual<K>& x) {
  const double t = kx::tanh(x.v);
  return chain(x, t, 1.0 - t * t);
}
template <int K>
KX_HD inline Dual<K> fabs(const Dual<K>& x) {
  return x.v < 0.0 ? -x : x;
}
template <int K>
KX_HD inline Dual<K> fmin(const Dual<K>& a, const Dual<K>& b) {
  return b.v < a.v ? b : a;
}
template <int K>
KX_HD inline Dual<K> fmax(const Dual<K>& a, const Dual<K>& b) {
  return b.v > a.v ? b : a;
}
template <int K>
KX_HD inline Dual<K> pow(const Dual<K>& x, const Dual<K>& y) {
  Dual<K> r;
  r.v = kx::pow(x.v, y.v);
  bool dy = false;
  for (int k = 0; k < K; ++k) dy = dy || y.d[k] != 0.0;
  // (a zero base, an underflowed fit, or x^y = 0 has no derivative: 0^y
  // is 0 or, at y = 0, 1, constant wherever the base stays 0; the base's
  // term as x^y y (dx / x), which stays finite at the smallest x)
  const bool live = r.v != 0.0 && x.v != 0.0;
  const double b = dy && live ? kx::log(x.v) : 0.0;
  for (int k = 0; k < K; ++k)
    r.d[k] = live ? r.v * ((x.d[k] != 0.0 ? y.v * (x.d[k] / x.v) : 0.0) +
                           (y.d[k] != 0.0 ? b * y.d[k] : 0.0))
                  : 0.0;
  return r;
}
template <int K>
KX_HD inline Dual<K> pow(const Dual<K>& x, double y) {
  return pow(x, Dual<K>(y));
}
template <int K>
KX_HD inline Dual<K> pow(double x, const Dual<K>& y) {
  return pow(Dual<K>(x), y);
}

}  // namespace kx
