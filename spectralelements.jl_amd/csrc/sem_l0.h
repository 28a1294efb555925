// Host-side 1-D spectral primitives shared by the library and the build-time
// table generator: GLL nodes/weights (FastGaussQuadrature.gausslobatto as
// called at mesh.jl:70-71), derivMat (derivMat.jl:9-35), interpMat
// (interp.jl:10-35), semmesh (semmesh.jl:9-27).
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace sem {
namespace {

void legendre(int N, long double x, long double* P, long double* dP) {
  long double p0 = 1.0L, p1 = x;
  if (N == 0) {
    *P = 1.0L;
    *dP = 0.0L;
    return;
  }
  for (int k = 2; k <= N; ++k) {
    const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *P = p1;
  *dP = N * (x * p1 - p0) / (x * x - 1.0L);
}

void gll(int n, double* z, double* w) {
  const int N = n - 1;
  std::vector<long double> zz(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; ++i) zz[i] = -cosl(pi * i / N);
  zz[0] = -1.0L;
  zz[N] = 1.0L;
  for (int i = 1; i < N; ++i) {
    long double x = zz[i];
    for (int it = 0; it < 100; ++it) {
      long double P, dP;
      legendre(N, x, &P, &dP);
      const long double f = (1.0L - x * x) * dP;
      const long double df = -(long double)N * (N + 1) * P;
      const long double step = f / df;
      x -= step;
      if (fabsl(step) < 1e-19L) break;
    }
    zz[i] = x;
  }
  for (int i = 1; i < N; ++i) {
    const long double s = 0.5L * (zz[i] - zz[N - i]);
    zz[i] = s;
  }
  if (n % 2 == 1) zz[N / 2] = 0.0L;
  for (int i = 0; i < n; ++i) {
    long double P, dP;
    legendre(N, zz[i], &P, &dP);
    z[i] = (double)zz[i];
    w[i] = (double)(2.0L / ((long double)N * (N + 1) * P * P));
  }
}

// derivMat.jl:9-35 (same loop order)
void derivmat(int n, const double* x, double* D) {
  std::vector<double> a(n, 1.0);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < i; ++j) a[i] = a[i] * (x[i] - x[j]);
    for (int j = i + 1; j < n; ++j) a[i] = a[i] * (x[i] - x[j]);
  }
  for (int i = 0; i < n; ++i) a[i] = 1.0 / a[i];
  auto at = [&](int i, int j) -> double& { return D[i + (size_t)n * j]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) at(i, j) = (i == j) ? 1.0 : 1.0 / (x[i] - x[j]);
  for (int i = 0; i < n; ++i) {
    at(i, i) = 0.0;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s = s + at(i, j);
    at(i, i) = s;
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      if (i != j) at(i, j) = a[j] / (a[i] * (x[i] - x[j]));
}

// interp.jl:10-35
void interpmat(int no, const double* xo, int ni, const double* xi, double* J) {
  std::vector<double> a(ni, 1.0), s(ni, 1.0), t(ni, 1.0);
  for (int i = 0; i < ni; ++i) {
    for (int j = 0; j < i; ++j) a[i] = a[i] * (xi[i] - xi[j]);
    for (int j = i + 1; j < ni; ++j) a[i] = a[i] * (xi[i] - xi[j]);
  }
  for (int i = 0; i < ni; ++i) a[i] = 1.0 / a[i];
  for (int i = 0; i < no; ++i) {
    const double x = xo[i];
    for (int j = 1; j < ni; ++j) {
      s[j] = s[j - 1] * (x - xi[j - 1]);
      t[ni - 1 - j] = t[ni - j] * (x - xi[ni - j]);
    }
    for (int j = 0; j < ni; ++j) J[i + (size_t)no * j] = a[j] * s[j] * t[j];
  }
}

// semmesh.jl:9-27
void semmesh(int E, int n, double* z, double* w) {
  std::vector<double> z0(n), w0(n);
  gll(n, z0.data(), w0.data());
  for (int i = 0; i < n; ++i) {
    z0[i] = 0.5 * (z0[i] + 1.0);
    w0[i] = 0.5 * w0[i];
  }
  std::vector<double> ze(E + 1);
  for (int k = 0; k <= E; ++k) ze[k] = (2.0 * k - E) / (double)E;
  for (int e = 0; e < E; ++e) {
    const double dz = ze[e + 1] - ze[e];
    for (int a = 0; a < n; ++a) {
      z[e * n + a] = dz * z0[a] + ze[e];
      if (w) w[e * n + a] = dz * w0[a];
    }
  }
}


}  // namespace
}  // namespace sem
