// TEST INFRASTRUCTURE ONLY.  C entry points over the reference's own
// src/utils.cc (compiled unmodified from /root/reference into oracle/_ref by
// oracle/Makefile) so tests can pin the oracle's samplers and log-pmfs
// against the reference's arithmetic.  The declarations below restate the
// signatures of src/utils.hh:29-43.
#include <string>
void sran(const int seed);
int poisson_sample(const double lam);
unsigned int binomial_sample(const double ratio, const unsigned int nn);
double normal_probability(const double x, const double mean, const double var);
double poisson_probability(const int i, const double lam);
double negative_binomial_probability(const unsigned int val, const double m, const double shape);
double gamma_probability(const double x, const double a, const double b);

extern "C" {
void ref_poisson_stream(int seed, double lam, int n, int* out) {
  sran(seed);
  for (int i = 0; i < n; i++) out[i] = poisson_sample(lam);
}
void ref_binomial_stream(int seed, double p, unsigned nn, int n, int* out) {
  sran(seed);
  for (int i = 0; i < n; i++) out[i] = (int)binomial_sample(p, nn);
}
double ref_normal_probability(double x, double mean, double var) { return normal_probability(x, mean, var); }
double ref_poisson_probability(int i, double lam) { return poisson_probability(i, lam); }
double ref_negative_binomial_probability(unsigned v, double m, double shape) {
  return negative_binomial_probability(v, m, shape);
}
double ref_gamma_probability(double x, double a, double b) { return gamma_probability(x, a, b); }
}
