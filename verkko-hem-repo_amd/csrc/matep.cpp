// GL coefficients of superfluid He-3 (matep/src/matep.cc, matep/inc/matep.h),
// restated with the reference's float-literal and float-interpolation
// rounding so that alpha, beta_i and the bulk gaps agree bit for bit.
#include <cmath>

#include "glhip.h"

namespace {

// pi and zeta(3) are float literals stored in doubles (matep.h:102-103, matep.cc:51)
const double kPi = (double)3.14159265358979323846264338328f;
const double kZeta3 = (double)1.2020569031595942f;
// strong-coupling corrections c_i(p), 0..34 bar in 2-bar steps (matep.cc:59-63)
const double kScc[5][18] = {
    {-0.0098, -0.0127, -0.0155, -0.0181, -0.0207, -0.0231, -0.0254, -0.0275, -0.0295, -0.0314, -0.0330, -0.0345,
     -0.0358, -0.0370, -0.0381, -0.0391, -0.0402, -0.0413},
    {-0.0419, -0.0490, -0.0562, -0.0636, -0.0711, -0.0786, -0.0861, -0.0936, -0.1011, -0.1086, -0.1160, -0.1233,
     -0.1306, -0.1378, -0.1448, -0.1517, -0.1583, -0.1645},
    {-0.0132, -0.0161, -0.0184, -0.0202, -0.0216, -0.0226, -0.0233, -0.0239, -0.0243, -0.0247, -0.0249, -0.0252,
     -0.0255, -0.0258, -0.0262, -0.0265, -0.0267, -0.0268},
    {-0.0047, -0.0276, -0.0514, -0.0760, -0.1010, -0.1260, -0.1508, -0.1751, -0.1985, -0.2208, -0.2419, -0.2614,
     -0.2795, -0.2961, -0.3114, -0.3255, -0.3388, -0.3518},
    {-0.0899, -0.1277, -0.1602, -0.1880, -0.2119, -0.2324, -0.2503, -0.2660, -0.2801, -0.2930, -0.3051, -0.3167,
     -0.3280, -0.3392, -0.3502, -0.3611, -0.3717, -0.3815}};
// weak-coupling values in units of c_betai (matep.cc:149-201)
const double kWeak[5] = {-1.0, 2.0, 2.0, 2.0, -2.0};

// piecewise-linear table lookup on left-closed 2-bar bins, returned through a
// float (matep.cc:365-406).  Returns NaN outside [0, 34) bar.
double table_lookup(const double *c, double p)
{
  if (!(p >= 0.0 && p < 34.0)) return NAN;
  int bin = 0;
  while (bin < 16 && p >= 2.0 * (bin + 1)) ++bin;
  const float p_lo = (float)(2 * bin);
  const float val = (float)(((c[bin + 1] - c[bin]) / 2.0) * (p - (double)p_lo) + c[bin]);
  return (double)val;
}

}  // namespace

extern "C" int glh_matep(double p, double t, int scc, double out[12])
{
  if (!out) return GLH_ERR_ARG;
  const double cb = (7.0 * kZeta3) / ((80.0 * kPi) * kPi);
  double beta[5];
  out[0] = 1.0 * (t - 1.0);
  bool bad = false;
  for (int n = 0; n < 5; ++n) {
    double corr = 0.0;
    if (scc) {
      corr = table_lookup(kScc[n], p);
      bad = bad || std::isnan(corr);
      beta[n] = cb * (kWeak[n] + t * corr);
    } else {
      beta[n] = cb * kWeak[n];
    }
    out[1 + n] = beta[n];
  }
  const double bA = (beta[1] + beta[3]) + beta[4];
  const double bB = (beta[0] + beta[1]) + (double)(1.f / 3.f) * ((beta[2] + beta[3]) + beta[4]);
  const double a = out[0];
  out[6] = bA;
  out[7] = bB;
  out[8] = (t <= 1.0) ? std::sqrt(-a / (2.0 * bA)) : 0.0;
  out[9] = (t <= 1.0) ? std::sqrt(-a / (2.0 * bB)) : 0.0;
  out[10] = (t <= 1.0) ? (-0.25 * std::pow(a, 2.0)) / bA : 0.0;
  out[11] = (t <= 1.0) ? (-0.25 * std::pow(a, 2.0)) / bB : 0.0;
  return bad ? GLH_ERR_RANGE : GLH_OK;
}
