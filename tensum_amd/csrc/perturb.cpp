// perturb.cpp -- the reference's pert() functions for param.in expressions
// (src/perturb.cc:16-209), evaluated on the host with the sample's random
// numbers.  pert(type, x, y, p): x is ignored by every type.
//   1: multi-modal, M = n/2 modes a_i = r_i, b_i = r_{M+i}:
//      sum_i (a_i / sum a) sin(2 pi y + b_i) inside the centred band of width
//      p (clamped to [0,1]) of [0,1], else 0               (src/perturb.cc:56-94)
//   2: piecewise constant: the band split into n cells, cell k gives
//      (-1)^k r_k                                         (src/perturb.cc:119-144)
//   3: as 1 with wave numbers i: sin(2 i pi y + b_i)       (src/perturb.cc:150-188)
//   4: the p-th random number, 1-based                     (src/perturb.cc:194-209)
#include <cmath>
#include <vector>

#include "../../include/tsm.h"

namespace {

double band_modes(const double* r, int n, double y, double frac_in, bool wave_number) {
  const int M = (int)std::floor(n / 2.0);
  const double frac = std::max(std::min(frac_in, 1.0), 0.0);
  const double llim = (1.0 - frac) * 0.5;
  const double ulim = (frac + 1.0) * 0.5;
  double sum = 0.0;
  for (int i = 0; i < M; ++i) sum += r[i];
  double v = 0.0;
  if (y > llim && y < ulim)
    for (int i = 0; i < M; ++i) {
      const double arg = wave_number ? 2 * i * M_PI * y + r[M + i] : 2 * M_PI * y + r[M + i];
      v += (r[i] / sum) * std::sin(arg);
    }
  return v;
}

}  // namespace

extern "C" int tsm_pert(int type, double x, double y, double p, const double* rnd, int n_rnd,
                        double* out) {
  (void)x;
  if (!out || n_rnd < 0 || (n_rnd > 0 && !rnd)) return TSM_E_ARG;
  switch (type) {
    case 1: *out = band_modes(rnd, n_rnd, y, p, false); return TSM_OK;
    case 3: *out = band_modes(rnd, n_rnd, y, p, true); return TSM_OK;
    case 2: {
      const int M = n_rnd;
      const double frac = std::max(std::min(p, 1.0), 0.0);
      const double llim = (1.0 - frac) * 0.5;
      const double ulim = (frac + 1.0) * 0.5;
      const double div_size = frac / (double)M;
      if (y > llim && y < ulim) {
        const double yy = y - llim;
        int div = (int)std::floor(yy / div_size);
        if (div == M) div--;
        if (div < 0 || div >= M) return TSM_E_ARG;
        const int sgn = (int)std::pow(-1, div);
        *out = rnd[div] * sgn;
      } else {
        *out = 0.0;
      }
      return TSM_OK;
    }
    case 4: {
      const int id = (int)p;
      *out = 0.0;
      if (n_rnd > 0) {
        if (id >= 1 && id <= n_rnd) *out = rnd[id - 1];
        else return TSM_E_ARG;  // "INSUFFICIENT SIZE OF RANDOM ARRAY FOR PERT4!"
      }
      return TSM_OK;
    }
    default: return TSM_E_ARG;  // "Unknown perturbation type"
  }
}
