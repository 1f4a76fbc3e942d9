// The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a
// small Init (3 wavelengths; and 15, illuminance mode), each of its passes
// on texel subsets, and the three sky lookups, at small LUT sizes. Exits 0
// when every result is finite; the sanitizers abort on the first error.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../oracle/oracle.h"

namespace {

bool finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

}  // namespace

int main() {
  or_set_threads(2);
  const or_sizes s{32, 8, 4, 16, 4, 4, 8, 4};
  std::vector<double> wl, solar, ray, mie_s, mie_e, abs, alb;
  for (int i = 0; i < 16; ++i) {
    const double l = 360.0 + 30.0 * i;
    wl.push_back(l);
    solar.push_back(1.5 - 0.001 * i);
    ray.push_back(1.24062e-6 * std::pow(l * 1e-3, -4.0));
    mie_e.push_back(4.44e-6);
    mie_s.push_back(3.996e-6);
    abs.push_back(1e-7 * (1.0 + std::sin(0.01 * l)));
    alb.push_back(0.1);
  }
  const or_layer r_layer{0.0, 1.0, -1.0 / 8000.0, 0.0, 0.0};
  const or_layer m_layer{0.0, 1.0, -1.0 / 1200.0, 0.0, 0.0};
  const or_layer oz[2] = {{25000.0, 0.0, 0.0, 1.0 / 15000.0, -2.0 / 3.0},
                          {0.0, 0.0, 0.0, -1.0 / 15000.0, 8.0 / 3.0}};
  int failures = 0;
  for (unsigned nlambda : {3u, 15u}) {
    or_model_params p{static_cast<int>(wl.size()), wl.data(), solar.data(), 0.00935,
                      6360e3, 6420e3, 1, &r_layer, ray.data(), 1, &m_layer, mie_s.data(),
                      mie_e.data(), 0.8, 2, oz, abs.data(), alb.data(),
                      102.0 / 180.0 * 3.14159265358979323846, 1000.0, nlambda, 0, 0};
    const size_t nt = 4u * s.transmittance_width * s.transmittance_height;
    const size_t ns = 4u * s.scattering_r * s.scattering_mu * s.scattering_mu_s *
                      s.scattering_nu;
    const size_t ne = 4u * s.irradiance_width * s.irradiance_height;
    std::vector<double> T(nt), S(ns), M(ns), E(ne);
    if (or_model_init(&p, &s, 3, T.data(), S.data(), M.data(), E.data()) != 0) ++failures;
    if (!finite(T) || !finite(S) || !finite(M) || !finite(E)) ++failures;

    or_atmosphere a;
    const double rgb[3] = {680.0, 550.0, 440.0};
    or_lower(&p, rgb, &a);
    // Passes on texel subsets, the final LUTs standing in for the deltas.
    const int n_texels = static_cast<int>(ns / 4);
    std::vector<int> idx;
    for (int t = 0; t < n_texels; t += 37) idx.push_back(t);
    const int k = static_cast<int>(idx.size());
    std::vector<double> o1(4u * k), o2(4u * k);
    or_texels_single_scattering(&a, &s, T.data(), k, idx.data(), o1.data(), o2.data());
    or_texels_scattering_density(&a, &s, T.data(), S.data(), M.data(), S.data(), E.data(), 2,
                                 k, idx.data(), o1.data());
    or_texels_scattering_density(&a, &s, T.data(), S.data(), M.data(), S.data(), E.data(), 3,
                                 k, idx.data(), o2.data());
    or_texels_multiple_scattering(&a, &s, T.data(), S.data(), k, idx.data(), o1.data());
    std::vector<double> II(ne);
    or_pass_indirect_irradiance(&a, &s, S.data(), M.data(), S.data(), 2, II.data());
    if (!finite(o1) || !finite(o2) || !finite(II)) ++failures;

    // Lookups from inside, at the edge of and outside the atmosphere.
    const int n = 64;
    std::vector<double> cam(3 * n), dir(3 * n), pt(3 * n), sun(3 * n), shadow(n), out1(3 * n),
        out2(3 * n);
    for (int i = 0; i < n; ++i) {
      // Cameras from 0.5 km to 72 km (above the 60 km atmosphere too),
      // points 1 to 13 km away along the view ray.
      const double h = (i % 7) * 12.0 + 0.5, az = 0.1 * i, el = -1.5 + 3.0 * i / n;
      cam[3 * i + 2] = 6360.0 + h;
      dir[3 * i] = std::cos(el) * std::cos(az);
      dir[3 * i + 1] = std::cos(el) * std::sin(az);
      dir[3 * i + 2] = std::sin(el);
      for (int c = 0; c < 3; ++c)
        pt[3 * i + c] = cam[3 * i + c] + (1.0 + (i % 5) * 3.0) * dir[3 * i + c];
      sun[3 * i] = std::sin(0.03 * i);
      sun[3 * i + 2] = std::cos(0.03 * i);
      shadow[i] = (i % 3) * 0.5;
    }
    for (int combined = 0; combined < 2; ++combined) {
      or_sky_radiance(&a, &s, T.data(), S.data(), M.data(), combined, n, cam.data(),
                      dir.data(), shadow.data(), sun.data(), out1.data(), out2.data());
      if (!finite(out1) || !finite(out2)) ++failures;
      or_sky_radiance_to_point(&a, &s, T.data(), S.data(), M.data(), combined, n,
                               cam.data(), pt.data(), shadow.data(), sun.data(), out1.data(),
                               out2.data());
      if (!finite(out1) || !finite(out2)) ++failures;
    }
    or_sun_and_sky_irradiance(&a, &s, T.data(), E.data(), n, pt.data(), dir.data(),
                              sun.data(), out1.data(), out2.data());
    if (!finite(out1) || !finite(out2)) ++failures;
  }
  std::printf("oracle driver: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
