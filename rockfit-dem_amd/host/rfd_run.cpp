// rfd_run: the uniaxial-compression or Brazilian test of the reference's
// main program on the reference-layout tier.  Reads the particles from a
// DATA.dat (LoadAllData, md_savealldata.cpp:40), builds the lists and bonds,
// runs the explicit steps of h_CalculationUniaxialCompression
// (md_calculationuniaxialcompression.cu:113-197) or h_CalculationBrazilTest
// (md_calculationbraziltest.cu:113-198) with the reference granite material,
// and writes Uniaxial.txt / Brazil.txt and the final DATA.dat.
//
//   rfd_run --mode uniaxial|brazil --D <mm> --H <mm> --in DATA.dat --out <dir>
//           [--steps N] [--vload-scale s] [--until-fracture]
//
// The summary line ends with the wall time of the step loop (after set-up,
// to the end of its last step's GPU work) per step.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "rfd_driver.hpp"

using namespace rfdh;

int main(int argc, char **argv) {
  std::string mode = "uniaxial", in, out = ".";
  double D = 0, H = 0, vscale = 1.0;
  unsigned steps = 1000;
  bool until_fracture = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--mode") mode = next();
    else if (a == "--D") D = std::atof(next());
    else if (a == "--H") H = std::atof(next());
    else if (a == "--in") in = next();
    else if (a == "--out") out = next();
    else if (a == "--steps") steps = (unsigned)std::strtoul(next(), nullptr, 10);
    else if (a == "--vload-scale") vscale = std::atof(next());
    else if (a == "--until-fracture") until_fracture = true;
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  if ((mode != "uniaxial" && mode != "brazil") || D <= 0 || H <= 0 || in.empty()) {
    std::fprintf(stderr,
                 "usage: rfd_run --mode uniaxial|brazil --D mm --H mm --in DATA.dat --out dir "
                 "[--steps N] [--vload-scale s] [--until-fracture]\n");
    return 2;
  }
  const bool uni = mode == "uniaxial";
  const Checkpoint c0 = load_all_data(in);
  const uint32_t K = c0.IonP ? c0.IonP : 32;
  const Material m = Material::reference_granite();
  // the cases of CalculateStepCG (mp_determineconjgradient.cu:96-121)
  const double a = 2.0 * m.a_farcut;
  const SampleBox box(D, D, H, {D + 16.0, D + 16.0, H + 12.0}, a, m.p_r);
  Compression cmp(uni, box.L, box.center, m.p_r, m.dt,
                  (uni ? -10.0 : -100.0) * 0.02 / kVelocity * vscale,
                  (uni ? 15e6 : 2.1e6) / kStress, 0.1);

  rfd_ctx *ctx = nullptr;
  check(nullptr, rfd_ctx_create(&ctx, 0, nullptr), "rfd_ctx_create");
  {
    Driver d(ctx, c0.R, m, box, cmp, K);
    d.setup();
    check(ctx, rfd_sync(ctx), "rfd_sync");
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = until_fracture ? d.run(steps) : ([&]() {
      for (unsigned k = 0; k < steps; ++k) d.step();
      return (uint32_t)steps;
    })();
    check(ctx, rfd_sync(ctx), "rfd_sync");
    const double us = std::chrono::duration<double, std::micro>(
                          std::chrono::steady_clock::now() - t0).count();
    uint32_t flags = 0;
    check(ctx, rfd_read_flags(ctx, &flags), "rfd_read_flags");
    save_sum_forces_loading(out + (uni ? "/Uniaxial.txt" : "/Brazil.txt"), d.cmp);
    save_all_data(out + "/DATA.dat", checkpoint_of(d.s, (float)m.a_farcut));
    std::printf("%s: N=%u K=%u steps=%u stress=%g Pa peak=%g Pa strain=%g fracture=%d flags=%u "
                "us_per_step=%.3f\n",
                mode.c_str(), d.s.N, K, n, d.cmp.stress * kStress,
                d.cmp.fracture_stress * kStress, d.cmp.strain, (int)d.cmp.fracture, flags,
                n ? us / n : 0.0);
  }
  rfd_ctx_destroy(ctx);
  return 0;
}
