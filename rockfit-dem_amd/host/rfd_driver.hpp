// C++ host side of the reference-layout tier: the test-driver loops of
// h_CalculationUniaxialCompression (md_calculationuniaxialcompression.cu:113-197)
// and h_CalculationBrazilTest (md_calculationbraziltest.cu:113-198) over the
// C ABI of include/rfd.h, with the host parameter sets of md_data_types.h and
// the reference's DATA.dat / stress-strain files.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "rfd.h"

namespace rfdh {

// unit system (md_phys_constants.h:5-13): mm, mg, us
extern const double kLength, kMass, kTime, kDensity, kForce, kStress, kVelocity;

// potential_data (md_data_types.h:146-186), CalculateParameters :153-185
struct Material {
  double E = 0, nu = 0, rho = 0, p_r = 0, b_r = 0, b_d = 0, sc = 0, tc = 0;
  double p_e = 0.6, p_mu = 0.45, p_mur = 0.05, a_farcut = 0.0;
  double G = 0, b_S = 0, p_V = 0, p_m = 0, p_I = 0, b_dd = 0, hm_E = 0, hm_G = 0, p_A = 0;
  double aa_farcut = 0, T_rayleigh = 0, T_sim = 0, dt = 0, dt_d_m = 0, dt_d_I = 0;
  void calculate();
  // SetInitialParameters (mp_determineconjgradient.cu:15-33)
  static Material reference_granite();
};

// cell_data::CalculateParameters (md_data_types.h:42-55)
struct CellGrid {
  double a = 0;
  float inv_a = 0;
  uint32_t Nx = 0, Ny = 0, Nz = 0;
  CellGrid() = default;
  CellGrid(double a, double Bx, double By, double Bz);
  uint32_t cells() const { return Nx * Ny * Nz; }
};

// sample_data (md_data_types.h:78-122): a specimen of outer size L inside a
// padded space box
struct SampleBox {
  std::array<double, 3> A{}, B{}, center{}, L{}, sized2{}, hidden{};
  double cell_a = 0, p_r = 0;
  CellGrid grid;
  SampleBox(double Lx, double Ly, double Lz, std::array<double, 3> space, double cell_a,
            double p_r);
};

// compression_data (md_data_types.h:188-294) and SumForcesULoading
// (md_results.cu:34-74)
struct Compression {
  bool uniaxial = true;
  int axis = 2;
  std::array<double, 3> size{}, center{};
  double p_r = 0, dt = 0, v_load = 0, min_stress = 0, max_strain = 0;
  double size0 = 0, bottom = 0, bottom_r = 0, top = 0, top_r = 0, R = 0, RR = 0;
  double area = 0, inv_area = 0, zero_stress = 0, V = 0, stress = 0, strain = 0;
  bool loading = false, fracture = false;
  double fracture_stress = 0, fracture_strain = 0;
  std::vector<std::array<double, 3>> history;  // strain, bottom and top stress
  Compression(bool uniaxial, std::array<double, 3> size, std::array<double, 3> center,
              double p_r, double dt, double v_load, double min_stress, double max_strain);
  void sum_forces(double s0, double s1);
  void compress();  // CompressU / CompressB (md_data_types.h:223-244, 272-293)
};

// particle_data / cell_data / interaction_list_data (md_data_types.h:12-56,
// 124-144) in device memory, one method per launch site of the drivers
class RefLayout {
 public:
  RefLayout(rfd_ctx *ctx, const std::vector<float> &R_soa, uint32_t K, const CellGrid &grid);
  ~RefLayout();
  RefLayout(const RefLayout &) = delete;
  RefLayout &operator=(const RefLayout &) = delete;

  void cell_distribution();
  void il_construct(float aacut);
  void il_reconstruct_bpm(float aacut);
  void bonds_construct(float bdd);
  void delete_far_links(float aa_farcut);
  void zero_fm();
  void forces_bpm(const Material &m);
  void forces_contact(const Material &m);
  std::array<double, 2> platen(const Compression &c, float zcut);
  void increments(const Material &m);
  void cut_restriction(const SampleBox &box);

  uint32_t N, K;
  float *R = nullptr, *V = nullptr, *W = nullptr, *F = nullptr, *M = nullptr;
  uint32_t *CI = nullptr, *CIs = nullptr, *pnC = nullptr, *IL = nullptr;
  uint8_t *ILtype = nullptr;
  float *iL = nullptr;
  float *slot[7] = {};  // rij, Oijt, Fijn, Fijt, Mijn, Mijt, Mijadd (float3 per slot)

 private:
  rfd_ctx *ctx_;
  CellGrid g_;
};

// the per-step loop of either test
class Driver {
 public:
  Driver(rfd_ctx *ctx, const std::vector<float> &R_soa, const Material &m, const SampleBox &box,
         const Compression &cmp, uint32_t K);
  void setup();                   // md_calculationuniaxialcompression.cu:131-143
  std::array<double, 2> step();   // :169-189
  uint32_t run(uint32_t max_steps);
  RefLayout s;
  Material mat;
  SampleBox box;
  Compression cmp;
  float aacut, zcut;
};

// DATA.dat (SaveAllData / LoadAllData, md_savealldata.cpp:12-277) with
// 4-byte counts (uint_fast32_t under MSVC)
struct Checkpoint {
  std::array<float, 24> sample{};
  std::vector<float> F, V, R, M, W, Q;
  uint32_t IonP = 0;
  float acut = 0;
  std::vector<uint32_t> IL;
  std::vector<uint8_t> ILtype;
  std::vector<float> iL;
  std::vector<float> slot[7];
};
Checkpoint load_all_data(const std::string &path);
void save_all_data(const std::string &path, const Checkpoint &c);
Checkpoint checkpoint_of(const RefLayout &s, float acut);
// Uniaxial.txt / Brazil.txt (SaveSumForcesLoading, md_results.cu:119-131)
void save_sum_forces_loading(const std::string &path, const Compression &c);

// an rfd_* status other than 0 ends the program with the library's message
void check(rfd_ctx *ctx, int status, const char *what);

}  // namespace rfdh
