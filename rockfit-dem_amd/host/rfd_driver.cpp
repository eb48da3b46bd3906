// C++ host side of the reference-layout tier (see rfd_driver.hpp).
#include "rfd_driver.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>

namespace rfdh {

const double kLength = 1e-3, kMass = 1e-6, kTime = 1e-6;
const double kDensity = kMass / std::pow(kLength, 3.0);
const double kForce = kMass * kLength / std::pow(kTime, 2.0);
const double kStress = kForce / std::pow(kLength, 2.0);
const double kVelocity = kLength / kTime;

void check(rfd_ctx *ctx, int status, const char *what) {
  if (status == 0) return;
  std::fprintf(stderr, "%s failed with status %d: %s\n", what, status,
               ctx ? rfd_last_error(ctx) : "");
  std::exit(1);
}

static void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  std::exit(1);
}

// ------------------------------------------------------------ parameters
void Material::calculate() {
  if (a_farcut <= 0) a_farcut = 2.9 * p_r;
  G = 0.5 * E / (1.0 + nu);
  b_S = M_PI * b_r * b_r;
  p_V = 4.0 * (1.0 / 3.0) * M_PI * std::pow(p_r, 3.0);
  p_m = rho * p_V;
  b_dd = b_d * b_d;
  p_I = 0.4 * p_m * p_r * p_r;
  hm_E = 0.5 * E / (1.0 - nu * nu);
  hm_G = 0.5 * G / (2.0 - nu);
  const double le = std::log(p_e);
  p_A = le / std::sqrt(M_PI * M_PI + le * le);
  aa_farcut = a_farcut * a_farcut;
  // time step rule, md_data_types.h:166-176
  T_rayleigh = (M_PI * p_r * std::sqrt(2.0 * rho * (1.0 + nu))) /
               (std::sqrt(E) * (0.163 * nu + 0.8766));
  const double k = std::max(E * b_S / (2.0 * p_r), G * b_S / (2.0 * p_r));
  T_sim = 2.0 * std::sqrt(p_m / k);
  dt = 0.1 * std::min(T_rayleigh, T_sim);
  dt_d_m = dt / p_m;
  dt_d_I = dt / p_I;
}

Material Material::reference_granite() {
  Material m;
  m.p_r = 0.5 * 0.001 / kLength;
  m.E = 1.95255e11 / kStress;
  m.nu = 0.247;
  m.rho = 2610 / kDensity;
  m.b_r = 0.5 * 0.000330015 / kLength;
  m.b_d = 2.0 * m.p_r + 5.93169e-05 / kLength;
  m.sc = 5.1447e8 / kStress;
  m.tc = 3.7131e8 / kStress;
  m.calculate();
  return m;
}

CellGrid::CellGrid(double a_, double Bx, double By, double Bz) : a(a_) {
  inv_a = (float)(1.0 / (double)(float)a);
  const double ia = inv_a;
  Nx = (uint32_t)std::floor((double)(float)(Bx * ia)) + 1;
  Ny = (uint32_t)std::floor((double)(float)(By * ia)) + 1;
  Nz = (uint32_t)std::floor((double)(float)(Bz * ia)) + 3;
}

SampleBox::SampleBox(double Lx, double Ly, double Lz, std::array<double, 3> space, double a,
                     double pr)
    : cell_a(a), p_r(pr) {
  for (int d = 0; d < 3; ++d) {
    A[d] = 1.01 * a;
    B[d] = A[d] + space[d];
    center[d] = 0.5 * (A[d] + B[d]);
    sized2[d] = 0.5 * space[d];
  }
  L = {Lx - 2 * p_r, Ly - 2 * p_r, Lz - 2 * p_r};
  grid = CellGrid(a, B[0], B[1], B[2]);
  // CellDistributionInit (md_cell_distribution.cu:21-23)
  hidden = {A[0] + 0.5 * space[0], A[1] + 0.5 * space[1], (grid.Nz - 0.5) * a};
}

Compression::Compression(bool uni, std::array<double, 3> sz, std::array<double, 3> c,
                         double pr, double dt_, double vl, double ms, double mx)
    : uniaxial(uni), axis(uni ? 2 : 1), size(sz), center(c), p_r(pr), dt(dt_), v_load(vl),
      min_stress(ms), max_strain(mx) {
  size0 = size[axis];
  bottom = center[axis] - 0.5 * size[axis];
  bottom_r = bottom + p_r;
  top = center[axis] + 0.5 * size[axis];
  top_r = top - p_r;
  const double hx = 0.5 * size[0];
  if (uniaxial) {
    R = 1.5 * hx;
    area = M_PI * std::pow(hx + p_r, 2.0);
    zero_stress = 0.33 * min_stress;
  } else {
    const double hz = 0.5 * size[2];
    R = hx > hz ? hx : 1.05 * hz;
    area = 0.5 * 2 * M_PI * (hx + p_r) * (size[2] + 2.0 * p_r);
    zero_stress = 0.5 * min_stress;
  }
  RR = R * R;
  inv_area = 1.0 / area;
}

void Compression::sum_forces(double s0, double s1) {
  history.push_back({(size[axis] - size0) / size0, s0 * inv_area, s1 * inv_area});
  stress = 0.5 * (std::fabs(s0) + std::fabs(s1)) * inv_area;
  if (!loading && stress > min_stress) loading = true;
  if (fracture_stress < stress) {
    fracture_stress = stress;
    fracture_strain = (size0 - size[axis]) / size0;
  }
  if (loading && stress < zero_stress) fracture = true;
}

void Compression::compress() {
  V = stress < 1e-9 ? 10.0 * v_load : v_load;
  size[axis] += V * dt;
  top += V * dt;
  top_r = top - 2.0 * p_r;
  center[axis] += V * dt;
  strain = (size[axis] - size0) / size0;
}

// ------------------------------------------------------------ device layout
template <class T>
static T *dev_alloc(size_t n) {
  void *p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc");
  hip_check(hipMemset(p, 0, sizeof(T) * std::max<size_t>(n, 1)), "hipMemset");
  return static_cast<T *>(p);
}

RefLayout::RefLayout(rfd_ctx *ctx, const std::vector<float> &R_soa, uint32_t K_,
                     const CellGrid &g)
    : N((uint32_t)(R_soa.size() / 3)), K(K_), ctx_(ctx), g_(g) {
  const size_t S = (size_t)N * K;
  R = dev_alloc<float>(3 * (size_t)N);
  V = dev_alloc<float>(3 * (size_t)N);
  W = dev_alloc<float>(3 * (size_t)N);
  F = dev_alloc<float>(3 * (size_t)N);
  M = dev_alloc<float>(3 * (size_t)N);
  CI = dev_alloc<uint32_t>(N);
  CIs = dev_alloc<uint32_t>(2 * (size_t)N);
  pnC = dev_alloc<uint32_t>(2 * (size_t)g.cells());
  IL = dev_alloc<uint32_t>(S);
  ILtype = dev_alloc<uint8_t>(S);
  iL = dev_alloc<float>(S);
  for (auto &p : slot) p = dev_alloc<float>(3 * S);
  hip_check(hipMemset(IL, 0xFF, sizeof(uint32_t) * S), "hipMemset");
  hip_check(hipMemcpy(R, R_soa.data(), sizeof(float) * R_soa.size(), hipMemcpyHostToDevice),
            "hipMemcpy");
}

RefLayout::~RefLayout() {
  for (void *p : {(void *)R, (void *)V, (void *)W, (void *)F, (void *)M, (void *)CI,
                  (void *)CIs, (void *)pnC, (void *)IL, (void *)ILtype, (void *)iL})
    (void)hipFree(p);
  for (auto p : slot) (void)hipFree(p);
}

void RefLayout::cell_distribution() {
  check(ctx_, rfd_cell_distribution(ctx_, R, N, g_.inv_a, g_.Nx, g_.Ny, g_.Nz, CI, CIs, pnC),
        "rfd_cell_distribution");
}

void RefLayout::il_construct(float aacut) {
  check(ctx_, rfd_il_construct(ctx_, R, N, CIs, pnC, IL, K, (float)g_.a, g_.inv_a, aacut, g_.Nx,
                               g_.Ny, g_.Nz),
        "rfd_il_construct");
}

void RefLayout::il_reconstruct_bpm(float aacut) {
  check(ctx_, rfd_il_reconstruct_bpm(ctx_, R, N, CIs, pnC, IL, ILtype, K, (float)g_.a,
                                     g_.inv_a, aacut, g_.Nx, g_.Ny, g_.Nz, 1),
        "rfd_il_reconstruct_bpm");
}

void RefLayout::bonds_construct(float bdd) {
  check(ctx_, rfd_bonds_construct(ctx_, R, IL, ILtype, iL, slot[0], slot[1], slot[2], slot[3],
                                  slot[4], slot[5], slot[6], N, K, bdd),
        "rfd_bonds_construct");
}

void RefLayout::delete_far_links(float aa_farcut) {
  check(ctx_, rfd_delete_far_links(ctx_, R, N, IL, ILtype, slot[0], slot[1], slot[2], slot[3],
                                   slot[4], slot[5], slot[6], K, aa_farcut),
        "rfd_delete_far_links");
}

void RefLayout::zero_fm() { check(ctx_, rfd_zero_fm(ctx_, F, M, N), "rfd_zero_fm"); }

void RefLayout::forces_bpm(const Material &m) {
  check(ctx_, rfd_forces_bpm(ctx_, R, V, W, IL, ILtype, iL, slot[0], slot[1], slot[4], slot[5],
                             F, M, N, K, (float)m.dt, (float)m.E, (float)m.G, (float)m.b_r,
                             (float)m.sc, (float)m.tc),
        "rfd_forces_bpm");
}

void RefLayout::forces_contact(const Material &m) {
  check(ctx_, rfd_forces_contact(ctx_, R, V, W, IL, ILtype, slot[1], F, M, N, K, (float)m.dt,
                                 (float)m.hm_E, (float)m.hm_G, (float)m.p_A, (float)m.p_mu,
                                 (float)m.p_mur, (float)m.p_m, (float)m.p_r),
        "rfd_forces_contact");
}

std::array<double, 2> RefLayout::platen(const Compression &c, float zcut) {
  double sums[2] = {0, 0};
  if (c.uniaxial)
    check(ctx_, rfd_platen_uniaxial(ctx_, R, V, F, N, (float)c.center[0], (float)c.center[1],
                                    (float)c.RR, (float)c.top, (float)c.bottom, (float)c.top_r,
                                    (float)c.bottom_r, (float)c.V, zcut, sums),
          "rfd_platen_uniaxial");
  else
    check(ctx_, rfd_platen_brazil(ctx_, R, V, F, N, (float)c.center[0], (float)c.center[2],
                                  (float)c.RR, (float)c.top, (float)c.bottom, (float)c.top_r,
                                  (float)c.bottom_r, (float)c.V, zcut, sums),
          "rfd_platen_brazil");
  return {sums[0], sums[1]};
}

void RefLayout::increments(const Material &m) {
  check(ctx_, rfd_increments(ctx_, F, V, R, M, W, N, (float)m.dt_d_m, (float)m.dt,
                             (float)m.dt_d_I),
        "rfd_increments");
}

void RefLayout::cut_restriction(const SampleBox &box) {
  float c[3], h[3], p[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = (float)box.center[d];
    h[d] = (float)box.sized2[d];
    p[d] = (float)box.hidden[d];
  }
  check(ctx_, rfd_cut_restriction(ctx_, R, V, F, N, c, h, p), "rfd_cut_restriction");
}

// ------------------------------------------------------------ the driver
Driver::Driver(rfd_ctx *ctx, const std::vector<float> &R_soa, const Material &m,
               const SampleBox &b, const Compression &c, uint32_t K)
    : s(ctx, R_soa, K, CellGrid(b.grid)), mat(m), box(b), cmp(c) {
  const float af = (float)mat.a_farcut;
  aacut = af * af;
  zcut = (float)(box.center[2] + box.sized2[2]);
}

void Driver::setup() {
  s.cell_distribution();
  s.il_construct(aacut);
  s.bonds_construct((float)mat.b_dd);
  s.delete_far_links((float)mat.aa_farcut);
}

std::array<double, 2> Driver::step() {
  s.cell_distribution();
  s.il_reconstruct_bpm(aacut);
  s.zero_fm();
  s.forces_bpm(mat);
  s.forces_contact(mat);
  const auto sums = s.platen(cmp, zcut);
  cmp.sum_forces(sums[0], sums[1]);
  s.increments(mat);
  s.cut_restriction(box);
  s.delete_far_links((float)mat.aa_farcut);
  cmp.compress();
  return sums;
}

uint32_t Driver::run(uint32_t max_steps) {
  uint32_t n = 0;
  while (n < max_steps && cmp.strain < cmp.max_strain && !cmp.fracture) {
    step();
    ++n;
  }
  return n;
}

// ------------------------------------------------------------ files
static void put(std::ofstream &f, const void *p, size_t n) {
  f.write(static_cast<const char *>(p), (std::streamsize)n);
}

template <class T>
static void take(std::ifstream &f, std::vector<T> &v, size_t n) {
  v.resize(n);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(sizeof(T) * n));
  if (!f) throw std::runtime_error("DATA.dat ends early");
}

Checkpoint load_all_data(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  Checkpoint c;
  f.read(reinterpret_cast<char *>(c.sample.data()), sizeof(float) * 24);
  uint32_t N = 0, S = 0;
  f.read(reinterpret_cast<char *>(&N), 4);
  for (auto *v : {&c.F, &c.V, &c.R, &c.M, &c.W}) take(f, *v, 3 * (size_t)N);
  take(f, c.Q, 4 * (size_t)N);
  f.read(reinterpret_cast<char *>(&S), 4);
  f.read(reinterpret_cast<char *>(&c.IonP), 4);
  f.read(reinterpret_cast<char *>(&c.acut), 4);
  take(f, c.IL, S);
  take(f, c.ILtype, S);
  take(f, c.iL, S);
  for (auto &v : c.slot) take(f, v, 3 * (size_t)S);
  return c;
}

void save_all_data(const std::string &path, const Checkpoint &c) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const uint32_t N = (uint32_t)(c.R.size() / 3), S = (uint32_t)c.IL.size();
  put(f, c.sample.data(), sizeof(float) * 24);
  put(f, &N, 4);
  for (auto *v : {&c.F, &c.V, &c.R, &c.M, &c.W}) put(f, v->data(), sizeof(float) * v->size());
  put(f, c.Q.data(), sizeof(float) * c.Q.size());
  put(f, &S, 4);
  put(f, &c.IonP, 4);
  put(f, &c.acut, 4);
  put(f, c.IL.data(), 4 * (size_t)S);
  put(f, c.ILtype.data(), S);
  put(f, c.iL.data(), 4 * (size_t)S);
  for (auto &v : c.slot) put(f, v.data(), sizeof(float) * v.size());
}

template <class T>
static void download(std::vector<T> &v, const T *src, size_t n) {
  v.resize(n);
  hip_check(hipMemcpy(v.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost), "hipMemcpy");
}

Checkpoint checkpoint_of(const RefLayout &s, float acut) {
  Checkpoint c;
  const size_t N3 = 3 * (size_t)s.N, S = (size_t)s.N * s.K;
  download(c.F, s.F, N3);
  download(c.V, s.V, N3);
  download(c.R, s.R, N3);
  download(c.M, s.M, N3);
  download(c.W, s.W, N3);
  c.Q.assign(4 * (size_t)s.N, 0.0f);
  c.IonP = s.K;
  c.acut = acut;
  download(c.IL, s.IL, S);
  download(c.ILtype, s.ILtype, S);
  download(c.iL, s.iL, S);
  for (int k = 0; k < 7; ++k) download(c.slot[k], s.slot[k], 3 * S);
  return c;
}

void save_sum_forces_loading(const std::string &path, const Compression &c) {
  std::FILE *f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "EpsZ StressZ_bottom StressZ_top\n");
  for (const auto &h : c.history)
    std::fprintf(f, "%g %g %g\n", h[0], h[1] * kStress, h[2] * kStress);
  std::fclose(f);
}

}  // namespace rfdh
