// NPT molecular dynamics around the potential: the Martyna-Tobias-Klein barostat with
// Nose-Hoover chains on the particles and on the barostat (MTTK; G. J. Martyna, D. J.
// Tobias, M. L. Klein, J. Chem. Phys. 101, 4177 (1994); the measure-preserving splitting of
// M. E. Tuckerman et al., J. Phys. A 39, 5629 (2006); the reversible half-step order of
// W. Shinoda, M. Shiga, M. Mikami, Phys. Rev. B 69, 134103 (2004)).  The reference benchmark
// runs GPUMD's `ensemble npt_mttk temp T1 T2 aniso p1 p2 tperiod 50 pperiod 500`
// (docs/gpumd-sus2-sh-interface-plan.md:264-271); GPUMD's integrator itself is not part of the
// reference overlay, so this is the published scheme, checked through its conserved extended
// energy and its reduction to velocity Verlet.
//
// One step (dt):
//   barostat chain (dt/2); particle chain (dt/2, scales v); barostat velocities omega_dot
//   (dt/2, from (P - P0) V and the MTK kinetic term); v scaled by exp(-dt/2 (omega_dot +
//   mtk2)); half kick; box and x dilated by exp(dt/2 omega_dot); drift; dilated again;
//   forces; half kick; v scaled; omega_dot (dt/2); particle chain (dt/2); barostat chain
//   (dt/2).
// Uniform per-axis velocity scalings change the kinetic tensor by known factors, so the
// device is read once per step: after the second kick (kinetic tensor, potential energy,
// virial trace).  Coupling "iso" replaces the three diagonal pressures by their mean (the
// three axes then share one strain rate), "aniso" couples each axis to its own pressure.
// Units: A, fs, amu, eV, K; pressures in GPa at the interface.
#include <cmath>

#include "engine.hpp"

namespace sus2sh {

namespace {

constexpr double kBoltzmann = 8.617333262e-5;   // eV / K
constexpr double kEvPerA3PerGPa = 1.0 / 160.21766208;
constexpr int kChain = 3;                       // thermostat chain length (particles, barostat)

void hip_ok(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

// box matrix (rows x, y, z; columns = cell vectors) and its inverse into the GPUMD layout
void set_box(sus2sh_box& b, const double (&H)[3][3])
{
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) b.h[3 * r + c] = H[r][c];
  const double det = H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) -
                     H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
                     H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
  if (!(std::fabs(det) > 0.0)) throw ArgumentError("md_npt: the box became singular");
  const double inv[9] = {(H[1][1] * H[2][2] - H[1][2] * H[2][1]) / det,
                         (H[0][2] * H[2][1] - H[0][1] * H[2][2]) / det,
                         (H[0][1] * H[1][2] - H[0][2] * H[1][1]) / det,
                         (H[1][2] * H[2][0] - H[1][0] * H[2][2]) / det,
                         (H[0][0] * H[2][2] - H[0][2] * H[2][0]) / det,
                         (H[0][2] * H[1][0] - H[0][0] * H[1][2]) / det,
                         (H[1][0] * H[2][1] - H[1][1] * H[2][0]) / det,
                         (H[0][1] * H[2][0] - H[0][0] * H[2][1]) / det,
                         (H[0][0] * H[1][1] - H[0][1] * H[1][0]) / det};
  for (int k = 0; k < 9; ++k) b.h[9 + k] = inv[k];
}

// Nose-Hoover chain of `kChain` links on a kinetic energy `ke2` (twice the kinetic energy,
// i.e. sum m v^2 or sum W omega_dot^2) for `dof` degrees of freedom: one half step (dt/2) of
// the Trotter-split chain, with quarter steps on the chain velocities around the scaling of
// the thermostatted velocities.  Returns the velocity factor and advances the chain.
struct Chain {
  double x[kChain] = {0, 0, 0};       // positions eta
  double v[kChain + 1] = {0, 0, 0, 0};  // velocities eta_dot (v[kChain] = 0)
  double q[kChain] = {0, 0, 0};       // masses
  double a[kChain] = {0, 0, 0};       // accelerations

  void set_masses(double dof, double kT, double freq)
  {
    q[0] = dof * kT / (freq * freq);
    for (int i = 1; i < kChain; ++i) q[i] = kT / (freq * freq);
  }

  double half_step(double& ke2, double dof, double kT, double dt)
  {
    const double dt2 = 0.5 * dt, dt4 = 0.25 * dt, dt8 = 0.125 * dt;
    a[0] = q[0] > 0.0 ? (ke2 - dof * kT) / q[0] : 0.0;
    for (int i = kChain - 1; i > 0; --i) {
      a[i] = (q[i - 1] * v[i - 1] * v[i - 1] - kT) / q[i];  // v[i - 1] is not updated yet
      const double e = std::exp(-dt8 * v[i + 1]);
      v[i] = (v[i] * e + a[i] * dt4) * e;
    }
    double e = std::exp(-dt8 * v[1]);
    v[0] = (v[0] * e + a[0] * dt4) * e;
    const double s = std::exp(-dt2 * v[0]);
    ke2 *= s * s;
    a[0] = q[0] > 0.0 ? (ke2 - dof * kT) / q[0] : 0.0;
    for (int i = 0; i < kChain; ++i) x[i] += dt2 * v[i];
    v[0] = (v[0] * e + a[0] * dt4) * e;
    for (int i = 1; i < kChain; ++i) {
      e = std::exp(-dt8 * v[i + 1]);
      a[i] = (q[i - 1] * v[i - 1] * v[i - 1] - kT) / q[i];
      v[i] = (v[i] * e + a[i] * dt4) * e;
    }
    return s;
  }

  // thermostat part of the conserved energy: dof kT x_0 + kT sum_{i>0} x_i + sum q v^2 / 2
  double energy(double dof, double kT) const
  {
    double h = dof * kT * x[0] + 0.5 * q[0] * v[0] * v[0];
    for (int i = 1; i < kChain; ++i) h += kT * x[i] + 0.5 * q[i] * v[i] * v[i];
    return h;
  }
};

}  // namespace

void Engine::md_npt(sus2sh_box& box, int N, const int* type, const double* masses, double* pos,
                    double* vel, const NptParams& prm, int nsteps, double* pot, double* force,
                    double* virial, double* trace, hipStream_t s)
{
  if (N <= 1 || nsteps < 0) throw ArgumentError("md_npt: N must be > 1 and nsteps >= 0");
  if (!masses || !pos || !vel || !type || !pot || !force || !virial)
    throw ArgumentError("md_npt: null pointer");
  if (!(prm.dt > 0.0) || !(prm.tperiod > 0.0) || !(prm.pperiod > 0.0) || !(prm.t_start > 0.0) ||
      !(prm.t_stop > 0.0))
    throw ArgumentError("md_npt: dt, temperatures, tperiod and pperiod must be positive");
  if (prm.couple != 0 && prm.couple != 1) throw ArgumentError("md_npt: couple must be 0 (iso) or 1 (aniso)");
  for (int d = 0; d < 3; ++d)
    if (!box.pbc[d]) throw ArgumentError("md_npt: the barostat needs a periodic box");
  const int S = model_.species_count;
  for (int t = 0; t < S; ++t)
    if (!(masses[t] > 0.0)) throw ArgumentError("md_npt: masses must be positive");
  md_mass_.upload(masses, S * sizeof(double));
  md_red_.resize(8 * sizeof(double));
  md_partial_.resize(md_partials_size(N) * sizeof(double));
  if (!host_red_) hip_ok(hipHostMalloc(reinterpret_cast<void**>(&host_red_), 8 * sizeof(double), 0),
                         "hipHostMalloc");
  const double* mass = md_mass_.as<double>();
  double* red = md_red_.as<double>();
  const size_t n = static_cast<size_t>(N);

  double H[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) H[r][c] = box.h[3 * r + c];
  auto volume = [&]() {
    return std::fabs(H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) -
                     H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
                     H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]));
  };
  const double vol0 = volume();

  const double dt = prm.dt, dt2 = 0.5 * dt;
  const double tdof = 3.0 * N - 3.0;
  // barostat degrees of freedom: iso moves the three axes with one strain rate (their common
  // omega_dot has mass 3 W), aniso has three independent ones
  const double pdof = prm.couple == 0 ? 1.0 : 3.0;
  const double tfreq = 1.0 / (prm.tperiod * dt), pfreq = 1.0 / (prm.pperiod * dt);
  Chain tchain, pchain;
  double omega_dot[3] = {0, 0, 0}, wmass[3] = {0, 0, 0};
  double ke2[3] = {0, 0, 0};   // sum m v_a^2 (eV) of the device velocities times pending scales
  double pending[3] = {1, 1, 1};  // velocity scales not yet applied on the device
  double pe = 0.0, wvir[3] = {0, 0, 0};
  double mtk2 = 0.0;

  auto evaluate_and_reduce = [&](bool kick) {
    launch_fill_double(pot, 0.0, n, s);
    launch_fill_double(force, 0.0, 3 * n, s);
    launch_fill_double(virial, 0.0, 9 * n, s);
    compute(box, N, type, pos, pot, force, virial, s, nullptr);
    if (kick) launch_half_kick(N, dt2, type, mass, force, vel, s);
    launch_npt_reduce(N, type, mass, pot, vel, virial, red, md_partial_.as<double>(), s);
    hip_ok(hipMemcpyAsync(host_red_, red, 7 * sizeof(double), hipMemcpyDeviceToHost, s),
           "hipMemcpyAsync D2H");
    hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (int a = 0; a < 3; ++a) {
      ke2[a] = host_red_[a];
      wvir[a] = host_red_[4 + a];
    }
    pe = host_red_[3];
  };
  auto apply_pending = [&]() {
    if (pending[0] != 1.0 || pending[1] != 1.0 || pending[2] != 1.0)
      launch_scale3(N, pending[0], pending[1], pending[2], vel, s);
    pending[0] = pending[1] = pending[2] = 1.0;
  };
  // thermostat / barostat targets at the start of step k (linear ramps over the run)
  double kT = 0.0, p0 = 0.0;
  auto targets = [&](int k) {
    const double f = nsteps > 0 ? static_cast<double>(k) / nsteps : 0.0;
    kT = kBoltzmann * (prm.t_start + f * (prm.t_stop - prm.t_start));
    p0 = kEvPerA3PerGPa * (prm.p_start + f * (prm.p_stop - prm.p_start));
    tchain.set_masses(tdof, kT, tfreq);
    pchain.set_masses(pdof, kT, pfreq);
    for (int a = 0; a < 3; ++a) wmass[a] = (N + 1) * kT / (pfreq * pfreq);
  };
  auto particle_chain = [&]() {
    double k = ke2[0] + ke2[1] + ke2[2];
    const double f = tchain.half_step(k, tdof, kT, dt);
    for (int a = 0; a < 3; ++a) {
      ke2[a] *= f * f;
      pending[a] *= f;
    }
  };
  auto barostat_chain = [&]() {
    double k = 0.0;
    for (int a = 0; a < 3; ++a) k += wmass[a] * omega_dot[a] * omega_dot[a];
    const double f = pchain.half_step(k, pdof, kT, dt);
    for (int a = 0; a < 3; ++a) omega_dot[a] *= f;
  };
  auto barostat_velocity = [&]() {  // omega_dot += dt/2 ((P_a - P0) V + mtk1) / W
    const double V = volume();
    double p[3];
    for (int a = 0; a < 3; ++a) p[a] = (ke2[a] + wvir[a]) / V;
    if (prm.couple == 0) {
      const double m = (p[0] + p[1] + p[2]) / 3.0;
      p[0] = p[1] = p[2] = m;
    }
    const double mtk1 = (ke2[0] + ke2[1] + ke2[2]) / (3.0 * N);
    for (int a = 0; a < 3; ++a)
      omega_dot[a] += dt2 * ((p[a] - p0) * V + mtk1) / wmass[a];
    mtk2 = (omega_dot[0] + omega_dot[1] + omega_dot[2]) / (3.0 * N);
  };
  auto velocity_press = [&]() {  // v_a *= exp(-dt/2 (omega_dot_a + mtk2))
    for (int a = 0; a < 3; ++a) {
      const double f = std::exp(-dt2 * (omega_dot[a] + mtk2));
      ke2[a] *= f * f;
      pending[a] *= f;
    }
  };
  auto dilate = [&]() {  // box rows and positions by exp(dt/2 omega_dot)
    double d[3];
    for (int a = 0; a < 3; ++a) {
      d[a] = std::exp(dt2 * omega_dot[a]);
      for (int c = 0; c < 3; ++c) H[a][c] *= d[a];
    }
    set_box(box, H);
    launch_scale3(N, d[0], d[1], d[2], pos, s);
  };
  auto record = [&](int k) {
    if (!trace) return;
    const double V = volume();
    const double ke = 0.5 * (ke2[0] + ke2[1] + ke2[2]);
    double h = pe + ke + tchain.energy(tdof, kT) + pchain.energy(pdof, kT) + p0 * (V - vol0);
    for (int a = 0; a < 3; ++a) h += 0.5 * wmass[a] * omega_dot[a] * omega_dot[a];
    double* t = trace + 8 * static_cast<size_t>(k);
    t[0] = pe;
    t[1] = ke;
    t[2] = h;
    t[3] = V;
    t[4] = (ke2[0] + ke2[1] + ke2[2]) / (tdof * kBoltzmann);
    t[5] = (ke2[0] + ke2[1] + ke2[2] + wvir[0] + wvir[1] + wvir[2]) / (3.0 * V) / kEvPerA3PerGPa;
    t[6] = std::cbrt(V);
    t[7] = kT / kBoltzmann;
  };

  targets(0);
  evaluate_and_reduce(false);
  record(0);
  for (int k = 1; k <= nsteps; ++k) {
    targets(k - 1);
    barostat_chain();
    particle_chain();
    barostat_velocity();
    velocity_press();
    apply_pending();
    launch_half_kick(N, dt2, type, mass, force, vel, s);
    dilate();
    launch_drift(N, dt, vel, pos, s);
    dilate();
    evaluate_and_reduce(true);
    velocity_press();
    barostat_velocity();
    particle_chain();
    barostat_chain();
    apply_pending();
    record(k);
  }
  hip_ok(hipGetLastError(), "md_npt kernel launch");
  hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
}

}  // namespace sus2sh
