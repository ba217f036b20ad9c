// mfm_init.hip -- initial data and one-box periodic ghost fill on the device.
// These are caller-side utilities (the level driver's job in the
// reference), used to set up benchmarks and tests without host staging.
#include "mfm_device.hpp"
#include "mfm_kernels.hpp"

namespace mfm {

namespace {
constexpr int kB = 256;

__device__ __forceinline__ bool unravel(long t, const int n[3], int& a, int& b,
                                        int& c) {
  const long n01 = (long)n[0] * n[1];
  if (t >= n01 * n[2]) return false;
  c = (int)(t / n01);
  const long r = t - (long)c * n01;
  b = (int)(r / n[0]);
  a = (int)(r - (long)b * n[0]);
  return true;
}

struct Geo {
  int ulo[3], un[3], lo[3], n[3];
  long ucs;
};

// HELIOINIT (swlism/HeliosphericF.ChF:1150-1361), RIEMANNINIT
// (problems/RiemannF.ChF:65-127), ORSZAGINIT (problems/OrszagF.ChF:27-79)
__global__ __launch_bounds__(kB) void k_init(Consts C, double* U, Geo G,
                                             double dx) {
  int a, b, c;
  const long t = (long)blockIdx.x * kB + threadIdx.x;
  if (!unravel(t, G.n, a, b, c)) return;
  const int i = G.lo[0] + a, j = G.lo[1] + b, k = G.lo[2] + c;
  const long o = (long)(i - G.ulo[0]) +
                 (long)G.un[0] * ((long)(j - G.ulo[1]) + (long)G.un[1] * (k - G.ulo[2]));
  double u[kMaxComp];
  for (int q = 0; q < C.nc; ++q) u[q] = 0.0;
  const double hg = C.hgamma;
  const bool d3 = C.dims == 3;
  if (C.problem == MFM_PROBLEM_HELIO) {
    const double sinTILT = sin(C.sun_tilt), cosTILT = cos(C.sun_tilt);
    const double sunBeta = C.omega * kAU / (C.sun_v * C.lism_v_dim);
    const double x = (i + 0.5) * dx - C.sun_xc, y = (j + 0.5) * dx - C.sun_yc;
    const double z = d3 ? (k + 0.5) * dx - C.sun_zc : 0.0;
    double r = d3 ? x * x + y * y + z * z : x * x + y * y;
    r = sqrt(r);
    const double xy = sqrt(x * x + y * y);
    const double cosT = z / r, sinT = sqrt(1.0 - cosT * cosT);
    const double cosF = x / xy, sinF = y / xy;
    if (r > C.init_r) {
      double rr = C.init_r / r;
      rr = fmaxr(0.0, rr);
      rr = rr * rr * rr;
      const double vel = sinT * cosF * C.lism_ux + sinT * sinF * C.lism_uy + cosT * C.lism_uz;
      const double ux = C.lism_ux - 0.5 * rr * (3.0 * sinT * cosF * vel - C.lism_ux);
      const double uy = C.lism_uy - 0.5 * rr * (3.0 * sinT * sinF * vel - C.lism_uy);
      const double uz = C.lism_uz - 0.5 * rr * (3.0 * cosT * vel - C.lism_uz);
      const double p = 1.0 / (C.gamma * C.lism_m * C.lism_m);
      double bb = sinT * cosF * C.lism_bx + sinT * sinF * C.lism_by + cosT * C.lism_bz;
      const double Bx = C.lism_bx - 0.5 * rr * (3.0 * sinT * cosF * bb - C.lism_bx);
      const double By = C.lism_by - 0.5 * rr * (3.0 * sinT * sinF * bb - C.lism_by);
      const double Bz = C.lism_bz - 0.5 * rr * (3.0 * cosT * bb - C.lism_bz);
      bb = k1_4PI * (Bx * Bx + By * By + Bz * Bz);
      const double ke = ux * ux + uy * uy + uz * uz;
      u[0] = 1.0; u[1] = ux; u[2] = uy; u[3] = uz;
      u[4] = p / hg + 0.5 * (ke + bb);
      u[5] = Bx; u[6] = By; u[7] = Bz;
    } else {
      const double rr = r / C.sun_int_bc_radius;
      const double theta = acos(cosT);
      const double cosBr = cos(sunBeta * r), sinBr = sin(sunBeta * r);
      const double sinB = sinF * cosBr + cosF * sinBr;
      const double cosB = cosF * cosBr - sinF * sinBr;
      double thetaS = sinTILT * sinB / sqrt((cosTILT * cosB) * (cosTILT * cosB) + sinB * sinB);
      thetaS = asin(thetaS) + kPI_2;
      const double rho = C.sun_rho / (rr * rr);
      const double ux = C.sun_v * sinT * cosF, uy = C.sun_v * sinT * sinF, uz = C.sun_v * cosT;
      const double p = C.sun_p / pow(rr, 2.0 * C.gamma);
      const double br = C.dir_br_n * C.sun_b / (rr * rr);
      const double bp = -sunBeta * br * r * sinT;
      const double bx = br * sinT * cosF - bp * sinF;
      const double by = br * sinT * sinF + bp * cosF;
      const double bz = br * cosT;
      const double bsq = bx * bx + by * by + bz * bz;
      u[0] = rho; u[1] = rho * ux; u[2] = rho * uy; u[3] = rho * uz;
      u[4] = p / hg + 0.5 * (rho * C.sun_v * C.sun_v + k1_4PI * bsq);
      const bool pos = (C.sun_b_monopolar == 1) || (theta < thetaS);
      u[5] = pos ? bx : -bx; u[6] = pos ? by : -by; u[7] = pos ? bz : -bz;
    }
    if (C.nf > 1) {
      const int s = C.iRhoN;
      u[s] = C.net_rho; u[s + 1] = C.net_rho * C.lism_ux;
      u[s + 2] = C.net_rho * C.lism_uy; u[s + 3] = C.net_rho * C.lism_uz;
      u[s + 4] = C.net_p / hg + 0.5 * C.net_rho;
    }
    if (C.nf > 2) {
      const int s = C.iRhoN + 5;
      u[s] = kNet2Dens; u[s + 4] = kNet2Pres / hg;
    }
    if (C.nf > 3) {
      const int s = C.iRhoN + 10;
      u[s] = kNet3Dens; u[s + 4] = kNet3Pres / hg;
    }
  } else if (C.problem == MFM_PROBLEM_SWLISM) {
    // SWLISMINIT (swlism/SWLISMF.ChF:106-255), 2-D: LISM potential flow
    // outside initR, radial solar wind inside
    const double x = (i + 0.5) * dx - C.sun_xc, y = (j + 0.5) * dx - C.sun_yc;
    const double r = sqrt(x * x + y * y);
    const int s = C.iRhoN;
    if (r > C.init_r) {
      const double sinT = y / r, cosT = x / r;
      double h = C.init_r / r;
      h = h * h * h;
      const double ux = 1.0 - (1.0 - 1.5 * sinT * sinT) * h;
      const double uy = -1.5 * sinT * cosT * h;
      const double uz = 0.0;
      const double p = 1.0 / (C.gamma * C.lism_m * C.lism_m);
      const double Bx = C.lism_bx * ux, By = C.lism_bx * uy, Bz = 0.0;
      const double b = k1_4PI * (Bx * Bx + By * By + Bz * Bz);
      const double ke = ux * ux + uy * uy + uz * uz;
      u[0] = 1.0; u[1] = ux; u[2] = uy; u[3] = uz;
      u[4] = p / hg + 0.5 * (ke + b);
      u[5] = Bx; u[6] = By; u[7] = Bz;
      if (C.nf > 1) {
        u[s] = C.net_rho; u[s + 1] = C.net_rho * ux; u[s + 2] = C.net_rho * uy;
        u[s + 3] = C.net_rho * uz;
        u[s + 4] = C.net_p / hg + 0.5 * C.net_rho * ke;
      }
    } else {
      const double rho = C.sun_rho / (r * r);
      const double ux = C.sun_v * x / r, uy = C.sun_v * y / r;
      const double p = C.sun_p / pow(r, 2.0 * C.gamma);
      u[0] = rho; u[1] = rho * ux; u[2] = rho * uy; u[3] = 0.0;
      u[4] = p / hg + 0.5 * rho * C.sun_v * C.sun_v;
      u[5] = 0.0; u[6] = 0.0; u[7] = 0.0;
      if (C.nf > 1) {
        u[s] = C.net_rho; u[s + 1] = C.net_rho * 1.0; u[s + 2] = 0.0; u[s + 3] = 0.0;
        u[s + 4] = C.net_p / hg + 0.5 * C.net_rho;
      }
    }
    if (C.nf > 2) { u[s + 5] = kNet2Dens; u[s + 9] = kNet2Pres / hg; }
    if (C.nf > 3) { u[s + 10] = kNet3Dens; u[s + 14] = kNet3Pres / hg; }
  } else if (C.problem == MFM_PROBLEM_RIEMANN) {
    const double* L = C.rpL;
    const double* R = C.rpR;
    double bb = k1_4PI * (L[5] * L[5] + L[6] * L[6] + L[7] * L[7]);
    double ke = L[1] * L[1] + L[2] * L[2] + L[3] * L[3];
    const double eL = L[4] / hg + 0.5 * (L[0] * ke + bb);
    bb = k1_4PI * (R[5] * R[5] + R[6] * R[6] + R[7] * R[7]);
    ke = R[1] * R[1] + R[2] * R[2] + R[3] * R[3];
    const double eR = R[4] / hg + 0.5 * (R[0] * ke + bb);
    const double x = (i + 0.5) * dx;
    const bool left = x < C.rp_start_x;
    const double* S = left ? L : R;
    u[0] = S[0]; u[1] = S[0] * S[1]; u[2] = S[0] * S[2]; u[3] = S[0] * S[3];
    u[4] = left ? eL : eR;
    u[5] = S[5]; u[6] = S[6]; u[7] = S[7];
  } else {
    const double p0 = k1_4PI * C.gamma, r0 = p0 * C.gamma;
    const double x = (i + 0.5) * dx, y = (j + 0.5) * dx;
    const double Ux = -sin(k2PI * y), Uy = sin(k2PI * x);
    const double Bx = -sin(k2PI * y), By = sin((k2PI + k2PI) * x);
    const double bb = k1_4PI * (Bx * Bx + By * By);
    const double ke = Ux * Ux + Uy * Uy;
    u[0] = r0; u[1] = r0 * Ux; u[2] = r0 * Uy; u[3] = 0.0;
    u[4] = p0 / hg + 0.5 * (r0 * ke + bb);
    u[5] = Bx; u[6] = By; u[7] = 0.0;
  }
  for (int q = 0; q < C.nc; ++q) U[o + q * G.ucs] = u[q];
}

// Periodic images of a one-box level: every ghost cell copies the valid
// cell congruent to it modulo the box size.
__global__ __launch_bounds__(kB) void k_fill_periodic(double* U, Geo G, int nc,
                                                      int vlo0, int vlo1,
                                                      int vlo2, int vn0,
                                                      int vn1, int vn2,
                                                      int dims) {
  int a, b, c;
  const long t = (long)blockIdx.x * kB + threadIdx.x;
  if (!unravel(t, G.un, a, b, c)) return;
  int idx[3] = {G.ulo[0] + a, G.ulo[1] + b, G.ulo[2] + c};
  const int vlo[3] = {vlo0, vlo1, vlo2}, vn[3] = {vn0, vn1, vn2};
  int src[3] = {idx[0], idx[1], idx[2]};
  bool ghost = false;
  for (int d = 0; d < dims; ++d) {
    int o = idx[d] - vlo[d];
    o = ((o % vn[d]) + vn[d]) % vn[d];
    src[d] = vlo[d] + o;
    if (src[d] != idx[d]) ghost = true;
  }
  if (!ghost) return;
  const long dst = t;
  const long s = (long)(src[0] - G.ulo[0]) +
                 (long)G.un[0] * ((long)(src[1] - G.ulo[1]) + (long)G.un[1] * (src[2] - G.ulo[2]));
  for (int q = 0; q < nc; ++q) U[dst + q * G.ucs] = U[s + q * G.ucs];
}
}  // namespace

hipError_t launch_init(const Consts& C, double* U, const int ulo[3],
                       const int un[3], long ucs, const int lo[3],
                       const int n[3], double dx, hipStream_t s) {
  Geo G;
  for (int d = 0; d < 3; ++d) { G.ulo[d] = ulo[d]; G.un[d] = un[d]; G.lo[d] = lo[d]; G.n[d] = n[d]; }
  G.ucs = ucs;
  const long tot = (long)n[0] * n[1] * n[2];
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_init, dim3((unsigned)((tot + kB - 1) / kB)), dim3(kB), 0,
                     s, C, U, G, dx);
  return hipGetLastError();
}

hipError_t launch_fill_periodic(double* U, const int ulo[3], const int un[3],
                                long ucs, int nc, const int vlo[3],
                                const int vn[3], int dims, hipStream_t s) {
  Geo G;
  for (int d = 0; d < 3; ++d) { G.ulo[d] = ulo[d]; G.un[d] = un[d]; G.lo[d] = 0; G.n[d] = 0; }
  G.ucs = ucs;
  const long tot = (long)un[0] * un[1] * un[2];
  hipLaunchKernelGGL(k_fill_periodic, dim3((unsigned)((tot + kB - 1) / kB)),
                     dim3(kB), 0, s, U, G, nc, vlo[0], vlo[1], vlo[2], vn[0],
                     vn[1], vn[2], dims);
  return hipGetLastError();
}

}  // namespace mfm
